// Hybrid variant: picks an execution strategy and its index parameters per
// input.
//
// Reference: HausdorffDistanceHybrid::CalculateDistance
// (src/hd_impl/hausdorff_distance_hybrid.h:86-409) samples A with the EB
// kernel, runs EB only when the sampled iteration counts are small
// (p80 < eb_only_threshold, :171-199), and otherwise builds a grid and an
// OptiX BVH over it and hands points that exceed max_hit intersections to
// the EB kernel (:291-295, :377-400). With -auto_tune its n_points_cell,
// max_hit and eb_only_threshold come from gradient-boosted trees over 63
// input features (:708-778, src/models/tree_*.h), trained offline on a
// parameter sweep (src/autotune_hausdorff_distance.cu:181-296). Here:
//  * tiny problems (|A||B| <= 2^26) or few queries (|A| <= 1024): EB only;
//  * otherwise the Morton grid with early break, queries visiting more than
//    max_hit cells finished by the EB rounds (measured faster than the
//    LBVH on the clustered 2D and the 3D-surface workloads);
//  * eb_only_threshold > 0 and |B| < 500k: the reference's sampled pass
//    decides between the early-break scan and the index path (sample.hip);
//  * auto_tune: the parameters are chosen by timed trials on the actual
//    input instead of a model (the trees and their training data are not
//    available to this build): the strategy among grid, LBVH and EB (ALL),
//    then the grid's n_points_cell over {2, 4, 8, 16, 32} and its max_hit
//    over {64, 256, 1024, 4096}, one parameter at a time, each candidate
//    timed twice (the faster run counts). The choice is
//    cached in the context under a signature of the input's content (sizes,
//    MBRs and a strided sample of the points), so a caching allocator that
//    hands the same addresses to other data does not reuse it.
// Every candidate returns the same bits; only the time differs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

namespace {

constexpr uint32_t kDefaultNpc = 8;
constexpr uint32_t kDefaultMaxHit = 256;

int RuleBased(const DirectedArgs& a) {
  const double work = (double)a.n_a * (double)a.n_b;
  if (work <= 67108864.0 || a.n_a <= 1024) return HD_VARIANT_EB;
  // Measured on MI355X: the grid with early break beats the LBVH on
  // clustered 2D data and on noisy 3D surfaces alike.
  return HD_VARIANT_GRID;
}

int RunChoice(hd_context* ctx, const TunedChoice& c, const DirectedArgs& a,
              hd_directed_result* out) {
  int rc;
  if (c.variant == HD_VARIANT_EB) {
    rc = RunEarlyBreak(ctx, a, out);
  } else if (c.variant == HD_VARIANT_GRID) {
    DirectedArgs g = a;
    g.n_points_cell = c.n_points_cell;
    rc = RunGridHybrid(ctx, g, c.max_hit, out);
  } else {
    rc = RunLbvh(ctx, a, out);
  }
  out->variant_used = c.variant;
  ctx->last_choice = c;
  return rc;
}

// FNV-1a over bytes.
uint64_t Fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 0x100000001B3ull;
  }
  return h;
}

// Signature of the input's content: sizes, flags, MBRs, the tuning knobs
// and 256 points of each set at a fixed stride (one small gather kernel and
// one copy, ~20 us).
int Signature(hd_context* ctx, const DirectedArgs& a, std::string* sig) {
  uint64_t h = 0xCBF29CE484222325ull;
  const uint64_t dims[4] = {(uint64_t)a.n_dims, a.n_a, a.n_b, a.flags};
  h = Fnv(h, dims, sizeof(dims));
  h = Fnv(h, a.mbr_a, sizeof(a.mbr_a));
  h = Fnv(h, a.mbr_b, sizeof(a.mbr_b));
  const uint32_t knobs[3] = {(uint32_t)a.auto_tune, a.n_points_cell, a.max_hit};
  h = Fnv(h, knobs, sizeof(knobs));
  constexpr uint32_t kSample = 256;
  const uint32_t ma = (uint32_t)std::min<uint64_t>(a.n_a, kSample);
  const uint32_t mb = (uint32_t)std::min<uint64_t>(a.n_b, kSample);
  const size_t fa = (size_t)ma * a.n_dims, fb = (size_t)mb * a.n_dims;
  float* d_s = nullptr;
  HD_TRY(ContextArray(ctx, "hybrid_sig", fa + fb, &d_s));
  HD_TRY(SamplePoints(ctx, a.a, a.n_a, a.n_dims, ma, d_s));
  HD_TRY(SamplePoints(ctx, a.b, a.n_b, a.n_dims, mb, d_s + fa));
  void* pinned = nullptr;
  HD_TRY(ContextPinned(ctx, (fa + fb) * sizeof(float), &pinned));
  HD_HIP_TRY(hipMemcpyAsync(pinned, d_s, (fa + fb) * sizeof(float),
                            hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  h = Fnv(h, pinned, (fa + fb) * sizeof(float));
  char buf[64];
  snprintf(buf, sizeof(buf), "%016llx", (unsigned long long)h);
  *sig = buf;
  return HD_OK;
}

// Times one candidate twice and keeps the faster run (the first run of a
// candidate may grow the context's buffers); keeps the candidate in *best
// (and its result in *out) when it is the fastest so far.
int Trial(hd_context* ctx, const TunedChoice& c, const DirectedArgs& a,
          hipEvent_t e0, hipEvent_t e1, TunedChoice* best, float* best_ms,
          hd_directed_result* out, uint32_t* trials) {
  hd_directed_result r;
  float ms = 0.f;
  for (int rep = 0; rep < 2; ++rep) {
    HD_HIP_TRY(hipEventRecord(e0, ctx->stream));
    HD_TRY(RunChoice(ctx, c, a, &r));
    HD_HIP_TRY(hipEventRecord(e1, ctx->stream));
    HD_HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HD_HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    ms = rep == 0 ? t : std::min(ms, t);
  }
  ++*trials;
  if (best->variant < 0 || ms < *best_ms) {
    *best = c;
    *best_ms = ms;
    *out = r;
  }
  return HD_OK;
}

int Tune(hd_context* ctx, const DirectedArgs& args, TunedChoice* choice,
         hd_directed_result* out) {
  const bool all = (args.auto_tune & HD_AUTO_TUNE_ALL) != 0;
  const bool tune_npc = all || (args.auto_tune & HD_AUTO_TUNE_N_POINTS_CELL);
  const bool tune_hit = all || (args.auto_tune & HD_AUTO_TUNE_MAX_HIT);
  TunedChoice base;
  base.variant = HD_VARIANT_GRID;
  base.n_points_cell = args.n_points_cell ? args.n_points_cell : kDefaultNpc;
  base.max_hit = args.max_hit ? args.max_hit : kDefaultMaxHit;
  TunedChoice best;
  float best_ms = 0.f;
  uint32_t trials = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HD_HIP_TRY(hipEventCreate(&e0));
  int rc = hipEventCreate(&e1) == hipSuccess ? HD_OK
                                              : Fail(HD_EDEVICE, "hipEventCreate failed");
  auto trial = [&](const TunedChoice& c) {
    return rc == HD_OK ? (rc = Trial(ctx, c, args, e0, e1, &best, &best_ms, out, &trials))
                       : rc;
  };
  trial(base);
  if (all) {
    TunedChoice lbvh = base;
    lbvh.variant = HD_VARIANT_LBVH;
    trial(lbvh);
    // The early-break scan only where the index variants are not clearly
    // cheaper (beyond 2^40 pair work it never won).
    if ((double)args.n_a * (double)args.n_b <= 1.1e12) {
      TunedChoice eb = base;
      eb.variant = HD_VARIANT_EB;
      trial(eb);
    }
  }
  if (best.variant == HD_VARIANT_GRID) {
    if (tune_npc) {
      const TunedChoice from = best;
      for (uint32_t npc : {2u, 4u, 8u, 16u, 32u}) {
        if (npc == from.n_points_cell) continue;
        TunedChoice c = from;
        c.n_points_cell = npc;
        trial(c);
      }
    }
    if (tune_hit) {
      const TunedChoice from = best;
      for (uint32_t mh : {64u, 256u, 1024u, 4096u}) {
        if (mh == from.max_hit) continue;
        TunedChoice c = from;
        c.max_hit = mh;
        trial(c);
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc != HD_OK) return rc;
  best.tuned = true;
  best.trials = trials;
  *choice = best;
  return HD_OK;
}

}  // namespace

int RunHybrid(hd_context* ctx, const DirectedArgs& args,
              hd_directed_result* out) {
  TunedChoice rule;
  rule.n_points_cell = args.n_points_cell ? args.n_points_cell : kDefaultNpc;
  rule.max_hit = args.max_hit ? args.max_hit : kDefaultMaxHit;
  // With an EB-only threshold and B below the reference's sample threshold
  // (500k points, hausdorff_distance_hybrid.h:62,129), the sampled pass
  // decides: the early-break scan alone when the sampled points' 80th
  // percentile iteration count is below the threshold (:171-199), the
  // index path otherwise.
  if (args.eb_only_threshold > 0 && args.n_b < 500000u && !args.auto_tune) {
    std::vector<uint32_t> counts;
    const int st = StageBegin(ctx, "hybrid_sample");
    HD_TRY(SampledEbIterations(ctx, args, &counts));
    StageEnd(ctx, st);
    std::sort(counts.begin(), counts.end());
    const uint32_t p80 = counts[(size_t)((double)counts.size() * 0.8)];
    StageCount(ctx, "hybrid_sample", p80, counts.size());
    rule.variant = p80 < args.eb_only_threshold ? HD_VARIANT_EB : HD_VARIANT_GRID;
    return RunChoice(ctx, rule, args, out);
  }
  if (!args.auto_tune) {
    rule.variant = RuleBased(args);
    return RunChoice(ctx, rule, args, out);
  }
  std::string sig;
  HD_TRY(Signature(ctx, args, &sig));
  auto it = ctx->tuned.find(sig);
  if (it != ctx->tuned.end()) {
    TunedChoice c = it->second;
    c.trials = 0;
    return RunChoice(ctx, c, args, out);
  }
  TunedChoice choice;
  const int st = StageBegin(ctx, "hybrid_tune");
  HD_TRY(Tune(ctx, args, &choice, out));
  StageEnd(ctx, st);
  ctx->tuned[sig] = choice;
  ctx->last_choice = choice;
  out->variant_used = choice.variant;
  return HD_OK;
}

}  // namespace hd
