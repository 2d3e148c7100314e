// Early-break directed Hausdorff distance on gfx950.
//
// Reference behaviour (src/hd_impl/hausdorff_distance_early_break.h:44-219):
// for each a in A (random order) scan B (random order) keeping the running
// minimum of d2; stop as soon as that minimum is strictly below the shared
// cmax2; a point that finishes its scan raises cmax2 to its exact minimum.
// The value is independent of the visit order because every d2 is the same
// unfused FP32 expression and min / max are exact selections.
//
// MI355X design (three stages, one stream, no host round trip inside a
// stage):
//  1. seed:   the first S0 points of the visit order get exact minima from
//             the cooperative kernel (many waves per point), so cmax2 starts
//             at an attained value instead of 0 (the reference's cold start,
//             early_break.h:130).
//  2. slots:  the main kernel. A wave owns P "slots", each holding one A
//             point in scalar registers; the 64 lanes stream B in steps of 64
//             coalesced SoA points and every B point in registers is tested
//             against all P slots. A slot breaks when any lane's running
//             minimum is strictly below cmax2; a slot that scanned all of B
//             publishes its exact minimum; a slot that used its step budget
//             is escalated with its partial minimum. Slots refill from a
//             global cursor, each starting at the wave's current B position
//             (B is visited cyclically, so a slot is complete after n_bp/64
//             steps wherever it started).
//  3. escalation: escalated points, sorted by partial minimum descending
//             (the most likely new maxima first), go back through the
//             cooperative kernel, which splits each group of P points over
//             B segments processed by different waves.
// Break rule: strict '<', as the reference's CPU loop (early_break.h:77);
// the reference's GPU kernel uses '<=' (:182), which can lose the identity of
// tied maxima. With '<', a point whose exact minimum equals the final cmax2
// never breaks, so the packed key (d2 bits, ~index) names the smallest such
// index.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hd_device.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

using namespace dev;

namespace {

constexpr int kWaveSize = 64;
constexpr int kBlock = 256;  // 4 independent waves per workgroup

constexpr int kOrderArray = 0;     // order[pos]
constexpr int kOrderFeistel = 1;   // pseudo-random permutation of positions
constexpr int kOrderIdentity = 2;  // input order

HD_DEV uint32_t OrderAt(int mode, const uint32_t* order, const Feistel& perm,
                        uint32_t pos) {
  if (mode == kOrderFeistel) return Permute(perm, pos);
  if (mode == kOrderIdentity) return pos;
  return order[pos];
}

struct SlotArgs {
  const float* a;         // AoS A
  const uint32_t* order;  // visit order (kOrderArray)
  const float* list_min;  // optional partial minima per position: a position
                          // whose value is already below cmax2 is skipped
  Feistel perm_a;
  int order_mode;         // kOrder*
  uint32_t pos_end;       // order positions [cursor start, pos_end)
  const float* bx;
  const float* by;
  const float* bz;
  uint32_t n_steps;       // n_bp / 64
  uint32_t budget_steps;  // steps before a slot escalates
  unsigned long long* key;
  unsigned int* cursor;   // next order position
  uint32_t* esc_idx;      // escalated A indices
  float* esc_min;         // their partial minima
  unsigned int* esc_count;
  unsigned long long* evals;
};

template <int ND>
HD_DEV void LoadPoint(const float* __restrict__ a, uint32_t idx, float* p) {
  const float* src = a + (uint64_t)idx * ND;
#pragma unroll
  for (int d = 0; d < ND; ++d) p[d] = Uniform(src[d]);
}

HD_DEV float ReadLaneF(float v, uint32_t k) {
  return __uint_as_float(
      __builtin_amdgcn_readlane(__float_as_uint(v), (int)k));
}
HD_DEV uint32_t ReadLaneU(uint32_t v, uint32_t k) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)k);
}

// B step j (64 SoA points) into registers.
template <int ND>
HD_DEV void LoadStep(const float* __restrict__ bx, const float* __restrict__ by,
                     const float* __restrict__ bz, uint32_t step,
                     uint32_t lane, float* b) {
  const uint32_t j = (step << 6) + lane;
  b[0] = bx[j];
  b[1] = by[j];
  if constexpr (ND == 3) b[2] = bz[j];
}

template <int ND, int P>
HD_DEV void EvalStep(const float (&pa)[P][ND], uint32_t live_mask,
                     const float* b, float (&lmin)[P]) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    if (live_mask & (1u << p)) {
      float d;
      if constexpr (ND == 2)
        d = D2(pa[p][0], pa[p][1], b[0], b[1]);
      else
        d = D2(pa[p][0], pa[p][1], pa[p][2], b[0], b[1], b[2]);
      lmin[p] = fminf(lmin[p], d);
    }
  }
}

HD_DEV uint32_t NextStep(uint32_t s, uint32_t n) { return s + 1 == n ? 0 : s + 1; }

// Stage 2. One wave = P slots; block = 4 independent waves.
// Refill goes through a per-wave reservoir: 64 order positions claimed with
// one atomic, each lane loading one candidate point in parallel; slots then
// take points from the reservoir with readlane.
template <int ND, int P>
__global__ __launch_bounds__(kBlock) void EbSlotKernel(SlotArgs args) {
  const uint32_t lane = LaneId();
  const uint32_t wave_id = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const uint32_t n_steps = args.n_steps;
  uint32_t step = (uint32_t)(((uint64_t)wave_id * 7919u) % n_steps);

  float pa[P][ND];
  uint32_t aidx[P];
  uint32_t start[P];
  float lmin[P];
  uint32_t live = 0;  // bit p: slot p holds a point
#pragma unroll
  for (int p = 0; p < P; ++p) {
    aidx[p] = 0;
    start[p] = 0;
    lmin[p] = kFltMax;
#pragma unroll
    for (int d = 0; d < ND; ++d) pa[p][d] = 0.f;
  }
  // Reservoir (one candidate per lane) and its valid-lane mask.
  float rp[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) rp[d] = 0.f;
  uint32_t ridx = 0;
  unsigned long long rmask = 0;
  bool drained = false;
  uint32_t gstep = 0;
  unsigned long long evals = 0;
  // Escalations staged in lanes (lane k holds the k-th pending one) and
  // appended 64 at a time, so the global counter sees one atomic per 64.
  uint32_t stg_idx = 0;
  float stg_min = 0.f;
  uint32_t n_stg = 0;
  auto flush = [&]() {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(args.esc_count, n_stg);
    base = Uniform(base);
    if (lane < n_stg) {
      args.esc_idx[base + lane] = stg_idx;
      args.esc_min[base + lane] = stg_min;
    }
    n_stg = 0;
  };

  float cmax2 = LoadCmax2(args.key);
  float bcur[2][ND];
  LoadStep<ND>(args.bx, args.by, args.bz, step, lane, bcur[0]);
  LoadStep<ND>(args.bx, args.by, args.bz, NextStep(step, n_steps), lane,
               bcur[1]);

  while (true) {
    // Refill.
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (!(live & (1u << p))) {
        while (rmask == 0ull && !drained) {
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(args.cursor, (unsigned)kWaveSize);
          base = Uniform(base);
          if (base >= args.pos_end) {
            drained = true;
            break;
          }
          const uint32_t pos = base + lane;
          bool valid = pos < args.pos_end;
          if (valid && args.list_min != nullptr)
            valid = !(args.list_min[pos] < cmax2);
          if (valid) {
            ridx = OrderAt(args.order_mode, args.order, args.perm_a, pos);
            const float* src = args.a + (uint64_t)ridx * ND;
#pragma unroll
            for (int d = 0; d < ND; ++d) rp[d] = src[d];
          }
          rmask = __ballot(valid);
        }
        if (rmask != 0ull) {
          const uint32_t k = (uint32_t)__builtin_ctzll(rmask);
          rmask &= rmask - 1ull;
#pragma unroll
          for (int d = 0; d < ND; ++d) pa[p][d] = ReadLaneF(rp[d], k);
          aidx[p] = ReadLaneU(ridx, k);
          lmin[p] = kFltMax;
          start[p] = gstep;
          live |= 1u << p;
        }
      }
    }
    if (live == 0) break;
    const uint32_t n_live = __builtin_popcount(live);

    // Prefetch the next two steps and cmax2, then evaluate the current two.
    const uint32_t s2 = NextStep(NextStep(step, n_steps), n_steps);
    const uint32_t s3 = NextStep(s2, n_steps);
    float bnext[2][ND];
    LoadStep<ND>(args.bx, args.by, args.bz, s2, lane, bnext[0]);
    LoadStep<ND>(args.bx, args.by, args.bz, s3, lane, bnext[1]);
    const uint32_t* cm_ptr = reinterpret_cast<const uint32_t*>(args.key) + 1;
    const uint32_t cm_next =
        __hip_atomic_load(cm_ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    EvalStep<ND, P>(pa, live, bcur[0], lmin);
    EvalStep<ND, P>(pa, live, bcur[1], lmin);
    step = s2;
    gstep += 2;
    evals += (unsigned long long)n_live * 128ull;

#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (live & (1u << p)) {
        const uint32_t done = gstep - start[p];
        if (WaveAny(lmin[p] < cmax2)) {
          live &= ~(1u << p);  // cannot raise the maximum
        } else if (done >= n_steps) {
          // Whole of B scanned: exact minimum.
          const float m = WaveMin(lmin[p]);
          if (lane == 0 && m != kFltMax)
            atomicMax(args.key, PackKey(m, aidx[p]));
          live &= ~(1u << p);
        } else if (done >= args.budget_steps) {
          const float m = WaveMin(lmin[p]);
          if (lane == n_stg) {
            stg_idx = aidx[p];
            stg_min = m;
          }
          if (++n_stg == (uint32_t)kWaveSize) flush();
          live &= ~(1u << p);
        }
      }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      bcur[0][d] = bnext[0][d];
      bcur[1][d] = bnext[1][d];
    }
    cmax2 = __uint_as_float(Uniform(cm_next));
  }
  if (n_stg > 0) flush();
  if (lane == 0) atomicAdd(args.evals, evals);
}

struct CoopArgs {
  const float* a;          // AoS A
  const uint32_t* list;    // A indices, processed in groups of P
  const float* list_min;   // partial minima (null: none known)
  uint32_t n_list;
  const float* bx;
  const float* by;
  const float* bz;
  uint32_t n_steps;        // n_bp / 64
  uint32_t seg_steps;      // steps per segment
  uint32_t n_segs;         // segments per group
  unsigned long long n_items;  // groups * n_segs
  unsigned long long* claim;   // item cursor
  uint32_t* broken;        // per list entry: 1 once known not to be a max
  uint32_t* pmin;          // per list entry: FP32 bits of the running min
  uint32_t* gdone;         // per group: completed segments
  unsigned long long* key;
  unsigned long long* evals;
};

HD_DEV void MarkBroken(uint32_t* flag) {
  __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Stages 1, 3 and the tail: work item = (group of P list entries, B
// segment); items of one group are consecutive, so a group is spread over
// many waves at once. Entry data is loaded lane-parallel (lane p < P loads
// entry p) and handed to the uniform slots with readlane.
template <int ND, int P>
__global__ __launch_bounds__(kBlock) void EbCoopKernel(CoopArgs args) {
  const uint32_t lane = LaneId();
  const uint32_t n_steps = args.n_steps;
  float pa[P][ND];
  float lmin[P];
  unsigned long long evals = 0;

  while (true) {
    unsigned long long t = 0;
    if (lane == 0) t = atomicAdd(args.claim, 1ull);
    t = __builtin_amdgcn_readfirstlane((uint32_t)t) |
        ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(t >> 32))
         << 32);
    if (t >= args.n_items) break;
    const uint32_t g = (uint32_t)(t / args.n_segs);
    const uint32_t s = (uint32_t)(t % args.n_segs);

    float cmax2 = LoadCmax2(args.key);
    // Lane-parallel entry load.
    const uint32_t e_lane = g * P + lane;
    bool ok = lane < (uint32_t)P && e_lane < args.n_list;
    float ep[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) ep[d] = 0.f;
    if (ok) {
      const uint32_t flag = __hip_atomic_load(
          &args.broken[e_lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (flag != 0u) {
        ok = false;
      } else if (args.list_min != nullptr && args.list_min[e_lane] < cmax2) {
        MarkBroken(&args.broken[e_lane]);
        ok = false;
      } else {
        const float* src = args.a + (uint64_t)args.list[e_lane] * ND;
#pragma unroll
        for (int d = 0; d < ND; ++d) ep[d] = src[d];
      }
    }
    uint32_t live = (uint32_t)__ballot(ok);
    if (live == 0) {
      // Every entry of the group is decided: skip its remaining segments.
      if (lane == 0)
        atomicMax(args.claim, (unsigned long long)(g + 1) * args.n_segs);
      continue;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      lmin[p] = kFltMax;
#pragma unroll
      for (int d = 0; d < ND; ++d) pa[p][d] = ReadLaneF(ep[d], p);
    }

    const uint32_t st_begin = s * args.seg_steps;
    const uint32_t st_end = min(st_begin + args.seg_steps, n_steps);
    uint32_t since_flags = 0;
    float bcur[ND];
    LoadStep<ND>(args.bx, args.by, args.bz, st_begin, lane, bcur);
    const uint32_t* cm_ptr = reinterpret_cast<const uint32_t*>(args.key) + 1;
    for (uint32_t st = st_begin; st < st_end; ++st) {
      // Prefetch the next step (clamped inside the segment) and cmax2.
      float bnext[ND];
      LoadStep<ND>(args.bx, args.by, args.bz, st + 1 < st_end ? st + 1 : st,
                   lane, bnext);
      const uint32_t cm_next = __hip_atomic_load(
          cm_ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      EvalStep<ND, P>(pa, live, bcur, lmin);
      evals += (unsigned long long)__builtin_popcount(live) * 64ull;
      if (((st - st_begin) & 1u) == 1u || st + 1 == st_end) {
        // Break checks every two steps.
        const bool refresh = (++since_flags & 7u) == 0u;
        bool flagged = false;
        if (refresh && lane < (uint32_t)P && (live & (1u << lane)))
          flagged = __hip_atomic_load(&args.broken[g * P + lane],
                                      __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) != 0u;
        const uint32_t flagged_mask = (uint32_t)__ballot(flagged);
#pragma unroll
        for (int p = 0; p < P; ++p) {
          if (live & (1u << p)) {
            if (WaveAny(lmin[p] < cmax2)) {
              if (lane == 0) MarkBroken(&args.broken[g * P + p]);
              live &= ~(1u << p);
            } else if (flagged_mask & (1u << p)) {
              live &= ~(1u << p);  // another segment broke it
            }
          }
        }
        if (live == 0) break;
      }
#pragma unroll
      for (int d = 0; d < ND; ++d) bcur[d] = bnext[d];
      cmax2 = __uint_as_float(Uniform(cm_next));
    }

    // Publish segment minima, then count the segment as done.
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (live & (1u << p)) {
        const float m = WaveMin(lmin[p]);
        if (lane == 0) atomicMin(&args.pmin[g * P + p], __float_as_uint(m));
      }
    }
    uint32_t prev = 0;
    if (lane == 0)
      prev = __hip_atomic_fetch_add(&args.gdone[g], 1u, __ATOMIC_ACQ_REL,
                                    __HIP_MEMORY_SCOPE_AGENT);
    prev = Uniform(prev);
    if (prev + 1 == args.n_segs && lane < (uint32_t)P) {
      // Last segment of the group: every segment has published; entries
      // never marked broken were scanned over all of B.
      const uint32_t e = g * P + lane;
      if (e < args.n_list &&
          __hip_atomic_load(&args.broken[e], __ATOMIC_ACQUIRE,
                            __HIP_MEMORY_SCOPE_AGENT) == 0u) {
        const uint32_t bits = __hip_atomic_load(
            &args.pmin[e], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (bits != kFltMaxBits)
          atomicMax(args.key, PackKey(__uint_as_float(bits), args.list[e]));
      }
    }
  }
  if (lane == 0) atomicAdd(args.evals, evals);
}

// Seed list: the first n order positions as A indices.
__global__ void SeedListKernel(const uint32_t* order, Feistel perm,
                               int order_mode, uint32_t n, uint32_t* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    out[i] = OrderAt(order_mode, order, perm, i);
}

constexpr int kP = 8;  // slots per wave / points per cooperative group

// Cooperative kernel over `list` (groups of kP entries). Each group's B scan
// is split into n_segs segments processed by different waves: as many as
// keep every wave busy, but at least 64 steps each, so a group that breaks
// early wastes at most n_segs short partial scans.
template <int ND>
int LaunchCoop(hd_context* ctx, const float* a, const uint32_t* list,
               const float* list_min, uint32_t n_list, const float* bx,
               const float* by, const float* bz, uint32_t n_steps,
               unsigned long long* key, unsigned long long* evals,
               uint32_t min_segs = 1) {
  if (n_list == 0) return HD_OK;
  const uint32_t n_groups = (n_list + kP - 1) / kP;
  const uint32_t total_waves =
      (uint32_t)PersistentBlocks(ctx, kBlock, 8) * (kBlock / kWaveSize);
  uint32_t n_segs = std::max(min_segs, total_waves / n_groups);
  n_segs = std::min(n_segs, std::max(1u, n_steps / 64u));
  const uint32_t seg_steps = (n_steps + n_segs - 1) / n_segs;
  n_segs = (n_steps + seg_steps - 1) / seg_steps;
  unsigned long long* claim = nullptr;
  uint32_t *broken = nullptr, *pmin = nullptr, *gdone = nullptr;
  HD_TRY(ContextArray(ctx, "coop_claim", 1, &claim));
  HD_TRY(ContextArray(ctx, "coop_broken", n_list, &broken));
  HD_TRY(ContextArray(ctx, "coop_pmin", n_list, &pmin));
  HD_TRY(ContextArray(ctx, "coop_gdone", n_groups, &gdone));
  HD_TRY(FillU64(ctx, claim, 0ull, 1));
  HD_TRY(FillU32(ctx, broken, 0u, n_list));
  HD_TRY(FillU32(ctx, pmin, kFltMaxBits, n_list));
  HD_TRY(FillU32(ctx, gdone, 0u, n_groups));
  CoopArgs args;
  args.a = a;
  args.list = list;
  args.list_min = list_min;
  args.n_list = n_list;
  args.bx = bx;
  args.by = by;
  args.bz = bz;
  args.n_steps = n_steps;
  args.seg_steps = seg_steps;
  args.n_segs = n_segs;
  args.n_items = (unsigned long long)n_groups * n_segs;
  args.claim = claim;
  args.broken = broken;
  args.pmin = pmin;
  args.gdone = gdone;
  args.key = key;
  args.evals = evals;
  int blocks = PersistentBlocks(ctx, kBlock, 8);
  const unsigned long long waves_needed = args.n_items;
  if ((unsigned long long)blocks * 4 > waves_needed)
    blocks = (int)std::max<unsigned long long>(1, (waves_needed + 3) / 4);
  EbCoopKernel<ND, kP><<<blocks, kBlock, 0, ctx->stream>>>(args);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

// Slot kernel over order positions [pos_begin, pos_end); escalations are
// appended to (esc_idx, esc_min) and counted in *esc_count.
template <int ND>
int LaunchSlots(hd_context* ctx, const float* a, int order_mode,
                const uint32_t* order, const float* list_min,
                const Feistel& perm_a, uint32_t pos_begin, uint32_t pos_end,
                const float* bx, const float* by, const float* bz,
                uint32_t n_steps, uint32_t budget_steps,
                unsigned long long* key, unsigned long long* evals,
                unsigned int* counters /* cursor, esc_count */,
                uint32_t* esc_idx, float* esc_min) {
  if (pos_begin >= pos_end) return HD_OK;
  HD_TRY(FillU32(ctx, counters, pos_begin, 1));
  HD_TRY(FillU32(ctx, counters + 1, 0u, 1));
  SlotArgs sa;
  sa.a = a;
  sa.order = order;
  sa.list_min = list_min;
  sa.perm_a = perm_a;
  sa.order_mode = order_mode;
  sa.pos_end = pos_end;
  sa.bx = bx;
  sa.by = by;
  sa.bz = bz;
  sa.n_steps = n_steps;
  sa.budget_steps = budget_steps;
  sa.key = key;
  sa.cursor = counters;
  sa.esc_idx = esc_idx;
  sa.esc_min = esc_min;
  sa.esc_count = counters + 1;
  sa.evals = evals;
  const uint32_t waves_needed = (pos_end - pos_begin + kP - 1) / kP;
  int blocks = PersistentBlocks(ctx, kBlock, 8);
  if ((uint32_t)blocks * 4u > waves_needed)
    blocks = (int)std::max(1u, (waves_needed + 3u) / 4u);
  EbSlotKernel<ND, kP><<<blocks, kBlock, 0, ctx->stream>>>(sa);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int ReadCount(hd_context* ctx, const unsigned int* d_count, uint32_t* out) {
  void* pinned = nullptr;
  HD_TRY(ContextPinned(ctx, 64, &pinned));
  unsigned int* h = static_cast<unsigned int*>(pinned);
  HD_HIP_TRY(hipMemcpyAsync(h, d_count, 4, hipMemcpyDeviceToHost,
                            ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  *out = *h;
  return HD_OK;
}

// B repacked for the scans: SoA, pseudo-randomly permuted (unless the
// caller asks for input order), padded to whole 64-point steps with copies
// of existing points (duplicates never change a minimum).
template <int ND>
int PrepareB(hd_context* ctx, const float* b, uint32_t n_b, bool shuffle,
             int32_t seed, EbTarget* t) {
  t->n_bp = (n_b + 63u) & ~63u;
  t->n_steps = t->n_bp / 64u;
  t->bx = t->by = t->bz = nullptr;
  HD_TRY(ContextArray(ctx, "eb_bx", t->n_bp, &t->bx));
  HD_TRY(ContextArray(ctx, "eb_by", t->n_bp, &t->by));
  if (ND == 3) HD_TRY(ContextArray(ctx, "eb_bz", t->n_bp, &t->bz));
  const Feistel perm_b = MakeFeistel(n_b, (uint64_t)(int64_t)seed, 1);
  return GatherSoA(ctx, ND, b, n_b, t->n_bp, shuffle, perm_b, t->bx, t->by,
                   t->bz);
}

// Rounds over the points still open: slots with a growing step budget;
// after each round the escalated points are sorted by partial minimum
// (largest first) and the top ones are completed cooperatively, which
// raises cmax2 to an attained value before the next, longer round. Entries
// whose partial minimum is already below cmax2 are skipped. ev[0..4]:
// counters for rounds 1-3, top completions, tail.
template <int ND>
int ProcessOpen(hd_context* ctx, const float* a, uint32_t n_a,
                const EbTarget& t, int order_mode, const uint32_t* order,
                const float* order_min, const Feistel& perm_a,
                uint32_t begin, uint32_t end, uint32_t first_budget,
                unsigned long long* key, unsigned long long* ev,
                EbRoundStats* rs) {
  static const uint32_t kBudgets[] = {16u, 256u, 4096u};
  const int n_rounds = first_budget ? 1 : 3;
  unsigned int* counters = nullptr;
  HD_TRY(ContextArray(ctx, "eb_counters", 4, &counters));
  uint32_t *idx_in = nullptr, *idx_out = nullptr, *idx_sorted = nullptr;
  float *min_in = nullptr, *min_out = nullptr, *min_sorted = nullptr;
  HD_TRY(ContextArray(ctx, "eb_esc_idx_a", n_a, &idx_out));
  HD_TRY(ContextArray(ctx, "eb_esc_min_a", n_a, &min_out));
  HD_TRY(ContextArray(ctx, "eb_esc_idx_b", n_a, &idx_sorted));
  HD_TRY(ContextArray(ctx, "eb_esc_min_b", n_a, &min_sorted));
  int mode = order_mode;
  const uint32_t* list = order;
  const float* list_min = order_min;
  for (int r = 0; r < n_rounds && begin < end; ++r) {
    const uint32_t budget = first_budget ? first_budget : kBudgets[r];
    rs->round_points[r] += end - begin;
    const int st_slot = StageBegin(ctx, r == 0 ? "eb_slot" : "eb_slot_more");
    HD_TRY(LaunchSlots<ND>(ctx, a, mode, list, list_min, perm_a, begin, end,
                           t.bx, t.by, t.bz, t.n_steps, budget, key, ev + r,
                           counters, idx_out, min_out));
    StageEnd(ctx, st_slot);
    uint32_t n_esc = 0;
    HD_TRY(ReadCount(ctx, counters + 1, &n_esc));
    if (n_esc == 0) {
      begin = end = 0;
      break;
    }
    const int st_sort = StageBegin(ctx, "eb_sort");
    HD_TRY(SortPairsDescF32(ctx, min_out, min_sorted, idx_out, idx_sorted,
                            n_esc));
    StageEnd(ctx, st_sort);
    const uint32_t k_top = std::min<uint32_t>(n_esc, 256u);
    rs->top_points += k_top;
    const int st_top = StageBegin(ctx, "eb_top");
    HD_TRY((LaunchCoop<ND>(ctx, a, idx_sorted, min_sorted, k_top, t.bx, t.by,
                           t.bz, t.n_steps, key, ev + 3)));
    StageEnd(ctx, st_top);
    // The sorted escalations become the next round's list; the next round
    // writes its sort into the buffer this round read from.
    mode = kOrderArray;
    list = idx_sorted;
    list_min = min_sorted;
    begin = k_top;
    end = n_esc;
    std::swap(idx_sorted, idx_in);
    std::swap(min_sorted, min_in);
    if (idx_sorted == nullptr) {
      HD_TRY(ContextArray(ctx, "eb_esc_idx_c", n_a, &idx_sorted));
      HD_TRY(ContextArray(ctx, "eb_esc_min_c", n_a, &min_sorted));
    }
  }
  if (begin < end) {
    // What is left needed more than the last budget: cooperative scans,
    // each group split over at least 16 waves.
    rs->tail_points += end - begin;
    const int st_tail = StageBegin(ctx, "eb_tail");
    HD_TRY((LaunchCoop<ND>(ctx, a, list + begin,
                           list_min ? list_min + begin : nullptr, end - begin,
                           t.bx, t.by, t.bz, t.n_steps, key, ev + 4, 16u)));
    StageEnd(ctx, st_tail);
  }
  return HD_OK;
}

template <int ND>
int RunEarlyBreakImpl(hd_context* ctx, const DirectedArgs& in,
                      hd_directed_result* out) {
  const uint32_t n_a = (uint32_t)in.n_a;
  const uint32_t n_b = (uint32_t)in.n_b;
  const bool shuffle = (in.flags & HD_FLAG_NO_SHUFFLE) == 0;

  const int st_prep = StageBegin(ctx, "eb_prep");
  EbTarget t;
  HD_TRY(PrepareB<ND>(ctx, in.b, n_b, shuffle, in.seed, &t));
  // Visit order of A: learned (array), pseudo-random (Feistel) or input.
  const uint32_t* order = nullptr;
  const Feistel perm_a = MakeFeistel(n_a, (uint64_t)(int64_t)in.seed, 0);
  int order_mode = shuffle ? kOrderFeistel : kOrderIdentity;
  if (in.flags & HD_FLAG_LEARNED_ORDER) {
    HD_TRY(LearnedOrder(ctx, in, &order));
    order_mode = kOrderArray;
  }
  // key, then one evaluation counter per stage.
  unsigned long long* key = nullptr;
  HD_TRY(ContextArray(ctx, "eb_key", 8, &key));
  // counters: seed, rounds 1-3, top completions, tail
  unsigned long long* ev = key + 1;
  HD_TRY(FillU64(ctx, key, 0ull, 8));
  StageEnd(ctx, st_prep);

  // Stage 1: exact minima of the first s0 points of the order.
  const uint32_t s0 = std::min<uint32_t>(
      n_a, in.seed_points ? in.seed_points : 256u);
  uint32_t* seed_list = nullptr;
  HD_TRY(ContextArray(ctx, "eb_seed_list", s0, &seed_list));
  const int st_seed = StageBegin(ctx, "eb_seed");
  SeedListKernel<<<std::max(1u, (s0 + 255u) / 256u), 256, 0, ctx->stream>>>(
      order, perm_a, order_mode, s0, seed_list);
  HD_HIP_TRY(hipGetLastError());
  HD_TRY((LaunchCoop<ND>(ctx, in.a, seed_list, nullptr, s0, t.bx, t.by, t.bz,
                         t.n_steps, key, ev + 0)));
  StageEnd(ctx, st_seed);

  EbRoundStats rs;
  HD_TRY(ProcessOpen<ND>(ctx, in.a, n_a, t, order_mode, order, nullptr, perm_a,
                         s0, n_a, in.slot_budget_steps, key, ev + 1, &rs));

  const int st_fin = StageBegin(ctx, "eb_finish");
  HD_TRY(FinishDirected(ctx, ND, in.a, in.b, in.n_b, key, ev, 6, out));
  StageEnd(ctx, st_fin);
  if (ctx->profiling) {
    void* pinned = nullptr;
    HD_TRY(ContextPinned(ctx, 64, &pinned));
    unsigned long long* h = static_cast<unsigned long long*>(pinned);
    HD_HIP_TRY(hipMemcpyAsync(h, ev, 48, hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
    StageCount(ctx, "eb_seed", h[0], s0);
    StageCount(ctx, "eb_slot", h[1], rs.round_points[0]);
    StageCount(ctx, "eb_slot_more", h[2] + h[3],
               rs.round_points[1] + rs.round_points[2]);
    StageCount(ctx, "eb_top", h[4], rs.top_points);
    StageCount(ctx, "eb_tail", h[5], rs.tail_points);
  }
  return HD_OK;
}

}  // namespace

int RunEarlyBreak(hd_context* ctx, const DirectedArgs& args,
                  hd_directed_result* out) {
  if (args.n_dims == 2) return RunEarlyBreakImpl<2>(ctx, args, out);
  return RunEarlyBreakImpl<3>(ctx, args, out);
}

int EbPrepareTarget(hd_context* ctx, int n_dims, const float* b, uint32_t n_b,
                    bool shuffle, int32_t seed, EbTarget* t) {
  if (n_dims == 2) return PrepareB<2>(ctx, b, n_b, shuffle, seed, t);
  return PrepareB<3>(ctx, b, n_b, shuffle, seed, t);
}

int EbFinishList(hd_context* ctx, int n_dims, const float* a, uint32_t n_a,
                 const EbTarget& t, const uint32_t* list, const float* list_min,
                 uint32_t n_list, unsigned long long* key,
                 unsigned long long* ev, EbRoundStats* rs) {
  const Feistel none = MakeFeistel(1, 0, 0);
  if (n_dims == 2)
    return ProcessOpen<2>(ctx, a, n_a, t, kOrderArray, list, list_min, none, 0,
                          n_list, 0, key, ev, rs);
  return ProcessOpen<3>(ctx, a, n_a, t, kOrderArray, list, list_min, none, 0,
                        n_list, 0, key, ev, rs);
}

}  // namespace hd
