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
  const uint32_t* order;  // visit order, or null for the Feistel order
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

// Stage 2. One wave = P slots; block = 4 independent waves.
template <int ND, int P>
__global__ __launch_bounds__(kBlock) void EbSlotKernel(SlotArgs args) {
  const uint32_t lane = LaneId();
  const uint32_t wave_id = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  uint32_t step = (uint32_t)(((uint64_t)wave_id * 7919u) % args.n_steps);

  float pa[P][ND];
  uint32_t aidx[P];
  uint32_t done_steps[P];
  bool live[P];
  float lmin[P];
  uint32_t res_base = 0, res_left = 0;
  bool drained = false;
  unsigned long long evals = 0;

#pragma unroll
  for (int p = 0; p < P; ++p) {
    live[p] = false;
    lmin[p] = kFltMax;
    done_steps[p] = 0;
    aidx[p] = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) pa[p][d] = 0.f;
  }

  while (true) {
    // Refill empty slots from the wave's reservoir of order positions.
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (!live[p] && !drained) {
        if (res_left == 0) {
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(args.cursor, (unsigned)kWaveSize);
          base = Uniform(base);
          if (base >= args.pos_end) {
            drained = true;
          } else {
            res_base = base;
            res_left = min((uint32_t)kWaveSize, args.pos_end - base);
          }
        }
        if (!drained) {
          const uint32_t pos = res_base;
          res_base++;
          res_left--;
          const uint32_t idx = Uniform(
              OrderAt(args.order_mode, args.order, args.perm_a, pos));
          aidx[p] = idx;
          LoadPoint<ND>(args.a, idx, pa[p]);
          lmin[p] = kFltMax;
          done_steps[p] = 0;
          live[p] = true;
        }
      }
    }
    int n_live = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) n_live += live[p] ? 1 : 0;
    if (n_live == 0) break;

    // Two steps of 64 B points, then one break check.
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const uint32_t j = (step << 6) + lane;
      const float bxv = args.bx[j];
      const float byv = args.by[j];
      const float bzv = ND == 3 ? args.bz[j] : 0.f;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (live[p]) {
          float d;
          if constexpr (ND == 2)
            d = D2(pa[p][0], pa[p][1], bxv, byv);
          else
            d = D2(pa[p][0], pa[p][1], pa[p][2], bxv, byv, bzv);
          lmin[p] = fminf(lmin[p], d);
        }
      }
      step = (step + 1 == args.n_steps) ? 0 : step + 1;
    }
    evals += (unsigned long long)n_live * 128ull;

    const float cmax2 = LoadCmax2(args.key);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (live[p]) {
        done_steps[p] += 2;
        if (WaveAny(lmin[p] < cmax2)) {
          live[p] = false;  // cannot raise the maximum
        } else if (done_steps[p] >= args.n_steps) {
          // Whole of B scanned: exact minimum.
          const float m = WaveMin(lmin[p]);
          if (lane == 0 && m != kFltMax)
            atomicMax(args.key, PackKey(m, aidx[p]));
          live[p] = false;
        } else if (done_steps[p] >= args.budget_steps) {
          const float m = WaveMin(lmin[p]);
          if (lane == 0) {
            const unsigned slot = atomicAdd(args.esc_count, 1u);
            args.esc_idx[slot] = aidx[p];
            args.esc_min[slot] = m;
          }
          live[p] = false;
        }
      }
    }
  }
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
HD_DEV uint32_t LoadFlag(const uint32_t* flag) {
  return Uniform(
      __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Stages 1 and 3: work item = (group of P list entries, B segment); items of
// one group are consecutive, so a group is spread over many waves at once.
template <int ND, int P>
__global__ __launch_bounds__(kBlock) void EbCoopKernel(CoopArgs args) {
  const uint32_t lane = LaneId();
  float pa[P][ND];
  uint32_t aidx[P];
  bool live[P];
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
    int n_live = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const uint32_t e = g * P + p;
      live[p] = false;
      aidx[p] = 0;
      lmin[p] = kFltMax;
#pragma unroll
      for (int d = 0; d < ND; ++d) pa[p][d] = 0.f;
      if (e < args.n_list) {
        if (LoadFlag(&args.broken[e]) == 0u) {
          if (args.list_min != nullptr && Uniform(args.list_min[e]) < cmax2) {
            if (lane == 0) MarkBroken(&args.broken[e]);
          } else {
            aidx[p] = Uniform(args.list[e]);
            LoadPoint<ND>(args.a, aidx[p], pa[p]);
            live[p] = true;
            n_live++;
          }
        }
      }
    }
    if (n_live == 0) {
      // Every entry of the group is decided: skip its remaining segments.
      if (lane == 0)
        atomicMax(args.claim, (unsigned long long)(g + 1) * args.n_segs);
      continue;
    }

    const uint32_t st_begin = s * args.seg_steps;
    const uint32_t st_end = min(st_begin + args.seg_steps, args.n_steps);
    uint32_t since_flags = 0;
    for (uint32_t st = st_begin; st < st_end; st += 2) {
      const bool two = st + 1 < st_end;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if (u == 0 || two) {
          const uint32_t j = ((st + u) << 6) + lane;
          const float bxv = args.bx[j];
          const float byv = args.by[j];
          const float bzv = ND == 3 ? args.bz[j] : 0.f;
#pragma unroll
          for (int p = 0; p < P; ++p) {
            if (live[p]) {
              float d;
              if constexpr (ND == 2)
                d = D2(pa[p][0], pa[p][1], bxv, byv);
              else
                d = D2(pa[p][0], pa[p][1], pa[p][2], bxv, byv, bzv);
              lmin[p] = fminf(lmin[p], d);
            }
          }
        }
      }
      evals += (unsigned long long)n_live * (two ? 128ull : 64ull);
      cmax2 = LoadCmax2(args.key);
      const bool refresh = (++since_flags & 7u) == 0u;
      n_live = 0;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (live[p]) {
          const uint32_t e = g * P + p;
          if (WaveAny(lmin[p] < cmax2)) {
            if (lane == 0) MarkBroken(&args.broken[e]);
            live[p] = false;
          } else if (refresh && LoadFlag(&args.broken[e]) != 0u) {
            live[p] = false;  // another segment broke it
          } else {
            n_live++;
          }
        }
      }
      if (n_live == 0) break;
    }

    // Publish segment minima, then count the segment as done.
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (live[p]) {
        const float m = WaveMin(lmin[p]);
        if (lane == 0) atomicMin(&args.pmin[g * P + p], __float_as_uint(m));
      }
    }
    uint32_t prev = 0;
    if (lane == 0)
      prev = __hip_atomic_fetch_add(&args.gdone[g], 1u, __ATOMIC_ACQ_REL,
                                    __HIP_MEMORY_SCOPE_AGENT);
    prev = Uniform(prev);
    if (prev + 1 == args.n_segs && lane == 0) {
      // Last segment of the group: every segment has published; entries
      // never marked broken were scanned over all of B.
      for (int p = 0; p < P; ++p) {
        const uint32_t e = g * P + p;
        if (e >= args.n_list) break;
        if (__hip_atomic_load(&args.broken[e], __ATOMIC_ACQUIRE,
                              __HIP_MEMORY_SCOPE_AGENT) != 0u)
          continue;
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

template <int ND>
int LaunchCoop(hd_context* ctx, const float* a, const uint32_t* list,
               const float* list_min, uint32_t n_list, const float* bx,
               const float* by, const float* bz, uint32_t n_steps,
               unsigned long long* key, unsigned long long* evals) {
  if (n_list == 0) return HD_OK;
  const uint32_t seg_steps = 64;  // 4096 B points per work item
  const uint32_t n_segs = (n_steps + seg_steps - 1) / seg_steps;
  const uint32_t n_groups = (n_list + kP - 1) / kP;
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
  unsigned long long waves_needed = args.n_items;
  int blocks = PersistentBlocks(ctx, kBlock, 8);
  if ((unsigned long long)blocks * 4 > waves_needed)
    blocks = (int)std::max<unsigned long long>(1, (waves_needed + 3) / 4);
  EbCoopKernel<ND, kP><<<blocks, kBlock, 0, ctx->stream>>>(args);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

template <int ND>
int RunEarlyBreakImpl(hd_context* ctx, const DirectedArgs& in,
                      hd_directed_result* out) {
  const uint32_t n_a = (uint32_t)in.n_a;
  const uint32_t n_b = (uint32_t)in.n_b;
  const bool shuffle = (in.flags & HD_FLAG_NO_SHUFFLE) == 0;
  const uint32_t n_bp = (n_b + 63u) & ~63u;
  const uint32_t n_steps = n_bp / 64u;

  const int st_prep = StageBegin(ctx, "eb_prep");
  float *bx = nullptr, *by = nullptr, *bz = nullptr;
  HD_TRY(ContextArray(ctx, "eb_bx", n_bp, &bx));
  HD_TRY(ContextArray(ctx, "eb_by", n_bp, &by));
  if (ND == 3) HD_TRY(ContextArray(ctx, "eb_bz", n_bp, &bz));
  const Feistel perm_b = MakeFeistel(n_b, (uint64_t)(int64_t)in.seed, 1);
  HD_TRY(GatherSoA(ctx, ND, in.b, n_b, n_bp, shuffle, perm_b, bx, by, bz));

  // Visit order of A: learned (array), pseudo-random (Feistel) or input.
  const uint32_t* order = nullptr;
  const Feistel perm_a = MakeFeistel(n_a, (uint64_t)(int64_t)in.seed, 0);
  int order_mode = shuffle ? kOrderFeistel : kOrderIdentity;
  if (in.flags & HD_FLAG_LEARNED_ORDER) {
    HD_TRY(LearnedOrder(ctx, in, &order));
    order_mode = kOrderArray;
  }

  // key, then one evaluation counter per stage (seed, slots, escalation).
  unsigned long long* key = nullptr;
  HD_TRY(ContextArray(ctx, "eb_key", 4, &key));
  unsigned long long* ev_seed = key + 1;
  unsigned long long* ev_slot = key + 2;
  unsigned long long* ev_esc = key + 3;
  HD_TRY(FillU64(ctx, key, 0ull, 4));
  StageEnd(ctx, st_prep);

  // Stage 1: seed.
  const uint32_t s0 = std::min<uint32_t>(
      n_a, in.seed_points ? in.seed_points : 256u);
  uint32_t* seed_list = nullptr;
  HD_TRY(ContextArray(ctx, "eb_seed_list", s0, &seed_list));
  const int st_seed = StageBegin(ctx, "eb_seed");
  SeedListKernel<<<std::max(1u, (s0 + 255u) / 256u), 256, 0, ctx->stream>>>(
      order, perm_a, order_mode, s0, seed_list);
  HD_HIP_TRY(hipGetLastError());
  HD_TRY((LaunchCoop<ND>(ctx, in.a, seed_list, nullptr, s0, bx, by, bz,
                         n_steps, key, ev_seed)));
  StageEnd(ctx, st_seed);

  // Stage 2: slots over the remaining order positions.
  uint32_t n_esc = 0;
  if (s0 < n_a) {
    unsigned int *cursor = nullptr, *esc_count = nullptr;
    uint32_t* esc_idx = nullptr;
    float* esc_min = nullptr;
    HD_TRY(ContextArray(ctx, "eb_counters", 2, &cursor));
    esc_count = cursor + 1;
    HD_TRY(ContextArray(ctx, "eb_esc_idx", n_a, &esc_idx));
    HD_TRY(ContextArray(ctx, "eb_esc_min", n_a, &esc_min));
    HD_TRY(FillU32(ctx, cursor, s0, 1));
    HD_TRY(FillU32(ctx, esc_count, 0u, 1));
    SlotArgs sa;
    sa.a = in.a;
    sa.order = order;
    sa.perm_a = perm_a;
    sa.order_mode = order_mode;
    sa.pos_end = n_a;
    sa.bx = bx;
    sa.by = by;
    sa.bz = bz;
    sa.n_steps = n_steps;
    sa.budget_steps =
        in.slot_budget_steps ? in.slot_budget_steps : 512u;  // 32K evals
    sa.key = key;
    sa.cursor = cursor;
    sa.esc_idx = esc_idx;
    sa.esc_min = esc_min;
    sa.esc_count = esc_count;
    sa.evals = ev_slot;
    const uint32_t waves_needed = (n_a - s0 + kP - 1) / kP;
    int blocks = PersistentBlocks(ctx, kBlock, 8);
    if ((uint32_t)blocks * 4u > waves_needed)
      blocks = (int)std::max(1u, (waves_needed + 3u) / 4u);
    const int st_slot = StageBegin(ctx, "eb_slot");
    EbSlotKernel<ND, kP><<<blocks, kBlock, 0, ctx->stream>>>(sa);
    HD_HIP_TRY(hipGetLastError());
    StageEnd(ctx, st_slot);

    // Stage 3: escalated points, most promising first.
    void* pinned = nullptr;
    HD_TRY(ContextPinned(ctx, 64, &pinned));
    unsigned int* h_count = static_cast<unsigned int*>(pinned);
    HD_HIP_TRY(hipMemcpyAsync(h_count, esc_count, 4, hipMemcpyDeviceToHost,
                              ctx->stream));
    HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
    n_esc = *h_count;
    if (n_esc > 0) {
      float* sorted_min = nullptr;
      uint32_t* sorted_idx = nullptr;
      HD_TRY(ContextArray(ctx, "eb_esc_min_sorted", n_esc, &sorted_min));
      HD_TRY(ContextArray(ctx, "eb_esc_idx_sorted", n_esc, &sorted_idx));
      const int st_sort = StageBegin(ctx, "eb_sort");
      HD_TRY(SortPairsDescF32(ctx, esc_min, sorted_min, esc_idx, sorted_idx,
                              n_esc));
      StageEnd(ctx, st_sort);
      const int st_esc = StageBegin(ctx, "eb_escalate");
      HD_TRY((LaunchCoop<ND>(ctx, in.a, sorted_idx, sorted_min, n_esc, bx, by,
                             bz, n_steps, key, ev_esc)));
      StageEnd(ctx, st_esc);
    }
  }
  const int st_fin = StageBegin(ctx, "eb_finish");
  HD_TRY(FinishDirected(ctx, ND, in.a, in.b, in.n_b, key, key + 1, 3, out));
  StageEnd(ctx, st_fin);
  if (ctx->profiling) {
    void* pinned = nullptr;
    HD_TRY(ContextPinned(ctx, 64, &pinned));
    unsigned long long* h = static_cast<unsigned long long*>(pinned);
    HD_HIP_TRY(hipMemcpyAsync(h, key + 1, 24, hipMemcpyDeviceToHost,
                              ctx->stream));
    HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
    StageCount(ctx, "eb_seed", h[0], s0);
    StageCount(ctx, "eb_slot", h[1], n_a - s0);
    StageCount(ctx, "eb_escalate", h[2], n_esc);
  }
  return HD_OK;
}

}  // namespace

int RunEarlyBreak(hd_context* ctx, const DirectedArgs& args,
                  hd_directed_result* out) {
  if (args.n_dims == 2) return RunEarlyBreakImpl<2>(ctx, args, out);
  return RunEarlyBreakImpl<3>(ctx, args, out);
}

}  // namespace hd
