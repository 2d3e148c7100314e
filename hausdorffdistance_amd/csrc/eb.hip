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
#include <type_traits>

#include "hd_device.h"
#include "hd_internal.h"
#include "hd_lbvh.h"
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

// Segment scans. B's steps are dealt to n_segs segments: segment s takes
// steps s, s + n_segs, s + 2 n_segs, ... (about seg_steps of them), so the
// segments of an entry advance through B's visit order side by side. (With
// contiguous ranges instead, the points of one region, which the stratified
// order spreads evenly, sit at similar offsets in every range, and a hit
// comes no sooner than in one range: 10M GMM, learned order, 2% more
// evaluations, tail 2.46 against 2.34 ms.) Every segment has its own queue
// over the same list of order positions; a wave works on one segment at a
// time and
// keeps P slots, each holding one A point in scalar registers. The wave's
// lanes stream the segment's B steps cyclically (64 coalesced SoA points per
// step) and every B point in registers is tested against all P slots, so a
// slot that joins at any step has covered the whole segment seg_len steps
// later. Per slot:
//  * its running minimum falls strictly below cmax2: dropped (cannot raise
//    the maximum);
//  * n_segs == 1 and it scanned all of B: its exact minimum goes into the
//    packed key;
//  * n_segs == 1 and it used budget_steps first: escalated with its partial
//    minimum (esc_idx / esc_min);
//  * n_segs > 1 and it scanned its segment: the segment minimum goes into
//    pmin[pos] (atomic min) and gdone[pos] counts it; the wave completing the
//    last segment of an entry no segment dropped publishes pmin[pos];
//  * n_segs > 1 and another segment dropped the entry (broken[pos], polled
//    every 16 steps): dropped.
struct ScanArgs {
  const float* a;         // AoS A
  const uint32_t* order;  // visit order (kOrderArray)
  const float* list_min;  // optional partial minima per position: a position
                          // whose value is already below cmax2 is skipped
  Feistel perm_a;
  int order_mode;         // kOrder*
  uint32_t pos_end;       // positions [cursor start, pos_end)
  const float* bx;
  const float* by;
  const float* bz;
  uint32_t n_steps;       // n_bp / 64
  uint32_t seg_steps;     // steps per segment
  uint32_t n_segs;        // segments; 1 = whole-B scans
  uint32_t budget_steps;  // n_segs == 1: steps before a slot escalates
  uint32_t start_step;    // n_segs == 1: the step every wave starts at
  unsigned long long* key;
  unsigned int* cursors;  // per segment: next order position
  uint32_t claim;         // positions taken per cursor claim (<= 64)
  uint32_t* esc_idx;      // escalated A indices
  float* esc_min;         // their partial minima
  unsigned int* esc_count;
  uint32_t* broken;       // n_segs > 1, per position: 1 once dropped
  uint32_t* pmin;         // n_segs > 1, per position: FP32 bits of the min
  uint32_t* gdone;        // n_segs > 1, per position: segments completed
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

// B step `step` (64 SoA points) into registers. `step` is wave-uniform:
// each row is addressed through a scalar buffer descriptor (64 floats) and
// the lane's byte offset, so no per-lane address arithmetic is spent.
HD_DEV float LoadRow(const float* base, uint64_t row, uint32_t lane4) {
  __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(base + row), (short)0, 256, 0x00020000);
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)lane4, 0, 0));
}

template <int ND>
HD_DEV void LoadStep(const float* __restrict__ bx, const float* __restrict__ by,
                     const float* __restrict__ bz, uint32_t step,
                     uint32_t lane, float* b) {
  const uint64_t row = (uint64_t)Uniform(step) << 6;
  const uint32_t lane4 = lane << 2;
  b[0] = LoadRow(bx, row, lane4);
  b[1] = LoadRow(by, row, lane4);
  if constexpr (ND == 3) b[2] = LoadRow(bz, row, lane4);
}

// Running minima are FP32 bit patterns (MinBits).
template <int ND, int P, bool kAllLive>
HD_DEV void EvalStep(const float (&pa)[P][ND], uint32_t live_mask,
                     const float* b, uint32_t (&lmin)[P]) {
#pragma unroll
  for (int p = 0; p < P; ++p) {
    if (kAllLive || (live_mask & (1u << p))) {
      float d;
      if constexpr (ND == 2)
        d = D2(pa[p][0], pa[p][1], b[0], b[1]);
      else
        d = D2(pa[p][0], pa[p][1], pa[p][2], b[0], b[1], b[2]);
      lmin[p] = MinBits(lmin[p], __float_as_uint(d));
    }
  }
}

// Slot state of one wave over one segment; S = steps per iteration.
template <int ND, int P, int S>
struct ScanWave {
  static constexpr uint32_t kAll = (1u << P) - 1u;
  float pa[P][ND];
  uint32_t aidx[P];
  uint32_t lpos[P];
  uint32_t start[P];
  uint32_t lmin[P];
  uint32_t lane_pos;  // lane p < P: list position of slot p
  uint32_t live;
  float bcur[S][ND];
  uint32_t step;
  uint32_t gstep;
  unsigned long long evals;
};

// The step after st in a segment: every seg_stride-th step from seg_begin,
// cyclic below seg_end.
HD_DEV uint32_t NextInSeg(uint32_t st, uint32_t seg_begin, uint32_t seg_end,
                          uint32_t seg_stride) {
  return st + seg_stride >= seg_end ? seg_begin : st + seg_stride;
}

// Steps a wave's slots S B steps at a time until a slot breaks (hit), is
// flagged by another segment (gone), or reaches next_due.
template <int ND, int P, int S, bool kAllLive>
HD_DEV void ScanIterate(const ScanArgs& args, ScanWave<ND, P, S>& w,
                        uint32_t lane, uint32_t seg_begin, uint32_t seg_end,
                        uint32_t seg_stride, bool split, uint32_t next_due,
                        uint32_t& cmax2, uint32_t& hit, uint32_t& gone) {
  const uint32_t* cm_ptr = reinterpret_cast<const uint32_t*>(args.key) + 1;
  while (true) {
    float bnext[S][ND];
    uint32_t st = w.step;
#pragma unroll
    for (int k = 0; k < S; ++k)
      st = NextInSeg(st, seg_begin, seg_end, seg_stride);
    const uint32_t step_next = st;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      LoadStep<ND>(args.bx, args.by, args.bz, st, lane, bnext[k]);
      st = NextInSeg(st, seg_begin, seg_end, seg_stride);
    }
    const uint32_t cm_next =
        __hip_atomic_load(cm_ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Split scans poll the other segments' verdicts every iteration (S
    // steps; every 16 steps: the 10M GMM seeds and top completions took
    // 2-8% longer).
    const bool poll = split;
    uint32_t flag = 0;
    if (poll && lane < (uint32_t)P && ((w.live >> lane) & 1u))
      flag = __hip_atomic_load(&args.broken[w.lane_pos], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int k = 0; k < S; ++k)
      EvalStep<ND, P, kAllLive>(w.pa, w.live, w.bcur[k], w.lmin);
    w.gstep += S;
    w.evals += (unsigned long long)__builtin_popcount(w.live) * (64ull * S);
#pragma unroll
    for (int p = 0; p < P; ++p)
      if (WaveAny(w.lmin[p] < cmax2)) hit |= 1u << p;
    hit &= w.live;
    if (poll) gone = (uint32_t)__ballot(flag != 0u) & w.live & ~hit;
#pragma unroll
    for (int k = 0; k < S; ++k)
#pragma unroll
      for (int d = 0; d < ND; ++d) w.bcur[k][d] = bnext[k][d];
    w.step = step_next;
    cmax2 = Uniform(cm_next);
    if ((hit | gone) != 0u || w.gstep >= next_due) break;
  }
}

template <int ND, int P, int S>
HD_DEV void ScanSegment(const ScanArgs& args, uint32_t seg, uint32_t lane,
                        ScanWave<ND, P, S>& w,
                        uint32_t& stg_idx, float& stg_min, uint32_t& n_stg) {
  constexpr uint32_t kAll = ScanWave<ND, P, S>::kAll;
  const bool split = args.n_segs > 1;
  // Segment seg: steps seg, seg + n_segs, ... (n_segs == 1: all of B).
  const uint32_t seg_stride = args.n_segs;
  const uint32_t seg_begin = seg;
  const uint32_t seg_len =
      (args.n_steps - seg + args.n_segs - 1) / args.n_segs;
  const uint32_t seg_end = args.n_steps;
  const uint32_t limit = split ? seg_len : min(args.budget_steps, seg_len);
  unsigned int* cursor = args.cursors + seg;

  w.live = 0;
  w.gstep = 0;
  // All waves start at the same step and walk B together (shared L2
  // lines; 10M GMM: L2 hit rate 0.27 -> 0.66): the segment's first step,
  // or a per-launch step for whole-B scans, so that successive rounds do
  // not all re-scan the same first steps.
  w.step = split ? seg_begin : seg_begin + args.start_step % seg_len;
  uint32_t next_due = 0xFFFFFFFFu;
  // Reservoir: one candidate per lane, claimed 64 positions at a time.
  float rp[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) rp[d] = 0.f;
  uint32_t ridx = 0, rpos = 0;
  unsigned long long rmask = 0;
  bool drained = false;

  auto flush = [&]() __attribute__((always_inline)) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(args.esc_count, n_stg);
    base = Uniform(base);
    if (lane < n_stg) {
      args.esc_idx[base + lane] = stg_idx;
      args.esc_min[base + lane] = stg_min;
    }
    n_stg = 0;
  };

  uint32_t cmax2 = __float_as_uint(LoadCmax2(args.key));
  {
    uint32_t st = w.step;
#pragma unroll
    for (int k = 0; k < S; ++k) {
      LoadStep<ND>(args.bx, args.by, args.bz, st, lane, w.bcur[k]);
      st = NextInSeg(st, seg_begin, seg_end, seg_stride);
    }
  }
  while (true) {
    // Refill free slots from the reservoir.
    if (w.live != kAll) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (!(w.live & (1u << p))) {
          while (rmask == 0ull && !drained) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(cursor, args.claim);
            base = Uniform(base);
            if (base >= args.pos_end) {
              drained = true;
              break;
            }
            const uint32_t pos = base + lane;
            bool valid = lane < args.claim && pos < args.pos_end;
            if (valid && args.list_min != nullptr)
              valid = !(__float_as_uint(args.list_min[pos]) < cmax2);
            if (valid && split)
              valid = __hip_atomic_load(&args.broken[pos], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT) == 0u;
            if (valid) {
              ridx = OrderAt(args.order_mode, args.order, args.perm_a, pos);
              rpos = pos;
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
            for (int d = 0; d < ND; ++d) w.pa[p][d] = ReadLaneF(rp[d], k);
            w.aidx[p] = ReadLaneU(ridx, k);
            w.lpos[p] = ReadLaneU(rpos, k);
            if (lane == (uint32_t)p) w.lane_pos = w.lpos[p];
            w.lmin[p] = kFltMaxBits;
            w.start[p] = w.gstep;
            w.live |= 1u << p;
            next_due = min(next_due, w.gstep + limit);
          }
        }
      }
    }
    if (w.live == 0) break;

    // Iterate until a slot breaks, is due, or is flagged by another segment.
    // The loop is instantiated twice: every slot live (no per-slot
    // branches) and some slots free (the cursor has drained).
    uint32_t hit = 0, gone = 0;
    if (w.live == kAll)
      ScanIterate<ND, P, S, true>(args, w, lane, seg_begin, seg_end,
                                  seg_stride, split, next_due, cmax2, hit,
                                  gone);
    else
      ScanIterate<ND, P, S, false>(args, w, lane, seg_begin, seg_end,
                                   seg_stride, split, next_due, cmax2, hit,
                                   gone);

    w.live &= ~(hit | gone);
    if (w.gstep >= next_due) {
      uint32_t nd = 0xFFFFFFFFu;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (!(w.live & (1u << p))) continue;
        const uint32_t done = w.gstep - w.start[p];
        if (done < limit) {
          nd = min(nd, w.start[p] + limit);
          continue;
        }
        const uint32_t mb = WaveMinU32(w.lmin[p]);
        w.live &= ~(1u << p);
        if (done >= seg_len) {
          if (!split) {
            // Whole of B scanned: exact minimum.
            if (lane == 0 && mb != kFltMaxBits)
              atomicMax(args.key, PackKey(__uint_as_float(mb), w.aidx[p]));
          } else if (lane == 0) {
            // Segment minimum, then the completion count. The count's
            // increment depends on the value the atomic min returned, so it
            // is issued only after the min has been performed at the device
            // coherence point; the wave whose increment completes the count
            // then reads the minimum over every segment there. Segments that
            // broke or were dropped never count, so a full count means an
            // exact minimum. (Fences would also order this, but an acquire
            // invalidates the L2 that holds B.)
            const uint32_t e = w.lpos[p];
            const uint32_t was = atomicMin(&args.pmin[e], mb);
            const uint32_t prev = __hip_atomic_fetch_add(
                &args.gdone[e], 1u + (was >> 31), __ATOMIC_RELAXED,
                __HIP_MEMORY_SCOPE_AGENT);
            if (prev + 1u == args.n_segs) {
              const uint32_t bits = __hip_atomic_load(
                  &args.pmin[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              if (bits != kFltMaxBits)
                atomicMax(args.key, PackKey(__uint_as_float(bits), w.aidx[p]));
            }
          }
        } else {
          if (lane == n_stg) {
            stg_idx = w.aidx[p];
            stg_min = __uint_as_float(mb);
          }
          if (++n_stg == (uint32_t)kWaveSize) flush();
        }
      }
      next_due = nd;
    }
    if (split && hit != 0u && lane < (uint32_t)P && ((hit >> lane) & 1u)) {
      // Tell the other segments' waves this entry is decided.
      __hip_atomic_store(&args.broken[w.lane_pos], 1u, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    if ((hit | gone) != 0u) {
      // Recompute the earliest due step over the remaining slots.
      uint32_t nd = 0xFFFFFFFFu;
#pragma unroll
      for (int p = 0; p < P; ++p)
        if (w.live & (1u << p)) nd = min(nd, w.start[p] + limit);
      next_due = nd;
    }
  }
  if (n_stg > 0) flush();
}

template <int ND, int P, int S>
__global__ __launch_bounds__(kBlock) void EbScanKernel(ScanArgs args) {
  const uint32_t lane = LaneId();
  const uint32_t wave_id = Uniform((blockIdx.x * kBlock + threadIdx.x) >> 6);
  const uint32_t n_waves = gridDim.x * (kBlock / kWaveSize);
  ScanWave<ND, P, S> w;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    w.aidx[p] = 0;
    w.lpos[p] = 0;
    w.start[p] = 0;
    w.lmin[p] = kFltMaxBits;
#pragma unroll
    for (int d = 0; d < ND; ++d) w.pa[p][d] = 0.f;
  }
  w.evals = 0;
  w.lane_pos = 0;
  // Escalations staged in lanes (lane k holds the k-th pending one) and
  // appended 64 at a time.
  uint32_t stg_idx = 0, n_stg = 0;
  float stg_min = 0.f;
  // Wave k of the sweep takes segment k % n_segs; with fewer waves than
  // segments a wave takes segments wave_id, wave_id + n_waves, ...
  const uint32_t sweep = max(n_waves, args.n_segs);
  for (uint32_t k = wave_id; k < sweep; k += n_waves) {
    const uint32_t seg = k % args.n_segs;
    ScanSegment<ND, P, S>(args, seg, lane, w, stg_idx, stg_min, n_stg);
  }
  if (lane == 0) CounterAdd(args.evals, w.evals);
}

// Seed list: the first n order positions as A indices.
__global__ void SeedListKernel(const uint32_t* order, Feistel perm,
                               int order_mode, uint32_t n, uint32_t* out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    out[i] = OrderAt(order_mode, order, perm, i);
}

// ---------------------------------------------------------------------------
// Matrix-core filter for the budgeted rounds.
//
// With coordinates centred on c (p' = fl(p - c)), the 16x16x4 FP32 MFMA
// computes for 16 B points i and 16 A points j at once
//   2D: D[i][j] = x'_i (-2 x'_j) + y'_i (-2 y'_j) + |b'_i|^2 * 1 + 1 * |a'_j|^2
//   3D: D[i][j] = x'_i (-2 x'_j) + y'_i (-2 y'_j) + z'_i (-2 z'_j)
//                 + (|b'_i|^2 + C) * 1,   C = rho^2
// i.e. |a - b|^2 (3D: plus C - |a'_j|^2 >= 0, taken back through the
// threshold), with an error below E = 48 u rho^2 in 2D and 64 u rho^2 in 3D
// (u = 2^-24, rho = max |p - c| over A and B): products and sums inside the
// MFMA (at most 2u each, in any order: 8u times the sum of the magnitudes
// of the four terms, at most 4 rho^2), the norms and C (2-3u each), the
// centring (8u rho^2) and, in 3D, the threshold arithmetic (10u rho^2).
// Every D is >= -E, so running minima compare as signed integers on their
// bit patterns (any negative value is below the positive thresholds, which
// is all the test needs). A point is dropped only when its
// approximate running minimum is below T = cmax2 (1 - 2^-19) - E rounded
// down, which proves its exact FP32 minimum (the reference's unfused
// formula, within 4u d2 of |a - b|^2) is below cmax2. A step whose
// approximate minimum for a point lies in [T, cmax2 (1 + 2^-19) + E) may
// hold an exact hit: that point is then tested exactly against the step's
// 64 original B points, so points break at the same place as in the exact
// scan and the filter adds no evaluations. Escalated points carry approx +
// E rounded up, an upper bound of their exact partial minimum, so a later
// "below cmax2" test on it is still a proof.
//
// A wave holds 64 A points (4 tiles of 16) as MFMA B operands: lane l keeps
// component l/16 of point l%16 of each tile. A B tile (16 points) is one
// coalesced 256-byte load, lane l holding component l/16 of point l%16. The
// output lane l holds D for A point l%16 against B points 4(l/16)..+3, so a
// point's running minimum lives in 4 lanes of one register per tile. Slot
// metadata (A index, order position, first step) lives in lane s for slot
// s = 16 t + j.
//
// Grouped budgeted rounds (n_groups = G > 0): the budget is split over
// W = n_segs / G segments of seg_steps steps; position p belongs to group
// p mod G and is scanned by the W segments s with s mod G = p mod G (disjoint
// B ranges, W * seg_steps = budget steps in all), concurrently. A hit in any
// of them drops the entry (broken[p]); a segment that completes without one
// folds its bound (approx + E, rounded up) into pmin[p] and counts gdone[p];
// the W-th count escalates the entry with pmin[p]. The same work as one
// budgeted scan, in W times as many shorter items: the launch fills the
// chip and its drain is a quarter as long.
typedef float floatx4 __attribute__((ext_vector_type(4)));

// B steps (64 points each) per filter iteration: drops, escalations and
// refills are decided once per iteration.
constexpr int kFilterSteps = 4;

struct MfmaArgs {
  const float* a;         // AoS A (original coordinates)
  const uint32_t* order;
  const float* list_min;  // exact-bound partial minima (skip below cmax2)
  Feistel perm_a;
  int order_mode;
  uint32_t pos_end;
  unsigned int* cursors;  // per segment: next order position (grouped:
                          // next index within the segment's group)
  uint32_t claim;         // positions per cursor claim (<= 64)
  uint32_t seg_steps;     // steps per segment
  uint32_t n_segs;        // 1: budgeted round over all of B
  uint32_t n_groups;      // 0: every segment serves every position; else a
                          // budgeted round split over n_segs / n_groups
                          // segments per position (see EbMfmaKernel)
  uint32_t pos_begin;     // grouped: first position
  uint32_t* broken;       // n_segs > 1, per position: 1 once a hit is known
  uint32_t* gdone;        // n_segs > 1, per position: segments completed
  uint32_t* pmin;         // grouped, per position: min of segment bounds
  uint32_t* cand;         // n_segs > 1: A indices that completed every
  unsigned int* cand_count;  // segment without a hit (exact scan follows)
  const float* bt;        // B tiles, 64 floats each, 4 per step
  const float* bx;        // the same points, original coordinates (SoA)
  const float* by;
  const float* bz;
  uint32_t n_steps;
  uint32_t budget_steps;  // < n_steps: no slot ever completes a full scan
  uint32_t start_step;    // n_segs == 1: the step every wave starts at
  float cx, cy, cz;       // centre
  float err;              // E
  float off;              // 3D: C
  unsigned long long* key;
  uint32_t* esc_idx;
  float* esc_min;
  unsigned int* esc_count;
  unsigned long long* evals;
};

// Lane's register of one B tile (lane l: component l/16 of point l%16).
HD_DEV float LoadTile(const float* bt, uint64_t tile, uint32_t lane4) {
  return LoadRow(bt, tile << 6, lane4);
}

// The four tiles of one step through one buffer descriptor (1 KiB of
// records; the tile offsets go into the instructions' immediate field).
HD_DEV void LoadStepTiles(const float* bt, uint32_t step, uint32_t lane4,
                          float* out) {
  __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(bt + ((uint64_t)Uniform(step) << 8)), (short)0, 1024,
      0x00020000);
#pragma unroll
  for (int q = 0; q < 4; ++q)
    out[q] = __uint_as_float(
        __builtin_amdgcn_raw_buffer_load_b32(r, (int)(lane4 + 256u * q), 0, 0));
}

// A point's MFMA operand lane (component kc) and, in 3D, C - |a'|^2: the
// point centred on c, then [-2x', -2y', 1, |a'|^2] (2D) or
// [-2x', -2y', -2z', 1] (3D), so that D = B tile x operand = |a - b|^2
// (3D: plus C - |a'|^2).
template <int ND>
HD_DEV void FilterOperand(float px, float py, float pz, float cx, float cy,
                          float cz, float off, uint32_t kc, float* val,
                          float* cna) {
  const float x = px - cx;
  const float y = py - cy;
  float z = 0.f, na;
  if constexpr (ND == 3) {
    z = pz - cz;
    na = x * x + y * y + z * z;
  } else {
    na = x * x + y * y;
  }
  if constexpr (ND == 2)
    *val = kc == 0u ? -2.f * x : kc == 1u ? -2.f * y : kc == 2u ? 1.f : na;
  else
    *val = kc == 0u ? -2.f * x : kc == 1u ? -2.f * y : kc == 2u ? -2.f * z : 1.f;
  *cna = off - na;
}

// Test hook of the filter's arithmetic: D for every (A point, B point) of
// 16-point tiles, from the same tiles, operands and MFMA as EbMfmaKernel.
// Wave w takes A tile w / n_btiles and B tile w % n_btiles; out[i * n_b +
// j] = D for A point i and B point j (rows past n_a / columns past n_b are
// not written); cna_out[i] = C - |a'_i|^2 (3D; D - cna approximates d2).
template <int ND>
__global__ __launch_bounds__(64) void FilterProbeKernel(
    const float* __restrict__ bt, uint32_t n_btiles, uint32_t n_b,
    const float* __restrict__ a, uint32_t n_a, float cx, float cy, float cz,
    float off, float* __restrict__ out, float* __restrict__ cna_out) {
  const uint32_t lane = LaneId();
  const uint32_t ta = blockIdx.x / n_btiles, tb = blockIdx.x % n_btiles;
  const uint32_t j = lane & 15u, kc = lane >> 4;
  const uint32_t ia = ta * 16u + j;
  const uint32_t src = ia < n_a ? ia : n_a - 1u;
  const float* p = a + (uint64_t)src * ND;
  float val, cn;
  FilterOperand<ND>(p[0], p[1], ND == 3 ? p[2] : 0.f, cx, cy, cz, off, kc, &val, &cn);
  const float btile = LoadTile(bt, tb, lane << 2);
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  const floatx4 d = __builtin_amdgcn_mfma_f32_16x16x4f32(btile, val, zero, 0, 0, 0);
  if (ia < n_a) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t ib = tb * 16u + 4u * kc + (uint32_t)k;
      if (ib < n_b) out[(uint64_t)ia * n_b + ib] = d[k];
    }
    if (ND == 3 && kc == 0u) cna_out[ia] = cn;
  }
}

// One v_min3_i32 (the compiler would split min(a, min(b, c)) into two
// instructions to shorten the chain on a; here the chain is on one
// register per tile and every VALU slot counts against the matrix core).
HD_DEV int32_t Min3I(int32_t a, int32_t b, int32_t c) {
  int32_t r;
  asm("v_min3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// Thresholds for the approximate minima from cmax2 bits: below the lower
// one a point is proved below cmax2, at or above the upper one a step
// cannot hold an exact hit.
HD_DEV float FilterThreshold(uint32_t cm_bits, float err) {
  const double t = (double)__uint_as_float(cm_bits) * (1.0 - 0x1p-19) - (double)err;
  return t > 0.0 ? __double2float_rd(t) : -kFltMax;
}
HD_DEV float FilterUpper(uint32_t cm_bits, float err) {
  const double t = (double)__uint_as_float(cm_bits) * (1.0 + 0x1p-19) + (double)err;
  return __double2float_ru(t);
}

template <int ND, bool kGrouped>
__global__ __launch_bounds__(kBlock) void EbMfmaKernel(MfmaArgs args) {
  const uint32_t lane = LaneId();
  const uint32_t lane4 = lane << 2;
  const uint32_t j = lane & 15u;   // A point within a tile
  const uint32_t kc = lane >> 4;   // operand component
  const uint32_t wave_id = Uniform((blockIdx.x * kBlock + threadIdx.x) >> 6);
  const uint32_t n_waves = gridDim.x * (kBlock / kWaveSize);
  const uint32_t n_steps = args.n_steps;
  const bool split = args.n_segs > 1;
  constexpr bool grouped = kGrouped;
  const uint32_t per_entry = grouped ? args.n_segs / args.n_groups : args.n_segs;

  // bop: operand lanes; mn: running minima (FP32 bits as signed integers);
  // cna: 3D, C - |a'|^2 of the lane's point.
  float bop[4], cna[4];
  int32_t mn[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    bop[t] = 0.f;
    mn[t] = (int32_t)kFltMaxBits;
    cna[t] = 0.f;
  }
  uint32_t m_idx = 0, m_start = 0, m_pos = 0;  // lane s: slot s
  float m_ax = 0.f, m_ay = 0.f, m_az = 0.f;  // lane s: slot s, original
  unsigned long long live = 0;
  float r_x = 0.f, r_y = 0.f, r_z = 0.f;
  uint32_t r_idx = 0, r_pos = 0;
  uint32_t rhead = 0, rcount = 0;
  bool drained = false;
  uint32_t gstep = 0;
  unsigned long long evals = 0;
  const uint32_t low_j = (1u << j) - 1u;
  uint32_t stg_idx = 0, n_stg = 0;
  float stg_min = 0.f;
  auto flush = [&]() __attribute__((always_inline)) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(args.esc_count, n_stg);
    base = Uniform(base);
    if (lane < n_stg) {
      args.esc_idx[base + lane] = stg_idx;
      args.esc_min[base + lane] = stg_min;
    }
    n_stg = 0;
  };

  // Nothing raises cmax2 while a filtered round runs (its points never
  // finish a scan), so the threshold is fixed for the launch.
  const uint32_t cm = __float_as_uint(LoadCmax2(args.key));
  const float cmax2 = __uint_as_float(cm);
  const float thr = FilterThreshold(cm, args.err);
  const float thr_hi = FilterUpper(cm, args.err);
  // Wave k of the sweep takes segment k % n_segs (one segment, all of B, in
  // a budgeted round); with fewer waves than segments a wave takes
  // segments wave_id, wave_id + n_waves, ...
  const uint32_t sweep = n_waves > args.n_segs ? n_waves : args.n_segs;
  for (uint32_t kk = wave_id; kk < sweep; kk += n_waves) {
    const uint32_t seg = kk % args.n_segs;
    // Segment seg: steps seg, seg + n_segs, ... (one segment: all of B);
    // a grouped round's segments are contiguous ranges of its budget.
    const uint32_t seg_stride = grouped ? 1u : args.n_segs;
    const uint32_t seg_begin = grouped ? seg * args.seg_steps : seg;
    const uint32_t seg_len =
        grouped ? min(args.seg_steps, n_steps - seg_begin)
                : (n_steps - seg + args.n_segs - 1) / args.n_segs;
    const uint32_t seg_end = grouped ? seg_begin + seg_len : n_steps;
    // Budgeted rounds escalate after budget_steps (< n_steps); segments are
    // complete after seg_len steps.
    const uint32_t limit = split ? seg_len : args.budget_steps;
    unsigned int* cursor = args.cursors + seg;
    // The cursor counts q in [.., q_end); position = q_off + q_stride q
    // (grouped: the group's positions g, g + G, ...; else q itself).
    uint32_t q_off = 0, q_stride = 1, q_end = args.pos_end;
    if (grouped) {
      q_off = args.pos_begin + seg % args.n_groups;
      q_stride = args.n_groups;
      q_end = args.pos_end > q_off
                  ? (args.pos_end - q_off + q_stride - 1) / q_stride
                  : 0u;
    }
    // Every wave of a segment starts at its first step: the waves then walk
    // B in roughly the same window, so the L2 of each XCD serves most of
    // the tile loads (10M GMM: L2 hit rate 0.12 with scattered starts,
    // 0.55 aligned; 2-5% less time per launch).
    uint32_t step = split ? seg_begin : seg_begin + args.start_step % seg_len;
    live = 0;
    rhead = rcount = 0;
    drained = false;
    gstep = 0;
    // An iteration covers kFilterSteps B steps (4 kFilterSteps tiles, 16
    // kFilterSteps MFMAs); the next iteration's tiles are requested at the
    // top of each iteration.
    auto next_step = [&](uint32_t st) __attribute__((always_inline)) {
      return st + seg_stride >= seg_end ? seg_begin : st + seg_stride;
    };
    float bcur[4 * kFilterSteps], bnext[4 * kFilterSteps];
    uint32_t nxt = step;  // first step of the next iteration
#pragma unroll
    for (int k = 0; k < kFilterSteps; ++k) {
      LoadStepTiles(args.bt, nxt, lane4, bcur + 4 * k);
      nxt = next_step(nxt);
    }

    // Two iterations per pass with the tile buffers' roles swapped: the
    // prefetched tiles become the current ones with no copy or reload
    // (10M GMM: learned 10.08 -> 9.99 ms, plain 12.94 -> 12.79 ms).
    auto iteration = [&](float (&bcur)[4 * kFilterSteps],
                         float (&bnx)[4 * kFilterSteps])
        __attribute__((always_inline)) -> bool {
      // Refill free slots from the compacted reservoir.
      unsigned long long free_m = ~live;
      while (free_m != 0ull) {
        if (rhead == rcount) {
          if (drained) break;
          uint32_t base = 0;
          if (lane == 0) base = atomicAdd(cursor, args.claim);
          base = Uniform(base);
          if (base >= q_end) {
            drained = true;
            break;
          }
          const uint32_t q = base + lane;
          bool valid = lane < args.claim && q < q_end;
          const uint32_t pos = valid ? q_off + q_stride * q : 0u;
          if (valid && args.list_min != nullptr)
            valid = !(__float_as_uint(args.list_min[pos]) < cm);
          if (valid && split)
            valid = __hip_atomic_load(&args.broken[pos], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) == 0u;
          uint32_t idx = 0;
          float px = 0.f, py = 0.f, pz = 0.f;
          if (valid) {
            idx = OrderAt(args.order_mode, args.order, args.perm_a, pos);
            const float* src = args.a + (uint64_t)idx * ND;
            px = src[0];
            py = src[1];
            if constexpr (ND == 3) pz = src[2];
          }
          const unsigned long long vm = __ballot(valid);
          const uint32_t cnt = (uint32_t)__popcll(vm);
          const uint32_t below_v = __builtin_amdgcn_mbcnt_hi(
              (uint32_t)(vm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)vm, 0u));
          const uint32_t dest = valid ? below_v : cnt + (lane - below_v);
          const int d4 = (int)(dest << 2);
          r_x = __int_as_float(__builtin_amdgcn_ds_permute(d4, __float_as_int(px)));
          r_y = __int_as_float(__builtin_amdgcn_ds_permute(d4, __float_as_int(py)));
          if constexpr (ND == 3)
            r_z = __int_as_float(__builtin_amdgcn_ds_permute(d4, __float_as_int(pz)));
          r_idx = (uint32_t)__builtin_amdgcn_ds_permute(d4, (int)idx);
          r_pos = (uint32_t)__builtin_amdgcn_ds_permute(d4, (int)pos);
          rhead = 0;
          rcount = cnt;
          continue;
        }
        const uint32_t navail = rcount - rhead;
        // Slot s (this lane as metadata holder) takes reservoir entry
        // rhead + (number of free slots below s) when that is available.
        const uint32_t rank_s = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(free_m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)free_m, 0u));
        const bool take_s = ((free_m >> lane) & 1ull) && rank_s < navail;
        const uint32_t src_s = (rhead + rank_s) & 63u;
        const int src_s4 = (int)(src_s << 2);
        const uint32_t new_idx = (uint32_t)__builtin_amdgcn_ds_bpermute(src_s4, (int)r_idx);
        const uint32_t new_pos = (uint32_t)__builtin_amdgcn_ds_bpermute(src_s4, (int)r_pos);
        const float new_x = __int_as_float(__builtin_amdgcn_ds_bpermute(src_s4, __float_as_int(r_x)));
        const float new_y = __int_as_float(__builtin_amdgcn_ds_bpermute(src_s4, __float_as_int(r_y)));
        float new_z = 0.f;
        if constexpr (ND == 3)
          new_z = __int_as_float(__builtin_amdgcn_ds_bpermute(src_s4, __float_as_int(r_z)));
        if (take_s) {
          m_idx = new_idx;
          m_pos = new_pos;
          m_start = gstep;
          m_ax = new_x;
          m_ay = new_y;
          m_az = new_z;
        }
        // Operand lanes of each tile.
  #pragma unroll
        for (int t = 0; t < 4; ++t) {
          const uint32_t ft = (uint32_t)(free_m >> (16 * t)) & 0xFFFFu;
          if (ft == 0u) continue;
          const uint32_t below_t =
              t == 0 ? 0u : (uint32_t)__popcll(free_m & ((1ull << (16 * t)) - 1ull));
          const uint32_t rank = below_t + (uint32_t)__builtin_popcount(ft & low_j);
          const bool take = ((ft >> j) & 1u) && rank < navail;
          const int src4 = (int)(((rhead + rank) & 63u) << 2);
          const float px = __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(r_x)));
          const float py = __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(r_y)));
          float pz = 0.f;
          if constexpr (ND == 3)
            pz = __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(r_z)));
          float val, cn;
          FilterOperand<ND>(px, py, pz, args.cx, args.cy, args.cz, args.off, kc, &val, &cn);
          if (take) {
            bop[t] = val;
            mn[t] = (int32_t)kFltMaxBits;
            cna[t] = cn;
          }
        }
        const uint32_t took = (uint32_t)__popcll(free_m) < navail ? (uint32_t)__popcll(free_m) : navail;
        const unsigned long long took_m = __ballot(take_s);
        live |= took_m;
        free_m &= ~took_m;
        rhead += took;
      }
      if (live == 0ull) return false;

      // Segmented scans poll the other segments' verdicts every iteration
      // (the tail and top completions: a segment stops at most 4 steps after
      // another found a hit; 10M GMM learned order: 2.4% fewer evaluations),
      // the grouped round every 8 steps. The flag load is issued before the
      // next steps' loads: loads complete in order, so waiting for the flag
      // does not wait for them.
      const bool poll =
          split && (!kGrouped || ((gstep + kFilterSteps) & 7u) == 0u);
      uint32_t flag = 0;
      if (poll && ((live >> lane) & 1ull))
        flag = __hip_atomic_load(&args.broken[m_pos], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
      // The next iteration's tiles, then kFilterSteps steps: 4 kFilterSteps
      // B tiles against the 4 A tiles (16 kFilterSteps MFMAs).
      const uint32_t cur_step = step;
      step = nxt;
  #pragma unroll
      for (int k = 0; k < kFilterSteps; ++k) {
        LoadStepTiles(args.bt, nxt, lane4, bnx + 4 * k);
        nxt = next_step(nxt);
      }
      int32_t sm[4];  // this iteration's approximate minima
      const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  #pragma unroll
      for (int t = 0; t < 4; ++t) {
        sm[t] = (int32_t)kFltMaxBits;
        // A tile without live slots (the cursor has drained) costs nothing.
        if (((live >> (16 * t)) & 0xFFFFull) == 0ull) continue;
  #pragma unroll
        for (int q = 0; q < 4 * kFilterSteps; ++q) {
          const floatx4 d = __builtin_amdgcn_mfma_f32_16x16x4f32(bcur[q], bop[t], zero, 0, 0, 0);
          sm[t] = Min3I(sm[t], __float_as_int(d[0]), __float_as_int(d[1]));
          sm[t] = Min3I(sm[t], __float_as_int(d[2]), __float_as_int(d[3]));
        }
        mn[t] = min(mn[t], sm[t]);
      }
      gstep += kFilterSteps;
      evals += (unsigned long long)__popcll(live) * (64ull * kFilterSteps);

      // Drops (proved below cmax2, or verified exactly) and escalations.
      unsigned long long hit = 0, maybe = 0;
  #pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float v = __int_as_float(sm[t]);
        const float lim = ND == 3 ? thr + cna[t] : thr;
        const float lim_hi = ND == 3 ? thr_hi + cna[t] : thr_hi;
        unsigned long long c = __ballot(v < lim);
        unsigned long long m = __ballot(v < lim_hi);
        c |= c >> 32;
        c |= c >> 16;
        m |= m >> 32;
        m |= m >> 16;
        hit |= (c & 0xFFFFull) << (16 * t);
        maybe |= (m & 0xFFFFull) << (16 * t);
      }
      hit &= live;
      maybe &= live & ~hit;
      if (maybe != 0ull) {
        // Exact test of this iteration's 64 kFilterSteps original B points
        // (the reference's formula, strict '<'), loaded only when a test is
        // due.
        float ob[kFilterSteps][ND];
        uint32_t st = cur_step;
  #pragma unroll
        for (int k = 0; k < kFilterSteps; ++k) {
          LoadStep<ND>(args.bx, args.by, args.bz, st, lane, ob[k]);
          st = next_step(st);
        }
        while (maybe != 0ull) {
          const uint32_t sl = (uint32_t)__builtin_ctzll(maybe);
          maybe &= maybe - 1ull;
          const float ax = ReadLaneF(m_ax, sl), ay = ReadLaneF(m_ay, sl);
          bool below = false;
          if constexpr (ND == 2) {
  #pragma unroll
            for (int k = 0; k < kFilterSteps; ++k)
              below = below || D2(ax, ay, ob[k][0], ob[k][1]) < cmax2;
          } else {
            const float az = ReadLaneF(m_az, sl);
  #pragma unroll
            for (int k = 0; k < kFilterSteps; ++k)
              below = below || D2(ax, ay, az, ob[k][0], ob[k][1], ob[k][2]) < cmax2;
          }
          if (WaveAny(below)) hit |= 1ull << sl;
        }
      }
      if (split && hit != 0ull) {
        // Tell the other segments' waves this entry is decided.
        if ((hit >> lane) & 1ull)
          __hip_atomic_store(&args.broken[m_pos], 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      }
      const unsigned long long gone =
          poll ? __ballot(flag != 0u) & live & ~hit : 0ull;
      unsigned long long due =
          __ballot(gstep - m_start >= limit) & live & ~(hit | gone);
      if (split && !kGrouped && due != 0ull) {
        // Segment scanned without a hit: count it; an entry that completes
        // every segment goes to the exact scan.
        if ((due >> lane) & 1ull) {
          const uint32_t prev = atomicAdd(&args.gdone[m_pos], 1u);
          if (prev + 1u == args.n_segs) args.cand[atomicAdd(args.cand_count, 1u)] = m_idx;
        }
      } else if (due != 0ull) {
        // Partial minimum of slot s = 16 t + j: lanes j, j+16, j+32, j+48 of
        // mn[t]; lane s picks its own tile's value.
        double mine = 0.0;
  #pragma unroll
        for (int t = 0; t < 4; ++t) {
          int32_t r = min(mn[t], __shfl_xor(mn[t], 16, 64));
          r = min(r, __shfl_xor(r, 32, 64));
          double v = (double)__int_as_float(r);
          if (ND == 3) v -= (double)cna[t];
          if (kc == (uint32_t)t) mine = v;
        }
        // approx + E bounds the real |a - b|^2 of the partial minimum; the
        // reference's FP32 d2 can exceed that real value by about 4u d2, so
        // the bound is widened by 2^-19 relative (32u) before a later
        // "partial minimum below cmax2" test relies on it.
        const double ub =
            ((mine > 0.0 ? mine : 0.0) + (double)args.err) * (1.0 + 0x1p-19);
        float esc_val = __double2float_ru(ub);
        unsigned long long esc_m = due;
        if constexpr (kGrouped) {
          // Fold this segment's bound into the entry's, then count the
          // segment. The count's increment depends on the value the atomic
          // min returned, so it is issued after the min has been performed
          // at the device coherence point, and the count that completes the
          // entry reads the minimum over all its segments there.
          bool last = false;
          if ((due >> lane) & 1ull) {
            const uint32_t was = atomicMin(&args.pmin[m_pos], __float_as_uint(esc_val));
            const uint32_t prev = __hip_atomic_fetch_add(
                &args.gdone[m_pos], 1u + (was >> 31), __ATOMIC_RELAXED,
                __HIP_MEMORY_SCOPE_AGENT);
            if (prev + 1u == per_entry) {
              last = true;
              esc_val = __uint_as_float(__hip_atomic_load(
                  &args.pmin[m_pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            }
          }
          esc_m = __ballot(last);
        }
        // Stage the escalations in lanes [n_stg, n_stg + n_esc) (a forward
        // permutation: escalating slots in order, the other lanes after
        // them) and append 64 at a time, so the shared counter sees one
        // atomic per 64.
        const uint32_t n_esc = (uint32_t)__popcll(esc_m);
        if (n_stg + n_esc > (uint32_t)kWaveSize) flush();
        const bool esc = (esc_m >> lane) & 1ull;
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(esc_m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)esc_m, 0u));
        const uint32_t dest =
            (esc ? n_stg + below : n_stg + n_esc + (lane - below)) & 63u;
        const int d4 = (int)(dest << 2);
        const uint32_t p_idx = (uint32_t)__builtin_amdgcn_ds_permute(d4, (int)m_idx);
        const float p_min = __int_as_float(__builtin_amdgcn_ds_permute(
            d4, __float_as_int(esc_val)));
        if (lane >= n_stg && lane < n_stg + n_esc) {
          stg_idx = p_idx;
          stg_min = p_min;
        }
        n_stg += n_esc;
        if (n_stg == (uint32_t)kWaveSize) flush();
      }
      live &= ~(hit | gone | due);
      return true;
    };
    while (iteration(bcur, bnext) && iteration(bnext, bcur)) {
    }
  }
  if (n_stg > 0) flush();
  if (lane == 0) CounterAdd(args.evals, evals);
}

// The step at which every wave of a whole-B launch starts: fixed by the
// launch's first position and budget, so that the rounds of one run start
// at unrelated steps (an escalated point then does not re-scan the steps
// its previous round covered).
inline uint32_t RoundStartStep(uint32_t pos_begin, uint32_t budget,
                               uint32_t n_steps) {
  const uint64_t h = (uint64_t)pos_begin * 0x9E3779B97F4A7C15ull +
                     (uint64_t)budget * 0xC2B2AE3D27D4EB4Full;
  return n_steps ? (uint32_t)((h >> 32) % n_steps) : 0u;
}

constexpr int kP = 8;  // slots per wave
constexpr int kS = 4;  // B steps per iteration

// Blocks of the scan kernel that are resident at once (all CUs at the
// kernel's occupancy), so that no block of a persistent launch waits for
// another to exit.
template <int ND>
int ResidentBlocks(hd_context* ctx) {
  static int per_cu[4] = {0, 0, 0, 0};
  if (per_cu[ND] == 0) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, EbScanKernel<ND, kP, kS>, kBlock, 0) != hipSuccess || n < 1)
      n = 1;
    per_cu[ND] = n;
  }
  return per_cu[ND] * ctx->cu_count;
}

template <int ND>
int LaunchScan(hd_context* ctx, ScanArgs& args, uint32_t pos_begin,
               uint32_t n_entries) {
  // Cursors from the context unless the caller placed them (the
  // allocation-free path).
  if (args.cursors == nullptr)
    HD_TRY(ContextArray(ctx, "eb_cursors", args.n_segs, &args.cursors));
  HD_TRY(FillU32(ctx, args.cursors, pos_begin, args.n_segs));
  const unsigned long long items =
      ((unsigned long long)n_entries + kP - 1) / kP * args.n_segs;
  int blocks = ResidentBlocks<ND>(ctx);
  if ((unsigned long long)blocks * 4 > items)
    blocks = (int)std::max<unsigned long long>(1, (items + 3) / 4);
  // Small lists: claim fewer positions at a time so that every wave of a
  // segment gets a share.
  const uint32_t n_waves = (uint32_t)blocks * (kBlock / kWaveSize);
  const uint32_t per_seg = std::max(1u, n_waves / args.n_segs);
  args.claim = std::min<uint32_t>(
      kWaveSize, std::max<uint32_t>(1u, (n_entries + per_seg - 1) / per_seg));
  StageDispatch(ctx);
  EbScanKernel<ND, kP, kS><<<blocks, kBlock, 0, ctx->stream>>>(args);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

// Exact minima of the entries of `list` with B split into segments scanned
// by different waves: as many as keep every slot busy, at least min_segs,
// at most 512 steps each, at least 64 steps each. Long segments leave the
// launch to a few slow waves, short ones add the steps a segment spends
// after another's hit: the 2048-point seed of the learned order on the 10M
// GMM takes 1.35 ms with 128-step segments, 1.08 ms with 256, 1.06 ms
// with 512 (1.7 ms with 1953-step contiguous ranges).
template <int ND>
int LaunchSplit(hd_context* ctx, const float* a, const uint32_t* list,
                const float* list_min, uint32_t n_list, const float* bx,
                const float* by, const float* bz, uint32_t n_steps,
                unsigned long long* key, unsigned long long* evals,
                uint32_t min_segs = 1) {
  if (n_list == 0) return HD_OK;
  const uint32_t total_slots =
      (uint32_t)ResidentBlocks<ND>(ctx) * (kBlock / kWaveSize) * kP;
  const uint32_t want = (4u * total_slots + n_list - 1) / n_list;
  uint32_t n_segs = std::max(std::max(min_segs, want), n_steps / 512u);
  n_segs = std::min(n_segs, std::max(1u, n_steps / 64u));
  const uint32_t seg_steps = (n_steps + n_segs - 1) / n_segs;
  n_segs = (n_steps + seg_steps - 1) / seg_steps;
  ScanArgs args = {};
  args.a = a;
  args.order = list;
  args.list_min = list_min;
  args.perm_a = MakeFeistel(1, 0, 0);
  args.order_mode = kOrderArray;
  args.pos_end = n_list;
  args.bx = bx;
  args.by = by;
  args.bz = bz;
  args.n_steps = n_steps;
  args.seg_steps = seg_steps;
  args.n_segs = n_segs;
  args.budget_steps = 0xFFFFFFFFu;
  args.key = key;
  args.evals = evals;
  if (n_segs > 1) {
    HD_TRY(ContextArray(ctx, "eb_broken", n_list, &args.broken));
    HD_TRY(ContextArray(ctx, "eb_pmin", n_list, &args.pmin));
    HD_TRY(ContextArray(ctx, "eb_gdone", n_list, &args.gdone));
    HD_TRY(FillU32(ctx, args.broken, 0u, n_list));
    HD_TRY(FillU32(ctx, args.pmin, kFltMaxBits, n_list));
    HD_TRY(FillU32(ctx, args.gdone, 0u, n_list));
  }
  return LaunchScan<ND>(ctx, args, 0u, n_list);
}

// Whole-B scans over order positions [pos_begin, pos_end) with a step
// budget; escalations are appended to (esc_idx, esc_min) and counted in
// counters[1].
template <int ND>
int LaunchSlots(hd_context* ctx, const float* a, int order_mode,
                const uint32_t* order, const float* list_min,
                const Feistel& perm_a, uint32_t pos_begin, uint32_t pos_end,
                const float* bx, const float* by, const float* bz,
                uint32_t n_steps, uint32_t budget_steps,
                unsigned long long* key, unsigned long long* evals,
                unsigned int* counters, uint32_t* esc_idx, float* esc_min,
                bool reset_esc = true) {
  if (pos_begin >= pos_end) return HD_OK;
  if (reset_esc) HD_TRY(FillU32(ctx, counters + 1, 0u, 1));
  ScanArgs args = {};
  args.a = a;
  args.order = order;
  args.list_min = list_min;
  args.perm_a = perm_a;
  args.order_mode = order_mode;
  args.pos_end = pos_end;
  args.bx = bx;
  args.by = by;
  args.bz = bz;
  args.n_steps = n_steps;
  args.seg_steps = n_steps;
  args.n_segs = 1;
  args.budget_steps = budget_steps;
  args.start_step = RoundStartStep(pos_begin, budget_steps, n_steps);
  args.key = key;
  args.esc_idx = esc_idx;
  args.esc_min = esc_min;
  args.esc_count = counters + 1;
  args.evals = evals;
  return LaunchScan<ND>(ctx, args, pos_begin, pos_end - pos_begin);
}

// Short exact rounds (budget <= kDenseSteps steps) run without slots: every
// point of the round is compared with one window of B, the budget's steps
// from the launch's start step (where the slot kernel's waves start; a
// point the slot kernel refills mid-way scans the steps after its refill
// instead). The round is then a dense product of its points against the
// window. Each wave holds the window in registers (lane l: point l of every
// window step) and takes the round's positions 8 at a time as wave-uniform
// values: 6 vector operations per evaluation and one cross-lane minimum per
// point. A point is dropped (its window minimum is below cmax2) or
// escalated with that exact partial minimum; escalations are staged in
// lanes and appended 64 at a time. Drops and escalated minima are exact for
// any window; the evaluations after a point's first hit are made too, and
// counted.
constexpr uint32_t kDenseSteps = 16u;
constexpr int kDenseGroup = 8;  // positions per pass

struct DenseArgs {
  const float* a;
  const uint32_t* order;
  const float* list_min;
  Feistel perm_a;
  int order_mode;
  uint32_t pos_begin, pos_end;
  const float* bx;
  const float* by;
  const float* bz;
  uint32_t n_steps;
  uint32_t start_step;
  uint32_t window_steps;
  const unsigned long long* key;
  uint32_t* esc_idx;
  float* esc_min;
  unsigned int* esc_count;
  unsigned long long* evals;
};

template <int ND>
__global__ __launch_bounds__(256) void EbDenseRoundKernel(DenseArgs args) {
  const uint32_t lane = LaneId();
  const uint32_t wave = Uniform((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const uint32_t n_waves = gridDim.x * (blockDim.x >> 6);
  const uint32_t ws = args.window_steps;
  float wb[kDenseSteps][ND];
#pragma unroll
  for (uint32_t k = 0; k < kDenseSteps; ++k) {
    uint32_t st = args.start_step + (k < ws ? k : 0u);
    st = st >= args.n_steps ? st - args.n_steps : st;
    LoadStep<ND>(args.bx, args.by, args.bz, st, lane, wb[k]);
  }
  const uint32_t cm = __float_as_uint(LoadCmax2(args.key));
  unsigned long long evals = 0;
  uint32_t stg_idx = 0, n_stg = 0;
  float stg_min = 0.f;
  auto flush = [&]() __attribute__((always_inline)) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(args.esc_count, n_stg);
    base = Uniform(base);
    if (lane < n_stg) {
      args.esc_idx[base + lane] = stg_idx;
      args.esc_min[base + lane] = stg_min;
    }
    n_stg = 0;
  };
  const uint32_t n_pos = args.pos_end - args.pos_begin;
  const uint32_t n_groups = (n_pos + kDenseGroup - 1) / kDenseGroup;
  for (uint32_t g = wave; g < n_groups; g += n_waves) {
    const uint32_t p0 = args.pos_begin + g * kDenseGroup;
    // Lanes 0..7 fetch the group's points; the rest of the wave reads them
    // as uniform values.
    const uint32_t pos = p0 + (lane & (kDenseGroup - 1));
    bool valid = lane < (uint32_t)kDenseGroup && pos < args.pos_end;
    if (valid && args.list_min != nullptr)
      valid = !(__float_as_uint(args.list_min[pos]) < cm);
    uint32_t idx = 0;
    float px = 0.f, py = 0.f, pz = 0.f;
    if (valid) {
      idx = OrderAt(args.order_mode, args.order, args.perm_a, pos);
      const float* src = args.a + (uint64_t)idx * ND;
      px = src[0];
      py = src[1];
      if constexpr (ND == 3) pz = src[2];
    }
    const uint32_t vm = (uint32_t)__ballot(valid);
    uint32_t mine = kFltMaxBits;  // lane p < 8: point p's window minimum
#pragma unroll
    for (int q = 0; q < kDenseGroup; ++q) {
      if (!((vm >> q) & 1u)) continue;
      const float ax = ReadLaneF(px, (uint32_t)q), ay = ReadLaneF(py, (uint32_t)q);
      const float az = ND == 3 ? ReadLaneF(pz, (uint32_t)q) : 0.f;
      uint32_t m = kFltMaxBits;
#pragma unroll
      for (uint32_t k = 0; k < kDenseSteps; ++k) {
        if (k >= ws) break;
        float d;
        if constexpr (ND == 2)
          d = D2(ax, ay, wb[k][0], wb[k][1]);
        else
          d = D2(ax, ay, az, wb[k][0], wb[k][1], wb[k][2]);
        m = MinBits(m, __float_as_uint(d));
      }
      m = WaveMinU32(m);
      if (lane == (uint32_t)q) mine = m;
    }
    evals += (unsigned long long)__popc(vm) * (64ull * ws);
    // Escalations into lanes [n_stg, n_stg + n_esc), appended 64 at a time.
    const bool esc = valid && !(mine < cm);
    const unsigned long long em = __ballot(esc);
    const uint32_t n_esc = (uint32_t)__popcll(em);
    if (n_esc == 0u) continue;
    if (n_stg + n_esc > 64u) flush();
    // A forward permutation: escalating lanes to [n_stg, n_stg + n_esc) in
    // order, the other lanes after them.
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(em >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)em, 0u));
    const uint32_t dest = esc ? n_stg + below : n_stg + n_esc + (lane - below);
    const int d4 = (int)((dest & 63u) << 2);
    const uint32_t p_idx = (uint32_t)__builtin_amdgcn_ds_permute(d4, esc ? (int)idx : 0);
    const uint32_t p_min = (uint32_t)__builtin_amdgcn_ds_permute(d4, esc ? (int)mine : 0);
    if (lane >= n_stg && lane < n_stg + n_esc) {
      stg_idx = p_idx;
      stg_min = __uint_as_float(p_min);
    }
    n_stg += n_esc;
    if (n_stg == 64u) flush();
  }
  if (n_stg > 0u) flush();
  if (lane == 0) CounterAdd(args.evals, evals);
}

template <int ND>
int LaunchDenseRound(hd_context* ctx, const float* a, int order_mode,
                     const uint32_t* order, const float* list_min,
                     const Feistel& perm_a, uint32_t pos_begin,
                     uint32_t pos_end, const float* bx, const float* by,
                     const float* bz, uint32_t n_steps, uint32_t budget_steps,
                     unsigned long long* key, unsigned long long* evals,
                     unsigned int* counters, uint32_t* esc_idx, float* esc_min,
                     bool reset_esc) {
  if (pos_begin >= pos_end) return HD_OK;
  if (budget_steps == 0 || budget_steps > kDenseSteps || budget_steps >= n_steps)
    return Fail(HD_EINVAL, "dense round: budget %u steps", budget_steps);
  if (reset_esc) HD_TRY(FillU32(ctx, counters + 1, 0u, 1));
  DenseArgs args = {};
  args.a = a;
  args.order = order;
  args.list_min = list_min;
  args.perm_a = perm_a;
  args.order_mode = order_mode;
  args.pos_begin = pos_begin;
  args.pos_end = pos_end;
  args.bx = bx;
  args.by = by;
  args.bz = bz;
  args.n_steps = n_steps;
  // The window the slot kernel's waves would scan: from the launch's start
  // step, budget_steps steps, wrapping at the end of B.
  args.start_step = RoundStartStep(pos_begin, budget_steps, n_steps);
  args.window_steps = budget_steps;
  args.key = key;
  args.esc_idx = esc_idx;
  args.esc_min = esc_min;
  args.esc_count = counters + 1;
  args.evals = evals;
  const uint64_t groups = (pos_end - pos_begin + kDenseGroup - 1) / kDenseGroup;
  // Enough waves for every SIMD, each with at least 16 groups.
  const int blocks = (int)std::max<uint64_t>(
      1, std::min<uint64_t>((groups + 63u) / 64u, (uint64_t)ctx->cu_count * 8u));
  StageDispatch(ctx);
  EbDenseRoundKernel<ND><<<blocks, 256, 0, ctx->stream>>>(args);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

// Centred B tiles for the matrix-core filter: tile q holds points
// 16q..16q+15 of the shuffled SoA copy as [x' x16, y' x16, |b'|^2 x16, 1 x16]
// (2D) or [x', y', z', |b'|^2 + C] (3D).
template <int ND>
__global__ void BuildTilesKernel(const float* __restrict__ bx,
                                 const float* __restrict__ by,
                                 const float* __restrict__ bz, uint32_t n_bp,
                                 float cx, float cy, float cz, float off,
                                 float* bt) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n_bp;
       p += gridDim.x * blockDim.x) {
    const float x = bx[p] - cx, y = by[p] - cy;
    const uint64_t o = (uint64_t)(p >> 4) * 64u + (p & 15u);
    bt[o] = x;
    bt[o + 16] = y;
    if constexpr (ND == 2) {
      bt[o + 32] = x * x + y * y;
      bt[o + 48] = 1.f;
    } else {
      const float z = bz[p] - cz;
      bt[o + 32] = z;
      bt[o + 48] = (x * x + y * y + z * z) + off;
    }
  }
}

// Centre, error bound and tiles of the filter (see EbMfmaKernel). Leaves
// t->bt null when the coordinates are too large for FP32 norms.
template <int ND>
int PrepareFilter(hd_context* ctx, const float* mbr_a, const float* mbr_b,
                  EbTarget* t) {
  double rho2 = 0.0;
  for (int d = 0; d < ND; ++d) {
    const double lo = std::min(mbr_a[d], mbr_b[d]);
    const double hi = std::max(mbr_a[ND + d], mbr_b[ND + d]);
    const float c = (float)((lo + hi) * 0.5);
    const double r = std::max(std::fabs(lo - (double)c), std::fabs(hi - (double)c));
    t->center[d] = c;
    rho2 += r * r;
  }
  rho2 *= 1.0 + 0x1p-20;
  if (!(rho2 < 1e30)) {
    t->bt = nullptr;
    return HD_OK;
  }
  t->err = (float)((ND == 2 ? 48.0 : 64.0) * 0x1p-24 * rho2 * (1.0 + 0x1p-20) +
                   0x1p-100);
  t->off = ND == 3 ? (float)(rho2 * (1.0 + 0x1p-20)) : 0.f;
  HD_TRY(ContextArray(ctx, "eb_tiles", (uint64_t)t->n_bp * 4u, &t->bt));
  const int grid = (int)std::min<uint64_t>((t->n_bp + 255u) / 256u,
                                           (uint64_t)ctx->cu_count * 8u);
  BuildTilesKernel<ND><<<grid, 256, 0, ctx->stream>>>(
      t->bx, t->by, t->bz, t->n_bp, t->center[0], t->center[1], t->center[2],
      t->off, t->bt);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

template <int ND, bool kGrouped = false>
int MfmaResidentBlocks(hd_context* ctx) {
  static int per_cu[4] = {0, 0, 0, 0};
  if (per_cu[ND] == 0) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, EbMfmaKernel<ND, kGrouped>, kBlock, 0) != hipSuccess || n < 1)
      n = 1;
    per_cu[ND] = n;
  }
  return per_cu[ND] * ctx->cu_count;
}

// Filtered budgeted round over order positions [pos_begin, pos_end); every
// point that is not proved below cmax2 within budget_steps is escalated
// with an upper bound of its partial minimum.
// Segment length of grouped budgeted rounds.
constexpr uint32_t kGroupSteps = 256u;

template <int ND>
int LaunchMfma(hd_context* ctx, const float* a, int order_mode,
               const uint32_t* order, const float* list_min,
               const Feistel& perm_a, uint32_t pos_begin, uint32_t pos_end,
               const EbTarget& t, uint32_t budget_steps,
               unsigned long long* key, unsigned long long* evals,
               unsigned int* counters, uint32_t* esc_idx, float* esc_min,
               bool reset_esc = true) {
  if (pos_begin >= pos_end) return HD_OK;
  HD_TRY(FillU32(ctx, counters, pos_begin, 1));
  if (reset_esc) HD_TRY(FillU32(ctx, counters + 1, 0u, 1));
  MfmaArgs args = {};
  args.a = a;
  args.order = order;
  args.list_min = list_min;
  args.perm_a = perm_a;
  args.order_mode = order_mode;
  args.pos_end = pos_end;
  args.cursors = counters;
  args.seg_steps = t.n_steps;
  args.n_segs = 1;
  args.bt = t.bt;
  args.bx = t.bx;
  args.by = t.by;
  args.bz = t.bz;
  args.n_steps = t.n_steps;
  args.budget_steps = budget_steps;
  args.start_step = RoundStartStep(pos_begin, budget_steps, t.n_steps);
  args.cx = t.center[0];
  args.cy = t.center[1];
  args.cz = t.center[2];
  args.err = t.err;
  args.off = t.off;
  args.key = key;
  args.esc_idx = esc_idx;
  args.esc_min = esc_min;
  args.esc_count = counters + 1;
  args.evals = evals;
  const uint32_t n_entries = pos_end - pos_begin;
  // Too few entries for 8 per slot of a full launch: a budget of W >= 2
  // segments of kGroupSteps steps runs grouped (see EbMfmaKernel), W items
  // per entry. (With enough entries, splitting would only add the steps
  // that segments spend before they see another's hit.)
  const uint32_t w = budget_steps / kGroupSteps;
  const uint64_t full_items =
      (uint64_t)MfmaResidentBlocks<ND, false>(ctx) * (kBlock / kWaveSize) * 512u;
  const bool grouped = w >= 2 && budget_steps % kGroupSteps == 0 &&
                       t.n_steps / budget_steps >= 1 && n_entries < full_items;
  uint32_t n_groups = 1, items = n_entries;
  if (grouped) {
    n_groups = t.n_steps / budget_steps;
    args.n_groups = n_groups;
    args.pos_begin = pos_begin;
    args.seg_steps = kGroupSteps;
    args.n_segs = n_groups * w;
    items = n_entries * w;
    HD_TRY(ContextArray(ctx, "eb_mfma_cursors", args.n_segs, &args.cursors));
    HD_TRY(FillU32(ctx, args.cursors, 0u, args.n_segs));
    HD_TRY(ContextArray(ctx, "eb_broken", pos_end, &args.broken));
    HD_TRY(ContextArray(ctx, "eb_gdone", pos_end, &args.gdone));
    HD_TRY(ContextArray(ctx, "eb_pmin", pos_end, &args.pmin));
    HD_TRY(FillU32(ctx, args.broken + pos_begin, 0u, n_entries));
    HD_TRY(FillU32(ctx, args.gdone + pos_begin, 0u, n_entries));
    HD_TRY(FillU32(ctx, args.pmin + pos_begin, kFltMaxBits, n_entries));
  }
  int blocks = grouped ? MfmaResidentBlocks<ND, true>(ctx)
                       : MfmaResidentBlocks<ND, false>(ctx);
  // At least 8 items per slot: with fewer, most of the round is the drain,
  // where tiles with a few live slots still cost all their MFMAs.
  const uint32_t waves_needed = (uint32_t)std::min<uint64_t>(
      0xFFFFFFFFull, ((uint64_t)items + 511u) / 512u);
  if ((uint32_t)blocks * 4u > waves_needed)
    blocks = (int)std::max(1u, (waves_needed + 3u) / 4u);
  const uint32_t n_waves = (uint32_t)blocks * (kBlock / kWaveSize);
  if (grouped) {
    // Per segment: the group's positions over the waves that share it.
    const uint32_t per_seg = std::max(1u, n_waves / args.n_segs);
    const uint32_t per_group = (n_entries + n_groups - 1) / n_groups;
    args.claim = std::min<uint32_t>(
        kWaveSize, std::max<uint32_t>(16u, (per_group + per_seg - 1) / per_seg));
  } else {
    args.claim = std::min<uint32_t>(
        kWaveSize, std::max<uint32_t>(16u, (n_entries + n_waves - 1) / n_waves));
  }
  StageDispatch(ctx);
  if (grouped)
    EbMfmaKernel<ND, true><<<blocks, kBlock, 0, ctx->stream>>>(args);
  else
    EbMfmaKernel<ND, false><<<blocks, kBlock, 0, ctx->stream>>>(args);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

// Reads a device counter and, optionally, the current cmax2 (one sync).
// Segmented filter scan of `list` (the tail): B split into segments that
// different waves scan, each (entry, segment) a slot's item; an entry is
// dropped as soon as any segment finds a hit, and the entries that complete
// every segment without one (exact minimum >= cmax2: candidates for the
// maximum) are written to cand / *cand_count for an exact scan.
template <int ND>
int LaunchMfmaSplit(hd_context* ctx, const float* a, const uint32_t* list,
                    const float* list_min, uint32_t n_list, const EbTarget& t,
                    unsigned long long* key, unsigned long long* evals,
                    uint32_t* cand, unsigned int* cand_count) {
  HD_TRY(FillU32(ctx, cand_count, 0u, 1));
  if (n_list == 0) return HD_OK;
  const int blocks = MfmaResidentBlocks<ND>(ctx);
  const uint32_t n_waves = (uint32_t)blocks * (kBlock / kWaveSize);
  // Segments of about 1024 steps (at least 16 of them): long items leave
  // the end of the launch to a few waves, short ones multiply the work
  // spent after a hit until the other segments see it. A short list gets
  // more segments (down to 128 steps), enough for two claims of 16 entries
  // per wave: a wave steps through its segment one dependent load after
  // another, so the launch takes about seg_steps load latencies.
  const uint32_t want = (uint32_t)std::min<uint64_t>(
      0xFFFFFFFFull, (2ull * n_waves * 16u + n_list - 1) / n_list);
  uint32_t n_segs = std::max(std::max(16u, t.n_steps / 1024u),
                             std::min(want, t.n_steps / 128u));
  n_segs = std::min(n_segs, std::max(1u, t.n_steps / 128u));
  const uint32_t seg_steps = (t.n_steps + n_segs - 1) / n_segs;
  n_segs = (t.n_steps + seg_steps - 1) / seg_steps;
  MfmaArgs args = {};
  args.a = a;
  args.order = list;
  args.list_min = list_min;
  args.perm_a = MakeFeistel(1, 0, 0);
  args.order_mode = kOrderArray;
  args.pos_end = n_list;
  HD_TRY(ContextArray(ctx, "eb_mfma_cursors", n_segs, &args.cursors));
  HD_TRY(FillU32(ctx, args.cursors, 0u, n_segs));
  args.seg_steps = seg_steps;
  args.n_segs = n_segs;
  HD_TRY(ContextArray(ctx, "eb_broken", n_list, &args.broken));
  HD_TRY(ContextArray(ctx, "eb_gdone", n_list, &args.gdone));
  HD_TRY(FillU32(ctx, args.broken, 0u, n_list));
  HD_TRY(FillU32(ctx, args.gdone, 0u, n_list));
  args.cand = cand;
  args.cand_count = cand_count;
  args.bt = t.bt;
  args.bx = t.bx;
  args.by = t.by;
  args.bz = t.bz;
  args.n_steps = t.n_steps;
  args.budget_steps = 0xFFFFFFFFu;
  args.cx = t.center[0];
  args.cy = t.center[1];
  args.cz = t.center[2];
  args.err = t.err;
  args.off = t.off;
  args.key = key;
  args.evals = evals;
  const uint32_t per_seg = std::max(1u, n_waves / n_segs);
  args.claim = std::min<uint32_t>(
      kWaveSize, std::max<uint32_t>(16u, (n_list + per_seg - 1) / per_seg));
  StageDispatch(ctx);
  EbMfmaKernel<ND, false><<<blocks, kBlock, 0, ctx->stream>>>(args);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int ReadCount(hd_context* ctx, const unsigned int* d_count, uint32_t* out,
              const unsigned long long* key = nullptr, float* cmax2 = nullptr) {
  void* pinned = nullptr;
  HD_TRY(ContextPinned(ctx, 256, &pinned));
  unsigned int* h = static_cast<unsigned int*>(pinned);
  if (d_count != nullptr)
    HD_HIP_TRY(hipMemcpyAsync(h, d_count, 4, hipMemcpyDeviceToHost,
                              ctx->stream));
  if (key != nullptr)
    HD_HIP_TRY(hipMemcpyAsync(h + 2, key, 8, hipMemcpyDeviceToHost,
                              ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (out != nullptr) *out = h[0];
  if (cmax2 != nullptr) {
    float f;
    std::memcpy(&f, &h[3], 4);
    *cmax2 = f;
  }
  return HD_OK;
}

// Stage names of one kind of completion: filtered, exact follow-up, and
// exact-only.
struct CompletionStages {
  const char* filtered;
  const char* exact;
  const char* plain;
};

// Completes the listed entries over all of B, raising *key. With the
// matrix-core filter (B long enough for segments of 1024 steps and its
// error bound below half of cmax2): segmented filtered scans, then exact
// segmented scans of the entries that no segment could drop. Otherwise
// exact segmented scans (at least min_segs segments per entry). ev[0], ev[1]
// count the filtered and the follow-up evaluations, ev[2] the exact-only
// ones. cmax2 is a value *key has reached (a lower bound of the current).
template <int ND>
int CompleteEntries(hd_context* ctx, const float* a, const uint32_t* list,
                    const float* list_min, uint32_t n, const EbTarget& t,
                    float cmax2, unsigned long long* key,
                    unsigned long long* ev_filt, unsigned long long* ev_exact,
                    unsigned long long* ev_plain, uint32_t min_segs,
                    const CompletionStages& names, bool* filtered,
                    uint64_t* n_exact) {
  *filtered = false;
  if (n == 0) return HD_OK;
  const bool filt =
      t.bt != nullptr && t.n_steps >= 4096u && t.err < 0.5f * cmax2;
  *filtered = filt;
  if (!filt) {
    const int st = StageBegin(ctx, names.plain);
    HD_TRY((LaunchSplit<ND>(ctx, a, list, list_min, n, t.bx, t.by, t.bz,
                           t.n_steps, key, ev_plain, min_segs)));
    StageEnd(ctx, st);
    return HD_OK;
  }
  uint32_t* cand = nullptr;
  unsigned int* cand_count = nullptr;
  HD_TRY(ContextArray(ctx, "eb_cand", n, &cand));
  HD_TRY(ContextArray(ctx, "eb_cand_count", 1, &cand_count));
  const int st = StageBegin(ctx, names.filtered);
  HD_TRY((LaunchMfmaSplit<ND>(ctx, a, list, list_min, n, t, key, ev_filt, cand,
                              cand_count)));
  StageEnd(ctx, st);
  uint32_t n_cand = 0;
  HD_TRY(ReadCount(ctx, cand_count, &n_cand));
  *n_exact += n_cand;
  const int st_x = StageBegin(ctx, names.exact);
  HD_TRY((LaunchSplit<ND>(ctx, a, cand, nullptr, n_cand, t.bx, t.by, t.bz,
                         t.n_steps, key, ev_exact, 16u)));
  StageEnd(ctx, st_x);
  return HD_OK;
}

// Profiling stage of budgeted round r (0-based) on the exact scan or the
// matrix-core filter.
const char* RoundStage(int r, bool filt) {
  static const char* kScan[] = {"eb_scan_r1", "eb_scan_r2", "eb_scan_r3", "eb_scan_r4"};
  static const char* kMfma[] = {"eb_mfma_r1", "eb_mfma_r2", "eb_mfma_r3", "eb_mfma_r4"};
  return filt ? kMfma[r] : kScan[r];
}

// The filter pays off while steps that need an exact test stay rare, i.e.
// while its error bound is small next to cmax2.
bool UseFilter(const EbTarget& t, uint32_t budget, float cmax2) {
  return t.bt != nullptr && budget < t.n_steps && cmax2 > 0.f &&
         t.err < 0.5f * cmax2;
}

// Visit orders of B in the scans.
enum class BOrder { kInput, kPseudoRandom, kStratified };

// The order the flags select. By default the learned pipeline takes the
// stratified order and the other schedules the pseudo-random one: the
// ranked plain pipeline ranks A by partial minima over the first steps of
// B, and an evenly spread sample ranks the points of sparse regions first,
// which rarely define the HD (10M GMM: 1.7x the evaluations of the
// pseudo-random order), while the learned order's seed reaches the final
// cmax2 before any round (10M GMM: 12.5% fewer evaluations).
BOrder BOrderFor(uint32_t flags, uint32_t n_b) {
  if (flags & HD_FLAG_NO_SHUFFLE) return BOrder::kInput;
  if (n_b > kMaxStratified || (flags & HD_FLAG_B_ORDER_RANDOM))
    return BOrder::kPseudoRandom;
  if (flags & HD_FLAG_B_ORDER_STRATIFIED) return BOrder::kStratified;
  return (flags & HD_FLAG_LEARNED_ORDER) && !(flags & HD_FLAG_SINGLE_PASS)
             ? BOrder::kStratified
             : BOrder::kPseudoRandom;
}

// B repacked for the scans: SoA in the visit order, padded to whole 64-point
// steps with copies of existing points (duplicates never change a minimum).
// The stratified order (StratifiedOrder) spreads the B points of every
// small region evenly over the sequence, so a scan from any step meets the
// points near its A point sooner than in a pseudo-random order, where they
// fall at random (1M GMM, CPU model with the final cmax2 and one scan per A
// point from a random start: 3.2k evaluations per A point against 4.0k).
template <int ND>
int PrepareB(hd_context* ctx, const float* b, uint32_t n_b, BOrder order,
             int32_t seed, const float* mbr_b, EbTarget* t) {
  t->n_bp = (n_b + 63u) & ~63u;
  t->n_steps = t->n_bp / 64u;
  t->bx = t->by = t->bz = nullptr;
  HD_TRY(ContextArray(ctx, "eb_bx", t->n_bp, &t->bx));
  HD_TRY(ContextArray(ctx, "eb_by", t->n_bp, &t->by));
  if (ND == 3) HD_TRY(ContextArray(ctx, "eb_bz", t->n_bp, &t->bz));
  const Feistel perm_b = MakeFeistel(n_b, (uint64_t)(int64_t)seed, 1);
  uint32_t* idx = nullptr;
  if (order == BOrder::kStratified)
    HD_TRY(StratifiedOrder(ctx, ND, b, n_b, mbr_b, &idx));
  return GatherSoA(ctx, ND, b, n_b, t->n_bp, order == BOrder::kPseudoRandom,
                   perm_b, t->bx, t->by, t->bz, idx);
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
                EbRoundStats* rs, int first_round = 0,
                bool complete_top = true) {
  static const uint32_t kRanked[] = {16u, 256u, 1024u};
  // Without top completions (the learned order, cmax2 final from its seed)
  // longer budgets escalate fewer points (10M GMM, stratified order of B:
  // 512/1536 10.1 ms, 384/1536 10.2, 768/1536 10.15, 512/1280 10.2,
  // 384/1024 10.65, 384/2048 10.8).
  static const uint32_t kFinal[] = {16u, 512u, 1536u};
  const uint32_t* kBudgets = complete_top ? kRanked : kFinal;
  constexpr int kRounds = (int)(sizeof(kRanked) / sizeof(kRanked[0]));
  static_assert(kRounds <= kMaxRounds, "round statistics");
  const int n_rounds = first_budget ? 1 : kRounds;
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
  float cmax2 = 0.f;
  // cmax2 holds the key's value as of the last launch that could raise it
  // (read together with the escalation count; only top completions change
  // the key between that read and the next round), so the round start and
  // the tail read it again only after top completions.
  bool cmax2_current = false;
  // A pseudo-random order of a large set ranks on a sample: the first round
  // (short budget, cmax2 still far from its final value, exact scan) only
  // visits the first eighth of the order, enough for its top completions
  // to lift cmax2; the rest of the order joins the second round, which can
  // then run on the matrix-core filter.
  uint32_t extra_begin = 0, extra_end = 0;
  if (!first_budget && order_mode != kOrderArray && end - begin >= (1u << 21)) {
    extra_begin = begin + std::max<uint32_t>(65536u, (end - begin) / 8u);
    extra_end = end;
    end = extra_begin;
  }
  for (int r = first_budget ? 0 : first_round;
       r < n_rounds && (begin < end || extra_begin < extra_end); ++r) {
    const uint32_t budget = first_budget ? first_budget : kBudgets[r];
    const bool with_extra = r == 1 && extra_begin < extra_end;
    rs->round_points[r] += end - begin;
    if (t.bt != nullptr && budget < t.n_steps && !cmax2_current)
      HD_TRY(ReadCount(ctx, nullptr, nullptr, key, &cmax2));
    // Matrix-core filter when its threshold keeps most of cmax2, else the
    // exact scan.
    const bool filt = UseFilter(t, budget, cmax2);
    if (filt) rs->mfma_rounds |= 1u << r;
    const int st_slot = StageBegin(ctx, RoundStage(r, filt));
    auto launch = [&](int m, const uint32_t* l, const float* lm, uint32_t b,
                      uint32_t e, bool reset, unsigned long long* evc) -> int {
      if (filt)
        return LaunchMfma<ND>(ctx, a, m, l, lm, perm_a, b, e, t, budget, key,
                              evc, counters, idx_out, min_out, reset);
      if (budget <= kDenseSteps && budget < t.n_steps)
        return LaunchDenseRound<ND>(ctx, a, m, l, lm, perm_a, b, e, t.bx, t.by,
                                    t.bz, t.n_steps, budget, key, evc,
                                    counters, idx_out, min_out, reset);
      return LaunchSlots<ND>(ctx, a, m, l, lm, perm_a, b, e, t.bx, t.by, t.bz,
                             t.n_steps, budget, key, evc, counters, idx_out,
                             min_out, reset);
    };
    bool reset = true;
    if (begin < end) {
      HD_TRY(launch(mode, list, list_min, begin, end, true, ev + r));
      reset = false;
    }
    StageEnd(ctx, st_slot);
    if (with_extra) {
      // The part of the order the first round did not visit (a stage of
      // its own: one kernel launch per stage call).
      const int st_x = StageBegin(ctx, filt ? "eb_mfma_rest" : "eb_scan_rest");
      HD_TRY(launch(order_mode, order, nullptr, extra_begin, extra_end, reset,
                    ev + kMaxRounds + 3));
      StageEnd(ctx, st_x);
      rs->rest_points += extra_end - extra_begin;
      rs->rest_filtered = filt;
      extra_begin = extra_end;
    }
    uint32_t n_esc = 0;
    HD_TRY(ReadCount(ctx, counters + 1, &n_esc, key, &cmax2));
    cmax2_current = true;
    if (n_esc == 0) {
      begin = end = 0;
      continue;
    }
    const int st_sort = StageBegin(ctx, "eb_sort");
    HD_TRY(SortPairsDescF32(ctx, min_out, min_sorted, idx_out, idx_sorted,
                            n_esc));
    StageEnd(ctx, st_sort);
    // Exact completions of the escalations with the largest partial minima
    // (the likeliest to raise cmax2). After the first round's short budget
    // the ranking is noisy, so it completes more of them: on the 10M GMM
    // the chance of including one of the few points that set the final
    // cmax2 decides whether the later rounds run at 0.5 or 0.9 of it
    // (plain-order mean over 10 runs: 19.8 ms with 256, 16.8 ms with 1024).
    // Without top completions (a learned order whose seed already reached
    // the final cmax2) every escalation joins the next round: completing
    // 256 entries apart runs at a fifth of a round's evaluation rate. A
    // seed that missed the final cmax2 shows as a round escalating more
    // than an eighth of A (10M GMM: 3.5% after round 2; 20M x 15M sphere,
    // whose HD comes from radial noise the network does not resolve: 75%,
    // at 0.31 of the final cmax2), and then the rounds complete their top
    // entries as in the ranked pipeline.
    const bool seed_missed = (uint64_t)n_esc * 8u > (uint64_t)n_a;
    const uint32_t k_top = !complete_top && !seed_missed ? 0u
                           : std::min<uint32_t>(n_esc, (r == 0 && !first_budget) ? 1024u : 256u);
    rs->top_points += k_top;
    bool top_filt = false;
    HD_TRY((CompleteEntries<ND>(
        ctx, a, idx_sorted, min_sorted, k_top, t, cmax2, key,
        ev + kMaxRounds + 4, ev + kMaxRounds + 5, ev + kMaxRounds, 1u,
        {"eb_mfma_top", "eb_top_exact", "eb_top"}, &top_filt,
        &rs->top_exact_points)));
    if (top_filt) rs->top_filtered_points += k_top;
    if (k_top > 0) cmax2_current = false;
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
    // What is left needed more than the last budget: segmented scans, each
    // entry spread over at least 16 waves. With the filter, only the
    // entries that no segment can drop get an exact scan afterwards.
    rs->tail_points += end - begin;
    const uint32_t n_tail = end - begin;
    const float* tail_min = list_min ? list_min + begin : nullptr;
    if (t.bt != nullptr && t.n_steps >= 4096u && !cmax2_current)
      HD_TRY(ReadCount(ctx, nullptr, nullptr, key, &cmax2));
    HD_TRY((CompleteEntries<ND>(
        ctx, a, list + begin, tail_min, n_tail, t, cmax2, key,
        ev + kMaxRounds + 1, ev + kMaxRounds + 2, ev + kMaxRounds + 1, 16u,
        {"eb_mfma_tail", "eb_tail_exact", "eb_tail"}, &rs->tail_filtered,
        &rs->exact_points)));
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
  HD_TRY(PrepareB<ND>(ctx, in.b, n_b, BOrderFor(in.flags, n_b), in.seed,
                      in.has_mbr ? in.mbr_b : nullptr, &t));
  if (in.has_mbr && t.n_steps > 16u)
    HD_TRY(PrepareFilter<ND>(ctx, in.mbr_a, in.mbr_b, &t));
  // Visit order of A: learned (array), pseudo-random (Feistel) or input.
  const uint32_t* order = nullptr;
  const Feistel perm_a = MakeFeistel(n_a, (uint64_t)(int64_t)in.seed, 0);
  int order_mode = shuffle ? kOrderFeistel : kOrderIdentity;
  if (in.flags & HD_FLAG_LEARNED_ORDER) {
    HD_TRY(LearnedOrder(ctx, in, &order));
    order_mode = kOrderArray;
  }
  // A counter block: the key, then one evaluation counter per stage (seed,
  // the rounds, top completions, tail).
  static_assert(2 + kEbCounters <= (int)kCounterRow, "counter block row");
  unsigned long long* key = nullptr;
  HD_TRY(ContextArray(ctx, "eb_key", kCounterWords, &key));
  unsigned long long* ev = key + 1;
  HD_TRY(FillU64(ctx, key, 0ull, kCounterWords));
  StageEnd(ctx, st_prep);

  if (in.flags & HD_FLAG_SINGLE_PASS) {
    // One pass: every point of the order scans B until it breaks, cmax2
    // starting at 0; no escalation (the budget covers all of B).
    unsigned int* counters = nullptr;
    uint32_t* esc_idx = nullptr;
    float* esc_min = nullptr;
    HD_TRY(ContextArray(ctx, "eb_counters", 4, &counters));
    HD_TRY(ContextArray(ctx, "eb_esc_idx_a", 64, &esc_idx));
    HD_TRY(ContextArray(ctx, "eb_esc_min_a", 64, &esc_min));
    const int st = StageBegin(ctx, "eb_single_pass");
    HD_TRY((LaunchSlots<ND>(ctx, in.a, order_mode, order, nullptr, perm_a, 0u,
                           n_a, t.bx, t.by, t.bz, t.n_steps, 0xFFFFFFFFu, key,
                           ev + 0, counters, esc_idx, esc_min)));
    StageEnd(ctx, st);
    const int st_fin = StageBegin(ctx, "eb_finish");
    HD_TRY(FinishDirected(ctx, ND, in.a, in.b, in.n_b, key, ev, 1, out));
    StageEnd(ctx, st_fin);
    StageCount(ctx, "eb_single_pass", out->evals, n_a);
    return HD_OK;
  }

  // Stage 1: exact minima of the first s0 points of the order.
  // With a learned order the first points are the ones predicted farthest
  // from B; completing 2048 of them usually reaches the final cmax2 before
  // the rounds start. A pseudo-random order gains little from more than a
  // few hundred.
  const uint32_t s0_default = order_mode == kOrderArray ? 2048u : 256u;
  const uint32_t s0 = std::min<uint32_t>(
      n_a, in.seed_points ? in.seed_points : s0_default);
  uint32_t* seed_list = nullptr;
  HD_TRY(ContextArray(ctx, "eb_seed_list", s0, &seed_list));
  const int st_seed = StageBegin(ctx, "eb_seed");
  SeedListKernel<<<std::max(1u, (s0 + 255u) / 256u), 256, 0, ctx->stream>>>(
      order, perm_a, order_mode, s0, seed_list);
  HD_HIP_TRY(hipGetLastError());
  HD_TRY((LaunchSplit<ND>(ctx, in.a, seed_list, nullptr, s0, t.bx, t.by, t.bz,
                         t.n_steps, key, ev + 0)));
  StageEnd(ctx, st_seed);

  EbRoundStats rs;
  // The first round ranks points for the top completions; a learned order
  // has ranked them already (its seed usually reaches the final cmax2), so
  // it starts with the second and skips the top completions (10M GMM:
  // 12.5 ms against 13.05; the hybrid's EB finish of escalated points
  // keeps them, its cmax2 is not final: 8.5 ms against 4.5 without).
  HD_TRY(ProcessOpen<ND>(ctx, in.a, n_a, t, order_mode, order, nullptr, perm_a,
                         s0, n_a, in.slot_budget_steps, key, ev + 1, &rs,
                         order_mode == kOrderArray ? 1 : 0,
                         order_mode != kOrderArray));

  const int st_fin = StageBegin(ctx, "eb_finish");
  HD_TRY(FinishDirected(ctx, ND, in.a, in.b, in.n_b, key, ev, 1 + kEbCounters,
                        out));
  StageEnd(ctx, st_fin);
  if (ctx->profiling) {
    unsigned long long sums[kCounterRow];
    HD_TRY(ReadCounters(ctx, key, sums));
    const unsigned long long* h = sums + 1;
    StageCount(ctx, "eb_seed", h[0], s0);
    for (int r = 0; r < kMaxRounds; ++r) {
      if (rs.round_points[r] == 0) continue;
      const bool filt = (rs.mfma_rounds >> r) & 1u;
      StageCount(ctx, RoundStage(r, filt), h[1 + r], rs.round_points[r]);
    }
    if (rs.top_points > rs.top_filtered_points)
      StageCount(ctx, "eb_top", h[1 + kMaxRounds],
                 rs.top_points - rs.top_filtered_points);
    if (rs.top_filtered_points != 0) {
      StageCount(ctx, "eb_mfma_top", h[5 + kMaxRounds], rs.top_filtered_points);
      StageCount(ctx, "eb_top_exact", h[6 + kMaxRounds], rs.top_exact_points);
    }
    if (rs.rest_points != 0)
      StageCount(ctx, rs.rest_filtered ? "eb_mfma_rest" : "eb_scan_rest",
                 h[4 + kMaxRounds], rs.rest_points);
    if (rs.tail_filtered) {
      StageCount(ctx, "eb_mfma_tail", h[2 + kMaxRounds], rs.tail_points);
      StageCount(ctx, "eb_tail_exact", h[3 + kMaxRounds], rs.exact_points);
    } else {
      StageCount(ctx, "eb_tail", h[2 + kMaxRounds], rs.tail_points);
    }
  }
  return HD_OK;
}

}  // namespace

namespace {

template <int ND>
int EbSinglePassScratchImpl(hd_context* ctx, const float* a, uint32_t n_a,
                            const float* b, uint32_t n_b, bool shuffle,
                            int32_t seed, const EbScratch& s,
                            hd_directed_result* out) {
  const uint32_t n_bp = (n_b + 63u) & ~63u;
  if (n_bp > s.max_bp) return Fail(HD_EINVAL, "scratch holds %u points", s.max_bp);
  // Everything that must start at zero is cleared by fill kernels: the
  // scratch keeps the previous call's contents.
  HD_TRY(FillU64(ctx, s.key, 0ull, kCounterWords));
  HD_TRY(FillU32(ctx, s.counters, 0u, 4));
  const Feistel perm_b = MakeFeistel(n_b, (uint64_t)(int64_t)seed, 1);
  HD_TRY(GatherSoA(ctx, ND, b, n_b, n_bp, shuffle, perm_b, s.bx, s.by, s.bz));
  ScanArgs args = {};
  args.a = a;
  args.perm_a = MakeFeistel(n_a, (uint64_t)(int64_t)seed, 0);
  args.order_mode = shuffle ? kOrderFeistel : kOrderIdentity;
  args.pos_end = n_a;
  args.bx = s.bx;
  args.by = s.by;
  args.bz = s.bz;
  args.n_steps = n_bp / 64u;
  args.seg_steps = args.n_steps;
  args.n_segs = 1;
  args.budget_steps = 0xFFFFFFFFu;
  args.key = s.key;
  args.cursors = s.cursors;
  args.esc_idx = s.esc_idx;
  args.esc_min = s.esc_min;
  args.esc_count = s.counters + 1;
  args.evals = s.key + 1;
  HD_TRY(LaunchScan<ND>(ctx, args, 0u, n_a));
  // The key and the evaluation counter (column 1 of the counter block,
  // summed over its rows); a host array, nothing is allocated.
  unsigned long long h[2] = {0, 0};
  {
    unsigned long long block[kCounterWords];
    HD_HIP_TRY(hipMemcpyAsync(block, s.key, sizeof(block), hipMemcpyDeviceToHost,
                              ctx->stream));
    HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
    h[0] = block[0];
    for (uint32_t r = 0; r < kCounterRows; ++r) h[1] += block[r * kCounterRow + 1];
  }
  std::memset(out, 0, sizeof(*out));
  out->variant_used = HD_VARIANT_EB;
  out->evals = h[1];
  out->cmax2_bits = (uint32_t)(h[0] >> 32);
  float c;
  std::memcpy(&c, &out->cmax2_bits, 4);
  out->hd = std::sqrt(c);
  out->argmax = out->nn = 0xFFFFFFFFu;
  if (h[0] == 0ull) return HD_OK;
  out->argmax = 0xFFFFFFFFu - (uint32_t)h[0];
  HD_TRY(LaunchNnIndex(ctx, ND, a, out->argmax, b, n_b, out->cmax2_bits, s.nn));
  uint32_t nn = 0xFFFFFFFFu;
  HD_HIP_TRY(hipMemcpyAsync(&nn, s.nn, 4, hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  out->nn = nn;
  return HD_OK;
}

}  // namespace

int EbSinglePassScratch(hd_context* ctx, int n_dims, const float* a,
                        uint32_t n_a, const float* b, uint32_t n_b,
                        bool shuffle, int32_t seed, const EbScratch& s,
                        hd_directed_result* out) {
  if (n_dims == 2)
    return EbSinglePassScratchImpl<2>(ctx, a, n_a, b, n_b, shuffle, seed, s, out);
  return EbSinglePassScratchImpl<3>(ctx, a, n_a, b, n_b, shuffle, seed, s, out);
}

int FilterProbe(hd_context* ctx, int n_dims, const float* a, uint32_t n_a,
                const float* b, uint32_t n_b, float* out, float* cna_out,
                float* params) {
  float mbr_a[6], mbr_b[6];
  HD_TRY(HostMbrPair(ctx, n_dims, a, n_a, mbr_a, b, n_b, mbr_b));
  EbTarget t;
  if (n_dims == 2) {
    HD_TRY(PrepareB<2>(ctx, b, n_b, BOrder::kInput, 0, nullptr, &t));
    HD_TRY(PrepareFilter<2>(ctx, mbr_a, mbr_b, &t));
  } else {
    HD_TRY(PrepareB<3>(ctx, b, n_b, BOrder::kInput, 0, nullptr, &t));
    HD_TRY(PrepareFilter<3>(ctx, mbr_a, mbr_b, &t));
  }
  if (t.bt == nullptr)
    return Fail(HD_EUNSUPPORTED, "coordinates too large for the filter");
  const uint32_t n_bt = (n_b + 15u) / 16u, n_at = (n_a + 15u) / 16u;
  const uint64_t waves = (uint64_t)n_at * n_bt;
  if (waves > 0x7FFFFFFFull) return Fail(HD_EINVAL, "probe too large");
  if (n_dims == 2)
    FilterProbeKernel<2><<<(unsigned)waves, 64, 0, ctx->stream>>>(
        t.bt, n_bt, n_b, a, n_a, t.center[0], t.center[1], t.center[2], t.off,
        out, cna_out);
  else
    FilterProbeKernel<3><<<(unsigned)waves, 64, 0, ctx->stream>>>(
        t.bt, n_bt, n_b, a, n_a, t.center[0], t.center[1], t.center[2], t.off,
        out, cna_out);
  HD_HIP_TRY(hipGetLastError());
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  params[0] = t.center[0];
  params[1] = t.center[1];
  params[2] = t.center[2];
  params[3] = t.err;
  params[4] = t.off;
  return HD_OK;
}

int RunEarlyBreak(hd_context* ctx, const DirectedArgs& args,
                  hd_directed_result* out) {
  if (args.n_dims == 2) return RunEarlyBreakImpl<2>(ctx, args, out);
  return RunEarlyBreakImpl<3>(ctx, args, out);
}

int EbPrepareTarget(hd_context* ctx, int n_dims, const float* b, uint32_t n_b,
                    uint32_t flags, int32_t seed, const float* mbr_b,
                    EbTarget* t) {
  const BOrder order = BOrderFor(flags, n_b);
  if (n_dims == 2) return PrepareB<2>(ctx, b, n_b, order, seed, mbr_b, t);
  return PrepareB<3>(ctx, b, n_b, order, seed, mbr_b, t);
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
