// Software LBVH over B for nearest-neighbour queries with early break.
//
// The reference answers the same query with OptiX on RT cores: B's grid
// cells become radius-inflated custom primitives (src/rt/rt_engine.h:956-1180)
// and an intersection program prunes by cell min/max distance and scans the
// cell's points with early break (src/rt/shaders/
// shaders_nn_uniform_grid_2d.cu:16-106). CDNA4 has no ray-tracing units, so
// here the hierarchy is built and traversed on the shader cores:
//  * build: 32-bit Morton codes of B (16 bits per axis in 2D, 10-11 in 3D;
//    equal codes ordered by index), one radix sort, leaves = runs of kLeaf consecutive sorted points,
//    internal nodes by Karras (HPG 2012) with index tie-breaking for equal
//    codes, bottom-up boxes with one arrival counter per node;
//  * node record: both child boxes inline plus the two child ids, so a visit
//    is one 64-byte load;
//  * query: one lane per A point, nearest child first, the other pushed on a
//    per-lane stack in LDS; children whose box is not closer than the best
//    d2 are pruned; a query stops when its best d2 is strictly below cmax2.
// Pruning is exact: a box's FP32 squared distance uses the same
// subtract / square / add sequence as the point d2 and IEEE rounding is
// monotone, so it never exceeds the d2 of any point inside the box.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "hd_device.h"
#include "hd_internal.h"
#include "hd_lbvh.h"
#include "hd_primitives.h"

namespace hd {

using namespace dev;

namespace {

constexpr uint32_t kLeafFlag = 0x80000000u;
constexpr int kStack = 64;

HD_DEV unsigned long long Spread2x32(uint32_t v) {
  unsigned long long x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
HD_DEV unsigned long long Spread3x21(uint32_t v) {
  unsigned long long x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x1F00000000FFFFull;
  x = (x | (x << 16)) & 0x1F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// Morton code of point p in frame f (64 bits in 2D, 63 in 3D).
template <int ND>
HD_DEV unsigned long long MortonCode(const float* p, const LbvhFrame& f) {
  uint32_t q[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const float t = (p[d] - f.lo[d]) * f.scale[d];
    q[d] = (uint32_t)fminf(fmaxf(t, 0.0f), f.qmax);
  }
  if constexpr (ND == 2)
    return Spread2x32(q[0]) | (Spread2x32(q[1]) << 1);
  else
    return Spread3x21(q[0]) | (Spread3x21(q[1]) << 1) | (Spread3x21(q[2]) << 2);
}

// The top 32 bits of the Morton code: 16 bits per axis in 2D, 10-11 in 3D
// (a cell of 2^-16 or 2^-10 of B's extent; points sharing one are ordered
// by index). Four radix passes instead of eight.
template <int ND>
HD_DEV uint32_t MortonCode32(const float* p, const LbvhFrame& f) {
  return (uint32_t)(MortonCode<ND>(p, f) >> (ND == 2 ? 32 : 31));
}

template <int ND>
__global__ void MortonKernel32(const float* __restrict__ b, uint32_t n,
                               LbvhFrame f, uint32_t* __restrict__ keys,
                               uint32_t* __restrict__ vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    keys[i] = MortonCode32<ND>(b + (uint64_t)i * ND, f);
    vals[i] = i;
  }
}

template <int ND>
__global__ void GatherSortedKernel(const float* __restrict__ b, uint32_t n,
                                   const uint32_t* __restrict__ vals,
                                   float* __restrict__ x, float* __restrict__ y,
                                   float* __restrict__ z,
                                   uint32_t* __restrict__ orig) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t src = vals[i];
    const float* p = b + (uint64_t)src * ND;
    x[i] = p[0];
    y[i] = p[1];
    if constexpr (ND == 3) z[i] = p[2];
    orig[i] = src;
  }
}

// Karras' prefix length between leaves i and j (codes of their first
// points), with the leaf index as tie-breaker for equal codes.
HD_DEV int Delta(const uint32_t* codes, int n_leaf, uint32_t leaf, int i,
                 int j) {
  if (j < 0 || j >= n_leaf) return -1;
  const uint32_t ci = codes[(uint64_t)i * leaf];
  const uint32_t cj = codes[(uint64_t)j * leaf];
  if (ci == cj) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(ci ^ cj);
}

__global__ void KarrasKernel(const uint32_t* __restrict__ codes,
                             int n_leaf, uint32_t leaf,
                             uint32_t* __restrict__ child,
                             uint32_t* __restrict__ node_parent,
                             uint32_t* __restrict__ leaf_parent) {
  const int n_int = n_leaf - 1;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_int;
       i += gridDim.x * blockDim.x) {
    const int d =
        (Delta(codes, n_leaf, leaf, i, i + 1) - Delta(codes, n_leaf, leaf, i, i - 1)) > 0
            ? 1
            : -1;
    const int dmin = Delta(codes, n_leaf, leaf, i, i - d);
    int lmax = 2;
    while (Delta(codes, n_leaf, leaf, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (Delta(codes, n_leaf, leaf, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = Delta(codes, n_leaf, leaf, i, j);
    int s = 0;
    int t = l;
    do {
      t = (t + 1) >> 1;
      if (Delta(codes, n_leaf, leaf, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const bool left_leaf = min(i, j) == gamma;
    const bool right_leaf = max(i, j) == gamma + 1;
    child[2 * i] = left_leaf ? ((uint32_t)gamma | kLeafFlag) : (uint32_t)gamma;
    child[2 * i + 1] =
        right_leaf ? ((uint32_t)(gamma + 1) | kLeafFlag) : (uint32_t)(gamma + 1);
    if (left_leaf) leaf_parent[gamma] = i;
    else node_parent[gamma] = i;
    if (right_leaf) leaf_parent[gamma + 1] = i;
    else node_parent[gamma + 1] = i;
  }
}

// Leaf boxes (lo[ND], hi[ND] at stride 6), before the refit pass.
template <int ND>
__global__ void LeafBoxKernel(int n_leaf, uint32_t leaf, uint32_t n,
                              const float* __restrict__ x,
                              const float* __restrict__ y,
                              const float* __restrict__ z,
                              float* __restrict__ leaf_box) {
  for (int li = blockIdx.x * blockDim.x + threadIdx.x; li < n_leaf;
       li += gridDim.x * blockDim.x) {
    float lo[3] = {kFltMax, kFltMax, kFltMax};
    float hi[3] = {-kFltMax, -kFltMax, -kFltMax};
    const uint32_t b0 = (uint32_t)li * leaf;
    const uint32_t b1 = min(b0 + leaf, n);
    for (uint32_t k = b0; k < b1; ++k) {
      const float v[3] = {x[k], y[k], ND == 3 ? z[k] : 0.f};
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        lo[d] = fminf(lo[d], v[d]);
        hi[d] = fmaxf(hi[d], v[d]);
      }
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      leaf_box[(uint64_t)li * 6 + d] = lo[d];
      leaf_box[(uint64_t)li * 6 + 3 + d] = hi[d];
    }
  }
}

// Up the tree from every leaf: the second child subtree to arrive at a node
// writes the node record (both child boxes inline) and continues upwards
// with the union. A subtree's box reaches its sibling's thread through the
// node's mailbox (pub, 2 ND words per child slot): written with atomic
// exchanges whose results the arrival count's increment depends on, so the
// count is raised only after the box is at the device coherence point, and
// read with atomic loads by the thread that saw the other arrival. No
// acquire or release fence: at device scope those write back and
// invalidate the L2.
template <int ND>
__global__ void RefitKernel(int n_leaf, const float* __restrict__ leaf_box,
                            const uint32_t* __restrict__ child,
                            const uint32_t* __restrict__ node_parent,
                            const uint32_t* __restrict__ leaf_parent,
                            uint32_t* __restrict__ arrive,
                            uint32_t* __restrict__ pub,
                            LbvhNode* __restrict__ nodes) {
  for (int li = blockIdx.x * blockDim.x + threadIdx.x; li < n_leaf;
       li += gridDim.x * blockDim.x) {
    if (n_leaf == 1) continue;
    float lo[ND], hi[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      lo[d] = leaf_box[(uint64_t)li * 6 + d];
      hi[d] = leaf_box[(uint64_t)li * 6 + 3 + d];
    }
    uint32_t me = (uint32_t)li | kLeafFlag;
    uint32_t node = leaf_parent[li];
    while (true) {
      const uint32_t c0 = child[2 * node], c1 = child[2 * node + 1];
      const int slot = c0 == me ? 0 : 1;
      uint32_t* mail = pub + ((uint64_t)node * 2 + slot) * (2 * ND);
      uint32_t dep = 0xFFFFFFFFu;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        dep &= __hip_atomic_exchange(&mail[d], __float_as_uint(lo[d]),
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        dep &= __hip_atomic_exchange(&mail[ND + d], __float_as_uint(hi[d]),
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      // dep is all ones only if every old word was a NaN pattern; either
      // increment works (only prev == 0 is tested), the dependency is what
      // orders the count after the exchanges.
      const uint32_t prev = __hip_atomic_fetch_add(
          &arrive[node], 1u + (dep == 0xFFFFFFFFu ? 1u : 0u), __ATOMIC_RELAXED,
          __HIP_MEMORY_SCOPE_AGENT);
      if (prev == 0) break;  // the sibling is not done yet
      const uint32_t* other = pub + ((uint64_t)node * 2 + (1 - slot)) * (2 * ND);
      float olo[ND], ohi[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        olo[d] = __uint_as_float(__hip_atomic_load(&other[d], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT));
        ohi[d] = __uint_as_float(__hip_atomic_load(
            &other[ND + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      }
      LbvhNode rec;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float l0 = d < ND ? (slot == 0 ? lo[d] : olo[d]) : 0.f;
        const float h0 = d < ND ? (slot == 0 ? hi[d] : ohi[d]) : 0.f;
        const float l1 = d < ND ? (slot == 0 ? olo[d] : lo[d]) : 0.f;
        const float h1 = d < ND ? (slot == 0 ? ohi[d] : hi[d]) : 0.f;
        rec.lo[0][d] = l0;
        rec.hi[0][d] = h0;
        rec.lo[1][d] = l1;
        rec.hi[1][d] = h1;
      }
      rec.child[0] = c0;
      rec.child[1] = c1;
      rec.pad[0] = rec.pad[1] = 0u;
      nodes[node] = rec;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        lo[d] = fminf(lo[d], olo[d]);
        hi[d] = fmaxf(hi[d], ohi[d]);
      }
      if (node == 0) break;
      me = node;
      node = node_parent[node];
    }
  }
}

// Lower bound of each query's Morton code among B's sorted codes, queries
// listed in Morton order: lanes of a wave hold neighbouring codes, so their
// binary searches read the same cache lines.
template <int ND>
__global__ void StartLeafKernel(const float* __restrict__ q,
                                const uint32_t* __restrict__ order,
                                uint32_t n_q, LbvhFrame f,
                                const uint32_t* __restrict__ bcodes,
                                uint32_t n_b, uint32_t leaf, uint32_t n_leaf,
                                uint32_t* __restrict__ start) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_q;
       i += gridDim.x * blockDim.x) {
    const uint32_t c = MortonCode32<ND>(q + (uint64_t)order[i] * ND, f);
    uint32_t lo = 0, hi = n_b;  // first index with bcodes[k] >= c
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (bcodes[mid] < c) lo = mid + 1;
      else hi = mid;
    }
    start[i] = min(lo / leaf, n_leaf - 1u);
  }
}

// FP32 squared distance from q to a box, same operation order as D2.
template <int ND>
HD_DEV float BoxD2(const float* q, const float* lo, const float* hi) {
  float acc = 0.f;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    float diff = 0.f;
    if (q[d] < lo[d]) diff = lo[d] - q[d];
    else if (q[d] > hi[d]) diff = q[d] - hi[d];
    if (d == 0) acc = diff * diff;
    else acc = acc + diff * diff;
  }
  return acc;
}

template <int ND>
__global__ __launch_bounds__(64) void LbvhQueryKernel(LbvhQueryArgs args) {
  __shared__ uint32_t stack[kStack * 64];
  const LbvhDev& t = args.tree;
  const uint32_t lane = LaneId();
  const uint32_t lanes_total = gridDim.x * blockDim.x;
  const uint32_t wave_base = blockIdx.x * blockDim.x;
  const bool tie_index = args.nn_d2 != nullptr;
  unsigned long long evals = 0, nodes = 0;
  for (uint32_t base = args.pos_begin + wave_base; base < args.pos_end;
       base += lanes_total) {
    const uint32_t pos = args.stride > 1u ? (base + lane) * args.stride : base + lane;
    const bool active = base + lane < args.pos_end &&
                        !(args.skip_stride > 1u && pos % args.skip_stride == 0u &&
                          pos / args.skip_stride < args.skip_n);
    const float cmax2 = args.eb ? LoadCmax2(args.key) : 0.f;
    float best = kFltMax;
    uint32_t best_j = 0xFFFFFFFFu;
    bool broke = false, overflow = false;
    uint32_t idx = 0;
    if (active) {
      idx = args.order_mode == 1 ? Permute(args.perm, pos)
            : args.order_mode == 2 ? pos
                                   : args.order[pos];
      float q[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int d = 0; d < ND; ++d) q[d] = args.a[(uint64_t)idx * ND + d];
      // Exact scan of the leaves around the query's Morton position first:
      // on clustered data a point within cmax2 is usually there, and the
      // query ends without a traversal; otherwise the traversal starts with
      // their minimum as its bound.
      bool skip = false;
      if (args.start_leaf != nullptr) {
        const uint32_t s = args.start_leaf[pos];
        const uint32_t l0 = s > 0u ? s - 1u : 0u;
        const uint32_t l1 = min(s + 2u, t.n_leaf);
        for (uint32_t li = l0; li < l1; ++li) {
          const uint32_t b1 = min((li + 1u) * t.leaf, t.n);
          for (uint32_t k = li * t.leaf; k < b1; ++k) {
            float d2;
            if constexpr (ND == 2)
              d2 = D2(q[0], q[1], t.x[k], t.y[k]);
            else
              d2 = D2(q[0], q[1], q[2], t.x[k], t.y[k], t.z[k]);
            evals++;
            if (d2 < best) best = d2;
          }
        }
        if (args.eb && best < cmax2) {
          broke = true;
          skip = true;
        }
      }
      int sp = 0;
      uint32_t node = t.n_leaf == 1 ? (0u | kLeafFlag) : 0u;
      while (!skip) {
        // Child ids come from the build; an id out of range would mean a
        // malformed tree, so the query is handed to the exhaustive path
        // instead of reading outside the arrays.
        if ((node & kLeafFlag) ? (node & ~kLeafFlag) >= t.n_leaf
                               : node + 1 >= t.n_leaf) {
          overflow = true;
          break;
        }
        if (node & kLeafFlag) {
          const uint32_t li = node & ~kLeafFlag;
          const uint32_t b0 = li * t.leaf;
          const uint32_t b1 = min(b0 + t.leaf, t.n);
          for (uint32_t k = b0; k < b1; ++k) {
            float d2;
            if constexpr (ND == 2)
              d2 = D2(q[0], q[1], t.x[k], t.y[k]);
            else
              d2 = D2(q[0], q[1], q[2], t.x[k], t.y[k], t.z[k]);
            evals++;
            if (d2 < best) {
              best = d2;
              best_j = tie_index ? t.orig[k] : 0u;
            } else if (tie_index && d2 == best && t.orig[k] < best_j) {
              best_j = t.orig[k];
            }
          }
          if (args.eb && best < cmax2) {
            broke = true;
            break;
          }
          if (sp == 0) break;
          node = stack[(--sp) * 64 + lane];
          continue;
        }
        const LbvhNode rec = t.nodes[node];
        nodes++;
        const float d0 = BoxD2<ND>(q, rec.lo[0], rec.hi[0]);
        const float d1 = BoxD2<ND>(q, rec.lo[1], rec.hi[1]);
        const bool v0 = tie_index ? d0 <= best : d0 < best;
        const bool v1 = tie_index ? d1 <= best : d1 < best;
        if (v0 && v1) {
          const bool first0 = d0 <= d1;
          const uint32_t near = first0 ? rec.child[0] : rec.child[1];
          const uint32_t far = first0 ? rec.child[1] : rec.child[0];
          if (sp == kStack) {
            overflow = true;
            break;
          }
          stack[(sp++) * 64 + lane] = far;
          node = near;
        } else if (v0) {
          node = rec.child[0];
        } else if (v1) {
          node = rec.child[1];
        } else {
          if (sp == 0) break;
          node = stack[(--sp) * 64 + lane];
        }
      }
    }
    if (active && tie_index) {
      args.nn_d2[idx] = best;
      args.nn_idx[idx] = overflow ? 0xFFFFFFFEu : best_j;
    }
    // A stack overflow hands the point to the exact early-break path.
    const bool esc = active && overflow;
    const unsigned long long esc_mask = __ballot(esc);
    if (esc_mask != 0ull && args.esc_count != nullptr) {
      uint32_t slot_base = 0;
      if (lane == 0)
        slot_base = atomicAdd(args.esc_count, (unsigned)__popcll(esc_mask));
      slot_base = Uniform(slot_base);
      if (esc) {
        const uint32_t rank = __popcll(esc_mask & ((1ull << lane) - 1ull));
        args.esc_idx[slot_base + rank] = idx;
        args.esc_min[slot_base + rank] = best;
      }
    }
    if (args.key != nullptr) {
      const bool done = active && !broke && !overflow && best != kFltMax;
      const unsigned long long k = done ? PackKey(best, idx) : 0ull;
      const unsigned long long kmax = WaveMaxU64(k);
      if (lane == 0 && kmax != 0ull) KeyMax(args.key, kmax);
    }
  }
  unsigned long long e = evals, l = nodes;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    l += __shfl_xor(l, off, 64);
  }
  if (lane == 0 && args.evals != nullptr) CounterAdd(args.evals, e);
  if (lane == 0 && args.lookups != nullptr) CounterAdd(args.lookups, l);
}

}  // namespace

LbvhFrame MortonFrame(int n_dims, const float* mbr) {
  LbvhFrame f;
  std::memset(&f, 0, sizeof(f));
  // Largest quantisation level, exactly representable in FP32 so the
  // clamp in MortonKernel never converts an out-of-range float.
  const double levels = n_dims == 2 ? 4294967040.0 : 2097151.0;
  f.qmax = (float)levels;
  for (int d = 0; d < n_dims; ++d) {
    f.lo[d] = mbr[d];
    // A zero extent (a flat set whose widening rounded away) quantises to 0.
    const double ext = (double)mbr[n_dims + d] - (double)mbr[d];
    f.scale[d] = ext > 0 ? (float)(levels / ext) : 0.f;
  }
  return f;
}

int MortonOrder(hd_context* ctx, int n_dims, const float* pts, uint32_t n,
                const LbvhFrame& f, uint32_t** order_out, int begin_bit,
                const char* tag, uint32_t out_mul) {
  uint32_t *keys = nullptr, *keys_s = nullptr;
  uint32_t *vals = nullptr, *order = nullptr;
  const std::string t(tag);
  HD_TRY(ContextArray(ctx, (t + "keys").c_str(), n, &keys));
  HD_TRY(ContextArray(ctx, (t + "keys_s").c_str(), n, &keys_s));
  HD_TRY(ContextArray(ctx, (t + "vals").c_str(), n, &vals));
  HD_TRY(ContextArray(ctx, (t + "order").c_str(), n, &order));
  const int blocks = (int)std::min<uint64_t>((n + 255) / 256, 8192);
  if (n_dims == 2)
    MortonKernel32<2><<<blocks, 256, 0, ctx->stream>>>(pts, n, f, keys, vals);
  else
    MortonKernel32<3><<<blocks, 256, 0, ctx->stream>>>(pts, n, f, keys, vals);
  HD_HIP_TRY(hipGetLastError());
  HD_TRY(SortPairsU32(ctx, keys, keys_s, vals, order, n, 32, begin_bit,
                     out_mul));
  *order_out = order;
  return HD_OK;
}

int StratifiedOrder(hd_context* ctx, int n_dims, const float* pts, uint32_t n,
                    const float* mbr, uint32_t** perm_out) {
  if (n == 0 || n > kMaxStratified)
    return Fail(HD_EINVAL, "StratifiedOrder: %u points", n);
  float m[6];
  if (mbr)
    std::memcpy(m, mbr, sizeof(float) * 2 * n_dims);
  else
    HD_TRY(HostMbr(ctx, n_dims, pts, n, m));
  // Rank r goes to position r * step mod n (GoldenStep): the ranks of any
  // run (the points of one region) land evenly spread over the sequence.
  const uint32_t step = GoldenStep(n);
  // Ranks by the top 16 bits of the Morton code (cells of 2^-8 of the
  // extent per axis in 2D, about 2^-5 in 3D; two radix passes: a CPU model
  // of the 1M GMM pair gives the same evaluations as with 24 bits); the
  // sort's last pass writes each index straight to its position.
  HD_TRY(MortonOrder(ctx, n_dims, pts, n, MortonFrame(n_dims, m), perm_out,
                     16, "border_", n > 1 ? step : 0u));
  return HD_OK;
}

int StartLeaves(hd_context* ctx, const LbvhIndex& tree, const float* q,
                const uint32_t* order, uint32_t n, uint32_t* start) {
  if (n == 0) return HD_OK;
  const int blocks = (int)std::min<uint64_t>((n + 255) / 256, 8192);
  if (tree.n_dims == 2)
    StartLeafKernel<2><<<blocks, 256, 0, ctx->stream>>>(
        q, order, n, tree.frame, tree.codes, tree.dev.n, tree.dev.leaf,
        tree.dev.n_leaf, start);
  else
    StartLeafKernel<3><<<blocks, 256, 0, ctx->stream>>>(
        q, order, n, tree.frame, tree.codes, tree.dev.n, tree.dev.leaf,
        tree.dev.n_leaf, start);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int BuildLbvh(hd_context* ctx, int n_dims, const float* b, uint64_t n_b,
              uint32_t leaf_size, LbvhIndex* out, const float* mbr_b) {
  const int nd = n_dims;
  const uint32_t n = (uint32_t)n_b;
  const uint32_t leaf = leaf_size ? leaf_size : 8u;
  float mbr[6];
  if (mbr_b)
    std::memcpy(mbr, mbr_b, sizeof(float) * 2 * nd);
  else
    HD_TRY(HostMbr(ctx, nd, b, n_b, mbr));
  const LbvhFrame f = MortonFrame(nd, mbr);
  uint32_t *keys = nullptr, *keys_s = nullptr;
  uint32_t *vals = nullptr, *vals_s = nullptr;
  HD_TRY(ContextArray(ctx, "lbvh_keys", n, &keys));
  HD_TRY(ContextArray(ctx, "lbvh_keys_s", n, &keys_s));
  HD_TRY(ContextArray(ctx, "lbvh_vals", n, &vals));
  HD_TRY(ContextArray(ctx, "lbvh_vals_s", n, &vals_s));
  const int blocks = (int)std::min<uint64_t>((n + 255) / 256, 8192);
  if (nd == 2)
    MortonKernel32<2><<<blocks, 256, 0, ctx->stream>>>(b, n, f, keys, vals);
  else
    MortonKernel32<3><<<blocks, 256, 0, ctx->stream>>>(b, n, f, keys, vals);
  HD_HIP_TRY(hipGetLastError());
  HD_TRY(SortPairsU32(ctx, keys, keys_s, vals, vals_s, n, 32));

  LbvhDev& t = out->dev;
  std::memset(&t, 0, sizeof(t));
  float *x = nullptr, *y = nullptr, *z = nullptr;
  uint32_t* orig = nullptr;
  HD_TRY(ContextArray(ctx, "lbvh_x", n, &x));
  HD_TRY(ContextArray(ctx, "lbvh_y", n, &y));
  if (nd == 3) HD_TRY(ContextArray(ctx, "lbvh_z", n, &z));
  HD_TRY(ContextArray(ctx, "lbvh_orig", n, &orig));
  if (nd == 2)
    GatherSortedKernel<2><<<blocks, 256, 0, ctx->stream>>>(b, n, vals_s, x, y,
                                                           z, orig);
  else
    GatherSortedKernel<3><<<blocks, 256, 0, ctx->stream>>>(b, n, vals_s, x, y,
                                                           z, orig);
  HD_HIP_TRY(hipGetLastError());

  const uint32_t n_leaf = (n + leaf - 1) / leaf;
  const uint32_t n_int = n_leaf > 1 ? n_leaf - 1 : 1;
  uint32_t *child = nullptr, *node_parent = nullptr, *leaf_parent = nullptr,
           *arrive = nullptr, *pub = nullptr;
  float* leaf_box = nullptr;
  LbvhNode* nodes = nullptr;
  HD_TRY(ContextArray(ctx, "lbvh_child", 2 * (uint64_t)n_int, &child));
  HD_TRY(ContextArray(ctx, "lbvh_node_parent", n_int, &node_parent));
  HD_TRY(ContextArray(ctx, "lbvh_leaf_parent", n_leaf, &leaf_parent));
  HD_TRY(ContextArray(ctx, "lbvh_arrive", n_int, &arrive));
  HD_TRY(ContextArray(ctx, "lbvh_leaf_box", 6 * (uint64_t)n_leaf, &leaf_box));
  HD_TRY(ContextArray(ctx, "lbvh_pub", 4 * (uint64_t)nd * n_int, &pub));
  HD_TRY(ContextArray(ctx, "lbvh_nodes", n_int, &nodes));
  HD_TRY(FillU32(ctx, arrive, 0u, n_int));
  if (n_leaf > 1) {
    const int kb = (int)std::min<uint64_t>((n_int + 255) / 256, 8192);
    KarrasKernel<<<kb, 256, 0, ctx->stream>>>(keys_s, (int)n_leaf, leaf, child,
                                              node_parent, leaf_parent);
    HD_HIP_TRY(hipGetLastError());
  }
  const int lb = (int)std::min<uint64_t>((n_leaf + 255) / 256, 8192);
  if (nd == 2) {
    LeafBoxKernel<2><<<lb, 256, 0, ctx->stream>>>((int)n_leaf, leaf, n, x, y,
                                                  z, leaf_box);
    RefitKernel<2><<<lb, 256, 0, ctx->stream>>>((int)n_leaf, leaf_box, child,
                                                node_parent, leaf_parent,
                                                arrive, pub, nodes);
  } else {
    LeafBoxKernel<3><<<lb, 256, 0, ctx->stream>>>((int)n_leaf, leaf, n, x, y,
                                                  z, leaf_box);
    RefitKernel<3><<<lb, 256, 0, ctx->stream>>>((int)n_leaf, leaf_box, child,
                                                node_parent, leaf_parent,
                                                arrive, pub, nodes);
  }
  HD_HIP_TRY(hipGetLastError());
  t.nodes = nodes;
  t.x = x;
  t.y = y;
  t.z = z;
  t.orig = orig;
  t.n = n;
  t.n_leaf = n_leaf;
  t.leaf = leaf;
  out->n_dims = nd;
  out->frame = f;
  out->codes = keys_s;
  return HD_OK;
}

// One-wave blocks of the query kernel that a CU runs at once. The queries
// are split evenly over the blocks, so a launch of more blocks than run at
// once finishes with a late partial round. The occupancy query reports 10
// per CU on gfx950 (its 16 KB LDS stack), but launches behave as if 8 run
// at once (10M GMM: 8 per CU 2.53 ms, 10 2.77, 12 2.67, 16 2.54; 20M x
// 15M sphere: 16 per CU 5.29 ms, 10 5.71), so at most 8 are used.
constexpr int kLbvhBlocksPerCu = 8;

template <int ND>
int LbvhResidentBlocks(hd_context* ctx) {
  static int per_cu[4] = {0, 0, 0, 0};
  if (per_cu[ND] == 0) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, LbvhQueryKernel<ND>, 64,
                                                     0) != hipSuccess ||
        n < 1) {
      (void)hipGetLastError();
      n = 1;
    }
    per_cu[ND] = std::min(n, kLbvhBlocksPerCu);
  }
  return per_cu[ND] * ctx->cu_count;
}

int LaunchLbvhQuery(hd_context* ctx, const LbvhQueryArgs& args) {
  if (args.pos_begin >= args.pos_end) return HD_OK;
  const uint32_t n = args.pos_end - args.pos_begin;
  int blocks = args.tree_nd == 2 ? LbvhResidentBlocks<2>(ctx)
                                 : LbvhResidentBlocks<3>(ctx);
  blocks = (int)std::min<uint32_t>((uint32_t)blocks, (n + 63) / 64);
  blocks = std::max(blocks, 1);
  StageDispatch(ctx);
  if (args.tree_nd == 2)
    LbvhQueryKernel<2><<<blocks, 64, 0, ctx->stream>>>(args);
  else
    LbvhQueryKernel<3><<<blocks, 64, 0, ctx->stream>>>(args);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int LbvhNearestNeighbors(hd_context* ctx, int n_dims, const float* q,
                         uint64_t n_q, const float* ref, uint64_t n_ref,
                         float* out_d2, uint32_t* out_idx) {
  LbvhIndex tree;
  HD_TRY(BuildLbvh(ctx, n_dims, ref, n_ref, 8, &tree));
  unsigned int* esc_count = nullptr;
  uint32_t* esc_idx = nullptr;
  float* esc_min = nullptr;
  HD_TRY(ContextArray(ctx, "lbvh_nn_esc_count", 1, &esc_count));
  HD_TRY(ContextArray(ctx, "lbvh_nn_esc_idx", n_q, &esc_idx));
  HD_TRY(ContextArray(ctx, "lbvh_nn_esc_min", n_q, &esc_min));
  HD_TRY(FillU32(ctx, esc_count, 0u, 1));
  LbvhQueryArgs qa;
  std::memset(&qa, 0, sizeof(qa));
  qa.tree = tree.dev;
  qa.tree_nd = n_dims;
  qa.a = q;
  qa.order_mode = 2;
  qa.pos_begin = 0;
  qa.pos_end = (uint32_t)n_q;
  qa.nn_d2 = out_d2;
  qa.nn_idx = out_idx;
  qa.esc_idx = esc_idx;
  qa.esc_min = esc_min;
  qa.esc_count = esc_count;
  HD_TRY(LaunchLbvhQuery(ctx, qa));
  // Queries whose traversal stack overflowed get an exhaustive scan.
  void* pinned = nullptr;
  HD_TRY(ContextPinned(ctx, 64, &pinned));
  unsigned int* h = static_cast<unsigned int*>(pinned);
  HD_HIP_TRY(hipMemcpyAsync(h, esc_count, 4, hipMemcpyDeviceToHost,
                            ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return ExactMinForList(ctx, n_dims, q, esc_idx, *h, ref, n_ref, nullptr,
                         out_d2, out_idx);
}

namespace {

template <int ND>
int RunLbvhImpl(hd_context* ctx, const DirectedArgs& in,
                hd_directed_result* out) {
  const uint32_t n_a = (uint32_t)in.n_a;
  const int st_build = StageBegin(ctx, "lbvh_build");
  LbvhIndex tree;
  HD_TRY(BuildLbvh(ctx, ND, in.b, in.n_b, 8, &tree,
                   in.has_mbr ? in.mbr_b : nullptr));
  StageEnd(ctx, st_build);

  unsigned long long* key = nullptr;
  // key, evaluations of seed and query, then their node-record reads.
  HD_TRY(ContextArray(ctx, "lbvh_key", kCounterWords, &key));
  HD_TRY(FillU64(ctx, key, 0ull, kCounterWords));
  unsigned int* esc_count = nullptr;
  uint32_t* esc_idx = nullptr;
  float* esc_min = nullptr;
  HD_TRY(ContextArray(ctx, "lbvh_esc_count", 1, &esc_count));
  HD_TRY(ContextArray(ctx, "lbvh_esc_idx", n_a, &esc_idx));
  HD_TRY(ContextArray(ctx, "lbvh_esc_min", n_a, &esc_min));
  HD_TRY(FillU32(ctx, esc_count, 0u, 1));

  LbvhQueryArgs qa;
  std::memset(&qa, 0, sizeof(qa));
  qa.tree = tree.dev;
  qa.tree_nd = ND;
  qa.a = in.a;
  const bool shuffle = (in.flags & HD_FLAG_NO_SHUFFLE) == 0;
  qa.perm = MakeFeistel(n_a, (uint64_t)(int64_t)in.seed, 0);
  qa.order_mode = shuffle ? 1 : 2;
  if (in.flags & HD_FLAG_LEARNED_ORDER) {
    const uint32_t* order = nullptr;
    HD_TRY(LearnedOrder(ctx, in, &order));
    qa.order = order;
    qa.order_mode = 0;
  }
  qa.key = key;
  qa.esc_idx = esc_idx;
  qa.esc_min = esc_min;
  qa.esc_count = esc_count;

  // Visit the early-break queries in Morton order of A (in B's frame), so
  // lanes of a wave descend the same subtrees: coherent node loads.
  uint32_t* morton_order = nullptr;
  uint32_t* start_leaf = nullptr;
  if (!(in.flags & HD_FLAG_LEARNED_ORDER)) {
    const int st_ord = StageBegin(ctx, "lbvh_order");
    HD_TRY(MortonOrder(ctx, ND, in.a, n_a, tree.frame, &morton_order));
    HD_TRY(ContextArray(ctx, "lbvh_start_leaf", n_a, &start_leaf));
    HD_TRY(StartLeaves(ctx, tree, in.a, morton_order, n_a, start_leaf));
    StageEnd(ctx, st_ord);
  }

  // Seed with exact nearest neighbours of the first points of the order
  // (in Morton order: every (n_a / s0)-th point, a spatially stratified
  // sample whose start leaves give each traversal a tight first bound), then
  // early break for the rest.
  const uint32_t s0 = std::min<uint32_t>(n_a, 65536u);
  const int st_seed = StageBegin(ctx, "lbvh_seed");
  qa.pos_begin = 0;
  qa.pos_end = s0;
  qa.eb = 0;
  qa.evals = key + 1;
  qa.lookups = key + 4;
  if (morton_order != nullptr) {
    LbvhQueryArgs sa = qa;
    sa.order = morton_order;
    sa.order_mode = 0;
    sa.start_leaf = start_leaf;
    sa.stride = n_a / s0;
    HD_TRY(LaunchLbvhQuery(ctx, sa));
  } else {
    HD_TRY(LaunchLbvhQuery(ctx, qa));
  }
  StageEnd(ctx, st_seed);
  const int st_q = StageBegin(ctx, "lbvh_query");
  if (morton_order != nullptr) {
    // All points, spatially ordered, minus the seed's positions (with
    // s0 = n_a the seed took them all).
    qa.order = morton_order;
    qa.order_mode = 0;
    qa.pos_begin = n_a / s0 > 1 ? 0u : s0;
    qa.start_leaf = start_leaf;
    qa.skip_stride = n_a / s0;
    qa.skip_n = s0;
  } else {
    qa.pos_begin = s0;
  }
  qa.pos_end = n_a;
  qa.eb = 1;
  qa.evals = key + 2;
  qa.lookups = key + 5;
  HD_TRY(LaunchLbvhQuery(ctx, qa));
  StageEnd(ctx, st_q);
  // Stack overflows (deep trees) are finished exhaustively.
  void* pinned = nullptr;
  HD_TRY(ContextPinned(ctx, 64, &pinned));
  unsigned int* h = static_cast<unsigned int*>(pinned);
  HD_HIP_TRY(hipMemcpyAsync(h, esc_count, 4, hipMemcpyDeviceToHost,
                            ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const uint32_t n_over = *h;
  HD_TRY(ExactMinForList(ctx, ND, in.a, esc_idx, n_over, in.b, in.n_b, key,
                         nullptr, nullptr));
  const int st_fin = StageBegin(ctx, "lbvh_finish");
  HD_TRY(FinishDirected(ctx, ND, in.a, in.b, in.n_b, key, key + 1, 2, out));
  StageEnd(ctx, st_fin);
  if (ctx->profiling) {
    unsigned long long sums[kCounterRow];
    HD_TRY(ReadCounters(ctx, key, sums));
    const unsigned long long* hh = sums + 1;
    StageCount(ctx, "lbvh_seed", hh[0], s0, hh[3]);
    StageCount(ctx, "lbvh_query", hh[1], n_a - s0, hh[4]);
    StageCount(ctx, "lbvh_overflow", 0, n_over);
  }
  return HD_OK;
}

}  // namespace

int RunLbvh(hd_context* ctx, const DirectedArgs& args,
            hd_directed_result* out) {
  if (args.n_dims == 2) return RunLbvhImpl<2>(ctx, args, out);
  return RunLbvhImpl<3>(ctx, args, out);
}

}  // namespace hd
