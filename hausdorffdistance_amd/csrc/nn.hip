// Exact nearest neighbour of every query (hd_nn_device): the per-point
// minimum the reference's EB loop computes when it does not break
// (src/hd_impl/hausdorff_distance_early_break.h:69-76), with the index of the
// smallest reference point attaining it.
#include <algorithm>

#include "hd_device.h"
#include "hd_grid.h"
#include "hd_lbvh.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

using namespace dev;

namespace {

constexpr int kTile = 1024;

// One query per thread; reference points staged through LDS in index order,
// so a strict '<' keeps the smallest index among ties.
template <int ND>
__global__ __launch_bounds__(256) void BruteNnKernel(
    const float* __restrict__ q, uint64_t n_q, const float* __restrict__ r,
    uint64_t n_r, float* __restrict__ out_d2, uint32_t* __restrict__ out_idx) {
  __shared__ float tile[kTile * ND];
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  float pq[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) pq[d] = i < n_q ? q[i * ND + d] : 0.f;
  float best = kFltMax;
  uint32_t bi = 0xFFFFFFFFu;
  for (uint64_t base = 0; base < n_r; base += kTile) {
    const uint32_t cnt = (uint32_t)min((uint64_t)kTile, n_r - base);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < cnt * ND; k += blockDim.x)
      tile[k] = r[base * ND + k];
    __syncthreads();
    for (uint32_t k = 0; k < cnt; ++k) {
      const float* pr = &tile[k * ND];
      float d = ND == 2 ? D2(pq[0], pq[1], pr[0], pr[1])
                        : D2(pq[0], pq[1], pq[2], pr[0], pr[1], pr[2]);
      if (d < best) {
        best = d;
        bi = (uint32_t)(base + k);
      }
    }
  }
  if (i < n_q) {
    out_d2[i] = best;
    out_idx[i] = bi;
  }
}

}  // namespace

int BruteNearestNeighbors(hd_context* ctx, int n_dims, const float* q,
                          uint64_t n_q, const float* r, uint64_t n_r,
                          float* out_d2, uint32_t* out_idx) {
  const uint64_t blocks = (n_q + 255) / 256;
  if (n_dims == 2)
    BruteNnKernel<2><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
        q, n_q, r, n_r, out_d2, out_idx);
  else
    BruteNnKernel<3><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
        q, n_q, r, n_r, out_d2, out_idx);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int NearestNeighbors(hd_context* ctx, int index, int n_dims, const float* q,
                     uint64_t n_q, const float* ref, uint64_t n_ref,
                     float* out_d2, uint32_t* out_idx) {
  if (index == HD_VARIANT_EB)
    return BruteNearestNeighbors(ctx, n_dims, q, n_q, ref, n_ref, out_d2,
                                 out_idx);
  if (index == HD_VARIANT_GRID)
    return GridNearestNeighbors(ctx, n_dims, q, n_q, ref, n_ref, out_d2,
                                out_idx);
  if (index == HD_VARIANT_LBVH)
    return LbvhNearestNeighbors(ctx, n_dims, q, n_q, ref, n_ref, out_d2,
                                out_idx);
  return Fail(HD_EUNSUPPORTED, "index %d not built yet", index);
}

namespace {

// Packed maximum (d2 bits << 32 | ~index) over the per-point NN distances:
// the largest d2, and among equal d2 the smallest index. One atomic per
// block (blockDim 256).
__global__ __launch_bounds__(256) void MaxOfMinKernel(
    const float* __restrict__ d2, uint64_t n, unsigned long long* key) {
  __shared__ unsigned long long part[4];
  unsigned long long best = 0ull;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const float v = d2[i];
    if (v != kFltMax) best = max(best, PackKey(v, (uint32_t)i));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(best, off, 64);
    best = max(best, o);
  }
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    best = max(max(part[0], part[1]), max(part[2], part[3]));
    if (best != 0ull) dev::KeyMax(key, best);
  }
}

}  // namespace

// Reference hausdorff_distance_nearest_neighbor_search.h:107-153 (GPU):
// a kd-tree over B, the closest point of every A point, the maximum. Here
// the uniform grid's exact NN query (21 ms on the 10M clustered pair, the
// LBVH's 55 ms), then a packed-key maximum; the result bits and the arg-max
// are those of every other variant.
int RunNearestNeighborSearch(hd_context* ctx, const DirectedArgs& args,
                             hd_directed_result* out) {
  float* d2 = nullptr;
  uint32_t* idx = nullptr;
  unsigned long long* key = nullptr;
  HD_TRY(ContextArray(ctx, "nns_d2", args.n_a, &d2));
  HD_TRY(ContextArray(ctx, "nns_idx", args.n_a, &idx));
  HD_TRY(ContextArray(ctx, "nns_key", kCounterWords, &key));
  HD_TRY(FillU64(ctx, key, 0ull, kCounterWords));
  HD_TRY(FillU64(ctx, key + 1, (unsigned long long)args.n_a, 1));
  HD_TRY(GridNearestNeighbors(ctx, args.n_dims, args.a, args.n_a, args.b,
                              args.n_b, d2, idx));
  const int grid = (int)std::min<uint64_t>((args.n_a + 255) / 256,
                                           (uint64_t)ctx->cu_count * 2);
  MaxOfMinKernel<<<grid, 256, 0, ctx->stream>>>(d2, args.n_a, key);
  HD_HIP_TRY(hipGetLastError());
  HD_TRY(FinishDirected(ctx, args.n_dims, args.a, args.b, args.n_b, key,
                        key + 1, 1, out));
  out->variant_used = HD_VARIANT_NN;
  return HD_OK;
}

}  // namespace hd

namespace hd {
namespace {

// One wave per listed query: lanes stride over all of ref, then a wave
// reduction keeps the smallest d2 and, among equal d2, the smallest index.
template <int ND>
__global__ __launch_bounds__(256) void ExactMinListKernel(
    const float* __restrict__ q, const uint32_t* __restrict__ list,
    uint32_t n_list, const float* __restrict__ r, uint64_t n_r,
    unsigned long long* key, float* out_d2, uint32_t* out_idx) {
  const uint32_t lane = dev::LaneId();
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  for (uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) / 64; w < n_list;
       w += waves) {
    const uint32_t idx = list[w];
    float pq[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) pq[d] = q[(uint64_t)idx * ND + d];
    float best = dev::kFltMax;
    uint32_t bj = 0xFFFFFFFFu;
    for (uint64_t j = lane; j < n_r; j += 64) {
      const float* pr = r + j * ND;
      const float d = ND == 2 ? dev::D2(pq[0], pq[1], pr[0], pr[1])
                              : dev::D2(pq[0], pq[1], pq[2], pr[0], pr[1], pr[2]);
      if (d < best) {
        best = d;
        bj = (uint32_t)j;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const float ob = __shfl_xor(best, off, 64);
      const uint32_t oj = (uint32_t)__shfl_xor((int)bj, off, 64);
      if (ob < best || (ob == best && oj < bj)) {
        best = ob;
        bj = oj;
      }
    }
    if (lane == 0) {
      if (out_d2 != nullptr) {
        out_d2[idx] = best;
        out_idx[idx] = bj;
      }
      if (key != nullptr && best != dev::kFltMax)
        dev::KeyMax(key, dev::PackKey(best, idx));
    }
  }
}

}  // namespace

int ExactMinForList(hd_context* ctx, int n_dims, const float* q,
                    const uint32_t* list, uint32_t n_list, const float* ref,
                    uint64_t n_ref, unsigned long long* key, float* out_d2,
                    uint32_t* out_idx) {
  if (n_list == 0) return HD_OK;
  const uint32_t blocks = std::min<uint32_t>((n_list + 3) / 4, 4096u);
  if (n_dims == 2)
    ExactMinListKernel<2><<<blocks, 256, 0, ctx->stream>>>(
        q, list, n_list, ref, n_ref, key, out_d2, out_idx);
  else
    ExactMinListKernel<3><<<blocks, 256, 0, ctx->stream>>>(
        q, list, n_list, ref, n_ref, key, out_d2, out_idx);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

}  // namespace hd
