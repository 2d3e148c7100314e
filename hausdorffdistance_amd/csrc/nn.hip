// Exact nearest neighbour of every query (hd_nn_device): the per-point
// minimum the reference's EB loop computes when it does not break
// (src/hd_impl/hausdorff_distance_early_break.h:69-76), with the index of the
// smallest reference point attaining it.
#include "hd_device.h"
#include "hd_grid.h"
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
  return Fail(HD_EUNSUPPORTED, "index %d not built yet", index);
}

}  // namespace hd
