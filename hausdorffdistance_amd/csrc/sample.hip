// The hybrid's sampled early-break pass (reference
// src/hd_impl/hausdorff_distance_hybrid.h:120-199, kernel :519-596): a
// sample of A is scanned against a shuffled B with early break, and the
// 80th percentile of the per-point iteration counts decides whether the
// whole input runs the early-break scan only (p80 < eb_only_threshold).
//
// The reference's counts come from racing blocks that break against a
// cmax2 other blocks are still raising; here they are fixed by the value
// the pass ends with: cmax2 = max(HD lower bound^2 from the MBRs, exact NN
// d2 of every sampled point). A sampled point's count is then the
// reference's unit: iterations of 256 B points, checked every 32
// iterations (with <=), i.e. 32k + 1 for the first check k whose prefix of
// B holds a point within cmax2, or ceil(|B| / 256) when no check breaks.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hd_device.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

using namespace dev;

namespace {

constexpr uint32_t kRefBlock = 256;   // MAX_BLOCK_SIZE, src/utils/util.h:43
constexpr uint32_t kRefCheck = 32;    // iterations between break checks

template <int ND>
__global__ void SampleGatherKernel(const float* __restrict__ a, Feistel perm,
                                   uint32_t n_s, float* __restrict__ out) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n_s;
       s += gridDim.x * blockDim.x) {
    const uint32_t i = Permute(perm, s);
#pragma unroll
    for (int d = 0; d < ND; ++d) out[(uint64_t)s * ND + d] = a[(uint64_t)i * ND + d];
  }
}

// One sampled point per wave: the first position of the shuffled B whose
// d2 is <= cmax2, turned into the reference's iteration count.
template <int ND>
__global__ __launch_bounds__(256) void SampleItersKernel(
    const float* __restrict__ q, uint32_t n_s, const float* __restrict__ bx,
    const float* __restrict__ by, const float* __restrict__ bz, uint32_t n_b,
    const float* __restrict__ cmax2_ptr, uint32_t* __restrict__ counts) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  const float cmax2 = *cmax2_ptr;
  const uint32_t full = (n_b + kRefBlock - 1) / kRefBlock;
  for (uint32_t s = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < n_s; s += n_waves) {
    float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < ND; ++d) p[d] = q[(uint64_t)s * ND + d];
    uint32_t pos = n_b;
    for (uint32_t base = 0; base < n_b; base += 64u) {
      const uint32_t j = base + lane;
      bool hit = false;
      if (j < n_b) {
        const float d2 = ND == 2 ? D2(p[0], p[1], bx[j], by[j])
                                 : D2(p[0], p[1], p[2], bx[j], by[j], bz[j]);
        hit = d2 <= cmax2;
      }
      const unsigned long long m = __ballot(hit);
      if (m != 0ull) {
        pos = base + (uint32_t)__builtin_ctzll(m);
        break;
      }
    }
    if (lane == 0) {
      uint32_t count = full;
      if (pos < n_b) {
        // first check k (at iteration 32k) whose prefix (32k + 1) * 256
        // covers pos
        const uint32_t k = pos < kRefBlock ? 0u
                                           : (pos + 1u - kRefBlock + kRefCheck * kRefBlock - 1u) /
                                                 (kRefCheck * kRefBlock);
        if ((uint64_t)kRefCheck * k < full) count = kRefCheck * k + 1u;
      }
      counts[s] = count;
    }
  }
}

__global__ void MaxD2Kernel(const float* __restrict__ d2, uint32_t n, float lb2,
                            float* out) {
  // one block: the maximum of lb2 and every sampled point's NN d2
  __shared__ float part[4];
  float m = lb2;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    if (d2[i] != kFltMax) m = fmaxf(m, d2[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
}

}  // namespace

int SampledEbIterations(hd_context* ctx, const DirectedArgs& args,
                        std::vector<uint32_t>* counts_out) {
  const int nd = args.n_dims;
  const uint32_t n_a = (uint32_t)args.n_a, n_b = (uint32_t)args.n_b;
  const float rate = args.sample_rate > 0.f ? args.sample_rate : 0.01f;
  const uint32_t n_s = (uint32_t)std::min<double>(
      n_a, std::max(1.0, std::ceil((double)n_a * (double)rate)));
  // B in the pseudo-random order: the iteration counts are compared with
  // the reference's threshold, which its shuffled order calibrates.
  EbTarget t;
  HD_TRY(EbPrepareTarget(ctx, nd, args.b, n_b, HD_FLAG_B_ORDER_RANDOM,
                         args.seed, nullptr, &t));
  float *q = nullptr, *d2 = nullptr, *cm = nullptr;
  uint32_t *nn = nullptr, *counts = nullptr;
  HD_TRY(ContextArray(ctx, "smp_q", (uint64_t)n_s * nd, &q));
  HD_TRY(ContextArray(ctx, "smp_d2", n_s, &d2));
  HD_TRY(ContextArray(ctx, "smp_nn", n_s, &nn));
  HD_TRY(ContextArray(ctx, "smp_counts", n_s, &counts));
  HD_TRY(ContextArray(ctx, "smp_cmax2", 1, &cm));
  // Sampled A points: the first n_s positions of A's pseudo-random order
  // (the reference draws them with curand, sampler.h:26-62).
  const Feistel perm = MakeFeistel(n_a, (uint64_t)(int64_t)args.seed, 2);
  const int gb = (int)std::min<uint64_t>((n_s + 255) / 256, 1024);
  if (nd == 2)
    SampleGatherKernel<2><<<gb, 256, 0, ctx->stream>>>(args.a, perm, n_s, q);
  else
    SampleGatherKernel<3><<<gb, 256, 0, ctx->stream>>>(args.a, perm, n_s, q);
  HD_HIP_TRY(hipGetLastError());
  HD_TRY(NearestNeighbors(ctx, HD_VARIANT_GRID, nd, q, n_s, args.b, n_b, d2, nn));
  // Seed of the pass: the MBR lower bound (hybrid.h:100-111).
  float mbr_a[6], mbr_b[6];
  if (args.has_mbr) {
    std::memcpy(mbr_a, args.mbr_a, sizeof(mbr_a));
    std::memcpy(mbr_b, args.mbr_b, sizeof(mbr_b));
  } else {
    HD_TRY(HostMbrPair(ctx, nd, args.a, n_a, mbr_a, args.b, n_b, mbr_b));
  }
  float lo_a[3], hi_a[3], lo_b[3], hi_b[3], lb2 = 0.f, ub2 = 0.f;
  NonEmptyMbr(nd, mbr_a, lo_a, hi_a);
  NonEmptyMbr(nd, mbr_b, lo_b, hi_b);
  HdBounds2(nd, lo_a, hi_a, lo_b, hi_b, &lb2, &ub2);
  MaxD2Kernel<<<1, 256, 0, ctx->stream>>>(d2, n_s, lb2, cm);
  HD_HIP_TRY(hipGetLastError());
  const int sb = (int)std::min<uint64_t>((n_s + 3) / 4, (uint64_t)ctx->cu_count * 8);
  if (nd == 2)
    SampleItersKernel<2><<<sb, 256, 0, ctx->stream>>>(q, n_s, t.bx, t.by, t.bz, n_b, cm,
                                                      counts);
  else
    SampleItersKernel<3><<<sb, 256, 0, ctx->stream>>>(q, n_s, t.bx, t.by, t.bz, n_b, cm,
                                                      counts);
  HD_HIP_TRY(hipGetLastError());
  counts_out->resize(n_s);
  HD_HIP_TRY(hipMemcpyAsync(counts_out->data(), counts, n_s * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return HD_OK;
}

}  // namespace hd

