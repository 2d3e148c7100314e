// Support kernels shared by every variant: fills, AoS -> SoA repacking with
// an optional pseudo-random permutation, MBR reduction with input
// validation, and the post-pass that names the arg-max's nearest neighbour.
#include <cmath>
#include <cstring>

#include "hd.h"
#include "hd_device.h"
#include "hd_gen.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

using namespace dev;

namespace {

__global__ void FillU32Kernel(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void FillU64Kernel(unsigned long long* p, unsigned long long v,
                              uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void IotaKernel(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    p[i] = i;
}

template <int ND>
__global__ void GatherSoAKernel(const float* __restrict__ aos, uint32_t n,
                                uint32_t n_pad, int shuffle, Feistel perm,
                                const uint32_t* __restrict__ idx,
                                float* __restrict__ x, float* __restrict__ y,
                                float* __restrict__ z) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n_pad;
       j += gridDim.x * blockDim.x) {
    uint32_t src = j < n ? j : j % n;
    if (idx != nullptr)
      src = idx[src];
    else if (shuffle)
      src = Permute(perm, src);
    const float* p = aos + (uint64_t)src * ND;
    x[j] = p[0];
    y[j] = p[1];
    if constexpr (ND == 3) z[j] = p[2];
  }
}

// Order-preserving map of FP32 to uint32 (negative values reversed).
HD_DEV uint32_t OrderedBits(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Per block: min/max per coordinate and a count of non-finite coordinates,
// folded into slot (block mod n_slots) of out, kMbrSlotWords words apart:
// slot[0..ND) = ordered min, slot[ND..2ND) = ordered max, slot[2ND] = bad
// count. Spreading the blocks' atomics over slots (cache lines) keeps them
// from queueing on one address; the host folds the slots.
// The AoS array is read as flat floats, 16 bytes per load when it is 16-byte
// aligned (value k of the array is coordinate k mod ND), four loads in
// flight per thread.
template <int ND>
HD_DEV void MbrTake(float v, int d, uint32_t (&lo)[ND], uint32_t (&hi)[ND],
                    uint32_t& bad) {
  if (!isfinite(v)) {
    bad++;
    return;
  }
  const uint32_t k = OrderedBits(v);
#pragma unroll
  for (int e = 0; e < ND; ++e)
    if (e == d) {
      lo[e] = min(lo[e], k);
      hi[e] = max(hi[e], k);
    }
}

template <int ND>
HD_DEV void MbrTake4(float4 v, uint64_t j, uint32_t (&lo)[ND],
                     uint32_t (&hi)[ND], uint32_t& bad) {
  // first coordinate index of the 4 values: 4 j mod ND
  const int d0 = ND == 2 ? 0 : (int)((4ull * j) % 3ull);
  MbrTake<ND>(v.x, d0, lo, hi, bad);
  MbrTake<ND>(v.y, (d0 + 1) % ND, lo, hi, bad);
  MbrTake<ND>(v.z, (d0 + 2) % ND, lo, hi, bad);
  MbrTake<ND>(v.w, (d0 + 3) % ND, lo, hi, bad);
}

constexpr uint32_t kMbrSlotWords = 32;  // 128 bytes per slot
constexpr uint32_t kMbrSlots = 32;

__global__ void MbrInitKernel(uint32_t* out, int n_dims, uint32_t n_slots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_slots * kMbrSlotWords) return;
  const uint32_t w = i % kMbrSlotWords;
  out[i] = w < (uint32_t)n_dims ? 0xFFFFFFFFu : 0u;
}

template <int ND>
__global__ __launch_bounds__(256) void MbrKernel(const float* __restrict__ aos,
                                                 uint64_t n, int aligned,
                                                 uint32_t n_slots,
                                                 uint32_t* __restrict__ out_all) {
  uint32_t* out = out_all + (blockIdx.x % n_slots) * kMbrSlotWords;
  uint32_t lo[ND], hi[ND];
  uint32_t bad = 0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    lo[d] = 0xFFFFFFFFu;
    hi[d] = 0u;
  }
  const uint64_t total = n * ND;
  const uint64_t n4 = aligned ? total / 4 : 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const float4* a4 = reinterpret_cast<const float4*>(aos);
  uint64_t j = gid;
  for (; j + 3 * stride < n4; j += 4 * stride) {
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = a4[j + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) MbrTake4<ND>(v[u], j + u * stride, lo, hi, bad);
  }
  for (; j < n4; j += stride) MbrTake4<ND>(a4[j], j, lo, hi, bad);
  for (uint64_t k = 4 * n4 + gid; k < total; k += stride)
    MbrTake<ND>(aos[k], (int)(k % ND), lo, hi, bad);
  __shared__ uint32_t s_lo[4][ND], s_hi[4][ND], s_bad[4];
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    lo[d] = WaveMinU32(lo[d]);
    uint32_t h = hi[d];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      h = max(h, (uint32_t)__shfl_xor((int)h, off, 64));
    hi[d] = h;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor((int)bad, off, 64);
  if (LaneId() == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      s_lo[wave][d] = lo[d];
      s_hi[wave][d] = hi[d];
    }
    s_bad[wave] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int w = 1; w < nw; ++w) {
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        lo[d] = min(lo[d], s_lo[w][d]);
        hi[d] = max(hi[d], s_hi[w][d]);
      }
      bad += s_bad[w];
    }
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      atomicMin(&out[d], lo[d]);
      atomicMax(&out[ND + d], hi[d]);
    }
    if (bad) atomicAdd(&out[2 * ND], bad);
  }
}

float FromOrderedBits(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

template <int ND>
__global__ void NnIndexKernel(const float* __restrict__ a, uint32_t argmax,
                              const float* __restrict__ b, uint64_t n_b,
                              uint32_t target_bits, uint32_t* out) {
  float pa[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) pa[d] = a[(uint64_t)argmax * ND + d];
  uint32_t best = 0xFFFFFFFFu;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n_b;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const float* pb = b + j * ND;
    float d2 = ND == 2 ? D2(pa[0], pa[1], pb[0], pb[1])
                       : D2(pa[0], pa[1], pa[2], pb[0], pb[1], pb[2]);
    if (__float_as_uint(d2) == target_bits) {
      best = (uint32_t)j;
      break;  // j only grows along this thread's stride
    }
  }
  best = WaveMinU32(best);
  if (LaneId() == 0 && best != 0xFFFFFFFFu) atomicMin(out, best);
}

// out[k] = point floor(k * n / m) of pts, k < m (AoS, nd floats each).
__global__ void SamplePointsKernel(const float* __restrict__ pts, uint64_t n,
                                   int nd, uint32_t m, float* __restrict__ out) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m;
       k += gridDim.x * blockDim.x) {
    const uint64_t src = (uint64_t)k * n / m;
    for (int d = 0; d < nd; ++d) out[(uint64_t)k * nd + d] = pts[src * nd + d];
  }
}

inline int GridFor(uint64_t n, int threads, int cap) {
  uint64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > (uint64_t)cap) b = cap;
  return (int)b;
}

}  // namespace

Feistel MakeFeistel(uint32_t n, uint64_t seed, uint32_t stream_id) {
  Feistel f;
  f.n = n ? n : 1;
  uint32_t bits = 1;
  while (bits < 32 && (1ull << bits) < f.n) ++bits;
  f.half = (bits + 1) / 2;
  if (f.half < 1) f.half = 1;
  f.mask = (1u << f.half) - 1u;
  for (int i = 0; i < 4; ++i)
    f.key[i] = (uint32_t)hd_splitmix64_at(seed, stream_id * 4u + i);
  return f;
}

uint32_t HostPermute(const Feistel& f, uint32_t i) { return dev::Permute(f, i); }

int SamplePoints(hd_context* ctx, const float* pts, uint64_t n, int n_dims,
                 uint32_t m, float* d_out) {
  if (n == 0 || m == 0) return HD_OK;
  SamplePointsKernel<<<(m + 255) / 256, 256, 0, ctx->stream>>>(pts, n, n_dims,
                                                               m, d_out);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int FillU32(hd_context* ctx, uint32_t* p, uint32_t value, uint64_t n) {
  if (n == 0) return HD_OK;
  FillU32Kernel<<<GridFor(n, 256, 4096), 256, 0, ctx->stream>>>(p, value, n);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int IotaU32(hd_context* ctx, uint32_t* p, uint32_t n) {
  if (n == 0) return HD_OK;
  IotaKernel<<<GridFor(n, 256, 4096), 256, 0, ctx->stream>>>(p, n);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int FillU64(hd_context* ctx, unsigned long long* p, unsigned long long value,
            uint64_t n) {
  if (n == 0) return HD_OK;
  FillU64Kernel<<<GridFor(n, 256, 4096), 256, 0, ctx->stream>>>(p, value, n);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

uint32_t GoldenStep(uint32_t n) {
  if (n <= 2) return 1u;
  uint64_t step = (uint64_t)std::llround((double)n * 0.6180339887498949);
  auto gcd = [](uint64_t x, uint64_t y) {
    while (y) {
      const uint64_t t = x % y;
      x = y;
      y = t;
    }
    return x;
  };
  while (gcd(step, n) != 1) ++step;  // ends below n: gcd(n - 1, n) = 1
  return (uint32_t)step;
}

int GatherSoA(hd_context* ctx, int n_dims, const float* aos, uint32_t n,
              uint32_t n_pad, bool shuffle, const Feistel& perm, float* x,
              float* y, float* z, const uint32_t* idx) {
  if (n == 0 || n_pad < n) return Fail(HD_EINVAL, "GatherSoA: bad sizes");
  const int grid = GridFor(n_pad, 256, 8192);
  if (n_dims == 2)
    GatherSoAKernel<2><<<grid, 256, 0, ctx->stream>>>(aos, n, n_pad, shuffle,
                                                      perm, idx, x, y, z);
  else
    GatherSoAKernel<3><<<grid, 256, 0, ctx->stream>>>(aos, n, n_pad, shuffle,
                                                      perm, idx, x, y, z);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

namespace {

// Queues the MBR kernel of one set into d_out (n_slots slots of
// kMbrSlotWords words).
int LaunchMbr(hd_context* ctx, int n_dims, const float* aos, uint64_t n,
              uint32_t* d_out, uint32_t n_slots) {
  MbrInitKernel<<<(n_slots * kMbrSlotWords + 255) / 256, 256, 0, ctx->stream>>>(
      d_out, n_dims, n_slots);
  const int grid = GridFor(n * n_dims / 4 + 1, 256, ctx->cu_count * 8);
  const int aligned = (reinterpret_cast<uintptr_t>(aos) & 15u) == 0;
  if (n_dims == 2)
    MbrKernel<2><<<grid, 256, 0, ctx->stream>>>(aos, n, aligned, n_slots, d_out);
  else
    MbrKernel<3><<<grid, 256, 0, ctx->stream>>>(aos, n, aligned, n_slots, d_out);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

// Host MBR from the slots LaunchMbr wrote; the reference's ToNonemptyMBR
// widening (src/geoms/mbr.h:288-296).
int DecodeMbr(int n_dims, const uint32_t* h, uint32_t n_slots, float* mbr) {
  uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0, 0, 0};
  uint64_t bad = 0;
  for (uint32_t k = 0; k < n_slots; ++k) {
    const uint32_t* w = h + k * kMbrSlotWords;
    for (int d = 0; d < n_dims; ++d) {
      lo[d] = std::min(lo[d], w[d]);
      hi[d] = std::max(hi[d], w[n_dims + d]);
    }
    bad += w[2 * n_dims];
  }
  if (bad != 0)
    return Fail(HD_EINVAL, "%llu non-finite coordinates in input",
                (unsigned long long)bad);
  for (int d = 0; d < n_dims; ++d) {
    mbr[d] = FromOrderedBits(lo[d]);
    mbr[n_dims + d] = FromOrderedBits(hi[d]);
  }
  for (int d = 0; d < n_dims; ++d)
    if (mbr[n_dims + d] - mbr[d] == 0.0f)
      mbr[n_dims + d] = (float)((double)mbr[d] + 1e-5);  // mbr.h:288-296
  return HD_OK;
}

// Slots per set: one in a caller-provided workspace (kMbrSlotWords words per
// set, the allocation-free path), kMbrSlots in the context's buffer.
uint32_t MbrSlots(const uint32_t* d_work) { return d_work ? 1u : kMbrSlots; }

}  // namespace

int HostMbr(hd_context* ctx, int n_dims, const float* aos, uint64_t n,
            float* mbr, uint32_t* d_work) {
  if (n == 0) return Fail(HD_EINVAL, "empty point set");
  // A caller-provided workspace (the allocation-free path) holds one slot.
  const uint32_t slots = MbrSlots(d_work);
  uint32_t* d_out = d_work;
  if (d_out == nullptr)
    HD_TRY(ContextArray(ctx, "mbr_keys", 2 * kMbrSlots * kMbrSlotWords, &d_out));
  HD_TRY(LaunchMbr(ctx, n_dims, aos, n, d_out, slots));
  // A host array on the stack: this path allocates nothing (it also serves
  // the allocation-free entry point).
  uint32_t h[kMbrSlots * kMbrSlotWords];
  HD_HIP_TRY(hipMemcpyAsync(h, d_out, sizeof(uint32_t) * slots * kMbrSlotWords,
                            hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return DecodeMbr(n_dims, h, slots, mbr);
}

int HostMbrPair(hd_context* ctx, int n_dims, const float* a, uint64_t n_a,
                float* mbr_a, const float* b, uint64_t n_b, float* mbr_b,
                uint32_t* d_work) {
  if (n_a == 0 || n_b == 0) return Fail(HD_EINVAL, "empty point set");
  const uint32_t slots = MbrSlots(d_work);
  const uint32_t words = slots * kMbrSlotWords;
  uint32_t* d_out = d_work;
  if (d_out == nullptr)
    HD_TRY(ContextArray(ctx, "mbr_keys", 2 * kMbrSlots * kMbrSlotWords, &d_out));
  HD_TRY(LaunchMbr(ctx, n_dims, a, n_a, d_out, slots));
  HD_TRY(LaunchMbr(ctx, n_dims, b, n_b, d_out + words, slots));
  uint32_t h[2 * kMbrSlots * kMbrSlotWords];
  HD_HIP_TRY(hipMemcpyAsync(h, d_out, sizeof(uint32_t) * 2 * words,
                            hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  HD_TRY(DecodeMbr(n_dims, h, slots, mbr_a));
  return DecodeMbr(n_dims, h + words, slots, mbr_b);
}

int LaunchNnIndex(hd_context* ctx, int n_dims, const float* a,
                  uint32_t argmax, const float* b, uint64_t n_b,
                  uint32_t cmax2_bits, uint32_t* d_out) {
  HD_TRY(FillU32(ctx, d_out, 0xFFFFFFFFu, 1));
  const int grid = GridFor(n_b, 256, ctx->cu_count * 8);
  if (n_dims == 2)
    NnIndexKernel<2><<<grid, 256, 0, ctx->stream>>>(a, argmax, b, n_b,
                                                    cmax2_bits, d_out);
  else
    NnIndexKernel<3><<<grid, 256, 0, ctx->stream>>>(a, argmax, b, n_b,
                                                    cmax2_bits, d_out);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int ReadCounters(hd_context* ctx, const unsigned long long* d_block,
                 unsigned long long* sums) {
  // Through the context's pinned buffer (a request within kPinnedMinBytes
  // never replaces it under a caller): a pageable destination costs a staged
  // copy.
  static_assert(kCounterWords * sizeof(unsigned long long) <= kPinnedMinBytes,
                "counter block fits the pinned minimum");
  void* pinned = nullptr;
  HD_TRY(ContextPinned(ctx, kCounterWords * sizeof(unsigned long long), &pinned));
  const unsigned long long* h = static_cast<const unsigned long long*>(pinned);
  HD_HIP_TRY(hipMemcpyAsync(pinned, d_block, kCounterWords * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (uint32_t c = 0; c < kCounterRow; ++c) {
    sums[c] = 0;
    for (uint32_t r = 0; r < kCounterRows; ++r) sums[c] += h[r * kCounterRow + c];
  }
  return HD_OK;
}

int FinishDirected(hd_context* ctx, int n_dims, const float* a,
                   const float* b, uint64_t n_b,
                   const unsigned long long* key_dev,
                   const unsigned long long* evals_dev, int n_counters,
                   hd_directed_result* out) {
  const ptrdiff_t col = evals_dev ? evals_dev - key_dev : 1;
  if (n_counters < 0 || col < 1 || col + n_counters > (ptrdiff_t)kCounterRow)
    return Fail(HD_EINVAL, "counters outside the key's counter block");
  unsigned long long h[kCounterRow];
  HD_TRY(ReadCounters(ctx, key_dev, h));
  out->evals = 0;
  if (evals_dev)
    for (int i = 0; i < n_counters; ++i) out->evals += h[col + i];
  const unsigned long long key = h[0];
  out->cmax2_bits = (uint32_t)(key >> 32);
  float c;
  std::memcpy(&c, &out->cmax2_bits, 4);
  out->hd = std::sqrt(c);
  if (key == 0) {
    out->argmax = 0xFFFFFFFFu;
    out->nn = 0xFFFFFFFFu;
    return HD_OK;
  }
  out->argmax = 0xFFFFFFFFu - (uint32_t)key;
  uint32_t* d_nn = nullptr;
  HD_TRY(ContextArray(ctx, "nn_index", 1, &d_nn));
  HD_TRY(LaunchNnIndex(ctx, n_dims, a, out->argmax, b, n_b, out->cmax2_bits,
                       d_nn));
  void* pinned = nullptr;
  HD_TRY(ContextPinned(ctx, 64, &pinned));
  uint32_t* hn = static_cast<uint32_t*>(pinned);
  HD_HIP_TRY(hipMemcpyAsync(hn, d_nn, 4, hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  out->nn = *hn;
  return HD_OK;
}

}  // namespace hd
