// Device helpers shared by the HIP translation units (gfx950, wave64).
#ifndef HD_DEVICE_H
#define HD_DEVICE_H

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "hd_internal.h"

// The squared distance must round exactly like the reference's unfused FP32
// expression; contraction into FMA is disabled for every TU including this
// header (the build also passes -ffp-contract=off).
#pragma clang fp contract(off)

#define HD_DEV __device__ __forceinline__
// Also callable from host code (the CPU early break visits A and B in the
// same pseudo-random order as the device pipelines).
#define HD_HOST_DEV __host__ __device__ __forceinline__

namespace hd {
namespace dev {

constexpr float kFltMax = FLT_MAX;
constexpr uint32_t kFltMaxBits = 0x7F7FFFFFu;

// Reference src/geoms/distance.h:17-24: left to right, unfused.
HD_DEV float D2(float ax, float ay, float bx, float by) {
  return (ax - bx) * (ax - bx) + (ay - by) * (ay - by);
}
HD_DEV float D2(float ax, float ay, float az, float bx, float by, float bz) {
  return (ax - bx) * (ax - bx) + (ay - by) * (ay - by) + (az - bz) * (az - bz);
}
template <int ND>
HD_DEV float D2(const float* a, float bx, float by, float bz) {
  if constexpr (ND == 2) {
    return D2(a[0], a[1], bx, by);
  } else {
    return D2(a[0], a[1], a[2], bx, by, bz);
  }
}

// d2 is +0 or positive (never -0; never NaN for finite inputs), so its FP32
// bit pattern orders like the value. Running minima are kept as bits and
// reduced with an unsigned min: one instruction, where fminf also has to
// quiet signalling NaNs first.
HD_DEV uint32_t MinBits(uint32_t a, uint32_t b) { return a < b ? a : b; }

// 64-bit key whose unsigned order is (d2 ascending, index descending); the
// maximum key is the largest d2 at the smallest index. d2 >= +0, so the FP32
// bit pattern orders like the value.
HD_DEV unsigned long long PackKey(float d2, uint32_t idx) {
  return (static_cast<unsigned long long>(__float_as_uint(d2)) << 32) |
         static_cast<unsigned long long>(0xFFFFFFFFu - idx);
}

// Current cmax2: the high word of the key (little-endian: word 1).
HD_DEV float LoadCmax2(const unsigned long long* key) {
  const uint32_t* hi = reinterpret_cast<const uint32_t*>(key) + 1;
  uint32_t bits = __hip_atomic_load(hi, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
  return __uint_as_float(__builtin_amdgcn_readfirstlane(bits));
}

HD_DEV uint32_t LaneId() { return threadIdx.x & 63u; }

// Raises the packed key to at least v. The key only grows, so whatever a
// load returns is a lower bound of its final value: a candidate that does
// not exceed it cannot change the result, and its atomic (which would queue
// behind every other wave's on the one word) is skipped.
HD_DEV void KeyMax(unsigned long long* key, unsigned long long v) {
  if (v > __hip_atomic_load(key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(key, v);
}

// Event counters (evaluations, lookups) live in blocks of kCounterRows rows
// of kCounterRow 64-bit words: a kernel adds its totals into row
// (blockIdx.x mod kCounterRows) of its counter's column, so the final
// atomics of thousands of waves spread over as many cache lines instead of
// queueing on one address (a device-scope atomic on one address is served
// one at a time: 16k of them cost the 10M grid query 0.6 ms). The host sums
// the rows (ReadCounters, FinishDirected). A counter block is allocated and
// cleared whole; its counters are columns of row 0, and column 0 of row 0
// may hold the packed result key, which only atomicMax touches.
constexpr uint32_t kCounterRow = 16;
constexpr uint32_t kCounterRows = 32;
constexpr uint32_t kCounterWords = kCounterRow * kCounterRows;

HD_DEV void CounterAdd(unsigned long long* c, unsigned long long v) {
  atomicAdd(c + kCounterRow * (blockIdx.x % kCounterRows), v);
}

HD_DEV uint32_t Uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
HD_DEV float Uniform(float v) {
  return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}

// Minimum over the 64 lanes of a wave (all lanes active).
HD_DEV float WaveMin(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
HD_DEV uint32_t WaveMinU32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
  return v;
}
HD_DEV unsigned long long WaveMaxU64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// True on any lane of the wave.
HD_DEV bool WaveAny(bool c) { return __ballot(c) != 0ull; }

// Feistel permutation of [0, n) (struct Feistel in hd_internal.h): a 4-round
// Feistel network on 2*half bits (domain < 4n) with cycle walking back into
// range.
HD_HOST_DEV uint32_t Hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

HD_HOST_DEV uint32_t FeistelOnce(const Feistel& f, uint32_t x) {
  uint32_t l = x >> f.half, r = x & f.mask;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t nl = r;
    r = l ^ (Hash32(r ^ f.key[i]) & f.mask);
    l = nl;
  }
  return (l << f.half) | r;
}

HD_HOST_DEV uint32_t Permute(const Feistel& f, uint32_t i) {
  uint32_t y = FeistelOnce(f, i);
  while (y >= f.n) y = FeistelOnce(f, y);
  return y;
}

}  // namespace dev
}  // namespace hd

#endif  // HD_DEVICE_H
