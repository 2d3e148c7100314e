// Host early break: execution = CPU, and the independent answer of the
// -check option.
//
// The reference's CPU path (src/hd_impl/hausdorff_distance_early_break.h:
// 44-115): visit A and B in a pseudo-random order, split A statically over
// `parallelism` threads (ceil(|A| / T) points each, :58-65), scan B for
// every A point with a running minimum that starts at FLT_MAX, stop as soon
// as the minimum is strictly below the shared cmax2 (:70-79), and raise
// cmax2 with a compare-and-swap maximum (:80-82, :229-234). Its time
// (ReportedTime) excludes the shuffle (:101-112).
//
// Written for the host CPU rather than restated:
//   * inputs stay const: the visit order is the library's Feistel
//     permutation (the same one the device pipelines use), applied while A
//     and B are gathered into private structure-of-arrays copies; B is padded
//     to whole blocks with copies of its first visited point (a duplicate
//     never changes a minimum);
//   * d2 is evaluated eight (AVX2) or four (SSE2) B points per instruction,
//     with the reference's unfused formula (this file is compiled with
//     -ffp-contract=off and no FMA intrinsic is used), and the break test
//     runs once per block of 32 B points instead of after every point. A
//     point that breaks has a minimum below an attained cmax2 either way
//     and a point that does not break gets its exact minimum, so cmax2 is
//     the same; only the evaluation count differs (it is counted per
//     visited block, up to |B|);
//   * cmax2 lives in the high word of a 64-bit key whose low word is
//     ~index, so the maximum also names the smallest index attaining it
//     (the arg-max), as on the device.
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "hd.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

namespace {

constexpr uint32_t kBlock = 32;  // B points per break test

struct HostSoA {
  std::vector<float> x, y, z;
  std::vector<uint32_t> id;  // source index of each position (A only)
  uint32_t n = 0;
  uint32_t n_pad = 0;
};

double NowMs() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch())
      .count();
}

// Runs f(t, begin, end) on `threads` threads over [0, n) split into
// contiguous chunks of ceil(n / threads).
template <typename F>
void ParallelChunks(uint64_t n, int threads, F f) {
  const uint64_t avg = (n + threads - 1) / threads;
  if (threads == 1) {
    f(0, uint64_t{0}, n);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (int t = 0; t < threads; ++t) {
    const uint64_t begin = std::min<uint64_t>(n, t * avg);
    const uint64_t end = std::min<uint64_t>(n, begin + avg);
    pool.emplace_back([=, &f] { f(t, begin, end); });
  }
  for (auto& th : pool) th.join();
}

// Gathers `n` AoS points into visit order (position i holds source index
// perm(i), or i without shuffling), padded to n_pad with position 0.
void Gather(int nd, const float* pts, uint32_t n, uint32_t n_pad, bool shuffle,
            const Feistel& perm, bool keep_ids, int threads, HostSoA* out) {
  out->n = n;
  out->n_pad = n_pad;
  out->x.resize(n_pad);
  out->y.resize(n_pad);
  out->z.resize(nd == 3 ? n_pad : 0);
  out->id.resize(keep_ids ? n : 0);
  ParallelChunks(n, threads, [&](int, uint64_t begin, uint64_t end) {
    for (uint64_t i = begin; i < end; ++i) {
      const uint32_t src = shuffle ? HostPermute(perm, (uint32_t)i) : (uint32_t)i;
      const float* p = pts + (uint64_t)src * nd;
      out->x[i] = p[0];
      out->y[i] = p[1];
      if (nd == 3) out->z[i] = p[2];
      if (keep_ids) out->id[i] = src;
    }
  });
  for (uint32_t i = n; i < n_pad; ++i) {
    out->x[i] = out->x[0];
    out->y[i] = out->y[0];
    if (nd == 3) out->z[i] = out->z[0];
  }
}

inline float HighWord(uint64_t key) {
  const uint32_t hi = (uint32_t)(key >> 32);
  float f;
  std::memcpy(&f, &hi, 4);
  return f;
}

inline uint64_t PackKeyHost(float d2, uint32_t idx) {
  uint32_t bits;
  std::memcpy(&bits, &d2, 4);
  return ((uint64_t)bits << 32) | (uint64_t)(0xFFFFFFFFu - idx);
}

inline void RaiseKey(std::atomic<uint64_t>* key, uint64_t v) {
  uint64_t prev = key->load(std::memory_order_relaxed);
  while (prev < v &&
         !key->compare_exchange_weak(prev, v, std::memory_order_relaxed)) {
  }
}

// One A point against B: the running minimum over whole blocks, stopping
// after the first block that leaves it below cmax2. *blocks = blocks
// visited.
template <int ND>
__attribute__((target("avx2"))) float ScanAvx2(float px, float py, float pz,
                                               const HostSoA& b,
                                               const std::atomic<uint64_t>& key,
                                               uint32_t* blocks) {
  const __m256 ax = _mm256_set1_ps(px), ay = _mm256_set1_ps(py),
               az = _mm256_set1_ps(pz);
  const float* bx = b.x.data();
  const float* by = b.y.data();
  const float* bz = ND == 3 ? b.z.data() : nullptr;
  float cmin = FLT_MAX;
  uint32_t j = 0;
  while (j < b.n_pad) {
    __m256 m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t o = j + 8 * k;
      const __m256 dx = _mm256_sub_ps(ax, _mm256_loadu_ps(bx + o));
      const __m256 dy = _mm256_sub_ps(ay, _mm256_loadu_ps(by + o));
      __m256 d = _mm256_add_ps(_mm256_mul_ps(dx, dx), _mm256_mul_ps(dy, dy));
      if constexpr (ND == 3) {
        const __m256 dz = _mm256_sub_ps(az, _mm256_loadu_ps(bz + o));
        d = _mm256_add_ps(d, _mm256_mul_ps(dz, dz));
      }
      m[k] = d;
    }
    const __m256 m8 = _mm256_min_ps(_mm256_min_ps(m[0], m[1]),
                                    _mm256_min_ps(m[2], m[3]));
    __m128 v = _mm_min_ps(_mm256_castps256_ps128(m8),
                          _mm256_extractf128_ps(m8, 1));
    v = _mm_min_ps(v, _mm_movehl_ps(v, v));
    v = _mm_min_ss(v, _mm_shuffle_ps(v, v, 1));
    const float bm = _mm_cvtss_f32(v);
    if (bm < cmin) cmin = bm;
    j += kBlock;
    if (cmin < HighWord(key.load(std::memory_order_relaxed))) break;
  }
  *blocks = j / kBlock;
  return cmin;
}

template <int ND>
float ScanSse2(float px, float py, float pz, const HostSoA& b,
               const std::atomic<uint64_t>& key, uint32_t* blocks) {
  const __m128 ax = _mm_set1_ps(px), ay = _mm_set1_ps(py), az = _mm_set1_ps(pz);
  const float* bx = b.x.data();
  const float* by = b.y.data();
  const float* bz = ND == 3 ? b.z.data() : nullptr;
  float cmin = FLT_MAX;
  uint32_t j = 0;
  while (j < b.n_pad) {
    __m128 m = _mm_set1_ps(FLT_MAX);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t o = j + 4 * k;
      const __m128 dx = _mm_sub_ps(ax, _mm_loadu_ps(bx + o));
      const __m128 dy = _mm_sub_ps(ay, _mm_loadu_ps(by + o));
      __m128 d = _mm_add_ps(_mm_mul_ps(dx, dx), _mm_mul_ps(dy, dy));
      if constexpr (ND == 3) {
        const __m128 dz = _mm_sub_ps(az, _mm_loadu_ps(bz + o));
        d = _mm_add_ps(d, _mm_mul_ps(dz, dz));
      }
      m = _mm_min_ps(d, m);
    }
    m = _mm_min_ps(m, _mm_movehl_ps(m, m));
    m = _mm_min_ss(m, _mm_shuffle_ps(m, m, 1));
    const float bm = _mm_cvtss_f32(m);
    if (bm < cmin) cmin = bm;
    j += kBlock;
    if (cmin < HighWord(key.load(std::memory_order_relaxed))) break;
  }
  *blocks = j / kBlock;
  return cmin;
}

template <int ND>
void ScanRange(const HostSoA& a, const HostSoA& b, uint64_t begin,
               uint64_t end, bool avx2, std::atomic<uint64_t>* key,
               uint64_t* evals) {
  uint64_t local = 0;
  for (uint64_t i = begin; i < end; ++i) {
    const float pz = ND == 3 ? a.z[i] : 0.f;
    uint32_t blocks = 0;
    const float cmin = avx2 ? ScanAvx2<ND>(a.x[i], a.y[i], pz, b, *key, &blocks)
                            : ScanSse2<ND>(a.x[i], a.y[i], pz, b, *key, &blocks);
    local += std::min<uint64_t>((uint64_t)blocks * kBlock, b.n);
    // A point none of whose pairs is below FLT_MAX is skipped (:80).
    if (cmin != FLT_MAX) RaiseKey(key, PackKeyHost(cmin, a.id[i]));
  }
  *evals = local;
}

// d2 of the reference's formula (src/geoms/distance.h:17-24), unfused.
inline float HostD2(int nd, const float* p, const float* q) {
  float d = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]);
  if (nd == 3) d = d + (p[2] - q[2]) * (p[2] - q[2]);
  return d;
}

}  // namespace

void HostPointsMbr(int n_dims, const float* pts, uint64_t n, float* mbr) {
  for (int d = 0; d < n_dims; ++d) {
    mbr[d] = FLT_MAX;
    mbr[n_dims + d] = -FLT_MAX;
  }
  for (uint64_t i = 0; i < n; ++i)
    for (int d = 0; d < n_dims; ++d) {
      const float v = pts[i * n_dims + d];
      mbr[d] = std::min(mbr[d], v);
      mbr[n_dims + d] = std::max(mbr[n_dims + d], v);
    }
}

int HostThreads(int parallelism) {
  if (parallelism > 0) return parallelism;
  const unsigned n = std::thread::hardware_concurrency();
  return n > 0 ? (int)n : 1;
}

int CpuEarlyBreak(int n_dims, const float* a, uint64_t n_a, const float* b,
                  uint64_t n_b, uint32_t flags, int32_t seed, int parallelism,
                  hd_directed_result* out, CpuRunInfo* info) {
  if (n_dims != 2 && n_dims != 3)
    return Fail(HD_EINVAL, "n_dims must be 2 or 3, got %d", n_dims);
  if (!a || !b || !out) return Fail(HD_EINVAL, "null pointer");
  if (n_a == 0 || n_b == 0)
    return Fail(HD_EINVAL, "empty point set (|A|=%llu, |B|=%llu)",
                (unsigned long long)n_a, (unsigned long long)n_b);
  if (n_a > 0xFFFFFFFEull || n_b > 0xFFFFFFFEull)
    return Fail(HD_EINVAL, "point sets are limited to 2^32-2 points");
  if (flags & HD_FLAG_LEARNED_ORDER)
    return Fail(HD_EUNSUPPORTED,
                "the learned visit order needs the GPU (execution = GPU)");
  const int threads = HostThreads(parallelism);
  const bool shuffle = !(flags & HD_FLAG_NO_SHUFFLE);
  std::memset(out, 0, sizeof(*out));
  out->variant_used = HD_VARIANT_EB;

  // Non-finite coordinates are rejected, as on the device.
  for (const auto& set : {std::make_pair(a, n_a), std::make_pair(b, n_b)}) {
    std::atomic<bool> bad{false};
    ParallelChunks(set.second * n_dims, threads,
                   [&](int, uint64_t begin, uint64_t end) {
                     for (uint64_t i = begin; i < end; ++i)
                       if (!std::isfinite(set.first[i])) {
                         bad.store(true, std::memory_order_relaxed);
                         return;
                       }
                   });
    if (bad.load())
      return Fail(HD_EINVAL, "non-finite coordinate in the input");
  }

  const double t0 = NowMs();
  HostSoA va, vb;
  try {
    Gather(n_dims, a, (uint32_t)n_a, (uint32_t)n_a, shuffle,
           MakeFeistel((uint32_t)n_a, (uint64_t)(int64_t)seed, 0), true, threads,
           &va);
    const uint32_t n_pad = (uint32_t)((n_b + kBlock - 1) / kBlock * kBlock);
    Gather(n_dims, b, (uint32_t)n_b, n_pad, shuffle,
           MakeFeistel((uint32_t)n_b, (uint64_t)(int64_t)seed, 1), false,
           threads, &vb);
  } catch (const std::bad_alloc&) {
    return Fail(HD_ENOMEM, "host allocation for the CPU early break failed");
  }
  const double t1 = NowMs();

  const bool avx2 = __builtin_cpu_supports("avx2");
  std::atomic<uint64_t> key{0};
  std::vector<uint64_t> evals(threads, 0);
  ParallelChunks(n_a, threads, [&](int t, uint64_t begin, uint64_t end) {
    if (n_dims == 2)
      ScanRange<2>(va, vb, begin, end, avx2, &key, &evals[t]);
    else
      ScanRange<3>(va, vb, begin, end, avx2, &key, &evals[t]);
  });
  const double t2 = NowMs();

  const uint64_t k = key.load();
  out->cmax2_bits = (uint32_t)(k >> 32);
  const float cmax2 = HighWord(k);
  out->hd = std::sqrt(cmax2);
  for (uint64_t e : evals) out->evals += e;
  out->time_ms = t2 - t1;
  if (k == 0) {
    out->argmax = 0xFFFFFFFFu;
    out->nn = 0xFFFFFFFFu;
  } else {
    out->argmax = 0xFFFFFFFFu - (uint32_t)k;
    out->nn = 0xFFFFFFFFu;
    const float* p = a + (uint64_t)out->argmax * n_dims;
    for (uint64_t j = 0; j < n_b; ++j) {
      const float d = HostD2(n_dims, p, b + j * n_dims);
      uint32_t bits;
      std::memcpy(&bits, &d, 4);
      if (bits == out->cmax2_bits) {
        out->nn = (uint32_t)j;
        break;
      }
    }
  }
  if (info) {
    info->threads = threads;
    info->shuffle_ms = t1 - t0;
    info->compute_ms = t2 - t1;
    info->simd = avx2 ? "avx2" : "sse2";
  }
  return HD_OK;
}

}  // namespace hd

extern "C" int hd_directed_host(int n_dims, const float* a, uint64_t n_a,
                                const float* b, uint64_t n_b, uint32_t flags,
                                int32_t seed, int32_t parallelism,
                                hd_directed_result* out) {
  hd::ClearError();
  if (!out) return hd::Fail(HD_EINVAL, "null result pointer");
  return hd::CpuEarlyBreak(n_dims, a, n_a, b, n_b, flags, seed, parallelism,
                           out, nullptr);
}
