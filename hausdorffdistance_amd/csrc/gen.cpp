// Deterministic synthetic generators (declared in include/hd_gen.h).
// Host-only code; built with -ffp-contract=off so the FP64 -> FP32 point
// coordinates are identical wherever the library runs on the same libm.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "hd.h"
#include "hd_gen.h"

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

inline uint64_t Mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Sequential view of one SplitMix64 stream starting at call `k`.
struct Stream {
  uint64_t state;
  Stream(uint64_t seed, uint64_t k) : state(seed + k * kGamma) {}
  uint64_t Next() {
    state += kGamma;
    return Mix(state);
  }
};

inline float U01(uint64_t x) {
  return static_cast<float>(x >> 40) * (1.0f / 16777216.0f);
}

inline void NormalPair(uint64_t x, uint64_t y, double* z0, double* z1) {
  const double u1 = (static_cast<double>(x >> 11) + 1.0) * 0x1.0p-53;
  const double u2 = static_cast<double>(y >> 11) * 0x1.0p-53;
  const double r = std::sqrt(-2.0 * std::log(u1));
  const double t = 2.0 * M_PI * u2;
  *z0 = r * std::cos(t);
  *z1 = r * std::sin(t);
}

// Runs fn(begin, end) over [0, n) on up to hardware_concurrency threads.
template <typename Fn>
void ParallelFor(uint64_t n, Fn fn) {
  unsigned nt = std::max(1u, std::thread::hardware_concurrency());
  nt = std::min<unsigned>(nt, 32);
  if (n < 65536 || nt == 1) {
    fn(uint64_t{0}, n);
    return;
  }
  std::vector<std::thread> threads;
  const uint64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const uint64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    threads.emplace_back([=] { fn(b, e); });
  }
  for (auto& th : threads) th.join();
}

// Partial Fisher-Yates over 0..n-1, first k draws.
void PartialFisherYates(uint64_t seed, uint64_t n, uint64_t k,
                        uint32_t* out) {
  std::vector<uint32_t> perm(n);
  for (uint64_t i = 0; i < n; ++i) perm[i] = static_cast<uint32_t>(i);
  Stream s(seed, 0);
  for (uint64_t i = 0; i < k; ++i) {
    const uint64_t j = i + s.Next() % (n - i);
    std::swap(perm[i], perm[j]);
    out[i] = perm[i];
  }
}

// One Gaussian-mixture set drawn from a sequential stream (batched pairs).
void GmmSetSequential(Stream& s, uint32_t n, double sigma, float* out) {
  float cx[4], cy[4];
  for (int c = 0; c < 4; ++c) {
    cx[c] = U01(s.Next());
    cy[c] = U01(s.Next());
  }
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t comp = static_cast<uint32_t>(s.Next() % 4);
    const uint64_t x = s.Next(), y = s.Next();
    double z0, z1;
    NormalPair(x, y, &z0, &z1);
    out[2 * i] = static_cast<float>(static_cast<double>(cx[comp]) + sigma * z0);
    out[2 * i + 1] =
        static_cast<float>(static_cast<double>(cy[comp]) + sigma * z1);
  }
}

inline uint32_t BatchSize(uint64_t x) {
  const double u = static_cast<double>(U01(x));
  return static_cast<uint32_t>(std::floor(32.0 * std::pow(2048.0, u)));
}

}  // namespace

extern "C" {

uint64_t hd_splitmix64_at(uint64_t seed, uint64_t k) {
  return Mix(seed + (k + 1) * kGamma);
}

int hd_gen_uniform(uint64_t seed, uint64_t n, int n_dims, float* out) {
  if ((n_dims != 2 && n_dims != 3) || (n > 0 && out == nullptr))
    return HD_EINVAL;
  ParallelFor(n, [&](uint64_t b, uint64_t e) {
    Stream s(seed, b * n_dims);
    for (uint64_t i = b; i < e; ++i)
      for (int d = 0; d < n_dims; ++d) out[i * n_dims + d] = U01(s.Next());
  });
  return HD_OK;
}

int hd_gen_gmm_pair(uint64_t n, uint32_t n_clusters, double sigma,
                    float shift_x, float shift_y, uint64_t n_outliers,
                    uint64_t seed_centres, uint64_t seed_points,
                    uint64_t seed_pick, uint64_t seed_outlier, float* a_out,
                    float* b_out) {
  if (n_clusters == 0 || n_outliers > n || n > 0xFFFFFFFFull ||
      (n > 0 && (a_out == nullptr || b_out == nullptr)))
    return HD_EINVAL;
  std::vector<float> cx(n_clusters), cy(n_clusters);
  {
    Stream s(seed_centres, 0);
    for (uint32_t k = 0; k < n_clusters; ++k) {
      cx[k] = U01(s.Next());
      cy[k] = U01(s.Next());
    }
  }
  ParallelFor(n, [&](uint64_t b, uint64_t e) {
    Stream s(seed_points, b * 3);
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t comp = static_cast<uint32_t>(s.Next() % n_clusters);
      const uint64_t x = s.Next(), y = s.Next();
      double z0, z1;
      NormalPair(x, y, &z0, &z1);
      const float ax =
          static_cast<float>(static_cast<double>(cx[comp]) + sigma * z0);
      const float ay =
          static_cast<float>(static_cast<double>(cy[comp]) + sigma * z1);
      a_out[2 * i] = ax;
      a_out[2 * i + 1] = ay;
      b_out[2 * i] = ax + shift_x;
      b_out[2 * i + 1] = ay + shift_y;
    }
  });
  if (n_outliers > 0) {
    std::vector<uint32_t> picked(n_outliers);
    PartialFisherYates(seed_pick, n, n_outliers, picked.data());
    Stream s(seed_outlier, 0);
    for (uint64_t i = 0; i < n_outliers; ++i) {
      const uint32_t j = picked[i];
      b_out[2ull * j] = U01(s.Next());
      b_out[2ull * j + 1] = U01(s.Next());
    }
  }
  return HD_OK;
}

int hd_gen_sphere(uint64_t seed, uint64_t n, double noise, double rot_deg,
                  float* out) {
  if (n > 0 && out == nullptr) return HD_EINVAL;
  const double rad = rot_deg * M_PI / 180.0;
  const float c = static_cast<float>(std::cos(rad));
  const float sn = static_cast<float>(std::sin(rad));
  ParallelFor(n, [&](uint64_t b, uint64_t e) {
    Stream s(seed, b * 4);
    for (uint64_t i = b; i < e; ++i) {
      const uint64_t x0 = s.Next(), x1 = s.Next(), x2 = s.Next(),
                     x3 = s.Next();
      double g0, g1, g2, g3;
      NormalPair(x0, x1, &g0, &g1);
      NormalPair(x2, x3, &g2, &g3);
      const double len = std::sqrt(g0 * g0 + g1 * g1 + g2 * g2);
      const double r = (1.0 + noise * g3) / len;
      float px = static_cast<float>(g0 * r);
      float py = static_cast<float>(g1 * r);
      const float pz = static_cast<float>(g2 * r);
      if (rot_deg != 0.0) {
        const float rx = c * px - sn * py;
        const float ry = sn * px + c * py;
        px = rx;
        py = ry;
      }
      out[3 * i] = px;
      out[3 * i + 1] = py;
      out[3 * i + 2] = pz;
    }
  });
  return HD_OK;
}

int hd_gen_batch_sizes(uint64_t seed, uint32_t n_pairs, uint32_t* sizes_a,
                       uint32_t* sizes_b) {
  if (n_pairs > 0 && (sizes_a == nullptr || sizes_b == nullptr))
    return HD_EINVAL;
  for (uint32_t p = 0; p < n_pairs; ++p) {
    Stream s(hd_splitmix64_at(seed, p), 0);
    sizes_a[p] = BatchSize(s.Next());
    sizes_b[p] = BatchSize(s.Next());
  }
  return HD_OK;
}

int hd_gen_batch_points(uint64_t seed, uint32_t n_pairs, double sigma,
                        const uint32_t* sizes_a, const uint32_t* sizes_b,
                        const uint64_t* offsets_a, const uint64_t* offsets_b,
                        float* out_a, float* out_b) {
  if (n_pairs > 0 && (sizes_a == nullptr || sizes_b == nullptr ||
                      offsets_a == nullptr || offsets_b == nullptr ||
                      out_a == nullptr || out_b == nullptr))
    return HD_EINVAL;
  ParallelFor(n_pairs, [&](uint64_t b, uint64_t e) {
    for (uint64_t p = b; p < e; ++p) {
      Stream s(hd_splitmix64_at(seed, p), 2);
      GmmSetSequential(s, sizes_a[p], sigma, out_a + 2 * offsets_a[p]);
      GmmSetSequential(s, sizes_b[p], sigma, out_b + 2 * offsets_b[p]);
    }
  });
  return HD_OK;
}

int hd_gen_subsample(uint64_t seed, uint64_t n, uint64_t k, uint32_t* out) {
  if (k > n || n > 0xFFFFFFFFull || (k > 0 && out == nullptr))
    return HD_EINVAL;
  PartialFisherYates(seed, n, k, out);
  return HD_OK;
}

}  // extern "C"
