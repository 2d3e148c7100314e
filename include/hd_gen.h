/*
 * hd_gen.h — deterministic synthetic point-set generators.
 *
 * The reference loads real datasets from files (src/loaders/loader.h:26-234,
 * src/loaders/ply_loader.h:37-101); it has no synthetic generator, so these
 * entry points have no counterpart there. They exist so that the CLI, the
 * benchmark and the tests all draw bit-identical inputs for the five
 * workloads the hot path is measured on.
 *
 * Random source: SplitMix64 (Steele, Lea, Flood 2014). A stream is a 64-bit
 * state s; call k (k = 0, 1, ...) returns mix(s + (k + 1) * 0x9E3779B97F4A7C15).
 * Because call k depends only on k, every generator below is parallelised by
 * giving each point a fixed number of calls.
 *   u01(x)      = (x >> 40) * 2^-24                      (exact in FP32)
 *   normal pair = Box-Muller in FP64: u1 = ((x >> 11) + 1) * 2^-53,
 *                 u2 = (y >> 11) * 2^-53, r = sqrt(-2 ln u1),
 *                 (r cos 2 pi u2, r sin 2 pi u2)
 * Points are written AoS: x0 y0 [z0] x1 y1 [z1] ...
 *
 * All functions return 0 on success or a negative HD_E* code (hd.h).
 */
#ifndef HD_GEN_H
#define HD_GEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SplitMix64 output number `k` of the stream seeded with `seed`. */
uint64_t hd_splitmix64_at(uint64_t seed, uint64_t k);

/* n points, each coordinate u01, stream `seed`, 2 or 3 dims (n_dims calls
 * per point). Uniform workload: A = seed 1, B = seed 2, 100,000 points. */
int hd_gen_uniform(uint64_t seed, uint64_t n, int n_dims, float* out);

/* Gaussian-mixture pair (the 10M x 10M and 50M x 50M 2D workloads).
 *   centres: stream seed_centres, n_clusters x (u01, u01)
 *   A[i]:    stream seed_points, 3 calls per point: component = x0 % K,
 *            (zx, zy) = normal pair(x1, x2), a = (float)(c + sigma * z)
 *   B[i]     = A[i] + (shift_x, shift_y)           (FP32 add)
 *   outliers: n_outliers distinct indices from a partial Fisher-Yates over
 *            0..n-1 driven by stream seed_pick (j = i + x_i % (n - i)); the
 *            i-th picked index gets (u01, u01) from stream seed_outlier
 *            (2 calls per outlier).
 * a_out and b_out each hold 2 * n floats. */
int hd_gen_gmm_pair(uint64_t n, uint32_t n_clusters, double sigma,
                    float shift_x, float shift_y, uint64_t n_outliers,
                    uint64_t seed_centres, uint64_t seed_points,
                    uint64_t seed_pick, uint64_t seed_outlier, float* a_out,
                    float* b_out);

/* Noisy unit sphere (the 3D 20M x 15M workload). Per point 4 calls:
 * normal pairs (g0, g1), (g2, g3); p = g[0..2] / |g[0..2]| * (1 + noise*g3)
 * in FP64, cast to FP32. If rot_deg != 0 the FP32 point is then rotated about
 * z with c = (float)cos(rad), s = (float)sin(rad):
 * x' = c*x - s*y, y' = s*x + c*y (unfused FP32). out holds 3 * n floats. */
int hd_gen_sphere(uint64_t seed, uint64_t n, double noise, double rot_deg,
                  float* out);

/* Batched workload sizes: pair p uses the stream seeded with
 * hd_splitmix64_at(seed, p). Call 0 gives |A_p| = floor(32 * 2048^u01),
 * call 1 gives |B_p| likewise (log-uniform on [32, 65536)).
 * sizes_a / sizes_b hold n_pairs entries. */
int hd_gen_batch_sizes(uint64_t seed, uint32_t n_pairs, uint32_t* sizes_a,
                       uint32_t* sizes_b);

/* Batched workload points. After the two size calls, pair p's stream
 * draws set A then set B, each as: 4 centres (u01, u01) then per point
 * 3 calls (component = x % 4, normal pair) with sigma. Offsets are exclusive
 * prefix sums of the sizes, in points; out_a / out_b hold 2 * total floats. */
int hd_gen_batch_points(uint64_t seed, uint32_t n_pairs, double sigma,
                        const uint32_t* sizes_a, const uint32_t* sizes_b,
                        const uint64_t* offsets_a, const uint64_t* offsets_b,
                        float* out_a, float* out_b);

/* k indices drawn without replacement from 0..n-1 by the same partial
 * Fisher-Yates as the outlier pick (stream `seed`), in draw order. */
int hd_gen_subsample(uint64_t seed, uint64_t n, uint64_t k, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* HD_GEN_H */
