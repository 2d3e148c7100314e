/*
 * hd_oracle.h — CPU restatement of the reference's Hausdorff-distance path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, as the checker (and, in bench.py, as the timed CPU baseline).
 *
 * Built with gcc -O3 -ffp-contract=off (no -march), which matches the
 * reference's host build: x86-64 baseline code has no FMA, and the device
 * build uses --fmad=false "for consistency to the CPU results"
 * (reference CMakeLists.txt:122, src/CMakeLists.txt:29).
 */
#ifndef HD_ORACLE_H
#define HD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_result {
  uint32_t cmax2_bits;     /* FP32 bits of max over A of min over B of d2 */
  uint32_t argmax;         /* smallest A index whose exact NN d2 == cmax2 */
  uint32_t nn;             /* smallest B index at d2 == cmax2 from argmax */
  uint32_t n_threads;      /* threads actually used */
  uint64_t compared_pairs; /* d2 evaluations ("ComparedPairs") */
  double compute_ms;       /* scan time, shuffle excluded ("ReportedTime") */
  double shuffle_ms;       /* order set-up time ("ShuffleTime") */
} orc_result;

/* FP32 squared distance, left-to-right and unfused:
 * (ax-bx)*(ax-bx) + (ay-by)*(ay-by) [+ (az-bz)*(az-bz)].
 * Restates reference src/geoms/distance.h:17-24. */
float orc_d2(const float* p, const float* q, int n_dims);

/* Directed early-break HD h(A -> B) on the CPU. Restates
 * reference src/hd_impl/hausdorff_distance_early_break.h:44-115:
 * random order of A and B (:50-52), static contiguous partition of A over
 * n_threads threads (:58-65, :87-99), per point a running min from FLT_MAX
 * updated on strict '<' (:74-76), break as soon as the running min is
 * strictly below the shared cmax2 (:77-79), and a CAS max of the running min
 * when it is not FLT_MAX (:80-82, :229-234).
 * The shared maximum is kept as a 64-bit key (d2 bits << 32 | ~index) so the
 * arg-max index rides along; its high word is the reference's cmax2.
 * Order: when shuffle != 0, A and B are visited through Fisher-Yates
 * permutations driven by SplitMix64(seed) (the reference uses std::shuffle
 * with mt19937, whose exact permutation is library-specific; the order only
 * changes compared_pairs and time, never the result). Inputs are not
 * modified (the reference shuffles the caller's vectors in place).
 * n_threads <= 0 means hardware concurrency (run_hausdorff_distance.cu:222).
 * Returns 0, or -1 on invalid arguments. */
int orc_eb_directed(const float* a, uint64_t n_a, const float* b, uint64_t n_b,
                    int n_dims, int n_threads, uint64_t seed, int shuffle,
                    orc_result* out);

/* Directed HD by exhaustive max-min (no early break), with the arg-max and
 * NN tie rules stated in orc_result. */
int orc_brute_directed(const float* a, uint64_t n_a, const float* b,
                       uint64_t n_b, int n_dims, int n_threads,
                       orc_result* out);

/* Exact nearest neighbour of every query point: out_d2[i] = min_j d2 and
 * out_idx[i] = smallest j attaining it (UINT32_MAX and FLT_MAX when no pair
 * is below FLT_MAX, which is the reference's "not found" convention). */
int orc_nn_all(const float* q, uint64_t n_q, const float* r, uint64_t n_r,
               int n_dims, int n_threads, float* out_d2, uint32_t* out_idx);

/* Smallest j with d2(p, r_j) == target_d2 bits, or UINT32_MAX. */
uint32_t orc_first_at_d2(const float* p, const float* r, uint64_t n_r,
                         int n_dims, uint32_t target_bits);

/* HD lower / upper bounds between the MBRs of A and B, restating reference
 * src/geoms/hd_bounds.h:123-160 over src/geoms/mbr.h:193-266 (FP32).
 * mbr_a / mbr_b are (lower[0..d), upper[0..d)). Outputs squared bounds. */
void orc_hd_bounds2(const float* mbr_a, const float* mbr_b, int n_dims,
                    float* lb2, float* ub2);

/* MBR of a point set, with the reference's ToNonemptyMBR widening of a
 * zero extent by 1e-5 (src/geoms/mbr.h:288-296). out = lower, upper. */
void orc_mbr(const float* p, uint64_t n, int n_dims, float* out);

/* Two-hidden-layer MLP d_in -> H -> H -> 1 with ReLU, evaluated in FP64 from
 * FP32 parameters; returns the mean squared error over the batch and writes
 * d(loss)/d(param) for every parameter (same layout as params).
 * Parameter layout: W1[H][d_in], b1[H], W2[H][H], b2[H], w3[H], b3[1]. */
double orc_mlp_loss_grad(const float* params, int d_in, int hidden,
                         const float* x, const float* y, int batch,
                         double* grad);

#ifdef __cplusplus
}
#endif
#endif /* HD_ORACLE_H */
