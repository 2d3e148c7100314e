/*
 * hd.h — C ABI of the MI355X Hausdorff-distance library (libhd.so).
 *
 * Drop-in boundary for the reference's hot path:
 *   void hd::RunHausdorffDistance(const RunConfig&)
 *     reference src/run_hausdorff_distance.cuh:5, src/run_hausdorff_distance.cu:53-402
 *   HausdorffDistance<T,N>::CalculateDistance(stream, d_a, d_b)
 *     reference src/hausdorff_distance.h:15-42 (GPU overload :27-32)
 * Plain C types only; no exception or abort crosses this boundary. Every
 * entry point returns HD_OK (0) or a negative HD_E* code and leaves a
 * message for hd_last_error() on the calling thread.
 *
 * Exactness contract. A directed result is
 *   cmax2 = max over a in A of min over b in B of d2(a, b)
 * with d2 the reference's FP32 formula (ax-bx)*(ax-bx)+(ay-by)*(ay-by)
 * [+(az-bz)*(az-bz)], left to right and unfused (src/geoms/distance.h:17-24).
 * Pairs with d2 >= FLT_MAX never count, and a point none of whose pairs
 * counts is skipped, as the reference's FLT_MAX-initialised running min does
 * (src/hd_impl/hausdorff_distance_early_break.h:69-82). Every variant returns
 * the same cmax2 bits. The reference returns only sqrt(cmax2); this library
 * also returns the arg-max (the smallest source index whose exact NN squared
 * distance equals cmax2) and its NN (the smallest target index at that d2).
 */
#ifndef HD_H
#define HD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HD_ABI_VERSION 1u

/* Return codes. The reference aborts through glog CHECK / LOG(FATAL)
 * (src/main.cpp:93,105,134; src/run_hausdorff_distance.cu:115-116,333) and
 * throws hd::Exception on CUDA errors (src/utils/exception.h:104-160). */
#define HD_OK 0
#define HD_EINVAL (-1)        /* bad argument (n_dims, empty set, NaN, ...) */
#define HD_ENOMEM (-2)        /* device or host allocation failed */
#define HD_EDEVICE (-3)       /* HIP runtime or kernel launch error */
#define HD_EUNSUPPORTED (-4)  /* valid request this build does not serve */
#define HD_ETRUNC (-5)        /* stats JSON did not fit; result is valid */

/* Variants. Reference enum Variant (src/run_config.h:7-16) and the CLI
 * strings of src/main.cpp:76-94: "eb" -> EB, "rt" -> LBVH (the OptiX BVH
 * query becomes a software LBVH), "hybrid" -> HYBRID, "nn" -> NN (the
 * nearest-neighbour-search baseline, hausdorff_distance_nearest_neighbor_
 * search.h:107-153: every A point's exact NN in B, no early break, then the
 * maximum; the kd-tree becomes the uniform grid). GRID is the uniform grid index
 * queried on its own. */
#define HD_VARIANT_EB 0
#define HD_VARIANT_GRID 1
#define HD_VARIANT_LBVH 2
#define HD_VARIANT_HYBRID 3
#define HD_VARIANT_NN 4
/* "rt-hdist": the quantized-grid baseline (hausdorff_distance_rt_hdist.h,
 * shaders_nn_quantized_grid_{2,3}d.cu) without ray-tracing hardware. Its
 * result is at least the exact HD and equals it on ordinary inputs (each A
 * point's minimum is taken over the B points of the quantized cells within
 * three cell diagonals beyond its first hit). */
#define HD_VARIANT_RT_HDIST 5

/* Execution (src/run_config.h:18). CPU runs the early-break variant on the
 * host with `parallelism` threads (hd_directed_host); other variants return
 * HD_EUNSUPPORTED on the CPU, as the reference's base class aborts with
 * "CalculateDistance on the CPU is not implemented"
 * (src/hausdorff_distance.h:21-26). CPU execution needs no GPU. */
#define HD_EXEC_CPU 0
#define HD_EXEC_GPU 1

/* hd_run_config.flags */
#define HD_FLAG_SYMMETRIC 0x1u       /* H = max(h(A->B), h(B->A)) */
#define HD_FLAG_CHECK 0x2u           /* re-run the host early break with
                                        hardware_concurrency threads and
                                        compare cmax2 bits and arg-max
                                        (-check, run_hausdorff_distance.cu:
                                        366-388) */
#define HD_FLAG_LEARNED_ORDER 0x4u   /* visit A in descending MLP-predicted
                                        NN distance */
#define HD_FLAG_COUNT_EVALS 0x8u     /* count d2 evaluations */
#define HD_FLAG_INPUT_ON_DEVICE 0x10u/* a, b are device pointers */
#define HD_FLAG_NO_SHUFFLE 0x20u     /* visit A and B in input order */
#define HD_FLAG_SINGLE_PASS 0x40u    /* early break as one pass over A in the
                                        visit order: every point scans B
                                        until it breaks, cmax2 starts at 0
                                        (the reference's GPU kernel,
                                        early_break.h:117-219); no seed,
                                        rounds or matrix-core filter. EB
                                        variant only: a comparison
                                        baseline, same result */
#define HD_FLAG_B_ORDER_RANDOM 0x80u /* early break: visit B in a pseudo-
                                        random order, as the reference
                                        shuffles it (early_break.h:51-52,
                                        132-137) */
#define HD_FLAG_B_ORDER_STRATIFIED 0x100u /* early break: visit B in a
                                        spatially stratified order (Morton
                                        ranks spread evenly over the
                                        sequence). Neither B_ORDER flag: the
                                        learned order (HD_FLAG_LEARNED_ORDER)
                                        takes the stratified one unless it
                                        runs as a single pass, the other
                                        schedules the pseudo-random one.
                                        HD_FLAG_NO_SHUFFLE (input order)
                                        overrides both flags; above 2^31 B
                                        points the order is pseudo-random.
                                        The two B_ORDER flags exclude each
                                        other (HD_EINVAL). Same result in
                                        every order */

/* hd_run_config.auto_tune (reference -auto_tune, -auto_tune_n_points_cell,
 * -auto_tune_max_hit, src/flags.cc:28-31; the reference predicts them with
 * gradient-boosted trees, hausdorff_distance_hybrid.h:722-778). ALL also
 * picks the strategy (grid with early break, LBVH or the early-break scan).
 * The choice is cached per context for inputs with the same content. */
#define HD_AUTO_TUNE_ALL 0x1
#define HD_AUTO_TUNE_N_POINTS_CELL 0x2
#define HD_AUTO_TUNE_MAX_HIT 0x4

/* Reference RunConfig (src/run_config.h:20-61) restricted to the hot path,
 * with the defaults of src/flags.cc:3-57 (hd_run_config_init). */
typedef struct hd_run_config {
  uint32_t abi_version;    /* HD_ABI_VERSION */
  int32_t variant;         /* HD_VARIANT_*            (-variant, default eb) */
  int32_t execution;       /* HD_EXEC_*               (-execution, gpu) */
  int32_t n_dims;          /* 2 or 3                  (-n_dims) */
  int32_t repeat;          /* >= 1                    (-repeat, 5) */
  int32_t seed;            /* visit-order seed        (-seed, 0) */
  int32_t parallelism;     /* host threads, <=0 = all (-parallelism, -1) */
  uint32_t flags;          /* HD_FLAG_* */
  float sample_rate;       /* hybrid sampling rate    (-sample_rate, 0.01) */
  uint32_t n_points_cell;  /* grid points per cell    (-n_points_cell, 8) */
  uint32_t max_hit;        /* index cells a query may visit before the point
                              is handed to the EB kernel (-max_hit, 256) */
  uint32_t eb_only_threshold; /* hybrid: EB only if sampled p80 iterations
                              fall below this (-eb_only_threshold, 0) */
  int32_t auto_tune;       /* hybrid parameters chosen per input by timed
                              trials instead of the flags: HD_AUTO_TUNE_*
                              bits (-auto_tune, -auto_tune_n_points_cell,
                              -auto_tune_max_hit) */
  int32_t stats_n_points_cell; /* points per cell of the input-statistics
                              grid written to the stats JSON ("Input":
                              MBR, density, cell counts, Gini, percentiles;
                              run_hausdorff_distance.cu:151-216); 0 = none
                              (-stats_n_points_cell, 10) */
  const void* model;       /* optional hd_mlp* for HD_FLAG_LEARNED_ORDER */
} hd_run_config;

/* One direction. */
typedef struct hd_directed_result {
  uint32_t cmax2_bits;  /* FP32 bits of cmax2 */
  uint32_t argmax;      /* source index, UINT32_MAX if no point counted */
  uint32_t nn;          /* target index, UINT32_MAX if no point counted */
  int32_t variant_used; /* HD_VARIANT_* that produced it (hybrid picks) */
  float hd;             /* sqrtf(cmax2) */
  float reserved;
  uint64_t evals;       /* d2 evaluations (HD_FLAG_COUNT_EVALS), else 0 */
  double time_ms;       /* device time of this direction */
} hd_directed_result;

typedef struct hd_result {
  float hd;             /* HDResult: directed h(A->B), or symmetric H */
  uint32_t cmax2_bits;
  uint32_t argmax;      /* in set `direction` names */
  uint32_t nn;          /* in the other set */
  int32_t direction;    /* 0: A->B attains it, 1: B->A (ties: 0) */
  int32_t check_pass;   /* -1 not checked, 0 mismatch, 1 bit-identical */
  uint32_t n_repeats;
  uint32_t reserved;
  hd_directed_result ab;
  hd_directed_result ba; /* zeroed unless HD_FLAG_SYMMETRIC */
  double avg_time_ms;    /* Running.AvgTime: mean ReportedTime */
  double median_time_ms;
  double min_time_ms;
} hd_result;

typedef struct hd_context hd_context;
typedef struct hd_mlp hd_mlp;

/* Thread-local message of the last failing call on this thread. */
const char* hd_last_error(void);

/* Fills the reference defaults (src/flags.cc:3-57): variant eb, execution
 * gpu, repeat 5, seed 0, parallelism -1, sample_rate 0.01, n_points_cell 8,
 * max_hit 256, eb_only_threshold 0. n_dims is left 0 (must be set). */
void hd_run_config_init(hd_run_config* cfg);

/* A context owns one HIP stream on `device`, cached device buffers and
 * events; calls on one context are serialised by the caller. The reference
 * creates one stream per run (src/utils/stream.h:13-27). */
int hd_context_create(int device, hd_context** out);
void hd_context_destroy(hd_context* ctx);
/* The context's hipStream_t, as void*. */
void* hd_context_stream(hd_context* ctx);
/* Caps the device memory the context's work buffers may hold at `bytes`
 * (0 = no cap) and releases the buffers it holds now (they are allocated
 * again on demand; a cached batch plan is dropped with them). A call that
 * would grow the buffers past the cap fails with HD_ENOMEM before it
 * launches anything that needs the missing buffer, and leaves the context
 * usable. Trained ordering models are not counted. No counterpart in the
 * reference, which allocates through thrust and aborts when an allocation
 * fails; it serves callers that share the GPU and must bound this library,
 * and tests of the allocation-failure path. */
int hd_context_set_memory_limit(hd_context* ctx, uint64_t bytes);
/* Device bytes the context's work buffers hold now. */
uint64_t hd_context_memory_held(hd_context* ctx);

/* Replaces hd::RunHausdorffDistance (src/run_hausdorff_distance.cu:53-402)
 * for in-memory inputs: validates, uploads (unless INPUT_ON_DEVICE), runs
 * `repeat` timed repetitions, optionally checks, and writes the reference's
 * stats JSON keys (Input, Running{Repeats[{Repeat, Algorithm, Execution,
 * ReportedTime, ...}], AvgTime}, HDResult, Check{HDAnswer, Diff, Pass}) into
 * stats_json (NUL-terminated; HD_ETRUNC if stats_cap is too small).
 * a, b: AoS FP32, n_dims floats per point. ctx may be NULL (a temporary
 * context on device 0 is used). Inputs are never modified. */
int hd_run_hausdorff_distance(hd_context* ctx, const hd_run_config* cfg,
                              const float* a, uint64_t n_a, const float* b,
                              uint64_t n_b, hd_result* out, char* stats_json,
                              uint64_t stats_cap);

/* One directed h(A->B) on device-resident AoS inputs; synchronous. The GPU
 * overload CalculateDistance(stream, d_a, d_b) of the reference
 * (src/hausdorff_distance.h:27-32), without its in-place shuffle
 * (early_break.h:132-136): d_a and d_b are read only. `variant` is
 * HD_VARIANT_*, `flags` takes LEARNED_ORDER, COUNT_EVALS, NO_SHUFFLE,
 * SINGLE_PASS.
 * `model` (may be NULL) is used with LEARNED_ORDER; NULL trains one. */
int hd_directed_device(hd_context* ctx, int variant, int n_dims,
                       const float* d_a, uint64_t n_a, const float* d_b,
                       uint64_t n_b, uint32_t flags, int32_t seed,
                       const hd_mlp* model, hd_directed_result* out);

/* One directed h(A->B) on the host CPU: the reference's CPU overload
 * CalculateDistance(std::vector&, std::vector&)
 * (src/hd_impl/hausdorff_distance_early_break.h:44-115), without its
 * in-place shuffle: a and b are host AoS arrays and are read only. A and B
 * are visited in the library's pseudo-random order (NO_SHUFFLE: input
 * order), A is split statically over `parallelism` threads (<= 0:
 * std::thread::hardware_concurrency(), run_hausdorff_distance.cu:222-224),
 * and every point stops scanning once its running minimum is strictly below
 * the shared cmax2. out->time_ms is the scan time without the gather
 * (ReportedTime, :101-112); out->evals counts evaluated pairs per visited
 * block of 32. Same cmax2 bits, arg-max and NN as every GPU variant.
 * `flags` takes NO_SHUFFLE; LEARNED_ORDER returns HD_EUNSUPPORTED. */
int hd_directed_host(int n_dims, const float* a, uint64_t n_a, const float* b,
                     uint64_t n_b, uint32_t flags, int32_t seed,
                     int32_t parallelism, hd_directed_result* out);

/* Allocation-free directed h(A->B) on device-resident AoS inputs: no
 * counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The early break of hd_directed_device with
 * HD_FLAG_SINGLE_PASS and the same kernels, but every buffer (key, counters,
 * B repacked) lives in one device array of the library, 12 MiB + 5 KiB per
 * GPU, held until the process ends; the call makes no allocation of any
 * kind. Limits: |B| <= 1,048,576 points (HD_EINVAL above); ctx must be a
 * context made beforehand (NULL is HD_EINVAL); flags take NO_SHUFFLE,
 * SINGLE_PASS (implied) and COUNT_EVALS; LEARNED_ORDER is HD_EUNSUPPORTED.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (several threads, several contexts) race and must be serialised by the
 * caller. Synchronous; same cmax2 bits, arg-max and NN as hd_directed_device. */
int hd_directed_device_noalloc(hd_context* ctx, int n_dims, const float* d_a,
                               uint64_t n_a, const float* d_b, uint64_t n_b,
                               uint32_t flags, int32_t seed,
                               hd_directed_result* out);

/* Exact nearest neighbour of every query in `ref` on the device: the
 * per-point minimum the EB loop computes when it does not break. index:
 * HD_VARIANT_EB (exhaustive scan), HD_VARIANT_GRID or HD_VARIANT_LBVH.
 * d_out_d2[i] = FP32 min d2 (FLT_MAX if no pair counts), d_out_idx[i] = the
 * smallest ref index attaining it (UINT32_MAX if none). */
int hd_nn_device(hd_context* ctx, int index, int n_dims, const float* d_q,
                 uint64_t n_q, const float* d_ref, uint64_t n_ref,
                 float* d_out_d2, uint32_t* d_out_idx);

/* Batched directed HD: pair p is A[offs_a[p] .. offs_a[p+1]) against
 * B[offs_b[p] .. offs_b[p+1]), all device-resident (AoS, offsets in points,
 * n_pairs + 1 entries each). One persistent-kernel launch captured in a HIP
 * graph; tiles come from the autotuner. flags: HD_FLAG_COUNT_EVALS,
 * HD_FLAG_NO_SHUFFLE (each pair's B scanned in input order; by default its
 * groups of 8 consecutive points in the order g -> g * step mod n_g, n_g =
 * |B_p| / 8 whole groups, step coprime with n_g near n_g / phi, the last
 * |B_p| mod 8 points last). out: host array of n_pairs results (evals and time_ms are
 * per batch, reported in out[0]). No counterpart in the reference, which
 * runs one pair per process. */
int hd_batch_directed_device(hd_context* ctx, int n_dims, const float* d_a,
                             const uint64_t* d_offs_a, const float* d_b,
                             const uint64_t* d_offs_b, uint32_t n_pairs,
                             uint32_t flags, hd_directed_result* out);

/* ---- learned visit order (no reference counterpart: the reference's
 * selector uses gradient-boosted trees, src/models/tree_*.h) ----
 * MLP n_dims -> 64 -> 64 -> 1 (sine, ReLU, linear), BF16 weights with an FP32 master copy,
 * FP32 accumulation. Input: coordinates normalised by the training set's
 * MBR. Output: predicted log of the NN distance. */
typedef struct hd_mlp_train_config {
  uint32_t steps;        /* Adam steps (2000) */
  uint32_t batch;        /* samples per step (4096) */
  uint32_t n_samples;    /* subsample of A labelled exactly (1,000,000) */
  uint32_t seed;         /* init + subsample seed (14) */
  float lr;              /* Adam learning rate (3e-3) */
  int32_t label_index;   /* HD_VARIANT_GRID / LBVH / EB for the exact labels */
} hd_mlp_train_config;

void hd_mlp_train_config_init(hd_mlp_train_config* cfg);

/* Trains on exact NN distances of a subsample of d_a against d_b, on the
 * GPU with this library's forward / backward / Adam kernels. final_loss
 * (may be NULL) receives the mean loss of the last 100 steps. */
int hd_mlp_train_device(hd_context* ctx, int n_dims, const float* d_a,
                        uint64_t n_a, const float* d_b, uint64_t n_b,
                        const hd_mlp_train_config* cfg, hd_mlp** out,
                        float* final_loss);
/* d_out[i] = predicted log NN distance of d_pts[i] (FP32). */
int hd_mlp_predict_device(hd_context* ctx, const hd_mlp* model,
                          const float* d_pts, uint64_t n, float* d_out);
void hd_mlp_destroy(hd_mlp* model);
/* Parameter count and FP32 master parameters (layout W1[64][d], b1[64],
 * W2[64][64], b2[64], w3[64], b3[1]); also the input normalisation
 * (lower[d], inv_extent[d]) and label shift/scale. */
int hd_mlp_n_params(int n_dims);
int hd_mlp_get_params(const hd_mlp* model, float* params, float* norm_lower,
                      float* norm_inv_extent);
/* One forward + backward pass on the GPU for a host batch with host FP32
 * parameters: writes the mean-squared-error loss and the FP32 gradient.
 * Weights are rounded to BF16 first, as in training. Test hook for the
 * gradient check against an FP64 CPU evaluation. */
int hd_mlp_loss_grad_device(hd_context* ctx, int n_dims, const float* params,
                            const float* x, const float* y, uint32_t batch,
                            float* loss, float* grad);

/* Test hook of the early break's matrix-core filter (no reference
 * counterpart): for device AoS sets A and B (1..65536 points each), the
 * filter's approximate value D for every pair, computed exactly as the
 * filter does (centred 16-point tiles, v_mfma_f32_16x16x4_f32):
 * d_out[i * n_b + j] approximates d2(a_i, b_j) in 2D; in 3D it is
 * d2 + d_cna[i] (d_cna may be NULL in 2D). params (host, 5 floats) receives
 * the centre (3), the error bound E the filter assumes for |D - d2| (3D:
 * |D - cna - d2|, before its threshold arithmetic) and the 3D offset C. */
int hd_filter_probe_device(hd_context* ctx, int n_dims, const float* d_a,
                           uint64_t n_a, const float* d_b, uint64_t n_b,
                           float* d_out, float* d_cna, float* params);

/* Test hook of the library's stable radix sort of (key, u32 value) pairs
 * (the reference sorts Morton codes with thrust::sort_by_key,
 * src/index/uniform_grid.h, src/lbvh/lbvh/bvh.cuh): sorts n device pairs on
 * the context stream and waits. key_kind 0: u16 keys, ascending; 1: u32
 * keys, ascending on bits [begin_bit, end_bit) (other bits carried);
 * 2: finite f32 keys, descending. Equal keys keep their input order.
 * Outputs must not alias inputs. */
int hd_sort_pairs_probe_device(hd_context* ctx, int key_kind,
                               const void* d_keys_in, const uint32_t* d_vals_in,
                               uint32_t n, int begin_bit, int end_bit,
                               void* d_keys_out, uint32_t* d_vals_out);

/* Host-side pieces of the input statistics (stats_n_points_cell), as test
 * hooks: for the multiset whose value v occurs hist[v] times (the number of
 * grid cells holding v points), the Gini index
 * sum_i (2i - n + 1) x_i / (n sum x) over the sorted values
 * (src/running_stats.h:32-57, here in FP64), and the percentile list the
 * reference writes from its HdrHistogram dump (running_stats.h:106-140,
 * uniform_grid.h:675-692): [{"value", "percentile", "count"}, ...] at the
 * levels 0, 1/6, 1/3, 1/2, 7/12, ... (three per halving of the distance to
 * 1, rounded to 6 decimals), then 1. As in a histogram made for values
 * 1..(number of cells) with 3 significant figures, values from
 * 2048 * 2^k above that number are left out and values from 2048 up are
 * reported as the top of their sub-bucket. */
int hd_stats_gini(const uint64_t* hist, uint64_t n_values, double* out);
/* Lower and upper bounds of the directed HD A -> B from the two MBRs
 * (lower[n_dims], upper[n_dims] each, made non-empty as the stats do), as the
 * reference's hybrid computes them (src/geoms/hd_bounds.h:111-160 over the
 * MBRs' faces; hausdorff_distance_hybrid.h:95-106) and writes them into
 * each repeat of the stats JSON ("HDLowerBound", "HDUpperBound"). */
int hd_mbr_hd_bounds(const float* mbr_a, const float* mbr_b, int n_dims,
                     float* lower, float* upper);
/* The hybrid's sampled early-break pass (hausdorff_distance_hybrid.h:
 * 120-199), used when eb_only_threshold > 0 and |B| < 500k: for the first
 * ceil(|A| * sample_rate) points of A's pseudo-random order (seed), the
 * iteration count of the reference's sampling kernel (units of 256 B
 * points of a shuffled B, break checked every 32 iterations with <=
 * against cmax2 = max(MBR lower bound^2, the sampled points' exact NN d2)).
 * The hybrid runs the early-break scan alone when the 80th percentile
 * (sorted counts, index floor(0.8 n)) is below eb_only_threshold. Device
 * inputs; writes min(n, cap) counts, *n_out = n (HD_ETRUNC if n > cap). */
int hd_hybrid_sample_iterations(hd_context* ctx, int n_dims, const float* d_a,
                                uint64_t n_a, const float* d_b, uint64_t n_b,
                                int32_t seed, float sample_rate, uint32_t* counts,
                                uint64_t cap, uint64_t* n_out);
int hd_stats_percentiles(const uint64_t* hist, uint64_t n_values, char* json,
                         uint64_t cap);

/* Per-stage device timing. With profiling on, each pipeline stage (kernel or
 * library call) is bracketed by HIP events on the context stream;
 * hd_context_stage_times writes {"stage": {"ms": total, "calls": n,
 * "evals": d2 evaluations, "points": points processed}, ...} accumulated
 * since the last reset, then resets if `reset`. No reference counterpart
 * (the reference times whole calls with a host Stopwatch,
 * src/utils/stopwatch.h). */
int hd_context_set_profiling(hd_context* ctx, int enable);
int hd_context_stage_times(hd_context* ctx, char* json, uint64_t cap,
                           int reset);

/* Device / build information. */
int hd_device_count(void);
/* Writes a short JSON with device name, CU count, clock, build flags. */
int hd_device_info(int device, char* buf, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif /* HD_H */
