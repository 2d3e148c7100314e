// Host-side internals shared by the library's translation units.
#ifndef HD_INTERNAL_H
#define HD_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "hd.h"

namespace hd {

// Thread-local error message behind hd_last_error().
void SetError(const std::string& msg);
void ClearError();

// Formats and records an error, returns `code`.
int Fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define HD_HIP_TRY(expr)                                                    \
  do {                                                                      \
    hipError_t hd_err_ = (expr);                                            \
    if (hd_err_ != hipSuccess) {                                            \
      return ::hd::Fail(hd_err_ == hipErrorOutOfMemory ? HD_ENOMEM          \
                                                       : HD_EDEVICE,        \
                        "%s failed: %s (%s:%d)", #expr,                     \
                        hipGetErrorString(hd_err_), __FILE__, __LINE__);    \
    }                                                                       \
  } while (0)

#define HD_TRY(expr)             \
  do {                           \
    int hd_rc_ = (expr);         \
    if (hd_rc_ != HD_OK) return hd_rc_; \
  } while (0)

// Parameters of a bijective pseudo-random permutation of [0, n) evaluated on
// the device (hd_device.h: Permute).
struct Feistel {
  uint32_t n;
  uint32_t half;
  uint32_t mask;
  uint32_t key[4];
};
namespace dev {
using hd::Feistel;
}

// Strategy and index parameters the hybrid runs with (chosen by rules, by
// the sampled pass or by trial under auto_tune).
struct TunedChoice {
  int variant = -1;          // HD_VARIANT_EB / GRID / LBVH
  uint32_t n_points_cell = 0;
  uint32_t max_hit = 0;
  bool tuned = false;        // chosen by trial (auto_tune)
  uint32_t trials = 0;       // runs timed to choose it (0: cached or rules)
};

// Grow-only device buffer cached in a context under a name.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Per-stage device timing (HIP events on the context stream), accumulated
// across calls while profiling is on.
struct StageStat {
  double total_ms = 0;
  uint64_t calls = 0;
  uint64_t dispatches = 0;  // kernel dispatches the stage's calls made
  uint64_t evals = 0;
  uint64_t points = 0;
  uint64_t lookups = 0;  // index records read (grid cells, LBVH nodes)
};
struct PendingStage {
  std::string name;
  hipEvent_t begin, end;
  uint64_t dispatches = 0;
};

}  // namespace hd

namespace hd {
struct BatchPlan;
void DestroyBatchPlan(BatchPlan* plan);
}  // namespace hd

// Opaque handle of the C ABI.
struct hd_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  int max_waves_per_cu = 32;
  std::unordered_map<std::string, hd::DeviceBuffer> buffers;
  // Device bytes the buffers hold, and the most they may hold (0: no limit).
  uint64_t memory_held = 0;
  uint64_t memory_limit = 0;
  void* host_pinned = nullptr;  // small pinned staging area
  size_t host_pinned_bytes = 0;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  bool profiling = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<hd::PendingStage> pending_stages;
  int open_stage = -1;  // pending stage between StageBegin and StageEnd
  std::map<std::string, hd::StageStat> stage_stats;
  // Ordering model trained on demand for HD_FLAG_LEARNED_ORDER without an
  // explicit model, with the inputs it was trained for.
  // Hybrid selections made by trial (auto_tune), keyed by a signature of
  // the inputs' content (sizes, MBRs, a strided sample of the points), and
  // the choice of the last hybrid call.
  std::map<std::string, hd::TunedChoice> tuned;
  hd::TunedChoice last_choice;
  // Batched-mode plan and HIP graph for the last batch signature.
  hd::BatchPlan* batch_plan = nullptr;
  hd_mlp* cached_model = nullptr;
  const float* cached_a = nullptr;
  const float* cached_b = nullptr;
  uint64_t cached_n_a = 0, cached_n_b = 0;
};

namespace hd {

// Returns a device buffer of at least `bytes` under `name` (contents are
// undefined; the buffer is reused across calls on the same context).
int ContextBuffer(hd_context* ctx, const char* name, size_t bytes, void** out);

template <typename T>
int ContextArray(hd_context* ctx, const char* name, size_t count, T** out) {
  void* p = nullptr;
  int rc = ContextBuffer(ctx, name, count * sizeof(T) + 256, &p);
  *out = static_cast<T*>(p);
  return rc;
}

// Pinned host staging of at least `bytes`. The buffer is never smaller than
// kPinnedMinBytes, so a request up to that size returns the same buffer as
// before (pointers that callers still hold stay valid); a larger request may
// replace it.
constexpr size_t kPinnedMinBytes = 4096;
int ContextPinned(hd_context* ctx, size_t bytes, void** out);

// Makes ctx->device current on the calling thread.
int ContextActivate(hd_context* ctx);

// Stage timing: StageBegin records a start event when profiling is on and
// returns a handle (-1 when off); StageEnd records the end event.
// StageCount adds work counters to a stage. Pending events are resolved
// lazily by ResolveStages (after the stream is synchronised).
int StageBegin(hd_context* ctx, const char* name);
void StageEnd(hd_context* ctx, int handle);
void StageCount(hd_context* ctx, const char* name, uint64_t evals,
                uint64_t points, uint64_t lookups = 0);
int ResolveStages(hd_context* ctx);
// Counts one kernel dispatch against the open stage (profiling only), so
// per-dispatch times divide by launches actually made.
inline void StageDispatch(hd_context* ctx) {
  if (ctx->open_stage >= 0 &&
      ctx->open_stage < (int)ctx->pending_stages.size())
    ctx->pending_stages[ctx->open_stage].dispatches += 1;
}

// Persistent grid size for wave-level kernels of `threads_per_block`.
inline int PersistentBlocks(const hd_context* ctx, int threads_per_block,
                            int waves_per_simd) {
  int waves_per_block = threads_per_block / 64;
  int waves = ctx->cu_count * 4 * waves_per_simd;
  int blocks = waves / waves_per_block;
  return blocks > 0 ? blocks : 1;
}

// Device-side HD pipelines (implemented in the .hip files).
struct DirectedArgs {
  int n_dims = 2;
  const float* a = nullptr;  // AoS device pointer
  uint64_t n_a = 0;
  const float* b = nullptr;  // AoS device pointer
  uint64_t n_b = 0;
  uint32_t flags = 0;
  int32_t seed = 0;
  const hd_mlp* model = nullptr;
  // tuning knobs (0 = library default)
  uint32_t n_points_cell = 0;
  uint32_t max_hit = 0;
  uint32_t slot_budget_steps = 0;
  uint32_t seed_points = 0;
  float sample_rate = 0.f;
  uint32_t eb_only_threshold = 0;
  int auto_tune = 0;
  // Bounding boxes (lo[nd], hi[nd]) of A and B when the caller measured
  // them; they enable the matrix-core filter of the early-break rounds.
  bool has_mbr = false;
  float mbr_a[6] = {0, 0, 0, 0, 0, 0};
  float mbr_b[6] = {0, 0, 0, 0, 0, 0};
};

int RunEarlyBreak(hd_context* ctx, const DirectedArgs& args,
                  hd_directed_result* out);
int RunGrid(hd_context* ctx, const DirectedArgs& args,
            hd_directed_result* out);
int RunLbvh(hd_context* ctx, const DirectedArgs& args,
            hd_directed_result* out);
int RunHybrid(hd_context* ctx, const DirectedArgs& args,
              hd_directed_result* out);
// The nearest-neighbour-search baseline: exact grid NN of every A point
// (no early break), then the maximum (nn.hip).
int RunNearestNeighborSearch(hd_context* ctx, const DirectedArgs& args,
                             hd_directed_result* out);
// The hybrid's sampled early-break pass (sample.hip): per sampled A point,
// the reference's iteration count (units of 256 B points).
int SampledEbIterations(hd_context* ctx, const DirectedArgs& args,
                        std::vector<uint32_t>* counts);
// The RT-HDIST baseline on a quantized grid (rthdist.hip).
int RunRtHdist(hd_context* ctx, const DirectedArgs& args, hd_directed_result* out);
// Grid query with early break whose queries that visit more than max_cells
// cells are finished by the early-break scan (the reference's hybrid hands
// such points from the RT query to its EB kernel, hybrid.h:291-295,377-400).
int RunGridHybrid(hd_context* ctx, const DirectedArgs& args,
                  uint32_t max_cells, hd_directed_result* out);

// B repacked for the early-break scans (eb.hip).
struct EbTarget {
  float* bx = nullptr;
  float* by = nullptr;
  float* bz = nullptr;
  uint32_t n_bp = 0;     // padded to whole 64-point steps
  uint32_t n_steps = 0;  // n_bp / 64
  // Matrix-core filter (null when not prepared): centred B tiles of 16
  // points, the centre and the error bound of the approximate d2.
  float* bt = nullptr;
  float center[3] = {0, 0, 0};
  float err = 0.f;
  float off = 0.f;  // 3D: C, which keeps the approximate values >= -err
};
// Budgeted rounds of the early-break pipeline, and the evaluation counters
// a directed run keeps for them: one per round, then the top completions,
// the tail, the tail's exact follow-up and the part of the order a sampled
// first round left for the second.
constexpr int kMaxRounds = 4;
// Evaluation counters of the early-break pipeline: one per budgeted round,
// then exact top completions, tail (filtered or exact-only), tail exact
// follow-up, the rest of a sampled first round, filtered top completions
// and their exact follow-up.
constexpr int kEbCounters = kMaxRounds + 6;
struct EbRoundStats {
  uint64_t round_points[kMaxRounds] = {0, 0, 0, 0};
  uint32_t mfma_rounds = 0;
  uint64_t top_points = 0;
  uint64_t top_filtered_points = 0;  // top completions on the filter
  uint64_t top_exact_points = 0;     // ... that no segment could drop
  uint64_t tail_points = 0;
  uint64_t exact_points = 0;    // tail entries no segment could drop
  bool tail_filtered = false;
  uint64_t rest_points = 0;     // order positions the sampled first round left
  bool rest_filtered = false;
};
// B repacked for the early-break scans in the visit order that `flags`
// select (HD_FLAG_NO_SHUFFLE: input order; HD_FLAG_RANDOM_B_ORDER or
// HD_FLAG_SINGLE_PASS: pseudo-random; else spatially stratified). mbr_b:
// B's MBR when the caller has it, else null.
int EbPrepareTarget(hd_context* ctx, int n_dims, const float* b, uint32_t n_b,
                    uint32_t flags, int32_t seed, const float* mbr_b,
                    EbTarget* t);
// Finishes the listed A points (indices, with optional partial minima over
// a subset of B) exactly with early break, raising *key; ev[0..kEbCounters)
// count evaluations of the rounds, the top completions and the tail.
int EbFinishList(hd_context* ctx, int n_dims, const float* a, uint32_t n_a,
                 const EbTarget& t, const uint32_t* list, const float* list_min,
                 uint32_t n_list, unsigned long long* key,
                 unsigned long long* ev, EbRoundStats* rs);

// Host early break (cpu_eb.cpp): the reference's CPU path on host AoS
// inputs with `parallelism` threads (<= 0: hardware_concurrency).
struct CpuRunInfo {
  int threads = 0;
  double shuffle_ms = 0;  // gather into visit order
  double compute_ms = 0;  // the scan (ReportedTime)
  const char* simd = "";
};
int HostThreads(int parallelism);
int CpuEarlyBreak(int n_dims, const float* a, uint64_t n_a, const float* b,
                  uint64_t n_b, uint32_t flags, int32_t seed, int parallelism,
                  hd_directed_result* out, CpuRunInfo* info);
// Raw MBR (lower[nd], upper[nd]) of host AoS points.
void HostPointsMbr(int n_dims, const float* pts, uint64_t n, float* mbr);

// Exact NN of queries against ref by index kind.
int NearestNeighbors(hd_context* ctx, int index, int n_dims, const float* q,
                     uint64_t n_q, const float* ref, uint64_t n_ref,
                     float* out_d2, uint32_t* out_idx);

// Builds the A visit order in descending predicted NN distance.
int LearnedOrder(hd_context* ctx, const DirectedArgs& args,
                 const uint32_t** order_out);

// Sums of the kCounterRow columns of the counter block at d_block (one
// synchronous copy of kCounterWords words).
int ReadCounters(hd_context* ctx, const unsigned long long* d_block,
                 unsigned long long* sums);

// Finishes a directed result from the device counter block key_dev: reads
// the key (word 0) and the n_counters evaluation counters at evals_dev (in
// the same block, summed over its rows), and finds the NN index of the
// arg-max.
int FinishDirected(hd_context* ctx, int n_dims, const float* a,
                   const float* b, uint64_t n_b,
                   const unsigned long long* key_dev,
                   const unsigned long long* evals_dev, int n_counters,
                   hd_directed_result* out);

// d_out[0] = the smallest index j of b with FP32 d2(a[argmax], b[j]) bits
// equal to cmax2_bits (UINT32_MAX if none); asynchronous, allocates nothing.
int LaunchNnIndex(hd_context* ctx, int n_dims, const float* a,
                  uint32_t argmax, const float* b, uint64_t n_b,
                  uint32_t cmax2_bits, uint32_t* d_out);

// Device memory of the single-pass early break when the caller provides it
// (the allocation-free entry point): no buffer is taken from the context.
struct EbScratch {
  unsigned long long* key;    // key, then the evaluation counter
  unsigned int* counters;     // 4: cursor of the round, escalation count, ...
  unsigned int* cursors;      // 1: the scan's segment cursor
  uint32_t* nn;               // 1
  uint32_t* esc_idx;          // 64 (never written: the budget covers all B)
  float* esc_min;             // 64
  float* bx;                  // max_bp each: B repacked (SoA, padded)
  float* by;
  float* bz;                  // 3D only
  uint32_t max_bp;
};
// One early-break pass over A in the visit order (HD_FLAG_SINGLE_PASS) with
// every buffer in `s`; synchronous; fills *out.
int EbSinglePassScratch(hd_context* ctx, int n_dims, const float* a,
                        uint32_t n_a, const float* b, uint32_t n_b,
                        bool shuffle, int32_t seed, const EbScratch& s,
                        hd_directed_result* out);

// The matrix-core filter's D for every pair (test hook, eb.hip): B in
// input order; params = centre[3], error bound E, 3D offset C.
int FilterProbe(hd_context* ctx, int n_dims, const float* a, uint32_t n_a,
                const float* b, uint32_t n_b, float* out, float* cna_out,
                float* params);

// Input statistics of the stats JSON (stats.hip): one point set on the
// device, AoS.
struct StatsInput {
  const float* pts;
  uint64_t n;
};
// {"NumPoints", "MBR", "Density", "Grid", "StatsGridNumPointsPerCell"} of the
// sets taken together; `mbr` is their raw MBR (lower[nd], upper[nd]). With
// n_points_cell 0 only NumPoints, MBR and Density; a non-null `density`
// replaces the sets' own.
int InputStatsJson(hd_context* ctx, int n_dims, const StatsInput* sets,
                   int n_sets, const float* mbr, uint32_t n_points_cell,
                   const float* density, std::string* json);
// Non-empty MBR rule of the stats (a zero extent becomes 1e-5) and the FP32
// volume.
void NonEmptyMbr(int n_dims, const float* mbr, float* lo, float* hi);
float MbrVolume(int n_dims, const float* lo, const float* hi);
// Squared lower and upper bounds of the directed HD A -> B from the MBRs
// (bounds.cpp).
void HdBounds2(int n_dims, const float* lo_a, const float* hi_a,
               const float* lo_b, const float* hi_b, float* lb2, float* ub2);

}  // namespace hd

#endif  // HD_INTERNAL_H
