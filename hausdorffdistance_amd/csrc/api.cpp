// C ABI of libhd.so (include/hd.h): argument validation, contexts, the
// RunHausdorffDistance driver with the reference's stats JSON, and dispatch
// to the variant pipelines.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "hd.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

namespace {
thread_local std::string g_last_error;

const char* AlgorithmName(int variant) {
  switch (variant) {
    case HD_VARIANT_EB:
      return "Early Break";
    case HD_VARIANT_GRID:
      return "Uniform Grid";
    case HD_VARIANT_LBVH:
      return "LBVH";
    case HD_VARIANT_HYBRID:
      return "Hybrid";
    case HD_VARIANT_NN:
      return "Nearest Neighbor Search";
    case HD_VARIANT_RT_HDIST:
      return "RT-HDIST";
  }
  return "Unknown";
}

std::string JsonEscape(const std::string& s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') {
      o.push_back('\\');
      o.push_back(c);
    } else if ((unsigned char)c < 0x20) {
      char buf[8];
      snprintf(buf, sizeof(buf), "\\u%04x", c);
      o += buf;
    } else {
      o.push_back(c);
    }
  }
  return o;
}

std::string Num(double v) {
  if (!std::isfinite(v)) return "null";
  char buf[64];
  snprintf(buf, sizeof(buf), "%.9g", v);
  return buf;
}

struct ScopedContext {
  hd_context* ctx = nullptr;
  bool owned = false;
  ~ScopedContext() {
    if (owned) hd_context_destroy(ctx);
  }
};

int AcquireContext(hd_context* given, ScopedContext* sc) {
  if (given) {
    sc->ctx = given;
    return ContextActivate(given);
  }
  hd_context* c = nullptr;
  int rc = hd_context_create(0, &c);
  if (rc != HD_OK) return rc;
  sc->ctx = c;
  sc->owned = true;
  return HD_OK;
}

bool ValidVariant(int v) { return v >= HD_VARIANT_EB && v <= HD_VARIANT_RT_HDIST; }

// The two visit orders of B exclude each other.
int ValidateBOrder(uint32_t flags) {
  if ((flags & HD_FLAG_B_ORDER_RANDOM) && (flags & HD_FLAG_B_ORDER_STRATIFIED))
    return Fail(HD_EINVAL, "B_ORDER_RANDOM and B_ORDER_STRATIFIED exclude each other");
  return HD_OK;
}

int ValidateSizes(uint64_t n_a, uint64_t n_b) {
  if (n_a == 0 || n_b == 0)
    return Fail(HD_EINVAL, "empty point set (|A|=%llu, |B|=%llu)",
                (unsigned long long)n_a, (unsigned long long)n_b);
  if (n_a > 0xFFFFFFFEull || n_b > 0xFFFFFFFEull)
    return Fail(HD_EINVAL, "point sets are limited to 2^32-2 points");
  return HD_OK;
}

int Dispatch(hd_context* ctx, int variant, const DirectedArgs& args,
             hd_directed_result* out) {
  std::memset(out, 0, sizeof(*out));
  out->variant_used = variant;
  switch (variant) {
    case HD_VARIANT_EB:
      return RunEarlyBreak(ctx, args, out);
    case HD_VARIANT_GRID:
      return RunGrid(ctx, args, out);
    case HD_VARIANT_LBVH:
      return RunLbvh(ctx, args, out);
    case HD_VARIANT_HYBRID:
      return RunHybrid(ctx, args, out);
    case HD_VARIANT_NN:
      return RunNearestNeighborSearch(ctx, args, out);
    case HD_VARIANT_RT_HDIST:
      return RunRtHdist(ctx, args, out);
  }
  return Fail(HD_EINVAL, "unknown variant %d", variant);
}

// Timed directed run on device inputs.
int TimedDirected(hd_context* ctx, int variant, const DirectedArgs& args,
                  hd_directed_result* out) {
  HD_HIP_TRY(hipEventRecord(ctx->ev_begin, ctx->stream));
  HD_TRY(Dispatch(ctx, variant, args, out));
  HD_HIP_TRY(hipEventRecord(ctx->ev_end, ctx->stream));
  HD_HIP_TRY(hipEventSynchronize(ctx->ev_end));
  float ms = 0.f;
  HD_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  out->time_ms = ms;
  return HD_OK;
}

// Check (run_hausdorff_distance.cu:366-388): the host early break with
// std::thread::hardware_concurrency() threads (:373) on host inputs; passes
// when cmax2 bits and arg-max are those of `out` in every direction run.
int CheckOnHost(const hd_run_config& cfg, const float* a, uint64_t n_a,
                const float* b, uint64_t n_b, hd_result* out,
                std::string* check_json) {
  const int nd = cfg.n_dims;
  const bool symmetric = (cfg.flags & HD_FLAG_SYMMETRIC) != 0;
  const uint32_t flags = cfg.flags & HD_FLAG_NO_SHUFFLE;
  hd_directed_result rab, rba;
  std::memset(&rba, 0, sizeof(rba));
  HD_TRY(CpuEarlyBreak(nd, a, n_a, b, n_b, flags, cfg.seed, 0, &rab, nullptr));
  if (symmetric)
    HD_TRY(CpuEarlyBreak(nd, b, n_b, a, n_a, flags, cfg.seed, 0, &rba, nullptr));
  const bool pass = rab.cmax2_bits == out->ab.cmax2_bits &&
                    rab.argmax == out->ab.argmax &&
                    (!symmetric || (rba.cmax2_bits == out->ba.cmax2_bits &&
                                    rba.argmax == out->ba.argmax));
  const hd_directed_result& cw =
      (symmetric && rba.cmax2_bits > rab.cmax2_bits) ? rba : rab;
  out->check_pass = pass ? 1 : 0;
  *check_json = ", \"Check\": {\"HDAnswer\": " + Num(cw.hd) +
                ", \"Diff\": " + Num((double)cw.hd - (double)out->hd) +
                ", \"Pass\": " + (pass ? "true" : "false") + "}";
  return HD_OK;
}

// execution = CPU: the early-break variant on the host (the reference's CPU
// overload, early_break.h:44-115), `repeat` times, no GPU touched. Other
// variants have no CPU overload in the reference either
// (hausdorff_distance.h:21-26).
int RunOnHost(const hd_run_config& cfg, const float* a, uint64_t n_a,
              const float* b, uint64_t n_b, hd_result* out, char* stats_json,
              uint64_t stats_cap) {
  if (cfg.variant != HD_VARIANT_EB)
    return Fail(HD_EUNSUPPORTED,
                "CalculateDistance on the CPU is not implemented for %s",
                AlgorithmName(cfg.variant));
  if (cfg.flags & HD_FLAG_INPUT_ON_DEVICE)
    return Fail(HD_EINVAL, "CPU execution takes host inputs");
  if (cfg.flags & HD_FLAG_LEARNED_ORDER)
    return Fail(HD_EUNSUPPORTED,
                "the learned visit order needs the GPU (execution = GPU)");
  const int nd = cfg.n_dims;
  const bool symmetric = (cfg.flags & HD_FLAG_SYMMETRIC) != 0;
  const uint32_t flags = cfg.flags & HD_FLAG_NO_SHUFFLE;
  std::vector<double> times;
  std::ostringstream reps;
  for (int r = 0; r < cfg.repeat; ++r) {
    hd_directed_result rab, rba;
    std::memset(&rba, 0, sizeof(rba));
    CpuRunInfo iab, iba;
    HD_TRY(CpuEarlyBreak(nd, a, n_a, b, n_b, flags, cfg.seed, cfg.parallelism,
                         &rab, &iab));
    double t = rab.time_ms, shuffle = iab.shuffle_ms;
    if (symmetric) {
      HD_TRY(CpuEarlyBreak(nd, b, n_b, a, n_a, flags, cfg.seed,
                           cfg.parallelism, &rba, &iba));
      t += rba.time_ms;
      shuffle += iba.shuffle_ms;
    }
    times.push_back(t);
    out->ab = rab;
    out->ba = rba;
    if (r) reps << ", ";
    reps << "{\"Repeat\": " << (r + 1)
         << ", \"Algorithm\": \"Early Break\", \"Execution\": \"CPU\""
         << ", \"Threads\": " << iab.threads << ", \"SIMD\": \"" << iab.simd
         << "\", \"ComparedPairs\": " << (rab.evals + rba.evals)
         << ", \"ShuffleTime\": " << Num(shuffle)
         << ", \"ComputeTime\": " << Num(t) << ", \"ReportedTime\": " << Num(t)
         << ", \"CMax2Bits\": " << rab.cmax2_bits << ", \"ArgMax\": "
         << rab.argmax << "}";
  }
  out->n_repeats = (uint32_t)cfg.repeat;
  const bool ba_wins = symmetric && out->ba.cmax2_bits > out->ab.cmax2_bits;
  const hd_directed_result& win = ba_wins ? out->ba : out->ab;
  out->hd = win.hd;
  out->cmax2_bits = win.cmax2_bits;
  out->argmax = win.argmax;
  out->nn = win.nn;
  out->direction = ba_wins ? 1 : 0;
  std::vector<double> sorted = times;
  std::sort(sorted.begin(), sorted.end());
  double sum = 0;
  for (double t : times) sum += t;
  out->avg_time_ms = sum / times.size();
  out->min_time_ms = sorted.front();
  out->median_time_ms = sorted[sorted.size() / 2];

  std::string check_json;
  if (cfg.flags & HD_FLAG_CHECK)
    HD_TRY(CheckOnHost(cfg, a, n_a, b, n_b, out, &check_json));

  if (stats_json && stats_cap > 0) {
    // Input object from host MBRs (the statistics grid needs the GPU and is
    // left out on this path).
    float mbr_a[6], mbr_b[6], mbr_ab[6], lo_a[3], hi_a[3], lo_b[3], hi_b[3],
        lo[3], hi[3];
    HostPointsMbr(nd, a, n_a, mbr_a);
    HostPointsMbr(nd, b, n_b, mbr_b);
    NonEmptyMbr(nd, mbr_a, lo_a, hi_a);
    NonEmptyMbr(nd, mbr_b, lo_b, hi_b);
    for (int d = 0; d < nd; ++d) {
      mbr_ab[d] = std::min(mbr_a[d], mbr_b[d]);
      mbr_ab[nd + d] = std::max(mbr_a[nd + d], mbr_b[nd + d]);
      lo[d] = std::min(lo_a[d], lo_b[d]);
      hi[d] = std::max(hi_a[d], hi_b[d]);
    }
    const float density = (float)(n_a + n_b) / MbrVolume(nd, lo, hi);
    const StatsInput sets[2] = {{a, n_a}, {b, n_b}};
    std::string files[2], both;
    HD_TRY(InputStatsJson(nullptr, nd, &sets[0], 1, mbr_a, 0, nullptr, &files[0]));
    HD_TRY(InputStatsJson(nullptr, nd, &sets[1], 1, mbr_b, 0, nullptr, &files[1]));
    HD_TRY(InputStatsJson(nullptr, nd, sets, 2, mbr_ab, 0, &density, &both));
    both.pop_back();
    std::ostringstream js;
    js << "{\"Input\": " << both << ", \"Files\": [" << files[0] << ", "
       << files[1] << "], \"NumDims\": " << nd << ", \"Type\": \"Float\"}"
       << ", \"Running\": {\"Seed\": " << cfg.seed
       << ", \"Parallelism\": " << HostThreads(cfg.parallelism)
       << ", \"Symmetric\": " << (symmetric ? "true" : "false")
       << ", \"Repeats\": [" << reps.str()
       << "], \"AvgTime\": " << Num(out->avg_time_ms)
       << ", \"MedianTime\": " << Num(out->median_time_ms) << "}"
       << ", \"HDResult\": " << Num(out->hd)
       << ", \"CMax2Bits\": " << out->cmax2_bits
       << ", \"ArgMax\": " << out->argmax << ", \"NN\": " << out->nn
       << ", \"Direction\": " << out->direction << check_json << "}";
    const std::string s = js.str();
    std::strncpy(stats_json, s.c_str(), stats_cap - 1);
    stats_json[stats_cap - 1] = '\0';
    if (s.size() + 1 > stats_cap)
      return Fail(HD_ETRUNC, "stats JSON needs %zu bytes", s.size() + 1);
  }
  return HD_OK;
}

}  // namespace

void SetError(const std::string& msg) { g_last_error = msg; }
void ClearError() { g_last_error.clear(); }

int Fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int ContextActivate(hd_context* ctx) {
  HD_HIP_TRY(hipSetDevice(ctx->device));
  return HD_OK;
}

int ContextBuffer(hd_context* ctx, const char* name, size_t bytes,
                  void** out) {
  DeviceBuffer& b = ctx->buffers[name];
  if (b.bytes < bytes) {
    const size_t want = std::max<size_t>(bytes, 256);
    if (ctx->memory_limit && ctx->memory_held - b.bytes + want > ctx->memory_limit)
      return Fail(HD_ENOMEM,
                  "context memory limit of %llu bytes reached: %s needs %zu bytes, "
                  "%llu held by other buffers",
                  (unsigned long long)ctx->memory_limit, name, want,
                  (unsigned long long)(ctx->memory_held - b.bytes));
    if (b.ptr) {
      // The stream may still use the old buffer.
      HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
      HD_HIP_TRY(hipFree(b.ptr));
      ctx->memory_held -= b.bytes;
      b.ptr = nullptr;
      b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
      b.ptr = nullptr;
      (void)hipGetLastError();
      return Fail(HD_ENOMEM, "hipMalloc(%zu) for %s failed: %s", want, name,
                  hipGetErrorString(e));
    }
    b.bytes = want;
    ctx->memory_held += want;
  }
  *out = b.ptr;
  return HD_OK;
}

int ContextPinned(hd_context* ctx, size_t bytes, void** out) {
  if (ctx->host_pinned_bytes < bytes) {
    if (ctx->host_pinned) {
      HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
      HD_HIP_TRY(hipHostFree(ctx->host_pinned));
      ctx->host_pinned = nullptr;
    }
    size_t want = std::max<size_t>(bytes, kPinnedMinBytes);
    HD_HIP_TRY(hipHostMalloc(&ctx->host_pinned, want, hipHostMallocDefault));
    ctx->host_pinned_bytes = want;
  }
  *out = ctx->host_pinned;
  return HD_OK;
}

int StageBegin(hd_context* ctx, const char* name) {
  if (!ctx->profiling) return -1;
  PendingStage st;
  st.name = name;
  for (hipEvent_t* ev : {&st.begin, &st.end}) {
    if (!ctx->event_pool.empty()) {
      *ev = ctx->event_pool.back();
      ctx->event_pool.pop_back();
    } else if (hipEventCreate(ev) != hipSuccess) {
      (void)hipGetLastError();
      return -1;
    }
  }
  if (hipEventRecord(st.begin, ctx->stream) != hipSuccess) {
    (void)hipGetLastError();
    ctx->event_pool.push_back(st.begin);
    ctx->event_pool.push_back(st.end);
    return -1;
  }
  ctx->pending_stages.push_back(st);
  ctx->open_stage = (int)ctx->pending_stages.size() - 1;
  return ctx->open_stage;
}

void StageEnd(hd_context* ctx, int handle) {
  if (handle < 0 || handle >= (int)ctx->pending_stages.size()) return;
  if (ctx->open_stage == handle) ctx->open_stage = -1;
  if (hipEventRecord(ctx->pending_stages[handle].end, ctx->stream) !=
      hipSuccess)
    (void)hipGetLastError();
}

void StageCount(hd_context* ctx, const char* name, uint64_t evals,
                uint64_t points, uint64_t lookups) {
  if (!ctx->profiling) return;
  StageStat& s = ctx->stage_stats[name];
  s.evals += evals;
  s.points += points;
  s.lookups += lookups;
}

int ResolveStages(hd_context* ctx) {
  if (ctx->pending_stages.empty()) return HD_OK;
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (PendingStage& st : ctx->pending_stages) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, st.begin, st.end) == hipSuccess) {
      StageStat& s = ctx->stage_stats[st.name];
      s.total_ms += ms;
      s.calls += 1;
      s.dispatches += st.dispatches;
    } else {
      (void)hipGetLastError();
    }
    ctx->event_pool.push_back(st.begin);
    ctx->event_pool.push_back(st.end);
  }
  ctx->pending_stages.clear();
  ctx->open_stage = -1;
  return HD_OK;
}

}  // namespace hd

using namespace hd;

extern "C" {

const char* hd_last_error(void) { return g_last_error.c_str(); }

void hd_run_config_init(hd_run_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->abi_version = HD_ABI_VERSION;
  cfg->variant = HD_VARIANT_EB;
  cfg->execution = HD_EXEC_GPU;
  cfg->n_dims = 0;
  cfg->repeat = 5;
  cfg->seed = 0;
  cfg->parallelism = -1;
  cfg->flags = 0;
  cfg->sample_rate = 0.01f;
  cfg->n_points_cell = 8;
  cfg->max_hit = 256;
  cfg->eb_only_threshold = 0;
  cfg->auto_tune = 0;
  cfg->stats_n_points_cell = 10;
  cfg->model = nullptr;
}

int hd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int hd_device_info(int device, char* buf, uint64_t cap) {
  ClearError();
  if (!buf || cap == 0) return Fail(HD_EINVAL, "null buffer");
  hipDeviceProp_t prop;
  HD_HIP_TRY(hipGetDeviceProperties(&prop, device));
  std::ostringstream os;
  os << "{\"Device\": " << device << ", \"name\": \"" << JsonEscape(prop.name)
     << "\", \"gcnArchName\": \"" << JsonEscape(prop.gcnArchName)
     << "\", \"multiProcessorCount\": " << prop.multiProcessorCount
     << ", \"clockRateKHz\": " << prop.clockRate
     << ", \"memoryClockRateKHz\": " << prop.memoryClockRate
     << ", \"memoryBusWidth\": " << prop.memoryBusWidth
     << ", \"totalGlobalMemBytes\": " << prop.totalGlobalMem << "}";
  std::string s = os.str();
  std::strncpy(buf, s.c_str(), cap - 1);
  buf[cap - 1] = '\0';
  return s.size() + 1 > cap ? HD_ETRUNC : HD_OK;
}

int hd_context_create(int device, hd_context** out) {
  ClearError();
  if (!out) return Fail(HD_EINVAL, "null output pointer");
  *out = nullptr;
  int n = hd_device_count();
  if (n <= 0) return Fail(HD_EDEVICE, "no HIP device is visible");
  if (device < 0 || device >= n)
    return Fail(HD_EINVAL, "device %d out of range (have %d)", device, n);
  std::unique_ptr<hd_context> ctx(new hd_context());
  ctx->device = device;
  HD_HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HD_HIP_TRY(hipGetDeviceProperties(&prop, device));
  ctx->cu_count = prop.multiProcessorCount;
  HD_HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HD_HIP_TRY(hipEventCreate(&ctx->ev_begin));
  HD_HIP_TRY(hipEventCreate(&ctx->ev_end));
  *out = ctx.release();
  return HD_OK;
}

void hd_context_destroy(hd_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->cached_model) hd_mlp_destroy(ctx->cached_model);
  if (ctx->batch_plan) DestroyBatchPlan(ctx->batch_plan);
  for (auto& kv : ctx->buffers)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  for (auto& st : ctx->pending_stages) {
    (void)hipEventDestroy(st.begin);
    (void)hipEventDestroy(st.end);
  }
  for (hipEvent_t ev : ctx->event_pool) (void)hipEventDestroy(ev);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  if (ctx->ev_begin) (void)hipEventDestroy(ctx->ev_begin);
  if (ctx->ev_end) (void)hipEventDestroy(ctx->ev_end);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int hd_context_set_memory_limit(hd_context* ctx, uint64_t bytes) {
  ClearError();
  if (!ctx) return Fail(HD_EINVAL, "null context");
  HD_TRY(ContextActivate(ctx));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  // The batch plan's graph holds pointers into the buffers released below.
  if (ctx->batch_plan) {
    DestroyBatchPlan(ctx->batch_plan);
    ctx->batch_plan = nullptr;
  }
  for (auto& kv : ctx->buffers)
    if (kv.second.ptr) HD_HIP_TRY(hipFree(kv.second.ptr));
  ctx->buffers.clear();
  ctx->memory_held = 0;
  ctx->memory_limit = bytes;
  return HD_OK;
}

uint64_t hd_context_memory_held(hd_context* ctx) { return ctx ? ctx->memory_held : 0; }

int hd_context_set_profiling(hd_context* ctx, int enable) {
  ClearError();
  if (!ctx) return Fail(HD_EINVAL, "null context");
  HD_TRY(ResolveStages(ctx));
  ctx->profiling = enable != 0;
  return HD_OK;
}

int hd_context_stage_times(hd_context* ctx, char* json, uint64_t cap,
                           int reset) {
  ClearError();
  if (!ctx || !json || cap == 0) return Fail(HD_EINVAL, "null argument");
  HD_TRY(ResolveStages(ctx));
  std::ostringstream os;
  os << "{";
  bool first = true;
  for (const auto& kv : ctx->stage_stats) {
    os << (first ? "" : ", ") << "\"" << JsonEscape(kv.first)
       << "\": {\"ms\": " << Num(kv.second.total_ms)
       << ", \"calls\": " << kv.second.calls
       << ", \"dispatches\": " << kv.second.dispatches
       << ", \"evals\": " << kv.second.evals
       << ", \"points\": " << kv.second.points
       << ", \"lookups\": " << kv.second.lookups << "}";
    first = false;
  }
  os << "}";
  if (reset) ctx->stage_stats.clear();
  const std::string s = os.str();
  std::strncpy(json, s.c_str(), cap - 1);
  json[cap - 1] = '\0';
  return s.size() + 1 > cap ? Fail(HD_ETRUNC, "stage JSON truncated") : HD_OK;
}

void* hd_context_stream(hd_context* ctx) {
  return ctx ? static_cast<void*>(ctx->stream) : nullptr;
}

int hd_directed_device(hd_context* ctx, int variant, int n_dims,
                       const float* d_a, uint64_t n_a, const float* d_b,
                       uint64_t n_b, uint32_t flags, int32_t seed,
                       const hd_mlp* model, hd_directed_result* out) {
  ClearError();
  if (!out) return Fail(HD_EINVAL, "null result pointer");
  std::memset(out, 0, sizeof(*out));
  if (n_dims != 2 && n_dims != 3)
    return Fail(HD_EINVAL, "n_dims must be 2 or 3, got %d", n_dims);
  if (!ValidVariant(variant))
    return Fail(HD_EINVAL, "unknown variant %d", variant);
  if ((flags & HD_FLAG_SINGLE_PASS) && variant != HD_VARIANT_EB)
    return Fail(HD_EINVAL, "SINGLE_PASS applies to the early-break variant");
  HD_TRY(ValidateBOrder(flags));
  HD_TRY(ValidateSizes(n_a, n_b));
  if (!d_a || !d_b) return Fail(HD_EINVAL, "null point array");
  ScopedContext sc;
  HD_TRY(AcquireContext(ctx, &sc));
  float mbr_a[6], mbr_b[6];
  HD_TRY(HostMbrPair(sc.ctx, n_dims, d_a, n_a, mbr_a, d_b, n_b, mbr_b));
  DirectedArgs args;
  args.n_dims = n_dims;
  args.a = d_a;
  args.n_a = n_a;
  args.b = d_b;
  args.n_b = n_b;
  args.flags = flags;
  args.seed = seed;
  args.model = model;
  args.has_mbr = true;
  std::memcpy(args.mbr_a, mbr_a, sizeof(mbr_a));
  std::memcpy(args.mbr_b, mbr_b, sizeof(mbr_b));
  return TimedDirected(sc.ctx, variant, args, out);
}

int hd_run_hausdorff_distance(hd_context* ctx, const hd_run_config* cfg,
                              const float* a, uint64_t n_a, const float* b,
                              uint64_t n_b, hd_result* out, char* stats_json,
                              uint64_t stats_cap) {
  ClearError();
  if (!cfg || !out) return Fail(HD_EINVAL, "null config or result");
  std::memset(out, 0, sizeof(*out));
  out->check_pass = -1;
  if (cfg->abi_version != HD_ABI_VERSION)
    return Fail(HD_EINVAL, "abi_version %u, library speaks %u",
                cfg->abi_version, HD_ABI_VERSION);
  if (cfg->n_dims != 2 && cfg->n_dims != 3)
    return Fail(HD_EINVAL,
                "Wrong number of dimensions, which can only be 2 or 3 (got %d)",
                cfg->n_dims);
  if (!ValidVariant(cfg->variant))
    return Fail(HD_EINVAL, "Unknown variant: %d", cfg->variant);
  if (cfg->execution != HD_EXEC_GPU && cfg->execution != HD_EXEC_CPU)
    return Fail(HD_EINVAL, "Unknown execution: %d", cfg->execution);
  if (cfg->repeat < 1)
    return Fail(HD_EINVAL, "repeat must be >= 1, got %d", cfg->repeat);
  HD_TRY(ValidateBOrder(cfg->flags));
  HD_TRY(ValidateSizes(n_a, n_b));
  if (!a || !b) return Fail(HD_EINVAL, "null point array");
  if (cfg->execution == HD_EXEC_CPU)
    return RunOnHost(*cfg, a, n_a, b, n_b, out, stats_json, stats_cap);
  const int nd = cfg->n_dims;

  ScopedContext sc;
  HD_TRY(AcquireContext(ctx, &sc));
  hd_context* c = sc.ctx;

  // Upload once, before the repeats (reference run_hausdorff_distance.cu:148).
  const float* d_a = a;
  const float* d_b = b;
  if (!(cfg->flags & HD_FLAG_INPUT_ON_DEVICE)) {
    float *ua = nullptr, *ub = nullptr;
    HD_TRY(ContextArray(c, "input_a", n_a * nd, &ua));
    HD_TRY(ContextArray(c, "input_b", n_b * nd, &ub));
    HD_HIP_TRY(hipMemcpyAsync(ua, a, n_a * nd * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
    HD_HIP_TRY(hipMemcpyAsync(ub, b, n_b * nd * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
    d_a = ua;
    d_b = ub;
  }
  float mbr_a[6], mbr_b[6];
  HD_TRY(HostMbrPair(c, nd, d_a, n_a, mbr_a, d_b, n_b, mbr_b));

  DirectedArgs ab;
  ab.n_dims = nd;
  ab.a = d_a;
  ab.n_a = n_a;
  ab.b = d_b;
  ab.n_b = n_b;
  ab.flags = cfg->flags;
  ab.seed = cfg->seed;
  ab.model = static_cast<const hd_mlp*>(cfg->model);
  ab.n_points_cell = cfg->n_points_cell;
  ab.max_hit = cfg->max_hit;
  ab.sample_rate = cfg->sample_rate;
  ab.eb_only_threshold = cfg->eb_only_threshold;
  ab.auto_tune = cfg->auto_tune;
  ab.has_mbr = true;
  std::memcpy(ab.mbr_a, mbr_a, sizeof(mbr_a));
  std::memcpy(ab.mbr_b, mbr_b, sizeof(mbr_b));
  DirectedArgs ba = ab;
  std::swap(ba.a, ba.b);
  std::swap(ba.n_a, ba.n_b);
  std::swap(ba.mbr_a, ba.mbr_b);
  const bool symmetric = (cfg->flags & HD_FLAG_SYMMETRIC) != 0;

  // MBR bounds of the directed HD (hausdorff_distance_hybrid.h:95-106),
  // written into every repeat.
  float hd_lb = 0.f, hd_ub = 0.f;
  {
    float lo_a[3], hi_a[3], lo_b[3], hi_b[3], lb2, ub2;
    NonEmptyMbr(nd, mbr_a, lo_a, hi_a);
    NonEmptyMbr(nd, mbr_b, lo_b, hi_b);
    HdBounds2(nd, lo_a, hi_a, lo_b, hi_b, &lb2, &ub2);
    hd_lb = std::sqrt(lb2);
    hd_ub = std::sqrt(ub2);
  }

  std::vector<double> times;
  std::ostringstream reps;
  for (int r = 0; r < cfg->repeat; ++r) {
    hd_directed_result rab, rba;
    std::memset(&rba, 0, sizeof(rba));
    HD_TRY(TimedDirected(c, cfg->variant, ab, &rab));
    const TunedChoice choice_ab = c->last_choice;
    TunedChoice choice_ba;
    double t = rab.time_ms;
    if (symmetric) {
      HD_TRY(TimedDirected(c, cfg->variant, ba, &rba));
      choice_ba = c->last_choice;
      t += rba.time_ms;
    }
    times.push_back(t);
    out->ab = rab;
    out->ba = rba;
    if (r) reps << ", ";
    reps << "{\"Repeat\": " << (r + 1) << ", \"Algorithm\": \""
         << AlgorithmName(rab.variant_used)
         << "\", \"Execution\": \"GPU\", \"ReportedTime\": " << Num(t)
         << ", \"ComparedPairs\": " << (rab.evals + rba.evals)
         << ", \"CMax2Bits\": " << rab.cmax2_bits
         << ", \"ArgMax\": " << rab.argmax
         << ", \"HDLowerBound\": " << Num(hd_lb)
         << ", \"HDUpperBound\": " << Num(hd_ub);
    if (cfg->variant == HD_VARIANT_HYBRID) {
      // The hybrid's parameters (the reference writes the predicted ones,
      // hausdorff_distance_hybrid.h:744-776): per direction, and the A->B
      // ones under the reference's keys.
      auto choice_json = [](const TunedChoice& ch) {
        std::ostringstream o;
        o << "{\"Variant\": \"" << AlgorithmName(ch.variant)
          << "\", \"NumPointsPerCell\": " << ch.n_points_cell
          << ", \"MaxHit\": " << ch.max_hit
          << ", \"Tuned\": " << (ch.tuned ? "true" : "false")
          << ", \"Trials\": " << ch.trials << "}";
        return o.str();
      };
      if (choice_ab.tuned)
        reps << ", \"PredictedNPointsCell\": " << choice_ab.n_points_cell
             << ", \"PredictedMaxHit\": " << choice_ab.max_hit;
      reps << ", \"HybridChoice\": {\"AB\": " << choice_json(choice_ab);
      if (symmetric) reps << ", \"BA\": " << choice_json(choice_ba);
      reps << "}";
    }
    reps << "}";
  }
  // Result.
  out->n_repeats = (uint32_t)cfg->repeat;
  const bool ba_wins = symmetric && out->ba.cmax2_bits > out->ab.cmax2_bits;
  const hd_directed_result& win = ba_wins ? out->ba : out->ab;
  out->hd = win.hd;
  out->cmax2_bits = win.cmax2_bits;
  out->argmax = win.argmax;
  out->nn = win.nn;
  out->direction = ba_wins ? 1 : 0;
  double sum = 0;
  for (double t : times) sum += t;
  std::vector<double> sorted = times;
  std::sort(sorted.begin(), sorted.end());
  out->avg_time_ms = sum / times.size();
  out->min_time_ms = sorted.front();
  out->median_time_ms = sorted[sorted.size() / 2];

  // Check (reference run_hausdorff_distance.cu:366-388): the host early
  // break with hardware_concurrency threads on the same inputs.
  std::string check_json;
  if (cfg->flags & HD_FLAG_CHECK) {
    const float* ha = a;
    const float* hb = b;
    std::vector<float> copy_a, copy_b;
    if (cfg->flags & HD_FLAG_INPUT_ON_DEVICE) {
      copy_a.resize(n_a * nd);
      copy_b.resize(n_b * nd);
      HD_HIP_TRY(hipMemcpy(copy_a.data(), d_a, n_a * nd * sizeof(float),
                           hipMemcpyDeviceToHost));
      HD_HIP_TRY(hipMemcpy(copy_b.data(), d_b, n_b * nd * sizeof(float),
                           hipMemcpyDeviceToHost));
      ha = copy_a.data();
      hb = copy_b.data();
    }
    HD_TRY(CheckOnHost(*cfg, ha, n_a, hb, n_b, out, &check_json));
  }

  if (stats_json && stats_cap > 0) {
    hipDeviceProp_t prop;
    HD_HIP_TRY(hipGetDeviceProperties(&prop, c->device));
    // Input statistics (run_hausdorff_distance.cu:151-204): per file and for
    // both files together; the top-level density is over the union of the
    // two non-empty MBRs.
    std::string files[2], both;
    float mbr_ab[6], lo_a[3], hi_a[3], lo_b[3], hi_b[3], lo[3], hi[3];
    NonEmptyMbr(nd, mbr_a, lo_a, hi_a);
    NonEmptyMbr(nd, mbr_b, lo_b, hi_b);
    for (int d = 0; d < nd; ++d) {
      mbr_ab[d] = std::min(mbr_a[d], mbr_b[d]);
      mbr_ab[nd + d] = std::max(mbr_a[nd + d], mbr_b[nd + d]);
      lo[d] = std::min(lo_a[d], lo_b[d]);
      hi[d] = std::max(hi_a[d], hi_b[d]);
    }
    const float density = (float)(n_a + n_b) / MbrVolume(nd, lo, hi);
    const StatsInput sets[2] = {{d_a, n_a}, {d_b, n_b}};
    const uint32_t spc = (uint32_t)std::max(cfg->stats_n_points_cell, 0);
    const int stage = StageBegin(c, "input_stats");
    HD_TRY(InputStatsJson(c, nd, &sets[0], 1, mbr_a, spc, nullptr, &files[0]));
    HD_TRY(InputStatsJson(c, nd, &sets[1], 1, mbr_b, spc, nullptr, &files[1]));
    HD_TRY(InputStatsJson(c, nd, sets, 2, mbr_ab, spc, &density, &both));
    StageEnd(c, stage);
    StageCount(c, "input_stats", 0, 2 * (n_a + n_b));
    both.pop_back();  // the union's object, extended below
    std::ostringstream js;
    js << "{\"GPU\": {\"Device\": " << c->device << ", \"name\": \""
       << JsonEscape(prop.name) << "\"}, \"Input\": " << both
       << ", \"Files\": [" << files[0]
       << ", " << files[1] << "], \"NumDims\": " << nd
       << ", \"Type\": \"Float\"}"
       << ", \"Running\": {\"Seed\": " << cfg->seed
       << ", \"SampleRate\": " << Num(cfg->sample_rate)
       << ", \"NumPointsPerCell\": " << cfg->n_points_cell
       << ", \"MaxHit\": " << cfg->max_hit
       << ", \"EBOnlyThreshold\": " << cfg->eb_only_threshold
       << ", \"AutoTune\": " << (cfg->auto_tune ? "true" : "false")
       << ", \"Symmetric\": " << (symmetric ? "true" : "false")
       << ", \"LearnedOrder\": "
       << ((cfg->flags & HD_FLAG_LEARNED_ORDER) ? "true" : "false")
       << ", \"Repeats\": [" << reps.str()
       << "], \"AvgTime\": " << Num(out->avg_time_ms)
       << ", \"MedianTime\": " << Num(out->median_time_ms) << "}"
       << ", \"HDResult\": " << Num(out->hd)
       << ", \"CMax2Bits\": " << out->cmax2_bits
       << ", \"ArgMax\": " << out->argmax << ", \"NN\": " << out->nn
       << ", \"Direction\": " << out->direction << check_json << "}";
    const std::string s = js.str();
    std::strncpy(stats_json, s.c_str(), stats_cap - 1);
    stats_json[stats_cap - 1] = '\0';
    if (s.size() + 1 > stats_cap) {
      Fail(HD_ETRUNC, "stats JSON needs %zu bytes", s.size() + 1);
      return HD_ETRUNC;
    }
  }
  return HD_OK;
}

int hd_filter_probe_device(hd_context* ctx, int n_dims, const float* d_a,
                           uint64_t n_a, const float* d_b, uint64_t n_b,
                           float* d_out, float* d_cna, float* params) {
  ClearError();
  if (n_dims != 2 && n_dims != 3)
    return Fail(HD_EINVAL, "n_dims must be 2 or 3, got %d", n_dims);
  if (!d_a || !d_b || !d_out || !params || (n_dims == 3 && !d_cna))
    return Fail(HD_EINVAL, "null pointer");
  if (n_a == 0 || n_b == 0 || n_a > (1u << 16) || n_b > (1u << 16))
    return Fail(HD_EINVAL, "the probe takes 1..65536 points per set");
  ScopedContext sc;
  HD_TRY(AcquireContext(ctx, &sc));
  return FilterProbe(sc.ctx, n_dims, d_a, (uint32_t)n_a, d_b, (uint32_t)n_b,
                     d_out, d_cna, params);
}

int hd_nn_device(hd_context* ctx, int index, int n_dims, const float* d_q,
                 uint64_t n_q, const float* d_ref, uint64_t n_ref,
                 float* d_out_d2, uint32_t* d_out_idx) {
  ClearError();
  if (n_dims != 2 && n_dims != 3)
    return Fail(HD_EINVAL, "n_dims must be 2 or 3, got %d", n_dims);
  if (index != HD_VARIANT_EB && index != HD_VARIANT_GRID &&
      index != HD_VARIANT_LBVH)
    return Fail(HD_EINVAL, "index must be EB, GRID or LBVH, got %d", index);
  HD_TRY(ValidateSizes(n_q, n_ref));
  if (!d_q || !d_ref || !d_out_d2 || !d_out_idx)
    return Fail(HD_EINVAL, "null pointer");
  ScopedContext sc;
  HD_TRY(AcquireContext(ctx, &sc));
  float mbr[6];
  HD_TRY(HostMbr(sc.ctx, n_dims, d_q, n_q, mbr));
  HD_TRY(HostMbr(sc.ctx, n_dims, d_ref, n_ref, mbr));
  HD_TRY(NearestNeighbors(sc.ctx, index, n_dims, d_q, n_q, d_ref, n_ref,
                          d_out_d2, d_out_idx));
  HD_HIP_TRY(hipStreamSynchronize(sc.ctx->stream));
  return HD_OK;
}

}  // extern "C"

extern "C" int hd_hybrid_sample_iterations(hd_context* ctx, int n_dims, const float* d_a,
                                           uint64_t n_a, const float* d_b, uint64_t n_b,
                                           int32_t seed, float sample_rate,
                                           uint32_t* counts, uint64_t cap, uint64_t* n_out) {
  ClearError();
  if (n_dims != 2 && n_dims != 3)
    return Fail(HD_EINVAL, "n_dims must be 2 or 3, got %d", n_dims);
  if (!d_a || !d_b || !n_out) return Fail(HD_EINVAL, "null pointer");
  HD_TRY(ValidateSizes(n_a, n_b));
  if (!(sample_rate >= 0.f && sample_rate <= 1.f))
    return Fail(HD_EINVAL, "sample_rate must be in [0, 1]");
  ScopedContext sc;
  HD_TRY(AcquireContext(ctx, &sc));
  DirectedArgs args;
  args.n_dims = n_dims;
  args.a = d_a;
  args.n_a = n_a;
  args.b = d_b;
  args.n_b = n_b;
  args.seed = seed;
  args.sample_rate = sample_rate;
  std::vector<uint32_t> v;
  HD_TRY(SampledEbIterations(sc.ctx, args, &v));
  *n_out = v.size();
  if (counts) std::memcpy(counts, v.data(), std::min<uint64_t>(cap, v.size()) * 4);
  if (counts && v.size() > cap)
    return Fail(HD_ETRUNC, "%zu counts, room for %llu", v.size(), (unsigned long long)cap);
  return HD_OK;
}
