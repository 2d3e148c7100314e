// The allocation-free directed early break (hd_directed_device_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops) get the single-pass early break with every buffer in
// one file-scope device array: the key and evaluation counter, the scan's
// cursor, the NN result, the MBR workspace and B repacked into padded SoA
// for at most kMaxB points. The kernels are the library's own (the exact
// scan of the early-break pipeline, the repack, the fills and the NN pass);
// they receive the array's device address as their buffer arguments.
//
// The array is one per GPU and lives until the process ends. Every call on a
// GPU uses the same array and leaves its contents behind, so each call first
// clears what must start at zero (with fill kernels), and calls that overlap
// on one GPU race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "hd.h"
#include "hd_device.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

namespace {

constexpr uint64_t kMaxB = 1ull << 20;  // B points (2^20 is a whole number of
                                        // 64-point steps: no padding beyond)
// Byte offsets inside the array.
constexpr size_t kOffKey = 0;         // counter block: u64[kCounterWords]
                                      // (key, evaluation counter), 4 KiB
constexpr size_t kOffCounters = 4096; // u32[4]
constexpr size_t kOffCursors = 4112;  // u32[1]
constexpr size_t kOffNn = 4116;       // u32[1]
constexpr size_t kOffMbr = 4224;      // u32[32] (one MBR slot)
constexpr size_t kOffEscIdx = 4352;   // u32[64]
constexpr size_t kOffEscMin = 4608;   // f32[64]
constexpr size_t kOffSoA = 5120;      // f32[kMaxB] x, y, z
static_assert(kOffKey + sizeof(unsigned long long) * dev::kCounterWords <= kOffCounters,
              "counter block");
constexpr size_t kScratchBytes = kOffSoA + 3 * kMaxB * sizeof(float);

// 12 MiB + 5 KiB per GPU. In host code this name stands for the array but is
// not its device address: kernels and copies use the address that
// hipGetSymbolAddress returns for the current device.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// Device address of the array on ctx's device (made current by the caller),
// looked up once per device.
int ScratchAddress(hd_context* ctx, unsigned char** out) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return Fail(HD_EINVAL, "device %d out of range", ctx->device);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[ctx->device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      return Fail(HD_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                  hipGetErrorString(e));
    }
    g_scratch[ctx->device] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch[ctx->device]);
  return HD_OK;
}

}  // namespace

}  // namespace hd

using namespace hd;

extern "C" int hd_directed_device_noalloc(hd_context* ctx, int n_dims,
                                          const float* d_a, uint64_t n_a,
                                          const float* d_b, uint64_t n_b,
                                          uint32_t flags, int32_t seed,
                                          hd_directed_result* out) {
  ClearError();
  if (!out) return Fail(HD_EINVAL, "null result pointer");
  std::memset(out, 0, sizeof(*out));
  if (!ctx)
    return Fail(HD_EINVAL,
                "the allocation-free entry point needs a context (making one "
                "would allocate)");
  if (n_dims != 2 && n_dims != 3)
    return Fail(HD_EINVAL, "n_dims must be 2 or 3, got %d", n_dims);
  if (!d_a || !d_b) return Fail(HD_EINVAL, "null point array");
  if (n_a == 0 || n_b == 0)
    return Fail(HD_EINVAL, "empty point set (|A|=%llu, |B|=%llu)",
                (unsigned long long)n_a, (unsigned long long)n_b);
  if (n_a > 0xFFFFFFFEull)
    return Fail(HD_EINVAL, "point sets are limited to 2^32-2 points");
  if (n_b > kMaxB)
    return Fail(HD_EINVAL,
                "|B| = %llu exceeds the allocation-free limit of %llu points "
                "(its scratch holds B repacked); use hd_directed_device",
                (unsigned long long)n_b, (unsigned long long)kMaxB);
  if (flags & HD_FLAG_LEARNED_ORDER)
    return Fail(HD_EUNSUPPORTED,
                "the learned order allocates; use hd_directed_device");
  if (flags & ~(HD_FLAG_NO_SHUFFLE | HD_FLAG_SINGLE_PASS | HD_FLAG_COUNT_EVALS))
    return Fail(HD_EINVAL, "unsupported flags 0x%x", flags);
  HD_TRY(ContextActivate(ctx));
  unsigned char* base = nullptr;
  HD_TRY(ScratchAddress(ctx, &base));

  // Non-finite coordinates are rejected as on the other entry points.
  float mbr[6];
  uint32_t* mbr_work = reinterpret_cast<uint32_t*>(base + kOffMbr);
  HD_TRY(HostMbr(ctx, n_dims, d_a, n_a, mbr, mbr_work));
  HD_TRY(HostMbr(ctx, n_dims, d_b, n_b, mbr, mbr_work));

  EbScratch s;
  s.key = reinterpret_cast<unsigned long long*>(base + kOffKey);
  s.counters = reinterpret_cast<unsigned int*>(base + kOffCounters);
  s.cursors = reinterpret_cast<unsigned int*>(base + kOffCursors);
  s.nn = reinterpret_cast<uint32_t*>(base + kOffNn);
  s.esc_idx = reinterpret_cast<uint32_t*>(base + kOffEscIdx);
  s.esc_min = reinterpret_cast<float*>(base + kOffEscMin);
  s.bx = reinterpret_cast<float*>(base + kOffSoA);
  s.by = s.bx + kMaxB;
  s.bz = n_dims == 3 ? s.by + kMaxB : nullptr;
  s.max_bp = (uint32_t)kMaxB;

  HD_HIP_TRY(hipEventRecord(ctx->ev_begin, ctx->stream));
  HD_TRY(EbSinglePassScratch(ctx, n_dims, d_a, (uint32_t)n_a, d_b,
                             (uint32_t)n_b, (flags & HD_FLAG_NO_SHUFFLE) == 0,
                             seed, s, out));
  HD_HIP_TRY(hipEventRecord(ctx->ev_end, ctx->stream));
  HD_HIP_TRY(hipEventSynchronize(ctx->ev_end));
  float ms = 0.f;
  HD_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end));
  out->time_ms = ms;
  return HD_OK;
}
