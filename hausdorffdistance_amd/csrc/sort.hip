// Device-wide sort and scan primitives on the context stream (rocPRIM).
// The kernels that matter for HD run elsewhere; these are the support steps:
// ordering escalated points, Morton-sorting index points, learned ordering.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

namespace {
template <typename Fn>
int WithTemp(hd_context* ctx, const char* what, Fn fn) {
  size_t bytes = 0;
  hipError_t e = fn(nullptr, bytes);
  if (e != hipSuccess)
    return Fail(HD_EDEVICE, "%s (size query): %s", what, hipGetErrorString(e));
  void* tmp = nullptr;
  HD_TRY(ContextBuffer(ctx, "prim_tmp", bytes + 256, &tmp));
  e = fn(tmp, bytes);
  if (e != hipSuccess)
    return Fail(HD_EDEVICE, "%s: %s", what, hipGetErrorString(e));
  return HD_OK;
}
}  // namespace

int SortPairsDescF32(hd_context* ctx, const float* k_in, float* k_out,
                     const uint32_t* v_in, uint32_t* v_out, uint32_t n) {
  if (n == 0) return HD_OK;
  return WithTemp(ctx, "radix_sort_pairs_desc<f32>", [&](void* t, size_t& b) {
    return rocprim::radix_sort_pairs_desc(t, b, k_in, k_out, v_in, v_out, n,
                                          0, 32, ctx->stream);
  });
}

int SortPairsU16(hd_context* ctx, const uint16_t* k_in, uint16_t* k_out,
                 const uint32_t* v_in, uint32_t* v_out, uint32_t n) {
  if (n == 0) return HD_OK;
  return WithTemp(ctx, "radix_sort_pairs<u16>", [&](void* t, size_t& b) {
    return rocprim::radix_sort_pairs(t, b, k_in, k_out, v_in, v_out, n, 0, 16,
                                     ctx->stream);
  });
}

int SortPairsU32(hd_context* ctx, const uint32_t* k_in, uint32_t* k_out,
                 const uint32_t* v_in, uint32_t* v_out, uint32_t n,
                 int end_bit) {
  if (n == 0) return HD_OK;
  return WithTemp(ctx, "radix_sort_pairs<u32>", [&](void* t, size_t& b) {
    return rocprim::radix_sort_pairs(t, b, k_in, k_out, v_in, v_out, n, 0,
                                     (unsigned)end_bit, ctx->stream);
  });
}

int SortPairsU64(hd_context* ctx, const unsigned long long* k_in,
                 unsigned long long* k_out, const uint32_t* v_in,
                 uint32_t* v_out, uint32_t n, int end_bit) {
  if (n == 0) return HD_OK;
  return WithTemp(ctx, "radix_sort_pairs<u64>", [&](void* t, size_t& b) {
    return rocprim::radix_sort_pairs(t, b, k_in, k_out, v_in, v_out, n, 0,
                                     (unsigned)end_bit, ctx->stream);
  });
}

int ExclusiveScanU32(hd_context* ctx, const uint32_t* in, uint32_t* out,
                     uint32_t n) {
  if (n == 0) return HD_OK;
  return WithTemp(ctx, "exclusive_scan<u32>", [&](void* t, size_t& b) {
    return rocprim::exclusive_scan(t, b, in, out, 0u, (size_t)n,
                                   rocprim::plus<uint32_t>(), ctx->stream);
  });
}

}  // namespace hd
