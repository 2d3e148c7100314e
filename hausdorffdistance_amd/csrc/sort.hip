// Stable LSD radix sort of (key, u32 value) pairs on the context stream.
//
// Callers: Morton-sorting the grid's, the LBVH's and RT-HDIST's points
// (u32 keys, 22-32 bits), ordering A by the learned model's 16-bit score
// (u16 keys), and ordering escalated points by their filter minimum (f32
// keys, descending). The reference sorts the same keys with thrust's radix
// sort (src/index/uniform_grid.h, src/lbvh/lbvh/bvh.cuh); the order of equal
// keys is the input order in both, so visit orders (and with them the
// evaluation counts) do not depend on the sort's implementation.
//
// One pass per 8-bit digit, three kernels a pass:
//  1. count: one 4096-key tile per 256-thread workgroup; a digit histogram
//     per wave in LDS, summed, written digit-major (counts[d][tile]);
//  2. scan: one workgroup per digit scans its row over the tiles (exclusive)
//     and writes the digit's total;
//  3. scatter: each workgroup ranks its tile again. Each wave owns 1024
//     consecutive keys and ranks them 64 at a time: the lanes holding the
//     same digit are found with eight ballots, a lane's rank is the number of
//     such lanes below it plus the wave's running count of that digit (LDS,
//     touched by this wave only, so no barrier). The tile is then laid out in
//     LDS in digit order and written back with consecutive lanes writing
//     consecutive addresses of one digit's run: the global position of key i
//     of digit d is base(d) + count of d in earlier tiles + (i - first of d in
//     the tile).
// Keys pass through a codec that maps them to unsigned radix keys (f32
// descending: the IEEE order flipped, then complemented) and back on the
// final write. The key traffic per pass is one read for the count, one read
// and one write for the scatter, plus the same read and write of the
// values; the scan's rows are 4 B per tile and digit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

namespace {

constexpr int kBlock = 256;   // 4 waves of 64
constexpr int kItems = 16;    // keys per thread and tile
constexpr int kTile = kBlock * kItems;
constexpr int kWaveKeys = kTile / 4;
constexpr int kRadix = 256;

struct CodecU16 {
  using Stored = uint16_t;
  __device__ static uint32_t Encode(uint16_t s) { return s; }
  __device__ static uint16_t Decode(uint32_t r) { return (uint16_t)r; }
};

struct CodecU32 {
  using Stored = uint32_t;
  __device__ static uint32_t Encode(uint32_t s) { return s; }
  __device__ static uint32_t Decode(uint32_t r) { return r; }
};

// Descending order of finite floats as ascending order of unsigned keys:
// non-negative values get the sign bit set, negative ones are complemented
// (the IEEE total order), and the result is complemented.
struct CodecF32Desc {
  using Stored = float;
  __device__ static uint32_t Encode(float s) {
    const uint32_t u = __float_as_uint(s);
    return ~((u & 0x80000000u) ? ~u : (u | 0x80000000u));
  }
  __device__ static float Decode(uint32_t r) {
    const uint32_t o = ~r;
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
  }
};

// Exclusive scan of one value per thread over the 256 threads of a
// workgroup; *total receives the sum. s_wave: 4 words of LDS.
__device__ uint32_t BlockExclusiveScan(uint32_t v, uint32_t* s_wave,
                                       uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  uint32_t before = 0, sum = 0;
  for (int i = 0; i < 4; ++i) {
    const uint32_t s = s_wave[i];
    before += i < w ? s : 0u;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return before + x - v;
}

template <class Codec>
__global__ __launch_bounds__(kBlock) void RadixCountKernel(
    const typename Codec::Stored* __restrict__ keys, uint32_t n, int shift,
    uint32_t mask, uint32_t n_tiles, uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[4][kRadix];
  const int t = threadIdx.x, w = t >> 6;
  for (int i = t; i < 4 * kRadix; i += kBlock) (&hist[0][0])[i] = 0u;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll 4
  for (int j = 0; j < kItems; ++j) {
    const uint64_t i = base + (uint64_t)j * kBlock + t;
    if (i < n) {
      const uint32_t d = (Codec::Encode(keys[i]) >> shift) & mask;
      atomicAdd(&hist[w][d], 1u);
    }
  }
  __syncthreads();
  counts[(size_t)t * n_tiles + blockIdx.x] =
      hist[0][t] + hist[1][t] + hist[2][t] + hist[3][t];
}

// Row d of counts (n_tiles entries) becomes its exclusive scan; totals[d]
// receives the row's sum.
__global__ __launch_bounds__(kBlock) void RadixScanKernel(
    uint32_t* __restrict__ counts, uint32_t n_tiles,
    uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_wave[4];
  uint32_t* row = counts + (size_t)blockIdx.x * n_tiles;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < n_tiles; c0 += kBlock * 4) {
    const uint32_t i0 = c0 + threadIdx.x * 4;
    uint32_t v[4], local = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = i0 + k < n_tiles ? row[i0 + k] : 0u;
      local += v[k];
    }
    uint32_t chunk = 0;
    uint32_t run = carry + BlockExclusiveScan(local, s_wave, &chunk);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (i0 + k < n_tiles) row[i0 + k] = run;
      run += v[k];
    }
    carry += chunk;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

template <class Codec>
__global__ __launch_bounds__(kBlock) void RadixScatterKernel(
    const typename Codec::Stored* __restrict__ keys_in,
    const uint32_t* __restrict__ vals_in,
    typename Codec::Stored* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, int shift, uint32_t mask,
    uint32_t n_tiles, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ totals, uint32_t val_mul,
    unsigned long long val_magic) {
  __shared__ uint32_t s_key[kTile];
  __shared__ uint32_t s_val[kTile];
  __shared__ uint32_t wcnt[4][kRadix];
  __shared__ uint32_t tile_first[kRadix];
  __shared__ uint32_t out_first[kRadix];
  __shared__ uint32_t s_wave[4];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int i = lane; i < kRadix; i += 64) wcnt[w][i] = 0u;
  const uint64_t tile0 = (uint64_t)blockIdx.x * kTile;
  const uint64_t seg0 = tile0 + (uint64_t)w * kWaveKeys;

  uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint64_t i = seg0 + (uint64_t)j * 64 + lane;
    const bool valid = i < n;
    key[j] = valid ? Codec::Encode(keys_in[i]) : 0u;
    val[j] = valid ? vals_in[i] : 0u;
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const bool valid = seg0 + (uint64_t)j * 64 + lane < n;
    const uint32_t d = (key[j] >> shift) & mask;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t before = (uint32_t)__popcll(peers & lt);
    const uint32_t run = wcnt[w][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && before == 0u) wcnt[w][d] = run + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rank[j] = run + before;
  }
  __syncthreads();
  // Thread t owns digit t: offsets of the waves' runs within the digit, the
  // digit's first position in the tile, and its first global position.
  {
    uint32_t c[4], sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      c[k] = wcnt[k][t];
      wcnt[k][t] = sum;
      sum += c[k];
    }
    uint32_t tile_total = 0;
    tile_first[t] = BlockExclusiveScan(sum, s_wave, &tile_total);
    uint32_t all = 0;
    const uint32_t digit_base = BlockExclusiveScan(totals[t], s_wave, &all);
    out_first[t] = digit_base + counts[(size_t)t * n_tiles + blockIdx.x];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    if (seg0 + (uint64_t)j * 64 + lane < n) {
      const uint32_t d = (key[j] >> shift) & mask;
      const uint32_t pos = tile_first[d] + wcnt[w][d] + rank[j];
      s_key[pos] = key[j];
      s_val[pos] = val[j];
    }
  }
  __syncthreads();
  const uint32_t tile_n = (uint32_t)std::min<uint64_t>(kTile, n - tile0);
#pragma unroll 4
  for (int j = 0; j < kItems; ++j) {
    const uint32_t i = (uint32_t)j * kBlock + t;
    if (i < tile_n) {
      const uint32_t k = s_key[i];
      const uint32_t d = (k >> shift) & mask;
      const uint32_t o = out_first[d] + (i - tile_first[d]);
      if (keys_out != nullptr) keys_out[o] = Codec::Decode(k);
      uint32_t dst = o;
      if (val_mul) {
        // o * val_mul mod n: q = mulhi(x, floor((2^64 - 1) / n)) is
        // floor(x / n) or one less (x < 2^62), so one subtraction fixes r.
        const unsigned long long x = (unsigned long long)o * val_mul;
        const unsigned long long q = __umul64hi(x, val_magic);
        unsigned long long r = x - q * n;
        if (r >= n) r -= n;
        dst = (uint32_t)r;
      }
      vals_out[dst] = s_val[i];
    }
  }
}

template <class Codec>
int RadixSortPairs(hd_context* ctx, const char* what,
                   const typename Codec::Stored* k_in,
                   typename Codec::Stored* k_out, const uint32_t* v_in,
                   uint32_t* v_out, uint32_t n, int end_bit, int begin_bit = 0,
                   uint32_t out_mul = 0) {
  using Stored = typename Codec::Stored;
  if (n == 0) return HD_OK;
  if (end_bit < 0 || end_bit > 32 || begin_bit < 0 || begin_bit > end_bit ||
      (k_in == k_out) || (v_in == v_out) || (out_mul && !k_out))
    return Fail(HD_EINVAL, "%s: bad arguments (bits %d..%d)", what, begin_bit,
                end_bit);
  const int passes = std::max(1, (end_bit - begin_bit + 7) / 8);
  const uint32_t n_tiles = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);
  uint32_t *counts = nullptr, *totals = nullptr;
  HD_TRY(ContextArray(ctx, "sort_counts", (size_t)kRadix * n_tiles, &counts));
  HD_TRY(ContextArray(ctx, "sort_totals", kRadix, &totals));
  Stored* tmp_k = nullptr;
  uint32_t* tmp_v = nullptr;
  if (passes > 1) {
    HD_TRY(ContextArray(ctx, "sort_tmp_keys", n, &tmp_k));
    HD_TRY(ContextArray(ctx, "sort_tmp_vals", n, &tmp_v));
  }
  const Stored* src_k = k_in;
  const uint32_t* src_v = v_in;
  for (int p = 0; p < passes; ++p) {
    // The last pass writes the output; earlier ones alternate with it.
    const bool to_out = ((passes - 1 - p) & 1) == 0;
    Stored* dst_k = to_out ? k_out : tmp_k;
    uint32_t* dst_v = to_out ? v_out : tmp_v;
    const int bits = std::min(8, end_bit - (begin_bit + 8 * p));
    // An empty bit range is one stable copy (digit 0 for every key).
    const int shift = bits > 0 ? begin_bit + 8 * p : 0;
    const uint32_t mask = bits > 0 ? (1u << bits) - 1u : 0u;
    RadixCountKernel<Codec><<<n_tiles, kBlock, 0, ctx->stream>>>(
        src_k, n, shift, mask, n_tiles, counts);
    RadixScanKernel<<<kRadix, kBlock, 0, ctx->stream>>>(counts, n_tiles,
                                                         totals);
    // With out_mul, the last pass writes value k to position k * out_mul
    // mod n of v_out, and no keys.
    const bool last = p + 1 == passes;
    RadixScatterKernel<Codec><<<n_tiles, kBlock, 0, ctx->stream>>>(
        src_k, src_v, last && out_mul ? nullptr : dst_k, dst_v, n, shift, mask,
        n_tiles, counts, totals, last ? out_mul : 0u, ~0ull / n);
    HD_HIP_TRY(hipGetLastError());
    src_k = dst_k;
    src_v = dst_v;
  }
  return HD_OK;
}

}  // namespace

int SortPairsDescF32(hd_context* ctx, const float* k_in, float* k_out,
                     const uint32_t* v_in, uint32_t* v_out, uint32_t n) {
  return RadixSortPairs<CodecF32Desc>(ctx, "sort<f32, descending>", k_in,
                                      k_out, v_in, v_out, n, 32);
}

int SortPairsU16(hd_context* ctx, const uint16_t* k_in, uint16_t* k_out,
                 const uint32_t* v_in, uint32_t* v_out, uint32_t n) {
  return RadixSortPairs<CodecU16>(ctx, "sort<u16>", k_in, k_out, v_in, v_out,
                                  n, 16);
}

int SortPairsU32(hd_context* ctx, const uint32_t* k_in, uint32_t* k_out,
                 const uint32_t* v_in, uint32_t* v_out, uint32_t n,
                 int end_bit, int begin_bit, uint32_t out_mul) {
  return RadixSortPairs<CodecU32>(ctx, "sort<u32>", k_in, k_out, v_in, v_out,
                                  n, end_bit, begin_bit, out_mul);
}

}  // namespace hd

// Test hook: sorts n (key, value) pairs on the context stream and waits.
// key_kind 0: u16 keys (16 bits), 1: u32 keys on bits [begin_bit, end_bit),
// 2: f32 keys, descending.
extern "C" int hd_sort_pairs_probe_device(hd_context* ctx, int key_kind,
                                          const void* d_keys_in,
                                          const uint32_t* d_vals_in,
                                          uint32_t n, int begin_bit,
                                          int end_bit, void* d_keys_out,
                                          uint32_t* d_vals_out) {
  using namespace hd;
  ClearError();
  if (!ctx || (n && (!d_keys_in || !d_vals_in || !d_keys_out || !d_vals_out)))
    return Fail(HD_EINVAL, "null argument");
  if (end_bit < 0 || end_bit > 32 || begin_bit < 0 || begin_bit > end_bit)
    return Fail(HD_EINVAL, "bits must satisfy 0 <= begin_bit <= end_bit <= 32");
  HD_TRY(ContextActivate(ctx));
  int rc;
  if (key_kind == 0)
    rc = SortPairsU16(ctx, static_cast<const uint16_t*>(d_keys_in),
                      static_cast<uint16_t*>(d_keys_out), d_vals_in,
                      d_vals_out, n);
  else if (key_kind == 1)
    rc = SortPairsU32(ctx, static_cast<const uint32_t*>(d_keys_in),
                      static_cast<uint32_t*>(d_keys_out), d_vals_in,
                      d_vals_out, n, end_bit, begin_bit);
  else if (key_kind == 2)
    rc = SortPairsDescF32(ctx, static_cast<const float*>(d_keys_in),
                          static_cast<float*>(d_keys_out), d_vals_in,
                          d_vals_out, n);
  else
    return Fail(HD_EINVAL, "key_kind must be 0 (u16), 1 (u32) or 2 (f32 desc)");
  if (rc != HD_OK) return rc;
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return HD_OK;
}
