// Host entry points of the device-wide support primitives (sort.hip,
// common.hip). All run on the context stream and are asynchronous.
#ifndef HD_PRIMITIVES_H
#define HD_PRIMITIVES_H

#include <cstdint>

#include "hd_internal.h"

namespace hd {

// Stable radix sorts of (key, u32 value) pairs (sort.hip): ascending on key
// bits [begin_bit, end_bit) for u32 (the other bits are carried), on all 16
// bits for u16, descending for finite f32 keys. Inputs and outputs must not
// alias. u32 with out_mul (coprime with n): the value of rank k goes to
// v_out[k * out_mul mod n] and k_out receives no sorted keys (it may still
// serve as scratch).
int SortPairsDescF32(hd_context* ctx, const float* k_in, float* k_out,
                     const uint32_t* v_in, uint32_t* v_out, uint32_t n);
int SortPairsU16(hd_context* ctx, const uint16_t* k_in, uint16_t* k_out,
                 const uint32_t* v_in, uint32_t* v_out, uint32_t n);
int SortPairsU32(hd_context* ctx, const uint32_t* k_in, uint32_t* k_out,
                 const uint32_t* v_in, uint32_t* v_out, uint32_t n,
                 int end_bit, int begin_bit = 0, uint32_t out_mul = 0);

// Host construction of a Feistel permutation of [0, n) (hd_device.h).
Feistel MakeFeistel(uint32_t n, uint64_t seed, uint32_t stream_id);
// The same permutation evaluated on the host: position i -> index.
uint32_t HostPermute(const Feistel& f, uint32_t i);

// d_out[k] = point floor(k n / m) of the AoS set pts, for k < m.
int SamplePoints(hd_context* ctx, const float* pts, uint64_t n, int n_dims,
                 uint32_t m, float* d_out);

// p[i] = i for i < n.
int IotaU32(hd_context* ctx, uint32_t* p, uint32_t n);

// Fills n 32-bit words with `value` (a kernel, so it is capturable).
int FillU32(hd_context* ctx, uint32_t* p, uint32_t value, uint64_t n);
int FillU64(hd_context* ctx, unsigned long long* p, unsigned long long value,
            uint64_t n);

// A step coprime with n near n / phi (1 for n <= 2): i -> i * step mod n
// is a bijection of [0, n) that spreads every run of consecutive i evenly
// over [0, n) (the gaps of an irrational rotation take at most three
// values).
uint32_t GoldenStep(uint32_t n);

// Repacks AoS points into SoA arrays of n_pad entries (n_pad >= n),
// visiting source index idx[j mod n] for destination j when idx (a device
// array of n indices) is given, else perm(j mod n) (identity if !shuffle).
// z may be null for 2D.
int GatherSoA(hd_context* ctx, int n_dims, const float* aos, uint32_t n,
              uint32_t n_pad, bool shuffle, const Feistel& perm,
              float* x, float* y, float* z, const uint32_t* idx = nullptr);

// Reads the MBR back (synchronous) and applies the reference's
// ToNonemptyMBR widening (src/geoms/mbr.h:288-296). Non-finite coordinates
// are an HD_EINVAL error. d_work: 32 device words, or null for a context
// buffer.
int HostMbr(hd_context* ctx, int n_dims, const float* aos, uint64_t n,
            float* mbr /* host, 2*n_dims */, uint32_t* d_work = nullptr);
// Both sets' MBRs with one synchronisation. d_work: 64 device words, or
// null for a context buffer.
int HostMbrPair(hd_context* ctx, int n_dims, const float* a, uint64_t n_a,
                float* mbr_a, const float* b, uint64_t n_b, float* mbr_b,
                uint32_t* d_work = nullptr);

}  // namespace hd

#endif  // HD_PRIMITIVES_H
