// Software LBVH (lbvh.hip): device records and query arguments.
#ifndef HD_LBVH_H
#define HD_LBVH_H

#include <cstdint>

#include "hd_internal.h"

namespace hd {

// Internal node: both child boxes inline (unused z is 0 in 2D) and the two
// children (bit 31 set: leaf index). 64 bytes, one load per visit.
struct LbvhNode {
  float lo[2][3];
  float hi[2][3];
  uint32_t child[2];
  uint32_t pad[2];
};
static_assert(sizeof(LbvhNode) == 64, "node record is one 64-byte line");

// Morton quantisation frame.
struct LbvhFrame {
  float lo[3];
  float scale[3];
  float qmax;
};

struct LbvhDev {
  const LbvhNode* nodes;  // n_leaf - 1 internal nodes, root at 0
  const float* x;         // points in Morton order (SoA)
  const float* y;
  const float* z;
  const uint32_t* orig;   // Morton position -> original index
  uint32_t n;             // points
  uint32_t n_leaf;        // leaves (runs of `leaf` points)
  uint32_t leaf;
};

struct LbvhIndex {
  LbvhDev dev;
  LbvhFrame frame;
  int n_dims = 2;
  const uint32_t* codes = nullptr;  // sorted 32-bit Morton codes of B
};

struct LbvhQueryArgs {
  LbvhDev tree;
  int tree_nd;
  const float* a;          // AoS queries
  const uint32_t* order;   // order_mode 0
  Feistel perm;            // order_mode 1
  int order_mode;          // 0 array, 1 Feistel, 2 identity
  uint32_t pos_begin, pos_end;
  int eb;
  unsigned long long* key;
  unsigned long long* evals;
  unsigned long long* lookups;  // optional: internal node records read
  float* nn_d2;            // optional per-query outputs (by A index)
  uint32_t* nn_idx;
  uint32_t* esc_idx;       // traversal-stack overflows
  float* esc_min;
  unsigned int* esc_count;
  const uint32_t* start_leaf;  // optional, by position: the leaf where the
                               // query's Morton code falls among B's codes
  uint32_t stride;             // > 1: query k takes position k * stride
  uint32_t skip_stride;        // > 1: positions k * skip_stride, k < skip_n,
  uint32_t skip_n;             // are skipped (done by a strided pass)
};

// mbr_b: B's MBR when the caller has it (host, 2 n_dims floats), else null.
int BuildLbvh(hd_context* ctx, int n_dims, const float* b, uint64_t n_b,
              uint32_t leaf_size, LbvhIndex* out, const float* mbr_b = nullptr);
int LaunchLbvhQuery(hd_context* ctx, const LbvhQueryArgs& args);
// Quantisation frame over an MBR (2 n_dims floats) for the Morton codes.
LbvhFrame MortonFrame(int n_dims, const float* mbr);
// Permutation of pts sorted by bits [begin_bit, 32) of their 32-bit Morton
// codes in frame f, equal keys in input order (device array owned by the
// context; `tag` prefixes the names of its buffers, so callers whose orders
// must coexist pass different tags).
int MortonOrder(hd_context* ctx, int n_dims, const float* pts, uint32_t n,
                const LbvhFrame& f, uint32_t** order_out, int begin_bit = 0,
                const char* tag = "lbvh_q", uint32_t out_mul = 0);
// A spatially stratified visit order of pts (perm[j] = index of the point
// visited j-th; device array owned by the context): the points ranked by
// the top 16 bits of their Morton code in their MBR (equal codes in input
// order), rank r at position r * step mod n with step coprime with n and
// near n / phi, so every run of consecutive ranks (the points of one
// region) is spread evenly over the sequence. mbr: pts' MBR when the caller
// has it, else null. (MortonOrder with out_mul: the order's entry of rank k
// is written to position k * out_mul mod n.)
constexpr uint32_t kMaxStratified = 1u << 31;
int StratifiedOrder(hd_context* ctx, int n_dims, const float* pts, uint32_t n,
                    const float* mbr, uint32_t** perm_out);
// start[p] = leaf of tree holding the first B code >= the Morton code of
// query q[order[p]] (clamped to the last leaf); order should be a Morton
// order of the queries, for coherent searches.
int StartLeaves(hd_context* ctx, const LbvhIndex& tree, const float* q,
                const uint32_t* order, uint32_t n, uint32_t* start);
int LbvhNearestNeighbors(hd_context* ctx, int n_dims, const float* q,
                         uint64_t n_q, const float* ref, uint64_t n_ref,
                         float* out_d2, uint32_t* out_idx);

// Exact minimum over all of `ref` for the queries listed in `list`
// (A indices): raises *key with each one's (d2, index) and, when out_d2 is
// given, writes out_d2[idx] / out_idx[idx] (smallest index among ties).
int ExactMinForList(hd_context* ctx, int n_dims, const float* q,
                    const uint32_t* list, uint32_t n_list, const float* ref,
                    uint64_t n_ref, unsigned long long* key, float* out_d2,
                    uint32_t* out_idx);

}  // namespace hd

#endif  // HD_LBVH_H
