// Uniform-grid index (grid.hip): host description and query arguments.
#ifndef HD_GRID_H
#define HD_GRID_H

#include <cstdint>

#include "hd_internal.h"

namespace hd {

// Device view of a built grid (passed by value to kernels).
struct GridDev {
  float lo[3], hi[3];     // MBR of the indexed set (after ToNonemptyMBR)
  uint32_t dims[3];       // cells per axis
  double lo_d[3], s_d[3]; // FP64 origin and cell side for the bounds
  double margin_d[3];     // worst-case FP32 cell-assignment error per axis
  double max_margin;
  int morton;             // sort key: Morton code (1) or row-major id (0)
  const float* x;         // indexed points in key order (SoA)
  const float* y;
  const float* z;
  const uint32_t* orig;   // key-order position -> original index
  const uint32_t* start;  // per row-major cell id: [start, end) in key order
  const uint32_t* end;
};

struct GridIndex {
  GridDev dev;
  int n_dims = 2;
  uint32_t n = 0;
  uint64_t n_cells = 0;
};

struct GridQueryArgs {
  GridDev grid;
  int grid_nd;
  const float* a;          // AoS queries
  const uint32_t* order;   // order_mode 0
  Feistel perm;            // order_mode 1
  int order_mode;          // 0 array, 1 Feistel, 2 identity
  uint32_t pos_begin, pos_end;
  int eb;                  // stop a query once its best d2 < cmax2
  uint32_t max_cells;      // 0: unlimited; else escalate after that many
                           // non-empty cells
  unsigned long long* key; // packed max key (may be null)
  unsigned long long* evals;
  unsigned long long* lookups;  // optional: cell table reads (start, end)
  float* nn_d2;            // optional per-query outputs (indexed by A index)
  uint32_t* nn_idx;
  uint32_t* esc_idx;       // escalations (max_cells exhausted)
  float* esc_min;
  unsigned int* esc_count;
  uint32_t stride;         // > 1: query k takes position k * stride
  uint32_t skip_stride;    // > 1: positions k * skip_stride, k < skip_n,
  uint32_t skip_n;         // are skipped (done by a strided pass)
};

// mbr_b: B's MBR when the caller has it (host, 2 n_dims floats), else null.
int BuildGrid(hd_context* ctx, int n_dims, const float* b, uint64_t n_b,
              uint32_t n_points_cell, GridIndex* out,
              const float* mbr_b = nullptr);
int LaunchGridQuery(hd_context* ctx, const GridQueryArgs& args);
int GridNearestNeighbors(hd_context* ctx, int n_dims, const float* q,
                         uint64_t n_q, const float* ref, uint64_t n_ref,
                         float* out_d2, uint32_t* out_idx);

}  // namespace hd

#endif  // HD_GRID_H
