// Uniform-grid index over B, rebuilt for gfx950 as a Morton-sorted cell
// table searched in expanding rings.
//
// Reference: src/index/uniform_grid.h (resolution :222-244, cell of a point
// :138-159, Insert :343-484). The reference marks occupied cells in a bitset,
// sorts the non-empty cell ids, counts points per cell with a binary search
// and atomics, and scatters point ids; its NN query is an OptiX traversal of
// radius-inflated cell boxes (src/rt/shaders/shaders_nn_uniform_grid_2d.cu).
// Here the build is: cell key per point (Morton code of the cell
// coordinates) -> one radix sort of (key, index) -> B gathered into SoA in
// key order -> a dense (start, end) table indexed by the row-major cell id.
// Points of a cell are contiguous, and cells that are close in space are
// close in memory.
//
// Query: one lane per A point. Rings r = 0, 1, 2, ... of cells at Chebyshev
// distance r from the point's cell are scanned (cells whose box is already
// farther than the best d2 are skipped). After ring r, every unvisited B
// point lies outside the box of cells [c-r, c+r]; the distance from the
// query to the outside of that box, computed in FP64 and shrunk by the
// worst-case error of the FP32 cell assignment, bounds all of them, so the
// search stops once its square reaches the best d2. Every d2 is the
// reference's FP32 expression; the pruning only decides which pairs are
// evaluated, never their values. With early break (the hybrid behaviour of
// shaders_nn_uniform_grid_2d.cu:100-103) a query stops as soon as its best
// d2 is strictly below cmax2.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hd_device.h"
#include "hd_grid.h"
#include "hd_internal.h"
#include "hd_lbvh.h"
#include "hd_primitives.h"

namespace hd {

using namespace dev;

namespace {

constexpr int kBlock = 256;

HD_DEV uint32_t CellCoord(const GridDev& g, int d, float v) {
  // reference uniform_grid.h:149-156, in FP32.
  const float norm = (v - g.lo[d]) / (g.hi[d] - g.lo[d]);
  const float c = fminf(fmaxf(norm * (float)g.dims[d], 0.0f),
                        (float)g.dims[d] - 1.0f);
  return (uint32_t)c;
}

HD_DEV uint32_t Spread2(uint32_t x) {  // 16 bits -> even bits
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
HD_DEV uint32_t Spread3(uint32_t x) {  // 10 bits -> every third bit
  x &= 0x3FFu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

template <int ND>
HD_DEV uint32_t SortKey(const GridDev& g, const uint32_t* c) {
  if (g.morton) {
    if constexpr (ND == 2) return Spread2(c[0]) | (Spread2(c[1]) << 1);
    else return Spread3(c[0]) | (Spread3(c[1]) << 1) | (Spread3(c[2]) << 2);
  }
  uint32_t lin = c[ND - 1];
#pragma unroll
  for (int d = ND - 2; d >= 0; --d) lin = lin * g.dims[d] + c[d];
  return lin;
}

template <int ND>
HD_DEV uint32_t Linear(const GridDev& g, const uint32_t* c) {
  uint32_t lin = c[ND - 1];
#pragma unroll
  for (int d = ND - 2; d >= 0; --d) lin = lin * g.dims[d] + c[d];
  return lin;
}

template <int ND>
__global__ void GridKeysKernel(GridDev g, const float* __restrict__ b,
                               uint32_t n, uint32_t* __restrict__ keys,
                               uint32_t* __restrict__ vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    uint32_t c[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) c[d] = CellCoord(g, d, b[(uint64_t)i * ND + d]);
    keys[i] = SortKey<ND>(g, c);
    vals[i] = i;
  }
}

template <int ND>
__global__ void GridFillKernel(GridDev g, const float* __restrict__ b,
                               uint32_t n, const uint32_t* __restrict__ keys,
                               const uint32_t* __restrict__ vals,
                               float* __restrict__ x, float* __restrict__ y,
                               float* __restrict__ z, uint32_t* __restrict__ orig,
                               uint32_t* __restrict__ start,
                               uint32_t* __restrict__ end) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t src = vals[i];
    const float* p = b + (uint64_t)src * ND;
    x[i] = p[0];
    y[i] = p[1];
    if constexpr (ND == 3) z[i] = p[2];
    orig[i] = src;
    const uint32_t k = keys[i];
    const bool first = i == 0 || keys[i - 1] != k;
    const bool last = i + 1 == n || keys[i + 1] != k;
    if (first || last) {
      uint32_t c[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) c[d] = CellCoord(g, d, p[d]);
      const uint32_t lin = Linear<ND>(g, c);
      if (first) start[lin] = i;
      if (last) end[lin] = i + 1;
    }
  }
}

// Squared distance from q to the box of cell c, in FP64, with the cell
// grown by the assignment margin; scaled down so that it never exceeds the
// FP32 d2 of a point the cell may hold.
template <int ND>
HD_DEV double CellMinDist2(const GridDev& g, const double* q,
                           const uint32_t* c) {
  double acc = 0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double blo = g.lo_d[d] + c[d] * g.s_d[d] - g.margin_d[d];
    const double bhi = g.lo_d[d] + (c[d] + 1.0) * g.s_d[d] + g.margin_d[d];
    double diff = 0;
    if (q[d] < blo) diff = blo - q[d];
    else if (q[d] > bhi) diff = q[d] - bhi;
    acc += diff * diff;
  }
  return acc * (1.0 - 0x1p-20);
}

// Lower bound of the FP32 d2 from q to any point outside the box of cells
// [qc - r, qc + r]; negative when the box covers the whole grid.
template <int ND>
HD_DEV double OutsideBound2(const GridDev& g, const double* q,
                            const uint32_t* qc, uint32_t r) {
  double best = 1e300;
  bool any = false;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    if (qc[d] > r) {  // cells below qc - r exist
      const double edge = g.lo_d[d] + (double)(qc[d] - r) * g.s_d[d];
      best = fmin(best, q[d] - edge);
      any = true;
    }
    if (qc[d] + r + 1 < g.dims[d]) {  // cells above qc + r exist
      const double edge = g.lo_d[d] + (double)(qc[d] + r + 1) * g.s_d[d];
      best = fmin(best, edge - q[d]);
      any = true;
    }
  }
  if (!any) return -1.0;
  double dist = best - g.max_margin;
  if (dist <= 0) return 0.0;
  return dist * dist * (1.0 - 0x1p-20);
}

struct QueryState {
  float best;
  uint32_t best_j;
  uint32_t evals;
  uint32_t cells;
  uint32_t lookups;  // cell table reads
  bool stop;     // early break or cell budget exhausted
  bool broke;    // best < cmax2
  bool over;     // cell budget exhausted
};

template <int ND>
HD_DEV void ScanCell(const GridDev& g, const float* qf, const double* qd,
                     const uint32_t* c, float cmax2, bool eb, bool tie_index,
                     uint32_t max_cells, QueryState& st) {
  if (st.stop) return;
  // With index tie-breaking a cell at exactly the best distance may still
  // hold a smaller index, so only strictly farther cells are skipped.
  const double cmin = CellMinDist2<ND>(g, qd, c);
  if (tie_index ? cmin > (double)st.best : cmin >= (double)st.best) return;
  if (max_cells != 0 && ++st.cells > max_cells) {
    // Too many cells for this query: hand it to the early-break scan.
    st.stop = true;
    st.over = true;
    return;
  }
  const uint32_t lin = Linear<ND>(g, c);
  st.lookups++;
  const uint32_t b0 = g.start[lin], b1 = g.end[lin];
  for (uint32_t k = b0; k < b1; ++k) {
    float d2;
    if constexpr (ND == 2)
      d2 = D2(qf[0], qf[1], g.x[k], g.y[k]);
    else
      d2 = D2(qf[0], qf[1], qf[2], g.x[k], g.y[k], g.z[k]);
    st.evals++;
    if (d2 < st.best) {
      st.best = d2;
      st.best_j = tie_index ? g.orig[k] : 0u;
    } else if (tie_index && d2 == st.best && g.orig[k] < st.best_j) {
      st.best_j = g.orig[k];
    }
    if (eb && st.best < cmax2) {
      st.stop = true;
      st.broke = true;
      return;
    }
  }
}

// All cells of ring r around qc that exist in the grid.
template <int ND>
HD_DEV void ScanRing(const GridDev& g, const float* qf, const double* qd,
                     const uint32_t* qc, uint32_t r, float cmax2, bool eb,
                     bool tie_index, uint32_t max_cells, QueryState& st) {
  const int ir = (int)r;
  if constexpr (ND == 2) {
    const int x0 = (int)qc[0] - ir, x1 = (int)qc[0] + ir;
    const int y0 = (int)qc[1] - ir, y1 = (int)qc[1] + ir;
    const int X = (int)g.dims[0], Y = (int)g.dims[1];
    if (r == 0) {
      uint32_t c[2] = {qc[0], qc[1]};
      ScanCell<2>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
      return;
    }
    for (int x = max(x0, 0); x <= min(x1, X - 1) && !st.stop; ++x) {
      if (y0 >= 0) {
        uint32_t c[2] = {(uint32_t)x, (uint32_t)y0};
        ScanCell<2>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
      }
      if (y1 < Y) {
        uint32_t c[2] = {(uint32_t)x, (uint32_t)y1};
        ScanCell<2>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
      }
    }
    for (int y = max(y0 + 1, 0); y <= min(y1 - 1, Y - 1) && !st.stop; ++y) {
      if (x0 >= 0) {
        uint32_t c[2] = {(uint32_t)x0, (uint32_t)y};
        ScanCell<2>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
      }
      if (x1 < X) {
        uint32_t c[2] = {(uint32_t)x1, (uint32_t)y};
        ScanCell<2>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
      }
    }
  } else {
    const int X = (int)g.dims[0], Y = (int)g.dims[1], Z = (int)g.dims[2];
    const int x0 = (int)qc[0] - ir, x1 = (int)qc[0] + ir;
    const int y0 = (int)qc[1] - ir, y1 = (int)qc[1] + ir;
    const int z0 = (int)qc[2] - ir, z1 = (int)qc[2] + ir;
    if (r == 0) {
      uint32_t c[3] = {qc[0], qc[1], qc[2]};
      ScanCell<3>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
      return;
    }
    for (int z = max(z0, 0); z <= min(z1, Z - 1) && !st.stop; ++z) {
      const bool zface = z == z0 || z == z1;
      for (int y = max(y0, 0); y <= min(y1, Y - 1) && !st.stop; ++y) {
        const bool yface = y == y0 || y == y1;
        if (zface || yface) {
          for (int x = max(x0, 0); x <= min(x1, X - 1) && !st.stop; ++x) {
            uint32_t c[3] = {(uint32_t)x, (uint32_t)y, (uint32_t)z};
            ScanCell<3>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
          }
        } else {
          if (x0 >= 0) {
            uint32_t c[3] = {(uint32_t)x0, (uint32_t)y, (uint32_t)z};
            ScanCell<3>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
          }
          if (x1 < X && !st.stop) {
            uint32_t c[3] = {(uint32_t)x1, (uint32_t)y, (uint32_t)z};
            ScanCell<3>(g, qf, qd, c, cmax2, eb, tie_index, max_cells, st);
          }
        }
      }
    }
  }
}

template <int ND>
__global__ __launch_bounds__(kBlock) void GridQueryKernel(GridQueryArgs args) {
  const GridDev& g = args.grid;
  const uint32_t lane = LaneId();
  const uint32_t lanes_total = gridDim.x * blockDim.x;
  const uint32_t wave_base = (blockIdx.x * blockDim.x + threadIdx.x) - lane;
  unsigned long long evals = 0, lookups = 0;
  for (uint32_t base = args.pos_begin + wave_base; base < args.pos_end;
       base += lanes_total) {
    const uint32_t pos = args.stride > 1u ? (base + lane) * args.stride : base + lane;
    const bool active = base + lane < args.pos_end &&
                        !(args.skip_stride > 1u && pos % args.skip_stride == 0u &&
                          pos / args.skip_stride < args.skip_n);
    const float cmax2 = args.eb ? LoadCmax2(args.key) : 0.f;
    QueryState st{kFltMax, 0xFFFFFFFFu, 0, 0, 0, false, false, false};
    uint32_t idx = 0;
    if (active) {
      idx = args.order_mode == 1 ? Permute(args.perm, pos)
            : args.order_mode == 2 ? pos
                                   : args.order[pos];
      float qf[ND];
      double qd[ND];
      uint32_t qc[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        qf[d] = args.a[(uint64_t)idx * ND + d];
        qd[d] = (double)qf[d];
        qc[d] = CellCoord(g, d, qf[d]);
      }
      for (uint32_t r = 0;; ++r) {
        ScanRing<ND>(g, qf, qd, qc, r, cmax2, args.eb != 0, args.nn_d2 != nullptr,
                     args.max_cells, st);
        if (st.stop) break;
        const double bound = OutsideBound2<ND>(g, qd, qc, r);
        if (bound < 0.0) break;
        if (args.nn_d2 != nullptr ? bound > (double)st.best
                                  : bound >= (double)st.best)
          break;
      }
    }
    evals += st.evals;
    lookups += st.lookups;
    if (active && args.nn_d2 != nullptr) {
      args.nn_d2[idx] = st.best;
      args.nn_idx[idx] = st.best_j;
    }
    // Escalations (cell budget exhausted) are appended wave-aggregated.
    const bool esc = active && st.over;
    const unsigned long long esc_mask = __ballot(esc);
    if (esc_mask != 0ull) {
      uint32_t slot_base = 0;
      if (lane == 0)
        slot_base = atomicAdd(args.esc_count, (unsigned)__popcll(esc_mask));
      slot_base = Uniform(slot_base);
      if (esc) {
        const uint32_t rank =
            __popcll(esc_mask & ((1ull << lane) - 1ull));
        args.esc_idx[slot_base + rank] = idx;
        args.esc_min[slot_base + rank] = st.best;
      }
    }
    if (args.key != nullptr) {
      const bool done = active && !st.stop && st.best != kFltMax;
      const unsigned long long k = done ? PackKey(st.best, idx) : 0ull;
      const unsigned long long kmax = WaveMaxU64(k);
      if (lane == 0 && kmax != 0ull) KeyMax(args.key, kmax);
    }
  }
  // Per-wave evaluation and lookup counts.
  unsigned long long e = evals, l = lookups;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    l += __shfl_xor(l, off, 64);
  }
  if (lane == 0 && args.evals != nullptr) CounterAdd(args.evals, e);
  if (lane == 0 && args.lookups != nullptr) CounterAdd(args.lookups, l);
}

}  // namespace

int BuildGrid(hd_context* ctx, int n_dims, const float* b, uint64_t n_b,
              uint32_t n_points_cell, GridIndex* out, const float* mbr_b) {
  const int nd = n_dims;
  float mbr[6];
  if (mbr_b)
    std::memcpy(mbr, mbr_b, sizeof(float) * 2 * nd);
  else
    HD_TRY(HostMbr(ctx, nd, b, n_b, mbr));
  GridDev& g = out->dev;
  std::memset(&g, 0, sizeof(g));
  // Resolution: reference uniform_grid.h:222-244 (cell side from the
  // volume and points per cell, dims = ceil(extent / side)).
  // An axis whose extent is zero (a flat set, or one whose +1e-5 widening
  // rounded away at large coordinates) gets one cell, and the side comes
  // from the other axes' volume in their own dimension; otherwise a zero
  // volume would give s = 0, which no growth below can leave.
  double volume = 1, max_ext = 0;
  int k = 0;
  for (int d = 0; d < nd; ++d) {
    const double ext = (double)(mbr[nd + d] - mbr[d]);
    max_ext = std::max(max_ext, ext);
    if (ext > 0) {
      volume *= ext;
      ++k;
    }
  }
  const double npc = n_points_cell ? n_points_cell : 8;
  double s = 1.0;
  if (k == nd)
    s = nd == 2 ? std::sqrt(volume * npc / (double)n_b)
                : std::cbrt(volume * npc / (double)n_b);
  else if (k > 0)
    s = std::pow(volume * npc / (double)n_b, 1.0 / k);
  // No axis may need more than 65535 cells.
  if (!(s > 0) || !std::isfinite(s)) s = max_ext > 0 ? max_ext : 1.0;
  s = std::max(s, max_ext / 65535.0);
  // Keep the dense table within budget: at most max(2^20, 4 |B|) cells and
  // per-axis sizes that fit the 32-bit cell id.
  const double cap = std::max<double>(1u << 20, 4.0 * (double)n_b);
  uint64_t n_cells = 0;
  for (int it = 0; it < 200; ++it) {
    n_cells = 1;
    bool fits = true;
    for (int d = 0; d < nd; ++d) {
      double ext = (double)(mbr[nd + d] - mbr[d]);
      double dim = std::max(1.0, std::ceil(ext / s));
      if (dim > 65535.0) fits = false;
      g.dims[d] = (uint32_t)std::min(dim, 65535.0);
      n_cells *= g.dims[d];
    }
    if (fits && (double)n_cells <= cap) break;
    s *= 1.1;
  }
  for (int d = 0; d < nd; ++d) {
    g.lo[d] = mbr[d];
    g.hi[d] = mbr[nd + d];
    g.lo_d[d] = (double)mbr[d];
    g.s_d[d] = ((double)mbr[nd + d] - (double)mbr[d]) / (double)g.dims[d];
    // FP32 cell assignment is off by at most ~3 ulps of dims * t.
    g.margin_d[d] = (3.0 * 0x1p-24 * g.dims[d] + 0x1p-20) * g.s_d[d] +
                    0x1p-22 * (std::fabs((double)mbr[d]) +
                               std::fabs((double)mbr[nd + d]));
    g.max_margin = std::max(g.max_margin, g.margin_d[d]);
  }
  uint32_t max_dim = 0;
  for (int d = 0; d < nd; ++d) max_dim = std::max(max_dim, g.dims[d]);
  g.morton = (nd == 2 && max_dim <= 65536u) || (nd == 3 && max_dim <= 1024u);
  out->n_dims = nd;
  out->n = (uint32_t)n_b;
  out->n_cells = n_cells;

  uint32_t *keys = nullptr, *vals = nullptr, *keys_s = nullptr,
           *vals_s = nullptr;
  HD_TRY(ContextArray(ctx, "grid_keys", n_b, &keys));
  HD_TRY(ContextArray(ctx, "grid_vals", n_b, &vals));
  HD_TRY(ContextArray(ctx, "grid_keys_s", n_b, &keys_s));
  HD_TRY(ContextArray(ctx, "grid_vals_s", n_b, &vals_s));
  float *x = nullptr, *y = nullptr, *z = nullptr;
  uint32_t *orig = nullptr, *start = nullptr, *end = nullptr;
  HD_TRY(ContextArray(ctx, "grid_x", n_b, &x));
  HD_TRY(ContextArray(ctx, "grid_y", n_b, &y));
  if (nd == 3) HD_TRY(ContextArray(ctx, "grid_z", n_b, &z));
  HD_TRY(ContextArray(ctx, "grid_orig", n_b, &orig));
  HD_TRY(ContextArray(ctx, "grid_start", n_cells, &start));
  HD_TRY(ContextArray(ctx, "grid_end", n_cells, &end));
  HD_TRY(FillU32(ctx, start, 0u, n_cells));
  HD_TRY(FillU32(ctx, end, 0u, n_cells));
  const int blocks = (int)std::min<uint64_t>((n_b + 255) / 256, 8192);
  if (nd == 2)
    GridKeysKernel<2><<<blocks, 256, 0, ctx->stream>>>(g, b, (uint32_t)n_b,
                                                       keys, vals);
  else
    GridKeysKernel<3><<<blocks, 256, 0, ctx->stream>>>(g, b, (uint32_t)n_b,
                                                       keys, vals);
  HD_HIP_TRY(hipGetLastError());
  int key_bits = 32;
  if (!g.morton) {
    key_bits = 1;
    while (key_bits < 32 && (1ull << key_bits) < n_cells) ++key_bits;
  }
  HD_TRY(SortPairsU32(ctx, keys, keys_s, vals, vals_s, (uint32_t)n_b,
                      key_bits));
  if (nd == 2)
    GridFillKernel<2><<<blocks, 256, 0, ctx->stream>>>(
        g, b, (uint32_t)n_b, keys_s, vals_s, x, y, z, orig, start, end);
  else
    GridFillKernel<3><<<blocks, 256, 0, ctx->stream>>>(
        g, b, (uint32_t)n_b, keys_s, vals_s, x, y, z, orig, start, end);
  HD_HIP_TRY(hipGetLastError());
  g.x = x;
  g.y = y;
  g.z = z;
  g.orig = orig;
  g.start = start;
  g.end = end;
  return HD_OK;
}

int LaunchGridQuery(hd_context* ctx, const GridQueryArgs& args) {
  if (args.pos_begin >= args.pos_end) return HD_OK;
  const uint32_t n = args.pos_end - args.pos_begin;
  int blocks = PersistentBlocks(ctx, kBlock, 8);
  blocks = (int)std::min<uint32_t>((uint32_t)blocks, (n + kBlock - 1) / kBlock);
  blocks = std::max(blocks, 1);
  StageDispatch(ctx);
  if (args.grid_nd == 2)
    GridQueryKernel<2><<<blocks, kBlock, 0, ctx->stream>>>(args);
  else
    GridQueryKernel<3><<<blocks, kBlock, 0, ctx->stream>>>(args);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

namespace {

// Non-empty cells a seed query may look up before it escalates.
constexpr uint32_t kSeedCells = 1024u;

template <int ND>
int RunGridImpl(hd_context* ctx, const DirectedArgs& in, uint32_t max_cells,
                hd_directed_result* out) {
  const uint32_t n_a = (uint32_t)in.n_a;
  const int st_build = StageBegin(ctx, "grid_build");
  GridIndex grid;
  HD_TRY(BuildGrid(ctx, ND, in.b, in.n_b, in.n_points_cell, &grid,
                   in.has_mbr ? in.mbr_b : nullptr));
  StageEnd(ctx, st_build);

  // key, seed and query counters, then the early-break counters of the
  // escalated queries.
  // key, evaluation counters (seed, query, EB finish), then the cell
  // lookups of the seed and of the query.
  constexpr int kKeys = 3 + kEbCounters;
  constexpr int kLookups = 2;
  unsigned long long* key = nullptr;
  static_assert(kKeys + kLookups <= (int)kCounterRow, "counter block row");
  HD_TRY(ContextArray(ctx, "grid_key", kCounterWords, &key));
  HD_TRY(FillU64(ctx, key, 0ull, kCounterWords));
  unsigned int* esc_count = nullptr;
  HD_TRY(ContextArray(ctx, "grid_esc_count", 1, &esc_count));
  HD_TRY(FillU32(ctx, esc_count, 0u, 1));
  // Escalations: queries that looked up too many cells (the hybrid's
  // max_hit, and every seed query past kSeedCells) finish on the EB rounds.
  uint32_t* esc_idx = nullptr;
  float* esc_min = nullptr;
  HD_TRY(ContextArray(ctx, "grid_esc_idx", n_a, &esc_idx));
  HD_TRY(ContextArray(ctx, "grid_esc_min", n_a, &esc_min));

  GridQueryArgs qa;
  std::memset(&qa, 0, sizeof(qa));
  qa.grid = grid.dev;
  qa.grid_nd = ND;
  qa.a = in.a;
  const bool shuffle = (in.flags & HD_FLAG_NO_SHUFFLE) == 0;
  qa.perm = MakeFeistel(n_a, (uint64_t)(int64_t)in.seed, 0);
  qa.order_mode = shuffle ? 1 : 2;
  if (in.flags & HD_FLAG_LEARNED_ORDER) {
    const uint32_t* order = nullptr;
    HD_TRY(LearnedOrder(ctx, in, &order));
    qa.order = order;
    qa.order_mode = 0;
  }
  qa.key = key;
  qa.esc_count = esc_count;
  qa.esc_idx = esc_idx;
  qa.esc_min = esc_min;

  // Queries in Morton order of A (in B's frame): lanes of a wave look up
  // neighbouring cells, so their table and point loads share cache lines
  // (a learned order is kept as given).
  const uint32_t s0 = std::min<uint32_t>(n_a, 65536u);
  uint32_t stride = 1;
  if (qa.order_mode != 0) {
    const int st_ord = StageBegin(ctx, "grid_order");
    LbvhFrame f;
    std::memset(&f, 0, sizeof(f));
    const double levels = ND == 2 ? 4294967040.0 : 2097151.0;
    f.qmax = (float)levels;
    for (int d = 0; d < ND; ++d) {
      f.lo[d] = grid.dev.lo[d];
      f.scale[d] = (float)(levels / ((double)grid.dev.hi[d] - (double)grid.dev.lo[d]));
    }
    uint32_t* mo = nullptr;
    // The grid's queries need locality, not the full code: the top 16 bits
    // (a 256 x 256 grid over the frame in 2D) in two radix passes instead of
    // four (10M GMM: grid 2.63 -> 2.51 ms). The LBVH keeps all 32 bits; its
    // traversals lose more than the sort saves with the coarse order
    // (2.97 -> 3.78 ms).
    HD_TRY(MortonOrder(ctx, ND, in.a, n_a, f, &mo, 16));
    StageEnd(ctx, st_ord);
    qa.order = mo;
    qa.order_mode = 0;
    stride = n_a / s0;
  }

  // Seed: exact nearest neighbours of s0 points of the order (in Morton
  // order: every stride-th point, a spatially stratified sample).
  const int st_seed = StageBegin(ctx, "grid_seed");
  qa.pos_begin = 0;
  qa.pos_end = s0;
  qa.stride = stride;
  qa.eb = 0;
  // Far seed queries (long runs of empty rings) escalate: the seed is a few
  // waves, so one far query would set its time.
  qa.max_cells = max_cells ? std::max<uint32_t>(max_cells, kSeedCells) : kSeedCells;
  qa.evals = key + 1;
  qa.lookups = key + kKeys;
  HD_TRY(LaunchGridQuery(ctx, qa));
  StageEnd(ctx, st_seed);
  // Everything else with early break against the running cmax2 (with a
  // stratified seed, the whole order minus the seed's positions).
  const int st_q = StageBegin(ctx, "grid_query");
  qa.pos_begin = stride > 1 ? 0u : s0;
  qa.pos_end = n_a;
  qa.stride = 1;
  qa.skip_stride = stride;
  qa.skip_n = s0;
  qa.eb = 1;
  qa.max_cells = max_cells;
  qa.evals = key + 2;
  qa.lookups = key + kKeys + 1;
  HD_TRY(LaunchGridQuery(ctx, qa));
  StageEnd(ctx, st_q);
  uint32_t n_esc = 0;
  {
    // Queries that needed too many cells: exact early-break scans, starting
    // from the partial minimum the grid found.
    void* pinned = nullptr;
    HD_TRY(ContextPinned(ctx, 256, &pinned));
    unsigned int* h = static_cast<unsigned int*>(pinned);
    HD_HIP_TRY(hipMemcpyAsync(h, esc_count, 4, hipMemcpyDeviceToHost,
                              ctx->stream));
    HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
    n_esc = *h;
    if (n_esc > 0) {
      const int st_eb = StageBegin(ctx, "hybrid_eb");
      EbTarget t;
      // B pseudo-randomly ordered unless the caller asks otherwise: the
      // escalated points are the far ones, which an evenly spread order
      // does not reach sooner, and the stratified order costs a sort of B.
      const uint32_t order_flags =
          in.flags & (HD_FLAG_NO_SHUFFLE | HD_FLAG_B_ORDER_STRATIFIED);
      HD_TRY(EbPrepareTarget(ctx, ND, in.b, (uint32_t)in.n_b, order_flags,
                             in.seed, in.has_mbr ? in.mbr_b : nullptr, &t));
      EbRoundStats rs;
      HD_TRY(EbFinishList(ctx, ND, in.a, n_a, t, esc_idx, esc_min, n_esc, key,
                          key + 3, &rs));
      StageEnd(ctx, st_eb);
    }
  }
  const int st_fin = StageBegin(ctx, "grid_finish");
  HD_TRY(FinishDirected(ctx, ND, in.a, in.b, in.n_b, key, key + 1, kKeys - 1,
                        out));
  StageEnd(ctx, st_fin);
  if (ctx->profiling) {
    unsigned long long sums[kCounterRow];
    HD_TRY(ReadCounters(ctx, key, sums));
    const unsigned long long* h = sums + 1;
    StageCount(ctx, "grid_seed", h[0], s0, h[kKeys - 1]);
    StageCount(ctx, "grid_query", h[1], n_a - s0, h[kKeys]);
    StageCount(ctx, "grid_build", 0, in.n_b);
    unsigned long long eb_evals = 0;
    for (int i = 2; i < kKeys - 1; ++i) eb_evals += h[i];
    StageCount(ctx, "hybrid_eb", eb_evals, n_esc);
  }
  return HD_OK;
}

}  // namespace

int RunGrid(hd_context* ctx, const DirectedArgs& args,
            hd_directed_result* out) {
  if (args.n_dims == 2) return RunGridImpl<2>(ctx, args, 0, out);
  return RunGridImpl<3>(ctx, args, 0, out);
}

int RunGridHybrid(hd_context* ctx, const DirectedArgs& args,
                  uint32_t max_cells, hd_directed_result* out) {
  if (args.n_dims == 2) return RunGridImpl<2>(ctx, args, max_cells, out);
  return RunGridImpl<3>(ctx, args, max_cells, out);
}

int GridNearestNeighbors(hd_context* ctx, int n_dims, const float* q,
                         uint64_t n_q, const float* ref, uint64_t n_ref,
                         float* out_d2, uint32_t* out_idx) {
  GridIndex grid;
  HD_TRY(BuildGrid(ctx, n_dims, ref, n_ref, 8, &grid));
  GridQueryArgs qa;
  std::memset(&qa, 0, sizeof(qa));
  qa.grid = grid.dev;
  qa.grid_nd = n_dims;
  qa.a = q;
  qa.order_mode = 2;
  qa.pos_begin = 0;
  qa.pos_end = (uint32_t)n_q;
  qa.nn_d2 = out_d2;
  qa.nn_idx = out_idx;
  return LaunchGridQuery(ctx, qa);
}

}  // namespace hd
