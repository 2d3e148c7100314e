// Input statistics for the run's stats JSON: MBR, density and a uniform
// "stats grid" per point set (src/run_hausdorff_distance.cu:151-204).
//
// The stats grid is its own dense table of per-cell point counts, not the
// query grid: its resolution follows uniform_grid.h:222-244 on the MBR made
// non-empty by mbr.h:288-296 (a zero extent becomes 1e-5), with no cell cap,
// and a point's cell is the reference's FP32 formula (uniform_grid.h:138-159):
// clamp(((v - lower) / (upper - lower)) * dims, 0, dims - 1), truncated. The
// union of A and B is counted by running the count kernel over each set into
// one table (same counts as counting their concatenation).
//
// Statistics (uniform_grid.h:343-404, :675-692): total and non-empty cells,
// the largest cell, the Gini index of the non-empty cells' counts
// (src/running_stats.h:32-57) and the percentile list the reference parses
// from its HdrHistogram dump (running_stats.h:106-140).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "hd_device.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {
namespace {

// Dense tables above this many cells are not built (the entry then reports
// the grid's size and why it was skipped).
constexpr uint64_t kMaxStatsCells = 1ull << 31;

struct StatsGridDev {
  float lo[3], hi[3];
  uint32_t dims[3];
};

template <int ND>
__global__ void StatsCellKernel(StatsGridDev g, const float* __restrict__ pts,
                                uint64_t n, unsigned int* __restrict__ counts) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t cell = 0, mul = 1;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const float v = pts[i * ND + d];
      const float norm = (v - g.lo[d]) / (g.hi[d] - g.lo[d]);
      const float gs = (float)g.dims[d];
      const float x = fminf(fmaxf(norm * gs, 0.0f), gs - 1.0f);
      cell += (uint64_t)(uint32_t)x * mul;
      mul *= g.dims[d];
    }
    atomicAdd(&counts[cell], 1u);
  }
}

// Non-empty cells and the largest count. Grid-stride over the table with a
// few blocks per CU; wave then block sums, one pair of atomics per block
// (an atomic per wave on two addresses serialises into the kernel's cost).
__global__ __launch_bounds__(256) void CellCountKernel(
    const unsigned int* __restrict__ counts, uint64_t n_cells,
    unsigned int* nonempty, unsigned int* max_count) {
  __shared__ unsigned int s_ne[4], s_mx[4];
  unsigned int ne = 0, mx = 0;
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       c < n_cells; c += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned int k = counts[c];
    ne += k > 0u;
    mx = max(mx, k);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    ne += __shfl_xor(ne, off, 64);
    mx = max(mx, (unsigned int)__shfl_xor(mx, off, 64));
  }
  const unsigned int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0) {
    s_ne[wave] = ne;
    s_mx[wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ne = s_ne[0] + s_ne[1] + s_ne[2] + s_ne[3];
    mx = max(max(s_mx[0], s_mx[1]), max(s_mx[2], s_mx[3]));
    if (ne) atomicAdd(nonempty, ne);
    if (mx) atomicMax(max_count, mx);
  }
}

// Cells per count. Most cells hold few points, so a handful of bins take
// nearly every update: counts below kHistBins go to a per-block LDS
// histogram that is added to the global one at the end; larger counts
// (rare) go to the global one directly. hist has max_count + 1 entries.
constexpr unsigned int kHistBins = 4096;

__global__ __launch_bounds__(256) void CellHistogramKernel(
    const unsigned int* __restrict__ counts, uint64_t n_cells,
    unsigned int* hist, unsigned int max_count) {
  __shared__ unsigned int bins[kHistBins];
  for (unsigned int i = threadIdx.x; i < kHistBins; i += blockDim.x) bins[i] = 0u;
  __syncthreads();
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
       c < n_cells; c += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned int k = counts[c];
    if (k == 0u) continue;
    if (k < kHistBins)
      atomicAdd(&bins[k], 1u);
    else
      atomicAdd(&hist[k], 1u);
  }
  __syncthreads();
  const unsigned int lim = min(max_count + 1u, kHistBins);
  for (unsigned int i = threadIdx.x; i < lim; i += blockDim.x)
    if (bins[i]) atomicAdd(&hist[i], bins[i]);
}

std::string Num(double v) {
  // JSON has no inf or nan (a flat set's density divides by a zero volume).
  if (!std::isfinite(v)) return "null";
  std::ostringstream o;
  o.precision(17);
  o << v;
  return o.str();
}

// A percentile as the reference's parser reads it back from "%12f":
// rounded to 6 decimals.
std::string Fixed6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.6f", v);
  std::string s(buf);
  while (s.size() > 2 && s.back() == '0' && s[s.size() - 2] != '.') s.pop_back();
  return s;
}

// HdrHistogram with lowest value 1 and 3 significant figures: values below
// 2048 have their own counter; above, a value shares a counter with the
// values that agree with it in the top 11 bits, and is reported as the
// largest of them.
uint64_t HdrHighestEquivalent(uint64_t v) {
  int shift = 0;
  while ((v >> shift) >= 2048u) ++shift;
  return ((v >> shift) << shift) + (1ull << shift) - 1;
}

// Values from this one up are outside a histogram made for `highest` (the
// smallest 2048 * 2^k above it); the reference's record call drops them.
uint64_t HdrUntrackable(uint64_t highest) {
  uint64_t limit = 2048;
  while (limit <= highest && limit < (1ull << 62)) limit <<= 1;
  return limit;
}

// Percentile list of HdrHistogram's percentile iterator with 3 ticks per
// half distance over (value, count) pairs in ascending value order: the
// level starts at 0; at each value whose cumulative share reaches the level
// the entry {value, level / 100, cumulative count} is written and the level
// advances by 100 / (3 * 2^(trunc(log2(100 / (100 - level))) + 1)); at the
// last value only one entry is written, then {last value, 1, total}.
std::string PercentileList(const std::vector<std::pair<uint64_t, uint64_t>>& vc) {
  uint64_t total = 0;
  for (const auto& p : vc) total += p.second;
  std::ostringstream o;
  o << "[";
  if (total == 0) {
    o << "]";
    return o.str();
  }
  double level = 0.0;
  uint64_t cum = 0, last_v = 0;
  bool first = true;
  auto emit = [&](uint64_t v, double pct, uint64_t count) {
    o << (first ? "" : ", ") << "{\"value\": " << v << ".0, \"percentile\": "
      << Fixed6(pct / 100.0) << ", \"count\": " << count << "}";
    first = false;
  };
  for (const auto& p : vc) {
    if (p.second == 0) continue;
    cum += p.second;
    last_v = p.first;
    const double reached = (100.0 * (double)cum) / (double)total;
    while (level <= reached) {
      emit(p.first, level, cum);
      const int64_t half = (int64_t)(std::log(100.0 / (100.0 - level)) / std::log(2.0)) + 1;
      const int64_t ticks = 3 * (int64_t)std::pow(2.0, (double)half);
      level += 100.0 / (double)ticks;
      if (cum == total) break;
    }
  }
  emit(last_v, 100.0, total);
  o << "]";
  return o.str();
}

}  // namespace

// Gini index of the multiset {v with multiplicity hist[v]}:
// sum_i (2i - n + 1) x_i / (n sum x) over the values sorted ascending
// (running_stats.h:32-57), evaluated exactly per run of equal values in
// FP64 (the reference sums in FP32).
double GiniFromHistogram(const std::vector<uint64_t>& hist) {
  double n = 0, total = 0;
  for (size_t v = 0; v < hist.size(); ++v) {
    n += (double)hist[v];
    total += (double)v * (double)hist[v];
  }
  if (n == 0 || total == 0) return 0.0;
  double weighted = 0, rank = 0;  // ranks [rank, rank + c) hold value v
  for (size_t v = 0; v < hist.size(); ++v) {
    const double c = (double)hist[v];
    if (c == 0) continue;
    // sum over i in [r, r + c) of (2i - n + 1) = c (2r + c - 1) - c (n - 1)
    weighted += (double)v * (c * (2 * rank + c - 1) - c * (n - 1));
    rank += c;
  }
  return weighted / (n * total);
}

// The reference's "Histogram" list for the non-empty cells' counts, given as
// hist[v] = number of cells holding v points (v >= 1): an HdrHistogram made
// for values 1..(number of cells) with 3 significant figures
// (uniform_grid.h:675-692), so counts at or above its untrackable limit are
// left out, and counts from 2048 up are merged into their sub-buckets.
std::string CellCountPercentiles(const std::vector<uint64_t>& hist) {
  uint64_t cells = 0;
  for (size_t v = 1; v < hist.size(); ++v) cells += hist[v];
  const uint64_t limit = HdrUntrackable(cells);
  std::vector<std::pair<uint64_t, uint64_t>> vc;
  for (size_t v = 1; v < hist.size() && v < limit; ++v) {
    if (hist[v] == 0) continue;
    const uint64_t key = HdrHighestEquivalent(v);
    if (!vc.empty() && vc.back().first == key)
      vc.back().second += hist[v];
    else
      vc.emplace_back(key, hist[v]);
  }
  return PercentileList(vc);
}

// MBR made non-empty (mbr.h:288-296): a zero extent's upper bound becomes
// FP32(lower + 1e-5).
void NonEmptyMbr(int nd, const float* mbr, float* lo, float* hi) {
  for (int d = 0; d < nd; ++d) {
    lo[d] = mbr[d];
    hi[d] = mbr[nd + d];
    if (hi[d] - lo[d] == 0.0f) hi[d] = (float)((double)lo[d] + 1e-5);
  }
}

// FP32 volume of an MBR (mbr.h:98-104).
float MbrVolume(int nd, const float* lo, const float* hi) {
  float v = 1.0f;
  for (int d = 0; d < nd; ++d) v *= hi[d] - lo[d];
  return v;
}

// The stats JSON object of the point sets `sets` taken together (device AoS;
// `mbr` is their raw MBR, lower then upper): {"NumPoints", "MBR", "Density",
// "Grid": {...}, "StatsGridNumPointsPerCell"}; npc 0 stops after "Density".
int InputStatsJson(hd_context* ctx, int nd, const StatsInput* sets, int n_sets,
                   const float* mbr, uint32_t npc, const float* density,
                   std::string* json) {
  uint64_t n = 0;
  for (int s = 0; s < n_sets; ++s) n += sets[s].n;
  StatsGridDev g;
  std::memset(&g, 0, sizeof(g));
  NonEmptyMbr(nd, mbr, g.lo, g.hi);
  // Resolution (uniform_grid.h:222-244): FP32 extents, FP64 volume and side.
  double volume = 1;
  for (int d = 0; d < nd; ++d) volume *= (double)(g.hi[d] - g.lo[d]);
  const double side = nd == 2 ? std::sqrt(volume * (double)npc / (double)(uint32_t)n)
                              : std::cbrt(volume * (double)npc / (double)(uint32_t)n);
  std::string skipped;
  uint64_t n_cells = 1;
  if (npc == 0) {
    skipped = "-";
  } else if (!(volume > 0) || !(side > 0)) {
    skipped = "zero-volume MBR";
  } else {
    for (int d = 0; d < nd; ++d) {
      const double dim = std::ceil((double)(g.hi[d] - g.lo[d]) / side);
      if (!(dim <= 4294967295.0)) {
        skipped = "grid size exceeds 32 bits per axis";
        break;
      }
      g.dims[d] = (uint32_t)std::max(dim, 1.0);
      n_cells = n_cells * g.dims[d];
      if (n_cells > kMaxStatsCells) {
        skipped = "more than 2^31 cells";
        break;
      }
    }
  }
  std::ostringstream o;
  o << "{\"NumPoints\": " << n << ", \"MBR\": [";
  for (int d = 0; d < nd; ++d)
    o << (d ? ", " : "") << "{\"Lower\": " << Num(g.lo[d])
      << ", \"Upper\": " << Num(g.hi[d]) << "}";
  o << "], \"Density\": "
    << Num(density ? *density : (float)n / MbrVolume(nd, g.lo, g.hi));
  if (npc == 0) {
    o << "}";
    *json = o.str();
    return HD_OK;
  }
  o << ", \"Grid\": {";
  if (!skipped.empty()) {
    o << "\"Skipped\": \"" << skipped << "\"}, \"StatsGridNumPointsPerCell\": "
      << npc << "}";
    *json = o.str();
    return HD_OK;
  }

  unsigned int* counts = nullptr;
  unsigned int* counters = nullptr;
  HD_TRY(ContextArray(ctx, "stats_counts", n_cells, &counts));
  HD_TRY(ContextArray(ctx, "stats_counters", 2, &counters));
  HD_TRY(FillU32(ctx, counts, 0u, n_cells));
  HD_TRY(FillU32(ctx, counters, 0u, 2));
  for (int s = 0; s < n_sets; ++s) {
    if (sets[s].n == 0) continue;
    const int pb = (int)std::min<uint64_t>((sets[s].n + 255) / 256, 8192);
    if (nd == 2)
      StatsCellKernel<2><<<pb, 256, 0, ctx->stream>>>(g, sets[s].pts, sets[s].n, counts);
    else
      StatsCellKernel<3><<<pb, 256, 0, ctx->stream>>>(g, sets[s].pts, sets[s].n, counts);
    HD_HIP_TRY(hipGetLastError());
  }
  const int cb = (int)std::min<uint64_t>((n_cells + 255) / 256,
                                         (uint64_t)std::max(1, ctx->cu_count) * 4);
  CellCountKernel<<<cb, 256, 0, ctx->stream>>>(counts, n_cells, counters,
                                               counters + 1);
  HD_HIP_TRY(hipGetLastError());
  unsigned int h[2] = {0, 0};
  HD_HIP_TRY(hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost,
                            ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const uint32_t nonempty = h[0], max_count = h[1];
  std::vector<uint64_t> hist(max_count + 1, 0);
  if (max_count > 0) {
    unsigned int* d_hist = nullptr;
    HD_TRY(ContextArray(ctx, "stats_hist", max_count + 1, &d_hist));
    HD_TRY(FillU32(ctx, d_hist, 0u, max_count + 1));
    CellHistogramKernel<<<cb, 256, 0, ctx->stream>>>(counts, n_cells, d_hist,
                                                     max_count);
    HD_HIP_TRY(hipGetLastError());
    std::vector<unsigned int> hh(max_count + 1);
    HD_HIP_TRY(hipMemcpyAsync(hh.data(), d_hist, sizeof(unsigned int) * hh.size(),
                              hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (size_t v = 0; v < hh.size(); ++v) hist[v] = hh[v];
  }
  o << "\"TotalCells\": " << n_cells << ", \"NonEmptyCells\": " << nonempty
    << ", \"MaxPoints\": " << max_count
    << ", \"GiniIndex\": " << Num(GiniFromHistogram(hist)) << ", \"GridSize\": [";
  for (int d = 0; d < nd; ++d) o << (d ? ", " : "") << g.dims[d];
  o << "], \"Histogram\": " << CellCountPercentiles(hist)
    << "}, \"StatsGridNumPointsPerCell\": " << npc << "}";
  *json = o.str();
  return HD_OK;
}

}  // namespace hd

using namespace hd;

extern "C" {

int hd_stats_gini(const uint64_t* hist, uint64_t n_values, double* out) {
  ClearError();
  if (!out || (!hist && n_values)) return Fail(HD_EINVAL, "null argument");
  *out = GiniFromHistogram(std::vector<uint64_t>(hist, hist + n_values));
  return HD_OK;
}

int hd_stats_percentiles(const uint64_t* hist, uint64_t n_values, char* json,
                         uint64_t cap) {
  ClearError();
  if (!json || cap == 0 || (!hist && n_values)) return Fail(HD_EINVAL, "null argument");
  const std::string s =
      CellCountPercentiles(std::vector<uint64_t>(hist, hist + n_values));
  std::strncpy(json, s.c_str(), cap - 1);
  json[cap - 1] = '\0';
  return s.size() + 1 > cap ? Fail(HD_ETRUNC, "percentile JSON truncated") : HD_OK;
}

}  // extern "C"
