// RT-HDIST, the quantized-grid baseline (reference
// src/hd_impl/hausdorff_distance_rt_hdist.h:45-222 with the OptiX programs of
// rt/shaders/shaders_nn_quantized_grid_{2,3}d.cu and the grid of
// src/index/quantized_grid.h), restated without ray-tracing hardware.
//
// The reference quantizes B onto a grid of 2^7 cells along the longest MBR
// axis, starts from radius ceil(half centre distance / cell diagonal) cell
// diagonals (in cell units) and re-launches every A point whose quantized
// position has no non-empty cell within the radius, growing the radius by
// one diagonal per launch. A point that hits is carried through three more
// launches; at the fourth it takes the exact minimum over the B points of
// every non-empty cell within that launch's radius, and the result is the
// maximum of these minima (at least the exact HD, equal to it whenever the
// nearest B point lies within that radius, as it does on ordinary inputs).
//
// Because the radius sequence r_k (FP32, r_{k+1} = r_k + diagonal) is the
// same for every point, a point's first hit depends only on the squared
// distance dmin2 from its quantized position to the nearest non-empty
// cell: k0 = first k with dmin2 <= r_k * r_k, and its scan radius is
// r_{k0+3}. So the launches collapse into: an exact nearest-cell query for
// every A point (the uniform grid over the non-empty cells' positions),
// then one pass that scans the cells within r_{k0+3}.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hd_device.h"
#include "hd_grid.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {

using namespace dev;

namespace {

constexpr int kQuantBits = 7;  // hausdorff_distance_rt_hdist.h:29

struct QGrid {
  float lo[3];
  float cell;      // cell side (FP32)
  int res[3];      // cells per axis
};

// Quantized position of p (quantized_grid.h:91-103): (int)((v - lower) /
// cell), truncated toward zero, not clamped; held to +-2^30 here so that
// far points stay representable.
template <int ND>
HD_DEV void Quantize(const QGrid& g, const float* p, int* q) {
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const float t = (p[d] - g.lo[d]) / g.cell;
    q[d] = (int)fminf(fmaxf(t, -1073741824.0f), 1073741824.0f);
  }
}

template <int ND>
__global__ void QuantizeBKernel(QGrid g, const float* __restrict__ b,
                                uint32_t n, uint32_t* __restrict__ keys,
                                uint32_t* __restrict__ vals) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    int q[ND];
    Quantize<ND>(g, b + (uint64_t)i * ND, q);
    uint32_t key = 0, mul = 1;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const int c = min(max(q[d], 0), g.res[d] - 1);
      key += (uint32_t)c * mul;
      mul *= (uint32_t)g.res[d];
    }
    keys[i] = key;
    vals[i] = i;
  }
}

// B in cell order (SoA) and each cell's [start, end); first-of-run threads
// also list the non-empty cells' positions as float points.
template <int ND>
__global__ void QuantizeFillKernel(QGrid g, const float* __restrict__ b,
                                   uint32_t n, const uint32_t* __restrict__ keys,
                                   const uint32_t* __restrict__ vals,
                                   float* __restrict__ x, float* __restrict__ y,
                                   float* __restrict__ z, uint32_t* __restrict__ orig,
                                   uint32_t* __restrict__ start,
                                   uint32_t* __restrict__ end,
                                   float* __restrict__ reps,
                                   unsigned int* __restrict__ n_reps) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t src = vals[i];
    const float* p = b + (uint64_t)src * ND;
    x[i] = p[0];
    y[i] = p[1];
    if constexpr (ND == 3) z[i] = p[2];
    orig[i] = src;
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) {
      start[k] = i;
      const unsigned int r = atomicAdd(n_reps, 1u);
      uint32_t rest = k;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        reps[(uint64_t)r * ND + d] = (float)(rest % (uint32_t)g.res[d]);
        rest /= (uint32_t)g.res[d];
      }
    }
    if (i + 1 == n || keys[i + 1] != k) end[k] = i + 1;
  }
}

template <int ND>
__global__ void QuantizeAKernel(QGrid g, const float* __restrict__ a,
                                uint32_t n, float* __restrict__ qa) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    int q[ND];
    Quantize<ND>(g, a + (uint64_t)i * ND, q);
#pragma unroll
    for (int d = 0; d < ND; ++d) qa[(uint64_t)i * ND + d] = (float)q[d];
  }
}

// Maximum of non-negative floats as bits; one atomic per block (256 threads).
__global__ __launch_bounds__(256) void MaxF32Kernel(const float* __restrict__ v, uint32_t n,
                                                    unsigned int* out_bits) {
  __shared__ float part[4];
  float m = 0.f;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    m = fmaxf(m, v[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(out_bits, __float_as_uint(fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]))));
}

// Scan of one A point per wave: first k with dmin2 <= r_k^2, radius
// r_{k+3}; the cells of the box around it are tested 64 at a time (a lane
// per cell), the non-empty ones within the radius are walked with the lanes
// striding over each cell's points, then a wave minimum.
template <int ND>
__global__ __launch_bounds__(256) void RtHdistScanKernel(
    QGrid g, const float* __restrict__ a, uint32_t n_a,
    const float* __restrict__ qa, const float* __restrict__ dmin2,
    const float* __restrict__ radius, uint32_t n_radius,
    const float* __restrict__ bx, const float* __restrict__ by,
    const float* __restrict__ bz, const uint32_t* __restrict__ start,
    const uint32_t* __restrict__ end, unsigned long long* key,
    unsigned long long* evals) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
  unsigned long long my_evals = 0, kmax = 0ull;
  for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n_a; i += n_waves) {
    const float dm = dmin2[i];
    uint32_t lo = 0, hi = n_radius - 4;  // radius[hi]^2 >= max dmin2
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (dm <= radius[mid] * radius[mid]) hi = mid; else lo = mid + 1;
    }
    const float r = radius[lo + 3];
    const float r2 = r * r;
    float q[3] = {0.f, 0.f, 0.f}, p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      q[d] = qa[(uint64_t)i * ND + d];
      p[d] = a[(uint64_t)i * ND + d];
    }
    int c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
    uint32_t w[3] = {1u, 1u, 1u};
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      c0[d] = (int)fmaxf(ceilf(q[d] - r), 0.f);
      c1[d] = (int)fminf(floorf(q[d] + r), (float)(g.res[d] - 1));
      w[d] = c1[d] >= c0[d] ? (uint32_t)(c1[d] - c0[d] + 1) : 0u;
    }
    const uint32_t n_box = w[0] * w[1] * (ND == 3 ? w[2] : 1u);
    float best = kFltMax;
    // The box's cells 64 at a time: lane l takes cell base + l, tests its
    // distance and loads its [start, end) (one round trip for 64 cells),
    // then the wave walks the non-empty ones with the lanes on their points.
    for (uint32_t base = 0; base < n_box; base += 64u) {
      const uint32_t k = base + lane;
      uint32_t cs = 0u, ce = 0u;
      if (k < n_box) {
        const int cx = c0[0] + (int)(k % w[0]);
        const int cy = c0[1] + (int)((k / w[0]) % w[1]);
        const int cz = ND == 3 ? c0[2] + (int)(k / (w[0] * w[1])) : 0;
        const float dx = (float)cx - q[0], dy = (float)cy - q[1];
        float dd = dx * dx + dy * dy;
        if constexpr (ND == 3) {
          const float dz = (float)cz - q[2];
          dd = dd + dz * dz;
        }
        if (dd <= r2) {
          const uint32_t cell =
              (uint32_t)cx + (uint32_t)g.res[0] *
                                 ((uint32_t)cy + (ND == 3 ? (uint32_t)g.res[1] * (uint32_t)cz : 0u));
          cs = start[cell];
          ce = end[cell];
        }
      }
      unsigned long long todo = __ballot(ce > cs);
      while (todo != 0ull) {
        // Four cells at a time, so that four independent loads are in
        // flight per lane (one cell at a time waits a memory round trip per
        // cell). An unused slot has s = e = 0 and reads index 0 unused.
        uint32_t cs4[4], ce4[4], maxlen = 0u;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          cs4[c] = ce4[c] = 0u;
          if (todo != 0ull) {
            const int src = __builtin_ctzll(todo);
            todo &= todo - 1ull;
            cs4[c] = (uint32_t)__shfl(cs, src, 64);
            ce4[c] = (uint32_t)__shfl(ce, src, 64);
          }
          maxlen = max(maxlen, ce4[c] - cs4[c]);
          my_evals += ce4[c] - cs4[c];
        }
        for (uint32_t off = lane; off < maxlen; off += 64u) {
          float px[4], py[4], pz[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const uint32_t j = cs4[c] + off < ce4[c] ? cs4[c] + off : cs4[c];
            px[c] = bx[j];
            py[c] = by[j];
            if constexpr (ND == 3) pz[c] = bz[j];
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            if (cs4[c] + off < ce4[c]) {
              const float d2 = ND == 2 ? D2(p[0], p[1], px[c], py[c])
                                       : D2(p[0], p[1], p[2], px[c], py[c], pz[c]);
              best = fminf(best, d2);
            }
          }
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) best = fminf(best, __shfl_xor(best, off, 64));
    if (best != kFltMax) kmax = max(kmax, PackKey(best, i));
  }
  // One atomic per wave (an atomic per A point on the one key serialises:
  // 1M of them cost ~10 ms). kmax and my_evals are the same in every lane.
  if (lane == 0 && kmax != 0ull) KeyMax(key, kmax);
  if (lane == 0 && my_evals) CounterAdd(evals, my_evals);
}

template <int ND>
__global__ __launch_bounds__(256) void SumKernel(const float* __restrict__ p, uint64_t n,
                                                 double* sums) {
  __shared__ double part[4][3];
  double s[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) s[d] = 0.0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int d = 0; d < ND; ++d) s[d] += (double)p[i * ND + d];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s[d] += __shfl_xor(s[d], off, 64);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6][d] = s[d];
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)ND)
    atomicAdd(&sums[threadIdx.x], part[0][threadIdx.x] + part[1][threadIdx.x] +
                                      part[2][threadIdx.x] + part[3][threadIdx.x]);
}

}  // namespace

int RunRtHdist(hd_context* ctx, const DirectedArgs& args, hd_directed_result* out) {
  const int nd = args.n_dims;
  const uint32_t n_a = (uint32_t)args.n_a, n_b = (uint32_t)args.n_b;
  const int stage_prep = StageBegin(ctx, "rthdist_grid");
  // Centres (primitive_utils.h:149-167; summed in FP64 here) and the MBR
  // of B (rt_hdist.h:78; made non-empty here, which changes only a B with
  // no extent along an axis, where the reference divides by zero).
  double* sums = nullptr;
  HD_TRY(ContextArray(ctx, "rth_sums", 6, &sums));
  HD_TRY(FillU64(ctx, reinterpret_cast<unsigned long long*>(sums), 0ull, 6));
  const int sb = (int)std::min<uint64_t>((std::max(n_a, n_b) + 255) / 256,
                                         (uint64_t)ctx->cu_count * 2);
  if (nd == 2) {
    SumKernel<2><<<sb, 256, 0, ctx->stream>>>(args.a, n_a, sums);
    SumKernel<2><<<sb, 256, 0, ctx->stream>>>(args.b, n_b, sums + 3);
  } else {
    SumKernel<3><<<sb, 256, 0, ctx->stream>>>(args.a, n_a, sums);
    SumKernel<3><<<sb, 256, 0, ctx->stream>>>(args.b, n_b, sums + 3);
  }
  HD_HIP_TRY(hipGetLastError());
  double h_sums[6];
  float mbr[6];
  HD_HIP_TRY(hipMemcpyAsync(h_sums, sums, sizeof(h_sums), hipMemcpyDeviceToHost, ctx->stream));
  HD_TRY(HostMbr(ctx, nd, args.b, n_b, mbr));  // synchronises the stream
  float ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) {
    ca[d] = (float)(h_sums[d] / (double)n_a);
    cb[d] = (float)(h_sums[3 + d] / (double)n_b);
  }
  float cd2 = 0.f;
  for (int d = 0; d < nd; ++d) {
    const float t = ca[d] - cb[d];
    cd2 = cd2 + t * t;
  }
  const float half_dist = std::sqrt(cd2) / 2.0f;

  // Grid (quantized_grid.h:142-158): cell = nextafter(longest / 128), cells
  // per axis = ceil(extent / cell), at least one.
  QGrid g;
  std::memset(&g, 0, sizeof(g));
  float longest = 0.f;
  for (int d = 0; d < nd; ++d) longest = std::max(longest, mbr[nd + d] - mbr[d]);
  g.cell = std::nextafter(longest / (float)(1 << kQuantBits), longest);
  if (!(g.cell > 0.f)) g.cell = 1.0f;  // coordinates so large that 1e-5 vanishes
  uint64_t n_cells = 1;
  for (int d = 0; d < nd; ++d) {
    g.lo[d] = mbr[d];
    g.res[d] = std::max(1, (int)std::ceil((mbr[nd + d] - mbr[d]) / g.cell));
    n_cells *= (uint64_t)g.res[d];
  }
  const float diag_q = (float)std::sqrt((double)nd);
  const float diag = (float)(std::sqrt((double)nd) * g.cell);
  const float n_diag = std::ceil(half_dist / diag);
  const float r0 = n_diag * diag_q;

  // B by cell.
  uint32_t *keys, *vals, *keys_s, *vals_s, *orig, *start, *end;
  float *bx, *by, *bz = nullptr, *reps, *qa, *dmin2;
  unsigned int* n_reps;
  HD_TRY(ContextArray(ctx, "rth_keys", n_b, &keys));
  HD_TRY(ContextArray(ctx, "rth_vals", n_b, &vals));
  HD_TRY(ContextArray(ctx, "rth_keys_s", n_b, &keys_s));
  HD_TRY(ContextArray(ctx, "rth_vals_s", n_b, &vals_s));
  HD_TRY(ContextArray(ctx, "rth_orig", n_b, &orig));
  HD_TRY(ContextArray(ctx, "rth_bx", n_b, &bx));
  HD_TRY(ContextArray(ctx, "rth_by", n_b, &by));
  if (nd == 3) HD_TRY(ContextArray(ctx, "rth_bz", n_b, &bz));
  HD_TRY(ContextArray(ctx, "rth_start", n_cells, &start));
  HD_TRY(ContextArray(ctx, "rth_end", n_cells, &end));
  HD_TRY(ContextArray(ctx, "rth_reps", (uint64_t)n_b * nd, &reps));
  HD_TRY(ContextArray(ctx, "rth_nreps", 2, &n_reps));
  HD_TRY(ContextArray(ctx, "rth_qa", (uint64_t)n_a * nd, &qa));
  HD_TRY(ContextArray(ctx, "rth_dmin2", n_a, &dmin2));
  HD_TRY(FillU32(ctx, start, 0u, n_cells));
  HD_TRY(FillU32(ctx, end, 0u, n_cells));
  HD_TRY(FillU32(ctx, n_reps, 0u, 2));
  const int bb = (int)std::min<uint64_t>((n_b + 255) / 256, 8192);
  const int ab = (int)std::min<uint64_t>((n_a + 255) / 256, 8192);
  int key_bits = 1;
  while (key_bits < 32 && (1ull << key_bits) < n_cells) ++key_bits;
  if (nd == 2) {
    QuantizeBKernel<2><<<bb, 256, 0, ctx->stream>>>(g, args.b, n_b, keys, vals);
  } else {
    QuantizeBKernel<3><<<bb, 256, 0, ctx->stream>>>(g, args.b, n_b, keys, vals);
  }
  HD_HIP_TRY(hipGetLastError());
  HD_TRY(SortPairsU32(ctx, keys, keys_s, vals, vals_s, n_b, key_bits));
  if (nd == 2) {
    QuantizeFillKernel<2><<<bb, 256, 0, ctx->stream>>>(g, args.b, n_b, keys_s, vals_s, bx, by,
                                                       nullptr, orig, start, end, reps, n_reps);
    QuantizeAKernel<2><<<ab, 256, 0, ctx->stream>>>(g, args.a, n_a, qa);
  } else {
    QuantizeFillKernel<3><<<bb, 256, 0, ctx->stream>>>(g, args.b, n_b, keys_s, vals_s, bx, by,
                                                       bz, orig, start, end, reps, n_reps);
    QuantizeAKernel<3><<<ab, 256, 0, ctx->stream>>>(g, args.a, n_a, qa);
  }
  HD_HIP_TRY(hipGetLastError());
  unsigned int h_reps = 0;
  HD_HIP_TRY(hipMemcpyAsync(&h_reps, n_reps, 4, hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  StageEnd(ctx, stage_prep);

  // Nearest non-empty cell of every quantized A point (exact on integer
  // positions below 2^24).
  const int stage_nn = StageBegin(ctx, "rthdist_first_hit");
  uint32_t* nn_idx = nullptr;
  HD_TRY(ContextArray(ctx, "rth_nnidx", n_a, &nn_idx));
  HD_TRY(GridNearestNeighbors(ctx, nd, qa, n_a, reps, h_reps, dmin2, nn_idx));
  HD_TRY(FillU32(ctx, n_reps + 1, 0u, 1));
  MaxF32Kernel<<<std::min(ab, std::max(1, ctx->cu_count) * 2), 256, 0, ctx->stream>>>(
      dmin2, n_a, n_reps + 1);
  HD_HIP_TRY(hipGetLastError());
  unsigned int max_bits = 0;
  HD_HIP_TRY(hipMemcpyAsync(&max_bits, n_reps + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  float max_dmin2;
  std::memcpy(&max_dmin2, &max_bits, 4);
  // Radius schedule as the launches grow it (rt_hdist.h:185-191), up to the
  // first radius every point hits, plus the three carried launches.
  std::vector<float> radius;
  float r = r0;
  radius.push_back(r);
  while (!(max_dmin2 <= r * r)) {
    r = r + diag_q;
    radius.push_back(r);
    if (radius.size() > (1u << 24))
      return Fail(HD_EUNSUPPORTED, "RT-HDIST: more than 2^24 radius steps");
  }
  for (int k = 0; k < 3; ++k) {
    r = r + diag_q;
    radius.push_back(r);
  }
  float* d_radius = nullptr;
  HD_TRY(ContextArray(ctx, "rth_radius", radius.size(), &d_radius));
  HD_HIP_TRY(hipMemcpyAsync(d_radius, radius.data(), radius.size() * sizeof(float),
                            hipMemcpyHostToDevice, ctx->stream));
  StageEnd(ctx, stage_nn);

  const int stage_scan = StageBegin(ctx, "rthdist_scan");
  unsigned long long* key = nullptr;
  HD_TRY(ContextArray(ctx, "rth_key", kCounterWords, &key));
  HD_TRY(FillU64(ctx, key, 0ull, kCounterWords));
  const int scan_blocks = (int)std::min<uint64_t>((n_a + 3) / 4, (uint64_t)ctx->cu_count * 8);
  if (nd == 2)
    RtHdistScanKernel<2><<<scan_blocks, 256, 0, ctx->stream>>>(
        g, args.a, n_a, qa, dmin2, d_radius, (uint32_t)radius.size(), bx, by, bz, start, end,
        key, key + 1);
  else
    RtHdistScanKernel<3><<<scan_blocks, 256, 0, ctx->stream>>>(
        g, args.a, n_a, qa, dmin2, d_radius, (uint32_t)radius.size(), bx, by, bz, start, end,
        key, key + 1);
  HD_HIP_TRY(hipGetLastError());
  StageEnd(ctx, stage_scan);
  HD_TRY(FinishDirected(ctx, nd, args.a, args.b, args.n_b, key, key + 1, 1, out));
  out->variant_used = HD_VARIANT_RT_HDIST;
  return HD_OK;
}

}  // namespace hd
