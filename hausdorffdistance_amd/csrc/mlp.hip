// Learned visit order for the early-break scan: a small MLP predicts each A
// point's nearest-neighbour distance so that A can be visited in descending
// predicted order (cmax2 rises early, more inner scans break). No
// counterpart in the reference, whose selector uses gradient-boosted trees
// exported as C (src/models/tree_*.h) for parameters, not for order.
//
// Network: d_in -> 64 -> 64 -> 1, sine first layer (h1 = sin(W1 x + b1)),
// ReLU second layer, linear output; BF16 weights (an FP32 master copy is
// what Adam updates), FP32 accumulation. The sine layer is a learned
// Fourier basis (as in SIREN, Sitzmann et al., NeurIPS 2020): NN distances
// of clustered data vary on the scale of a cluster, far finer than what 64
// ReLU hyperplanes over the whole extent resolve. Input: coordinates scaled
// by the training set's MBR; target: standardised log NN distance. Training
// labels are exact NN distances of a subsample (grid query). The exact
// search still decides every value: the order changes only how soon cmax2
// rises.
//
// Kernels: forward+backward with one lane per sample (weights in LDS as
// FP32 converted from BF16), a kernel that reduces the gradient over the
// batch (one block per 64 parameters, lanes on parameters, waves on
// samples) and applies Adam, and an inference kernel that
// runs the 64x64 layer on matrix cores (BF16 operands, FP32 accumulation;
// the sine activations are rounded to BF16 as MFMA operands).
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hd_device.h"
#include "hd_gen.h"
#include "hd_grid.h"
#include "hd_internal.h"
#include "hd_primitives.h"

namespace hd {
namespace {
constexpr int kH = 64;
}  // namespace
}  // namespace hd

struct hd_mlp {
  int n_dims = 2;
  int device = 0;
  int n_params = 0;
  std::vector<float> host_params;  // FP32 master (updated after training)
  float norm_lo[3] = {0, 0, 0};
  float norm_inv[3] = {1, 1, 1};
  float y_mean = 0, y_std = 1;
  // device state
  float* master = nullptr;     // FP32 master parameters
  uint16_t* wbf = nullptr;     // BF16 working copy
  float* adam_m = nullptr;
  float* adam_v = nullptr;
};

namespace hd {

using namespace dev;

namespace {

constexpr int kBatchMax = 8192;

// Offsets of the parameter blocks.
struct Layout {
  int w1, b1, w2, b2, w3, b3, n;
};
__host__ __device__ inline Layout MakeLayout(int nd) {
  Layout l;
  l.w1 = 0;
  l.b1 = l.w1 + kH * nd;
  l.w2 = l.b1 + kH;
  l.b2 = l.w2 + kH * kH;
  l.w3 = l.b2 + kH;
  l.b3 = l.w3 + kH;
  l.n = l.b3 + 1;
  return l;
}
inline Layout HostLayout(int nd) {
  Layout l;
  l.w1 = 0;
  l.b1 = l.w1 + kH * nd;
  l.w2 = l.b1 + kH;
  l.b2 = l.w2 + kH * kH;
  l.w3 = l.b2 + kH;
  l.b3 = l.w3 + kH;
  l.n = l.b3 + 1;
  return l;
}

HD_DEV float Bf16ToF32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even (inputs are finite).
HD_DEV uint16_t F32ToBf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

HD_DEV uint32_t SampleIndex(uint32_t step, uint32_t s, uint32_t n,
                            uint32_t seed) {
  return Hash32(Hash32(step * 0x9E3779B9u ^ seed) ^ (s * 0x85EBCA6Bu)) % n;
}

// Weights from BF16 into FP32 LDS.
template <int ND>
HD_DEV void LoadWeights(const uint16_t* __restrict__ wbf, float* s_w) {
  const Layout L = MakeLayout(ND);
  for (int i = threadIdx.x; i < L.n; i += blockDim.x) s_w[i] = Bf16ToF32(wbf[i]);
  __syncthreads();
}

// First layer pre-activation of unit j.
template <int ND>
HD_DEV float Layer1Pre(const float* s_w, const float* x, int j) {
  const Layout L = MakeLayout(ND);
  float acc = s_w[L.b1 + j];
#pragma unroll
  for (int k = 0; k < ND; ++k) acc = fmaf(s_w[L.w1 + j * ND + k], x[k], acc);
  return acc;
}

// First layer into registers (fully unrolled so h1 stays in VGPRs). The
// hardware sine (|argument| stays within a few hundred radians here).
template <int ND>
HD_DEV void Layer1(const float* s_w, const float* x, float (&h1)[kH]) {
#pragma unroll
  for (int j = 0; j < kH; ++j) h1[j] = __sinf(Layer1Pre<ND>(s_w, x, j));
}

// Hidden unit j of the second layer (pre-activation).
HD_DEV float Layer2Unit(const float* s_w, int w2, int b2, int j,
                        const float (&h1)[kH]) {
  float acc = s_w[b2 + j];
  const float* row = s_w + w2 + j * kH;
#pragma unroll
  for (int k = 0; k < kH; ++k) acc = fmaf(row[k], h1[k], acc);
  return acc;
}

struct TrainArgs {
  const float* x;      // [n_samples][ND] normalised inputs
  const float* y;      // [n_samples] standardised targets
  uint32_t n_samples;
  uint32_t batch;
  uint32_t seed;
  const uint32_t* step_ptr;  // device step counter (advanced by Adam)
  int identity;        // 1: sample s is row s (gradient check)
  const uint16_t* wbf;
  float* bx;           // [batch][ND] inputs of the batch
  float* h1;           // [batch][64]
  float* h2;
  float* g1;
  float* g2;
  float* go;           // [batch]
  float* loss;         // accumulated sum of squared errors for this step
};

template <int ND>
__global__ __launch_bounds__(256) void MlpForwardBackwardKernel(TrainArgs a) {
  __shared__ float s_w[kH * ND + kH + kH * kH + kH + kH + 1];
  LoadWeights<ND>(a.wbf, s_w);
  const Layout L = MakeLayout(ND);
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  float err2 = 0.f;
  if (s < a.batch) {
    const uint32_t step = *a.step_ptr;
    const uint32_t i = a.identity ? s : SampleIndex(step, s, a.n_samples, a.seed);
    float x[ND];
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      x[k] = a.x[(uint64_t)i * ND + k];
      a.bx[(uint64_t)s * ND + k] = x[k];
    }
    float h1[kH];
    Layer1<ND>(s_w, x, h1);
    float* h2s = a.h2 + (uint64_t)s * kH;
    float out = s_w[L.b3];
    for (int j = 0; j < kH; ++j) {
      const float h2 = fmaxf(Layer2Unit(s_w, L.w2, L.b2, j, h1), 0.f);
      h2s[j] = h2;
      out = fmaf(s_w[L.w3 + j], h2, out);
    }
    const float err = out - a.y[i];
    err2 = err * err;
    const float gout = 2.f * err / (float)a.batch;
    a.go[s] = gout;
    float g1[kH];
#pragma unroll
    for (int k = 0; k < kH; ++k) g1[k] = 0.f;
    for (int j = 0; j < kH; ++j) {
      const float g2 = h2s[j] > 0.f ? gout * s_w[L.w3 + j] : 0.f;
      a.g2[(uint64_t)s * kH + j] = g2;
      const float* row = s_w + L.w2 + j * kH;
#pragma unroll
      for (int k = 0; k < kH; ++k) g1[k] = fmaf(g2, row[k], g1[k]);
    }
#pragma unroll
    for (int k = 0; k < kH; ++k) {
      a.h1[(uint64_t)s * kH + k] = h1[k];
      a.g1[(uint64_t)s * kH + k] = g1[k] * __cosf(Layer1Pre<ND>(s_w, x, k));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) err2 += __shfl_xor(err2, off, 64);
  if (LaneId() == 0 && err2 != 0.f) atomicAdd(a.loss, err2);
}

struct AdamArgs {
  const float* bx;
  const float* h1;
  const float* h2;
  const float* g1;
  const float* g2;
  const float* go;
  uint32_t batch;
  float* master;
  uint16_t* wbf;
  float* m;
  float* v;
  float* grad_out;     // optional: raw gradient (gradient check)
  int apply;           // 0: gradient only
  float lr;
  uint32_t* step_ptr;  // advanced by thread 0 after the update
};

// Adam update of parameter p with gradient g (step t = *step_ptr + 1).
HD_DEV void AdamUpdate(const AdamArgs& a, int p, float g) {
  if (a.grad_out) a.grad_out[p] = g;
  if (!a.apply) return;
  const uint32_t t = *a.step_ptr + 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float m = b1 * a.m[p] + (1.f - b1) * g;
  const float v = b2 * a.v[p] + (1.f - b2) * g * g;
  a.m[p] = m;
  a.v[p] = v;
  const float mh = m / (1.f - powf(b1, (float)t));
  const float vh = v / (1.f - powf(b2, (float)t));
  const float w = a.master[p] - a.lr * mh / (sqrtf(vh) + eps);
  a.master[p] = w;
  a.wbf[p] = F32ToBf16(w);
}

constexpr int kGradWaves = 16;  // waves per block of the gradient reduction

// Gradients summed over the batch, then Adam. Each block owns 64
// parameters that share one per-sample factor (lane l: parameter l):
//   blocks 0..ND-1:   W1[l][k] = sum_s g1[s][l] x[s][k]   (k = block)
//   block  ND:        b1[l]    = sum_s g1[s][l]
//   blocks ND+1..+64: W2[j][l] = sum_s g2[s][j] h1[s][l]  (j = block-ND-1),
//                     and b2[j] = sum_s g2[s][j]
//   block  ND+65:     w3[l]    = sum_s go[s] h2[s][l],  and b3 = sum_s go[s]
// The waves of a block split the batch (wave w: samples w, w + 16, ...);
// lanes read 64 consecutive floats of a sample row, the shared factor is one
// broadcast load. Partial sums meet in LDS in a fixed order (deterministic).
template <int ND>
__global__ __launch_bounds__(64 * kGradWaves) void MlpGradAdamKernel(AdamArgs a) {
  __shared__ float red[kGradWaves][kH + 1];
  const Layout L = MakeLayout(ND);
  const int g = blockIdx.x;
  const int lane = (int)(threadIdx.x & 63u);
  const int wave = (int)(threadIdx.x >> 6);
  const uint32_t B = a.batch;
  float acc = 0.f, accb = 0.f;  // accb: the group's bias (one value)
  if (g <= ND) {
    const int k = g;
    for (uint32_t s = wave; s < B; s += kGradWaves) {
      const float f = k < ND ? a.bx[(uint64_t)s * ND + k] : 1.f;
      acc = fmaf(a.g1[(uint64_t)s * kH + lane], f, acc);
    }
  } else if (g <= ND + kH) {
    const int j = g - ND - 1;
    for (uint32_t s = wave; s < B; s += kGradWaves) {
      const float f = a.g2[(uint64_t)s * kH + j];
      acc = fmaf(f, a.h1[(uint64_t)s * kH + lane], acc);
      accb += f;
    }
  } else {
    for (uint32_t s = wave; s < B; s += kGradWaves) {
      const float f = a.go[s];
      acc = fmaf(f, a.h2[(uint64_t)s * kH + lane], acc);
      accb += f;
    }
  }
  red[wave][lane] = acc;
  if (lane == 0) red[wave][kH] = accb;
  __syncthreads();
  if (wave != 0) return;
  float sum = 0.f, sumb = 0.f;
#pragma unroll
  for (int w = 0; w < kGradWaves; ++w) {
    sum += red[w][lane];
    sumb += red[w][kH];
  }
  if (g < ND) {
    AdamUpdate(a, L.w1 + lane * ND + g, sum);
  } else if (g == ND) {
    AdamUpdate(a, L.b1 + lane, sum);
  } else if (g <= ND + kH) {
    const int j = g - ND - 1;
    AdamUpdate(a, L.w2 + j * kH + lane, sum);
    if (lane == 0) AdamUpdate(a, L.b2 + j, sumb);
  } else {
    AdamUpdate(a, L.w3 + lane, sum);
    if (lane == 0) AdamUpdate(a, L.b3, sumb);
  }
}

__global__ void StepAdvanceKernel(uint32_t* step_ptr) { *step_ptr += 1u; }

__global__ void RoundToBf16Kernel(const float* __restrict__ in,
                                  uint16_t* __restrict__ out, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    out[i] = F32ToBf16(in[i]);
}

// Training inputs: subsample rows, normalised coordinates and log labels.
template <int ND>
__global__ void MlpPrepKernel(const float* __restrict__ a,
                              const uint32_t* __restrict__ sub, uint32_t n,
                              const float* __restrict__ nn_d2, float lo0,
                              float lo1, float lo2, float inv0, float inv1,
                              float inv2, float* __restrict__ x,
                              float* __restrict__ y) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t src = sub[i];
    const float lo[3] = {lo0, lo1, lo2};
    const float inv[3] = {inv0, inv1, inv2};
#pragma unroll
    for (int k = 0; k < ND; ++k)
      x[(uint64_t)i * ND + k] = (a[(uint64_t)src * ND + k] - lo[k]) * inv[k];
    const float d2 = fmaxf(nn_d2[i], 1e-30f);
    y[i] = 0.5f * logf(d2);
  }
}

typedef short MlpShortx4 __attribute__((ext_vector_type(4)));
typedef float MlpFloatx4 __attribute__((ext_vector_type(4)));

// BF16 bits of two floats (round to nearest even), low half first.
HD_DEV uint32_t PackBf16(float lo, float hi) {
  return (uint32_t)F32ToBf16(lo) | ((uint32_t)F32ToBf16(hi) << 16);
}

// What inference writes: predictions, or 16-bit sort keys (ascending key =
// descending prediction, 2^-11 steps over [-16, 16)) with point indices.
struct PredictOut {
  float* pred;
  uint16_t* key;
  uint32_t* idx;
};

// Inference on matrix cores, 16 points per wave step. Lane l works on point
// l % 16 of the step and hidden units 16 kb + 4 (l / 16) + r (kb, r in
// 0..3): it evaluates those 16 sines, rounds them to BF16 and holds them as
// the B operand of v_mfma_f32_16x16x16_bf16 (B[k][j] at lane j + 16 (k / 4),
// element k % 4). The A operand is a 16x16 block of W2 (A[i][k] at lane
// i + 16 (k / 4)); the accumulator starts at b2 and ends as
// H2[16 ib + 4 (l / 16) + r][point l % 16]. The output layer is a dot
// product over each lane's 16 units (FP32) and a sum over the four lane
// groups. All weights stay in registers for the wave's lifetime.
template <int ND, bool kKeys>
__global__ __launch_bounds__(256) void MlpPredictKernel(
    const uint16_t* __restrict__ wbf, const float* __restrict__ pts,
    uint64_t n, float lo0, float lo1, float lo2, float inv0, float inv1,
    float inv2, PredictOut o) {
  const Layout L = MakeLayout(ND);
  const uint32_t lane = threadIdx.x & 63u, col = lane & 15u, grp = lane >> 4;
  const float lo[3] = {lo0, lo1, lo2};
  const float inv[3] = {inv0, inv1, inv2};
  // v_sin takes revolutions: the first layer is scaled by 1 / (2 pi).
  const float kRev = 0.15915494309189535f;
  float w1[16][ND], b1[16], b2[16], w3[16];
  MlpShortx4 w2[4][4];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int u = 16 * (q >> 2) + 4 * (int)grp + (q & 3);
    b1[q] = Bf16ToF32(wbf[L.b1 + u]) * kRev;
#pragma unroll
    for (int d = 0; d < ND; ++d)
      w1[q][d] = Bf16ToF32(wbf[L.w1 + u * ND + d]) * kRev;
    b2[q] = Bf16ToF32(wbf[L.b2 + u]);
    w3[q] = Bf16ToF32(wbf[L.w3 + u]);
  }
#pragma unroll
  for (int ib = 0; ib < 4; ++ib)
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        w2[ib][kb][r] = (short)wbf[L.w2 + (16 * ib + (int)col) * kH + 16 * kb +
                                   4 * (int)grp + r];
  const float b3 = Bf16ToF32(wbf[L.b3]);

  const uint64_t n_tiles = (n + 15) / 16;
  const uint64_t n_waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  for (uint64_t tile = (uint64_t)blockIdx.x * (blockDim.x >> 6) +
                       (threadIdx.x >> 6);
       tile < n_tiles; tile += n_waves) {
    const uint64_t p = tile * 16 + col;
    const uint64_t pc = p < n ? p : n - 1;
    float x[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) x[d] = (pts[pc * ND + d] - lo[d]) * inv[d];
    MlpShortx4 h1[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      float h[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float z = b1[4 * kb + r];
#pragma unroll
        for (int d = 0; d < ND; ++d) z = fmaf(w1[4 * kb + r][d], x[d], z);
        h[r] = __builtin_amdgcn_sinf(z);
      }
      const uint32_t p01 = PackBf16(h[0], h[1]), p23 = PackBf16(h[2], h[3]);
      h1[kb][0] = (short)(p01 & 0xFFFFu);
      h1[kb][1] = (short)(p01 >> 16);
      h1[kb][2] = (short)(p23 & 0xFFFFu);
      h1[kb][3] = (short)(p23 >> 16);
    }
    float s = 0.f;
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      MlpFloatx4 acc = {b2[4 * ib], b2[4 * ib + 1], b2[4 * ib + 2],
                        b2[4 * ib + 3]};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w2[ib][kb], h1[kb],
                                                         acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; ++r) s = fmaf(w3[4 * ib + r], fmaxf(acc[r], 0.f), s);
    }
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    s += b3;
    if (grp == 0 && p < n) {
      if (kKeys) {
        const float q = fminf(fmaxf((s + 16.f) * 2048.f, 0.f), 65535.f);
        o.key[p] = (uint16_t)(65535u - (uint32_t)q);
        o.idx[p] = (uint32_t)p;
      } else {
        o.pred[p] = s;
      }
    }
  }
}

// Subsample points with known indices into a contiguous AoS array.
template <int ND>
__global__ void GatherRowsKernel(const float* __restrict__ a,
                                 const uint32_t* __restrict__ idx, uint32_t n,
                                 float* __restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
#pragma unroll
    for (int k = 0; k < ND; ++k)
      out[(uint64_t)i * ND + k] = a[(uint64_t)idx[i] * ND + k];
}

int AllocModel(hd_mlp* m) {
  const int n = m->n_params;
  HD_HIP_TRY(hipMalloc(&m->master, sizeof(float) * n));
  HD_HIP_TRY(hipMalloc(&m->wbf, sizeof(uint16_t) * n));
  HD_HIP_TRY(hipMalloc(&m->adam_m, sizeof(float) * n));
  HD_HIP_TRY(hipMalloc(&m->adam_v, sizeof(float) * n));
  return HD_OK;
}

void FreeModel(hd_mlp* m) {
  if (!m) return;
  if (m->master) (void)hipFree(m->master);
  if (m->wbf) (void)hipFree(m->wbf);
  if (m->adam_m) (void)hipFree(m->adam_m);
  if (m->adam_v) (void)hipFree(m->adam_v);
  m->master = nullptr;
  m->wbf = nullptr;
  m->adam_m = m->adam_v = nullptr;
}

// Initialisation from SplitMix64 (+ Box-Muller). First layer: random
// hyperplanes through the unit box of the normalised inputs, with slopes of
// kOmega (first-layer weights uniform in [-kOmega/2, kOmega/2], biases in
// [-kOmega, kOmega]): sine features of up to ~kOmega/(2 pi) periods over the
// unit square, fine enough for cluster-scale structure (targets vary over
// ~1/50 of the extent); deeper layers He-normal; biases of layers 2 and 3
// zero.
constexpr double kOmega = 90.0;
void InitParams(int nd, uint32_t seed, std::vector<float>* p) {
  const Layout L = HostLayout(nd);
  p->assign(L.n, 0.f);
  uint64_t k = 0;
  auto uniform = [&]() {
    return (double)(hd_splitmix64_at(seed, k++) >> 11) * 0x1.0p-53;
  };
  auto normal = [&]() {
    const double u1 = uniform() + 0x1.0p-53;
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  };
  for (int j = 0; j < kH; ++j) {
    for (int i = 0; i < nd; ++i)
      (*p)[L.w1 + j * nd + i] = (float)((uniform() - 0.5) * kOmega);
    (*p)[L.b1 + j] = (float)((2.0 * uniform() - 1.0) * kOmega);
  }
  for (int i = 0; i < kH * kH; ++i) (*p)[L.w2 + i] = (float)(normal() * std::sqrt(2.0 / kH));
  for (int i = 0; i < kH; ++i) (*p)[L.w3 + i] = (float)(normal() * std::sqrt(1.0 / kH));
}

struct TrainBuffers {
  float *bx, *h1, *h2, *g1, *g2, *go, *loss;
  uint32_t* step;
};

int GetTrainBuffers(hd_context* ctx, int nd, uint32_t batch, uint32_t steps,
                    TrainBuffers* t) {
  HD_TRY(ContextArray(ctx, "mlp_bx", (size_t)batch * nd, &t->bx));
  HD_TRY(ContextArray(ctx, "mlp_h1", (size_t)batch * kH, &t->h1));
  HD_TRY(ContextArray(ctx, "mlp_h2", (size_t)batch * kH, &t->h2));
  HD_TRY(ContextArray(ctx, "mlp_g1", (size_t)batch * kH, &t->g1));
  HD_TRY(ContextArray(ctx, "mlp_g2", (size_t)batch * kH, &t->g2));
  HD_TRY(ContextArray(ctx, "mlp_go", batch, &t->go));
  HD_TRY(ContextArray(ctx, "mlp_loss", std::max<uint32_t>(steps, 1), &t->loss));
  HD_TRY(ContextArray(ctx, "mlp_step", 1, &t->step));
  return HD_OK;
}

template <int ND>
int LaunchStep(hd_context* ctx, hd_mlp* m, const TrainBuffers& tb,
               const float* x, const float* y, uint32_t n_samples,
               uint32_t batch, uint32_t seed, int identity, float* loss_slot,
               float lr, int apply, float* grad_out) {
  TrainArgs ta;
  ta.x = x;
  ta.y = y;
  ta.n_samples = n_samples;
  ta.batch = batch;
  ta.seed = seed;
  ta.step_ptr = tb.step;
  ta.identity = identity;
  ta.wbf = m->wbf;
  ta.bx = tb.bx;
  ta.h1 = tb.h1;
  ta.h2 = tb.h2;
  ta.g1 = tb.g1;
  ta.g2 = tb.g2;
  ta.go = tb.go;
  ta.loss = loss_slot;
  MlpForwardBackwardKernel<ND><<<(batch + 255) / 256, 256, 0, ctx->stream>>>(ta);
  HD_HIP_TRY(hipGetLastError());
  AdamArgs aa;
  aa.bx = tb.bx;
  aa.h1 = tb.h1;
  aa.h2 = tb.h2;
  aa.g1 = tb.g1;
  aa.g2 = tb.g2;
  aa.go = tb.go;
  aa.batch = batch;
  aa.master = m->master;
  aa.wbf = m->wbf;
  aa.m = m->adam_m;
  aa.v = m->adam_v;
  aa.grad_out = grad_out;
  aa.apply = apply;
  aa.lr = lr;
  aa.step_ptr = tb.step;
  // Blocks: ND columns of W1, b1, 64 rows of W2 (+ b2), w3 (+ b3).
  MlpGradAdamKernel<ND><<<ND + kH + 2, 64 * kGradWaves, 0, ctx->stream>>>(aa);
  HD_HIP_TRY(hipGetLastError());
  if (apply) {
    StepAdvanceKernel<<<1, 1, 0, ctx->stream>>>(tb.step);
    HD_HIP_TRY(hipGetLastError());
  }
  return HD_OK;
}

template <int ND>
int TrainImpl(hd_context* ctx, const float* d_a, uint64_t n_a,
              const float* d_b, uint64_t n_b, const hd_mlp_train_config& cfg,
              hd_mlp* m, float* final_loss) {
  const uint32_t n_s = (uint32_t)std::min<uint64_t>(cfg.n_samples, n_a);
  const uint32_t batch = std::min<uint32_t>(std::max(cfg.batch, 1u), kBatchMax);
  // Subsample of A (partial Fisher-Yates, as the workload generators).
  std::vector<uint32_t> sub(n_s);
  if (hd_gen_subsample(cfg.seed, n_a, n_s, sub.data()) != HD_OK)
    return Fail(HD_EINVAL, "subsample failed");
  uint32_t* d_sub = nullptr;
  float *d_pts = nullptr, *d_nn = nullptr, *d_x = nullptr, *d_y = nullptr;
  uint32_t* d_nn_idx = nullptr;
  HD_TRY(ContextArray(ctx, "mlp_sub", n_s, &d_sub));
  HD_TRY(ContextArray(ctx, "mlp_pts", (size_t)n_s * ND, &d_pts));
  HD_TRY(ContextArray(ctx, "mlp_nn", n_s, &d_nn));
  HD_TRY(ContextArray(ctx, "mlp_nn_idx", n_s, &d_nn_idx));
  HD_TRY(ContextArray(ctx, "mlp_x", (size_t)n_s * ND, &d_x));
  HD_TRY(ContextArray(ctx, "mlp_y", n_s, &d_y));
  HD_HIP_TRY(hipMemcpyAsync(d_sub, sub.data(), sizeof(uint32_t) * n_s,
                            hipMemcpyHostToDevice, ctx->stream));
  const int gb = (int)std::min<uint32_t>((n_s + 255) / 256, 4096);
  GatherRowsKernel<ND><<<gb, 256, 0, ctx->stream>>>(d_a, d_sub, n_s, d_pts);
  HD_HIP_TRY(hipGetLastError());
  // Exact NN distances of the subsample (labels).
  HD_TRY(NearestNeighbors(ctx, cfg.label_index, ND, d_pts, n_s, d_b, n_b,
                          d_nn, d_nn_idx));
  // Normalisation frame: MBR of A.
  float mbr[6];
  HD_TRY(HostMbr(ctx, ND, d_a, n_a, mbr));
  for (int k = 0; k < 3; ++k) {
    m->norm_lo[k] = k < ND ? mbr[k] : 0.f;
    m->norm_inv[k] = k < ND ? 1.0f / (mbr[ND + k] - mbr[k]) : 1.f;
  }
  MlpPrepKernel<ND><<<gb, 256, 0, ctx->stream>>>(
      d_a, d_sub, n_s, d_nn, m->norm_lo[0], m->norm_lo[1], m->norm_lo[2],
      m->norm_inv[0], m->norm_inv[1], m->norm_inv[2], d_x, d_y);
  HD_HIP_TRY(hipGetLastError());
  // Standardise targets on the host (n_s floats).
  std::vector<float> hy(n_s);
  HD_HIP_TRY(hipMemcpyAsync(hy.data(), d_y, sizeof(float) * n_s,
                            hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  double mean = 0, var = 0;
  for (float v : hy) mean += v;
  mean /= std::max<uint32_t>(n_s, 1);
  for (float v : hy) var += (v - mean) * (v - mean);
  var /= std::max<uint32_t>(n_s, 1);
  m->y_mean = (float)mean;
  m->y_std = (float)std::max(std::sqrt(var), 1e-6);
  for (float& v : hy) v = (float)((v - mean) / m->y_std);
  HD_HIP_TRY(hipMemcpyAsync(d_y, hy.data(), sizeof(float) * n_s,
                            hipMemcpyHostToDevice, ctx->stream));

  InitParams(ND, cfg.seed, &m->host_params);
  HD_HIP_TRY(hipMemcpyAsync(m->master, m->host_params.data(),
                            sizeof(float) * m->n_params,
                            hipMemcpyHostToDevice, ctx->stream));
  HD_TRY(FillU32(ctx, reinterpret_cast<uint32_t*>(m->adam_m), 0u, m->n_params));
  HD_TRY(FillU32(ctx, reinterpret_cast<uint32_t*>(m->adam_v), 0u, m->n_params));
  RoundToBf16Kernel<<<(m->n_params + 255) / 256, 256, 0, ctx->stream>>>(
      m->master, m->wbf, m->n_params);
  HD_HIP_TRY(hipGetLastError());
  TrainBuffers tb;
  HD_TRY(GetTrainBuffers(ctx, ND, batch, cfg.steps, &tb));
  HD_TRY(FillU32(ctx, tb.step, 0u, 1));
  HD_TRY(FillU32(ctx, reinterpret_cast<uint32_t*>(tb.loss), 0u,
                 std::max<uint32_t>(cfg.steps, 1)));
  for (uint32_t s = 0; s < cfg.steps; ++s)
    HD_TRY(LaunchStep<ND>(ctx, m, tb, d_x, d_y, n_s, batch, cfg.seed, 0,
                          tb.loss + s, cfg.lr, 1, nullptr));
  // Final loss: mean over the last 100 steps.
  const uint32_t tail = std::min<uint32_t>(cfg.steps, 100);
  std::vector<float> hl(tail);
  if (tail > 0)
    HD_HIP_TRY(hipMemcpyAsync(hl.data(), tb.loss + (cfg.steps - tail),
                              sizeof(float) * tail, hipMemcpyDeviceToHost,
                              ctx->stream));
  HD_HIP_TRY(hipMemcpyAsync(m->host_params.data(), m->master,
                            sizeof(float) * m->n_params,
                            hipMemcpyDeviceToHost, ctx->stream));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (final_loss) {
    double acc = 0;
    for (float v : hl) acc += v / batch;
    *final_loss = tail ? (float)(acc / tail) : 0.f;
  }
  return HD_OK;
}

}  // namespace

// Enough waves to fill the chip; each wave strides over 16-point steps.
template <int ND, bool kKeys>
int LaunchPredict(hd_context* ctx, const hd_mlp* m, const float* pts,
                  uint64_t n, PredictOut o) {
  if (n == 0) return HD_OK;
  const uint64_t n_tiles = (n + 15) / 16;
  const uint64_t blocks = std::min<uint64_t>((n_tiles + 3) / 4, 256u * 32u);
  MlpPredictKernel<ND, kKeys><<<(unsigned)blocks, 256, 0, ctx->stream>>>(
      m->wbf, pts, n, m->norm_lo[0], m->norm_lo[1], m->norm_lo[2],
      m->norm_inv[0], m->norm_inv[1], m->norm_inv[2], o);
  HD_HIP_TRY(hipGetLastError());
  return HD_OK;
}

int PredictInto(hd_context* ctx, const hd_mlp* m, const float* pts, uint64_t n,
                float* out) {
  const PredictOut o = {out, nullptr, nullptr};
  return m->n_dims == 2 ? LaunchPredict<2, false>(ctx, m, pts, n, o)
                        : LaunchPredict<3, false>(ctx, m, pts, n, o);
}

int TrainModel(hd_context* ctx, int n_dims, const float* d_a, uint64_t n_a,
               const float* d_b, uint64_t n_b, const hd_mlp_train_config& cfg,
               hd_mlp** out, float* final_loss) {
  hd_mlp* m = new hd_mlp();
  m->n_dims = n_dims;
  m->device = ctx->device;
  m->n_params = HostLayout(n_dims).n;
  int rc = AllocModel(m);
  if (rc == HD_OK)
    rc = n_dims == 2 ? TrainImpl<2>(ctx, d_a, n_a, d_b, n_b, cfg, m, final_loss)
                     : TrainImpl<3>(ctx, d_a, n_a, d_b, n_b, cfg, m, final_loss);
  if (rc != HD_OK) {
    FreeModel(m);
    delete m;
    return rc;
  }
  *out = m;
  return HD_OK;
}

// Visit order: A sorted by predicted NN distance, largest first.
int LearnedOrder(hd_context* ctx, const DirectedArgs& args,
                 const uint32_t** order_out) {
  const hd_mlp* model = args.model;
  if (model == nullptr) {
    // Train (once per context and input) with the default configuration.
    if (ctx->cached_model && ctx->cached_a == args.a &&
        ctx->cached_n_a == args.n_a && ctx->cached_b == args.b &&
        ctx->cached_n_b == args.n_b) {
      model = ctx->cached_model;
    } else {
      if (ctx->cached_model) {
        hd_mlp_destroy(ctx->cached_model);
        ctx->cached_model = nullptr;
      }
      hd_mlp_train_config cfg;
      hd_mlp_train_config_init(&cfg);
      hd_mlp* m = nullptr;
      HD_TRY(TrainModel(ctx, args.n_dims, args.a, args.n_a, args.b, args.n_b,
                        cfg, &m, nullptr));
      ctx->cached_model = m;
      ctx->cached_a = args.a;
      ctx->cached_n_a = args.n_a;
      ctx->cached_b = args.b;
      ctx->cached_n_b = args.n_b;
      model = m;
    }
  }
  if (model->n_dims != args.n_dims)
    return Fail(HD_EINVAL, "model has %d dims, input %d", model->n_dims,
                args.n_dims);
  const uint32_t n = (uint32_t)args.n_a;
  uint16_t *key = nullptr, *key_s = nullptr;
  uint32_t *iota = nullptr, *order = nullptr;
  HD_TRY(ContextArray(ctx, "lo_key", n, &key));
  HD_TRY(ContextArray(ctx, "lo_key_s", n, &key_s));
  HD_TRY(ContextArray(ctx, "lo_iota", n, &iota));
  HD_TRY(ContextArray(ctx, "lo_order", n, &order));
  const int st = StageBegin(ctx, "learned_order");
  // Predictions quantised to 16-bit keys: two radix passes instead of four,
  // and ties (within 2^-11 of a standard deviation) keep index order.
  const PredictOut o = {nullptr, key, iota};
  if (model->n_dims == 2)
    HD_TRY((LaunchPredict<2, true>(ctx, model, args.a, n, o)));
  else
    HD_TRY((LaunchPredict<3, true>(ctx, model, args.a, n, o)));
  HD_TRY(SortPairsU16(ctx, key, key_s, iota, order, n));
  StageEnd(ctx, st);
  *order_out = order;
  return HD_OK;
}

}  // namespace hd

using namespace hd;

extern "C" {

void hd_mlp_train_config_init(hd_mlp_train_config* c) {
  if (!c) return;
  c->steps = 2000;
  c->batch = 4096;
  c->n_samples = 1000000;
  c->seed = 14;
  c->lr = 3e-3f;
  c->label_index = HD_VARIANT_GRID;
}

int hd_mlp_train_device(hd_context* ctx, int n_dims, const float* d_a,
                        uint64_t n_a, const float* d_b, uint64_t n_b,
                        const hd_mlp_train_config* cfg, hd_mlp** out,
                        float* final_loss) {
  ClearError();
  if (!ctx || !out || !cfg) return Fail(HD_EINVAL, "null argument");
  *out = nullptr;
  if (n_dims != 2 && n_dims != 3) return Fail(HD_EINVAL, "n_dims must be 2 or 3");
  if (n_a == 0 || n_b == 0 || !d_a || !d_b)
    return Fail(HD_EINVAL, "empty point set");
  if (cfg->n_samples == 0 || cfg->batch == 0)
    return Fail(HD_EINVAL, "n_samples and batch must be positive");
  if (cfg->label_index != HD_VARIANT_EB && cfg->label_index != HD_VARIANT_GRID &&
      cfg->label_index != HD_VARIANT_LBVH)
    return Fail(HD_EINVAL, "label_index must be EB, GRID or LBVH");
  HD_TRY(ContextActivate(ctx));
  return TrainModel(ctx, n_dims, d_a, n_a, d_b, n_b, *cfg, out, final_loss);
}

int hd_mlp_predict_device(hd_context* ctx, const hd_mlp* model,
                          const float* d_pts, uint64_t n, float* d_out) {
  ClearError();
  if (!ctx || !model || !d_pts || !d_out) return Fail(HD_EINVAL, "null argument");
  HD_TRY(ContextActivate(ctx));
  HD_TRY(PredictInto(ctx, model, d_pts, n, d_out));
  HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return HD_OK;
}

void hd_mlp_destroy(hd_mlp* model) {
  if (!model) return;
  (void)hipSetDevice(model->device);
  FreeModel(model);
  delete model;
}

int hd_mlp_n_params(int n_dims) {
  if (n_dims != 2 && n_dims != 3) return -1;
  return HostLayout(n_dims).n;
}

int hd_mlp_get_params(const hd_mlp* model, float* params, float* norm_lower,
                      float* norm_inv_extent) {
  ClearError();
  if (!model) return Fail(HD_EINVAL, "null model");
  if (params)
    std::memcpy(params, model->host_params.data(),
                sizeof(float) * model->host_params.size());
  for (int k = 0; k < model->n_dims; ++k) {
    if (norm_lower) norm_lower[k] = model->norm_lo[k];
    if (norm_inv_extent) norm_inv_extent[k] = model->norm_inv[k];
  }
  return HD_OK;
}

int hd_mlp_loss_grad_device(hd_context* ctx, int n_dims, const float* params,
                            const float* x, const float* y, uint32_t batch,
                            float* loss, float* grad) {
  ClearError();
  if (!ctx || !params || !x || !y || !loss || !grad)
    return Fail(HD_EINVAL, "null argument");
  if (n_dims != 2 && n_dims != 3) return Fail(HD_EINVAL, "n_dims must be 2 or 3");
  if (batch == 0 || batch > kBatchMax)
    return Fail(HD_EINVAL, "batch must be in [1, %d]", kBatchMax);
  HD_TRY(ContextActivate(ctx));
  hd_mlp m;
  m.n_dims = n_dims;
  m.n_params = HostLayout(n_dims).n;
  int rc = AllocModel(&m);
  float *d_x = nullptr, *d_y = nullptr, *d_grad = nullptr;
  if (rc == HD_OK) rc = ContextArray(ctx, "mlpgc_x", (size_t)batch * n_dims, &d_x);
  if (rc == HD_OK) rc = ContextArray(ctx, "mlpgc_y", batch, &d_y);
  if (rc == HD_OK) rc = ContextArray(ctx, "mlpgc_grad", m.n_params, &d_grad);
  TrainBuffers tb;
  if (rc == HD_OK) rc = GetTrainBuffers(ctx, n_dims, batch, 1, &tb);
  auto body = [&]() -> int {
    HD_HIP_TRY(hipMemcpyAsync(m.master, params, sizeof(float) * m.n_params,
                              hipMemcpyHostToDevice, ctx->stream));
    HD_HIP_TRY(hipMemcpyAsync(d_x, x, sizeof(float) * batch * n_dims,
                              hipMemcpyHostToDevice, ctx->stream));
    HD_HIP_TRY(hipMemcpyAsync(d_y, y, sizeof(float) * batch,
                              hipMemcpyHostToDevice, ctx->stream));
    RoundToBf16Kernel<<<(m.n_params + 255) / 256, 256, 0, ctx->stream>>>(
        m.master, m.wbf, m.n_params);
    HD_HIP_TRY(hipGetLastError());
    HD_TRY(FillU32(ctx, tb.step, 0u, 1));
    HD_TRY(FillU32(ctx, reinterpret_cast<uint32_t*>(tb.loss), 0u, 1));
    if (n_dims == 2)
      HD_TRY(LaunchStep<2>(ctx, &m, tb, d_x, d_y, batch, batch, 0, 1, tb.loss,
                           0.f, 0, d_grad));
    else
      HD_TRY(LaunchStep<3>(ctx, &m, tb, d_x, d_y, batch, batch, 0, 1, tb.loss,
                           0.f, 0, d_grad));
    float h_loss = 0.f;
    HD_HIP_TRY(hipMemcpyAsync(&h_loss, tb.loss, sizeof(float),
                              hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP_TRY(hipMemcpyAsync(grad, d_grad, sizeof(float) * m.n_params,
                              hipMemcpyDeviceToHost, ctx->stream));
    HD_HIP_TRY(hipStreamSynchronize(ctx->stream));
    *loss = h_loss / batch;
    return HD_OK;
  };
  if (rc == HD_OK) rc = body();
  (void)hipStreamSynchronize(ctx->stream);
  FreeModel(&m);
  return rc;
}

}  // extern "C"
