/*
 * hd_oracle.c — CPU restatement of the reference's Hausdorff-distance path.
 * TEST INFRASTRUCTURE ONLY (see hd_oracle.h): the product never links this.
 *
 * Every function names the reference lines it follows. Build:
 *   gcc -O3 -ffp-contract=off -fPIC -shared -pthread (oracle/Makefile).
 */
#define _GNU_SOURCE
#include "hd_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

static float u2f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static int resolve_threads(int n_threads) {
  if (n_threads > 0) return n_threads;
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}

/* reference src/geoms/distance.h:17-24 (float2 / float3 overloads). */
float orc_d2(const float* p, const float* q, int n_dims) {
  if (n_dims == 2) {
    return (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]);
  }
  return (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
         (p[2] - q[2]) * (p[2] - q[2]);
}

/* SplitMix64 (Steele, Lea, Flood 2014). */
static uint64_t splitmix_next(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Fisher-Yates permutation of 0..n-1 (stands in for std::shuffle). */
static void shuffle_ids(uint32_t* ids, uint64_t n, uint64_t* state) {
  for (uint64_t i = n; i > 1; --i) {
    uint64_t j = splitmix_next(state) % i;
    uint32_t t = ids[i - 1];
    ids[i - 1] = ids[j];
    ids[j] = t;
  }
}

static uint64_t pack_key(float d2, uint32_t idx) {
  return ((uint64_t)f2u(d2) << 32) | (uint64_t)(0xFFFFFFFFu - idx);
}

/* reference early_break.h:229-234, on the 64-bit key. */
static void update_maximum(_Atomic uint64_t* key, uint64_t value) {
  uint64_t prev = atomic_load(key);
  while (prev < value && !atomic_compare_exchange_weak(key, &prev, value)) {
  }
}

/* Copies points into visit order so the hot loop reads contiguously, as the
 * reference's in-place shuffle does. */
static float* gather(const float* p, const uint32_t* ids, uint64_t n,
                     int n_dims) {
  float* out = (float*)malloc(sizeof(float) * (n ? n : 1) * n_dims);
  if (!out) return NULL;
  for (uint64_t i = 0; i < n; ++i)
    memcpy(out + i * n_dims, p + (uint64_t)ids[i] * n_dims,
           sizeof(float) * n_dims);
  return out;
}

typedef struct {
  const float* a; /* visit-ordered A */
  const uint32_t* a_ids;
  uint64_t n_a;
  const float* b; /* visit-ordered B */
  uint64_t n_b;
  int n_dims;
  uint64_t begin, end;
  _Atomic uint64_t* key;
  _Atomic uint64_t* compared;
} eb_task;

/* reference early_break.h:63-85 (the per-thread lambda). */
static void* eb_worker(void* arg) {
  eb_task* t = (eb_task*)arg;
  const int nd = t->n_dims;
  uint64_t local_compared = 0;
  for (uint64_t i = t->begin; i < t->end; ++i) {
    const float* pa = t->a + i * nd;
    float cmin2 = FLT_MAX;
    for (uint64_t j = 0; j < t->n_b; ++j) {
      float d = orc_d2(pa, t->b + j * nd, nd);
      local_compared++;
      if (d < cmin2) cmin2 = d;
      float cmax2 = u2f((uint32_t)(atomic_load_explicit(
                            t->key, memory_order_relaxed) >> 32));
      if (cmin2 < cmax2) break;
    }
    if (cmin2 != FLT_MAX) update_maximum(t->key, pack_key(cmin2, t->a_ids[i]));
  }
  atomic_fetch_add(t->compared, local_compared);
  return NULL;
}

uint32_t orc_first_at_d2(const float* p, const float* r, uint64_t n_r,
                         int n_dims, uint32_t target_bits) {
  for (uint64_t j = 0; j < n_r; ++j)
    if (f2u(orc_d2(p, r + j * n_dims, n_dims)) == target_bits)
      return (uint32_t)j;
  return 0xFFFFFFFFu;
}

static void finish_result(uint64_t key, const float* a, const float* b,
                          uint64_t n_b, int n_dims, orc_result* out) {
  out->cmax2_bits = (uint32_t)(key >> 32);
  uint32_t low = (uint32_t)key;
  if (key == 0) {
    out->argmax = 0xFFFFFFFFu;
    out->nn = 0xFFFFFFFFu;
    return;
  }
  out->argmax = 0xFFFFFFFFu - low;
  out->nn = orc_first_at_d2(a + (uint64_t)out->argmax * n_dims, b, n_b, n_dims,
                            out->cmax2_bits);
}

int orc_eb_directed(const float* a, uint64_t n_a, const float* b, uint64_t n_b,
                    int n_dims, int n_threads, uint64_t seed, int shuffle,
                    orc_result* out) {
  if (!out || (n_dims != 2 && n_dims != 3) || n_a == 0 || n_b == 0 || !a ||
      !b || n_a > 0xFFFFFFFFull || n_b > 0xFFFFFFFFull)
    return -1;
  memset(out, 0, sizeof(*out));
  int nt = resolve_threads(n_threads);

  double t0 = now_ms();
  uint32_t* ida = (uint32_t*)malloc(sizeof(uint32_t) * n_a);
  uint32_t* idb = (uint32_t*)malloc(sizeof(uint32_t) * n_b);
  if (!ida || !idb) {
    free(ida);
    free(idb);
    return -1;
  }
  for (uint64_t i = 0; i < n_a; ++i) ida[i] = (uint32_t)i;
  for (uint64_t i = 0; i < n_b; ++i) idb[i] = (uint32_t)i;
  if (shuffle) {
    uint64_t state = seed; /* one engine for both, as :50-52 */
    shuffle_ids(ida, n_a, &state);
    shuffle_ids(idb, n_b, &state);
  }
  float* va = gather(a, ida, n_a, n_dims);
  float* vb = gather(b, idb, n_b, n_dims);
  free(idb);
  if (!va || !vb) {
    free(va);
    free(vb);
    free(ida);
    return -1;
  }
  out->shuffle_ms = now_ms() - t0;

  double t1 = now_ms();
  _Atomic uint64_t key = 0;
  _Atomic uint64_t compared = 0;
  uint64_t avg = (n_a + nt - 1) / nt; /* :60 */
  eb_task* tasks = (eb_task*)calloc(nt, sizeof(eb_task));
  pthread_t* th = (pthread_t*)calloc(nt, sizeof(pthread_t));
  for (int t = 0; t < nt; ++t) {
    uint64_t begin = (uint64_t)t * avg;
    uint64_t end = begin + avg < n_a ? begin + avg : n_a;
    if (begin > end) begin = end;
    tasks[t] = (eb_task){va, ida, n_a, vb, n_b, n_dims, begin, end,
                         &key, &compared};
  }
  if (nt == 1) {
    eb_worker(&tasks[0]);
  } else {
    for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, eb_worker, &tasks[t]);
    for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
  }
  out->compute_ms = now_ms() - t1;
  out->compared_pairs = atomic_load(&compared);
  out->n_threads = (uint32_t)nt;
  finish_result(atomic_load(&key), a, b, n_b, n_dims, out);
  free(tasks);
  free(th);
  free(va);
  free(vb);
  free(ida);
  return 0;
}

typedef struct {
  const float* q;
  const float* r;
  uint64_t n_r;
  int n_dims;
  uint64_t begin, end;
  float* d2;
  uint32_t* idx;
} nn_task;

static void* nn_worker(void* arg) {
  nn_task* t = (nn_task*)arg;
  const int nd = t->n_dims;
  for (uint64_t i = t->begin; i < t->end; ++i) {
    const float* p = t->q + i * nd;
    float best = FLT_MAX;
    uint32_t bi = 0xFFFFFFFFu;
    for (uint64_t j = 0; j < t->n_r; ++j) {
      float d = orc_d2(p, t->r + j * nd, nd);
      if (d < best) {
        best = d;
        bi = (uint32_t)j;
      }
    }
    t->d2[i] = best;
    t->idx[i] = bi;
  }
  return NULL;
}

int orc_nn_all(const float* q, uint64_t n_q, const float* r, uint64_t n_r,
               int n_dims, int n_threads, float* out_d2, uint32_t* out_idx) {
  if ((n_dims != 2 && n_dims != 3) || !q || !r || !out_d2 || !out_idx ||
      n_r > 0xFFFFFFFFull)
    return -1;
  int nt = resolve_threads(n_threads);
  if ((uint64_t)nt > n_q) nt = n_q ? (int)n_q : 1;
  nn_task* tasks = (nn_task*)calloc(nt, sizeof(nn_task));
  pthread_t* th = (pthread_t*)calloc(nt, sizeof(pthread_t));
  uint64_t avg = (n_q + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    uint64_t begin = (uint64_t)t * avg;
    uint64_t end = begin + avg < n_q ? begin + avg : n_q;
    if (begin > end) begin = end;
    tasks[t] = (nn_task){q, r, n_r, n_dims, begin, end, out_d2, out_idx};
  }
  for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, nn_worker, &tasks[t]);
  for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
  free(tasks);
  free(th);
  return 0;
}

int orc_brute_directed(const float* a, uint64_t n_a, const float* b,
                       uint64_t n_b, int n_dims, int n_threads,
                       orc_result* out) {
  if (!out || n_a == 0 || n_b == 0) return -1;
  memset(out, 0, sizeof(*out));
  float* d2 = (float*)malloc(sizeof(float) * n_a);
  uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * n_a);
  if (!d2 || !idx) {
    free(d2);
    free(idx);
    return -1;
  }
  double t0 = now_ms();
  int rc = orc_nn_all(a, n_a, b, n_b, n_dims, n_threads, d2, idx);
  if (rc == 0) {
    uint64_t key = 0;
    for (uint64_t i = 0; i < n_a; ++i) {
      if (d2[i] == FLT_MAX) continue; /* the reference skips FLT_MAX mins */
      uint64_t k = pack_key(d2[i], (uint32_t)i);
      if (k > key) key = k;
    }
    finish_result(key, a, b, n_b, n_dims, out);
    out->compared_pairs = n_a * n_b;
    out->n_threads = (uint32_t)resolve_threads(n_threads);
  }
  out->compute_ms = now_ms() - t0;
  free(d2);
  free(idx);
  return rc;
}

/* reference src/geoms/mbr.h:193-213: Mbr::GetMinDist2(const Mbr&). */
static float mbr_min_dist2(const float* lo, const float* hi, const float* olo,
                           const float* ohi, int nd) {
  float dist2 = 0;
  for (int i = 0; i < nd; ++i) {
    float diff = 0;
    if (ohi[i] < lo[i]) {
      diff = lo[i] - ohi[i];
    } else if (hi[i] < olo[i]) {
      diff = olo[i] - hi[i];
    }
    dist2 += diff * diff;
  }
  return dist2;
}

/* reference src/geoms/mbr.h:249-266: Mbr::GetMaxDist2(const Mbr&). */
static float mbr_max_dist2(const float* lo, const float* hi, const float* olo,
                           const float* ohi, int nd) {
  float dist2 = 0;
  for (int i = 0; i < nd; ++i) {
    float d1 = fabsf(lo[i] - ohi[i]);
    float d2 = fabsf(hi[i] - olo[i]);
    float diff = d1 > d2 ? d1 : d2;
    dist2 += diff * diff;
  }
  return dist2;
}

/* Faces of an MBR as degenerate boxes, in the order of reference
 * src/geoms/hd_bounds.h:38-46 (2D: bottom, right, top, left) and :95-103
 * (3D: bottom y=lo, top y=hi, left x=lo, right x=hi, front z=lo, back z=hi).
 * face_lo / face_hi hold n_faces * n_dims floats. */
static int mbr_faces(const float* lo, const float* hi, int nd, float* face_lo,
                     float* face_hi) {
  if (nd == 2) {
    const float fl[4][2] = {{lo[0], lo[1]}, {hi[0], lo[1]}, {lo[0], hi[1]},
                            {lo[0], lo[1]}};
    const float fh[4][2] = {{hi[0], lo[1]}, {hi[0], hi[1]}, {hi[0], hi[1]},
                            {lo[0], hi[1]}};
    memcpy(face_lo, fl, sizeof(fl));
    memcpy(face_hi, fh, sizeof(fh));
    return 4;
  }
  /* corners p0..p7 as hd_bounds.h:65-90 */
  float p[8][3] = {{lo[0], lo[1], lo[2]}, {hi[0], lo[1], lo[2]},
                   {lo[0], hi[1], lo[2]}, {hi[0], hi[1], lo[2]},
                   {lo[0], lo[1], hi[2]}, {hi[0], lo[1], hi[2]},
                   {lo[0], hi[1], hi[2]}, {hi[0], hi[1], hi[2]}};
  const int pairs[6][2] = {{0, 5}, {2, 7}, {0, 6}, {1, 7}, {0, 3}, {4, 7}};
  for (int f = 0; f < 6; ++f) {
    memcpy(face_lo + f * 3, p[pairs[f][0]], sizeof(float) * 3);
    memcpy(face_hi + f * 3, p[pairs[f][1]], sizeof(float) * 3);
  }
  return 6;
}

/* reference src/geoms/hd_bounds.h:123-130 (GetLowerBound2) and :136-156
 * (GetUpperBound2), region_ = faces of MBR(A), other = MBR(B). */
void orc_hd_bounds2(const float* mbr_a, const float* mbr_b, int n_dims,
                    float* lb2, float* ub2) {
  const int nd = n_dims;
  float fa_lo[18], fa_hi[18], fb_lo[18], fb_hi[18];
  int nf = mbr_faces(mbr_a, mbr_a + nd, nd, fa_lo, fa_hi);
  mbr_faces(mbr_b, mbr_b + nd, nd, fb_lo, fb_hi);
  float lb = 0;
  for (int i = 0; i < nf; ++i) {
    float d = mbr_min_dist2(fa_lo + i * nd, fa_hi + i * nd, mbr_b, mbr_b + nd,
                            nd);
    lb = lb > d ? lb : d;
  }
  float max_dist2 = 0;
  for (int i = 0; i < nf; ++i) {
    float min_dist2 = FLT_MAX;
    for (int j = 0; j < nf; ++j) {
      float d = mbr_max_dist2(fa_lo + i * nd, fa_hi + i * nd, fb_lo + j * nd,
                              fb_hi + j * nd, nd);
      if (d <= max_dist2) break;
      min_dist2 = min_dist2 < d ? min_dist2 : d;
    }
    if (min_dist2 != FLT_MAX) max_dist2 = max_dist2 > min_dist2 ? max_dist2 : min_dist2;
  }
  *lb2 = lb;
  *ub2 = max_dist2;
}

/* reference src/geoms/mbr.h:164-173 (Expand) and :288-296 (ToNonemptyMBR). */
void orc_mbr(const float* p, uint64_t n, int n_dims, float* out) {
  for (int d = 0; d < n_dims; ++d) {
    out[d] = FLT_MAX;
    out[n_dims + d] = -FLT_MAX;
  }
  for (uint64_t i = 0; i < n; ++i)
    for (int d = 0; d < n_dims; ++d) {
      float v = p[i * n_dims + d];
      if (v < out[d]) out[d] = v;
      if (v > out[n_dims + d]) out[n_dims + d] = v;
    }
  for (int d = 0; d < n_dims; ++d)
    if (out[n_dims + d] - out[d] == 0) out[n_dims + d] = (float)((double)out[d] + 1e-5);
}

/* FP64 forward/backward of the ordering MLP (no reference counterpart: the
 * reference's selector uses gradient-boosted trees, src/models/tree_*.h):
 * h1 = sin(W1 x + b1), h2 = relu(W2 h1 + b2), f = w3 . h2 + b3.
 * Loss = mean over the batch of (f(x) - y)^2. */
double orc_mlp_loss_grad(const float* params, int d_in, int hidden,
                         const float* x, const float* y, int batch,
                         double* grad) {
  const int H = hidden;
  const float* W1 = params;
  const float* b1 = W1 + H * d_in;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* w3 = b2 + H;
  const float* b3 = w3 + H;
  const size_t n_params = (size_t)H * d_in + H + (size_t)H * H + H + H + 1;
  double* gW1 = grad;
  double* gb1 = gW1 + H * d_in;
  double* gW2 = gb1 + H;
  double* gb2 = gW2 + H * H;
  double* gw3 = gb2 + H;
  double* gb3 = gw3 + H;
  memset(grad, 0, sizeof(double) * n_params);
  double* z1 = (double*)malloc(sizeof(double) * H);
  double* h1 = (double*)malloc(sizeof(double) * H);
  double* h2 = (double*)malloc(sizeof(double) * H);
  double* g2 = (double*)malloc(sizeof(double) * H);
  double* g1 = (double*)malloc(sizeof(double) * H);
  double loss = 0;
  for (int s = 0; s < batch; ++s) {
    const float* xs = x + (size_t)s * d_in;
    for (int j = 0; j < H; ++j) {
      double acc = b1[j];
      for (int k = 0; k < d_in; ++k) acc += (double)W1[j * d_in + k] * xs[k];
      z1[j] = acc;
      h1[j] = sin(acc);
    }
    for (int j = 0; j < H; ++j) {
      double acc = b2[j];
      for (int k = 0; k < H; ++k) acc += (double)W2[j * H + k] * h1[k];
      h2[j] = acc > 0 ? acc : 0;
    }
    double out = b3[0];
    for (int k = 0; k < H; ++k) out += (double)w3[k] * h2[k];
    double err = out - y[s];
    loss += err * err;
    double go = 2.0 * err / batch;
    gb3[0] += go;
    for (int k = 0; k < H; ++k) {
      gw3[k] += go * h2[k];
      g2[k] = h2[k] > 0 ? go * w3[k] : 0;
    }
    for (int j = 0; j < H; ++j) {
      gb2[j] += g2[j];
      for (int k = 0; k < H; ++k) gW2[j * H + k] += g2[j] * h1[k];
    }
    for (int k = 0; k < H; ++k) {
      double acc = 0;
      for (int j = 0; j < H; ++j) acc += g2[j] * W2[j * H + k];
      g1[k] = acc * cos(z1[k]);
    }
    for (int j = 0; j < H; ++j) {
      gb1[j] += g1[j];
      for (int k = 0; k < d_in; ++k) gW1[j * d_in + k] += g1[j] * xs[k];
    }
  }
  free(z1);
  free(h1);
  free(h2);
  free(g2);
  free(g1);
  return loss / batch;
}
