// gs_learned.hip -- C ABI of the learned Schur-complement preconditioner
// (include/gale_learned.h): network objects, training pairs recorded by the
// Uzawa loop, device training, and the preconditioner application used by
// the flexible Uzawa iteration.
#include <hip/hip_runtime.h>
#include <string.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "gs_learned.h"
#include "gs_net.h"
#include "gs_reduce.h"

using namespace gs;

#define LCK(x)                                                                                              \
  do {                                                                                                      \
    hipError_t e_ = (x);                                                                                    \
    if (e_ != hipSuccess) {                                                                                 \
      std::fprintf(stderr, "gale_stokes(learned): HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                   __LINE__);                                                                               \
      return GS_E_HIP;                                                                                      \
    }                                                                                                       \
  } while (0)

struct gs_lpc {
  net::Net net;
  hipStream_t st = nullptr;
  // inference buffers (one grid)
  long long infer_V = 0;
  float* in2 = nullptr;
  float* out = nullptr;
  // training batch buffers
  long long train_n = 0;
  float* bin2 = nullptr;
  float* btgt = nullptr;
  float* bout = nullptr;
  float* bgout = nullptr;
  float* bw = nullptr;
  int* bproj = nullptr;
  int train_B = 0;
};

struct gs_lpc_dataset {
  int nx = 0, ny = 0, nz = 0;
  long long V = 0;
  int cap = 0, size = 0;
  float* in2 = nullptr;  // [cap][V][2]
  float* tgt = nullptr;  // [cap][V]
  float* w = nullptr;    // [cap]
  int* proj = nullptr;   // [cap]
};

namespace gs {

bool lpc_shape_ok(const Geo& g) {
  const int m = net::kGridMultiple;
  return g.dim == 3 && g.res[0] % m == 0 && g.res[1] % m == 0 && g.res[2] % m == 0 && g.res[0] >= m &&
         g.res[1] >= m && g.res[2] >= m;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_elem_logeta(const double4* __restrict__ pk, const int* __restrict__ off, long long E, double vol,
                              float* __restrict__ out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int p = off[e]; p < off[e + 1]; p++) s += pk[p].w;
    const double eta = s / vol;
    out[e] = eta > 0.0 ? (float)(log10(eta) / 3.0) : 0.f;
  }
}

// t = Qinv r (into z); S[S_LP_TT] = sum t^2
__global__ void k_lp_prep(const double* __restrict__ r, const double* __restrict__ Qinv, double* __restrict__ z,
                          long long n, double* S, Reducer rd) {
  double v[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double t = Qinv[i] * r[i];
    z[i] = t;
    v[0] += t * t;
  }
  block_partials<1, false>(v, rd);
}

__global__ void k_lp_input(const double* __restrict__ t, const float* __restrict__ logeta, float* __restrict__ in2,
                           long long n, const double* S) {
  const double rms = sqrt(S[S_LP_TT] / (double)n);
  const double inv = rms > 0.0 ? 1.0 / rms : 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    in2[2 * i] = (float)(t[i] * inv);
    in2[2 * i + 1] = logeta[i];
  }
}

// z = t + rms c; S_ZR = z.r; S_LP_ZQ = z.qprev; S_UBETA = -alpha (z.qprev) / zr_prev
__global__ void k_lp_combine(const double* __restrict__ r, const double* __restrict__ qprev, double* __restrict__ z,
                             const float* __restrict__ c, long long n, double* S, Reducer rd) {
  const double rms = sqrt(S[S_LP_TT] / (double)n);
  double v[2] = {0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double zi = z[i] + rms * (double)c[i];
    z[i] = zi;
    v[0] += zi * r[i];
    if (qprev) v[1] += zi * qprev[i];
  }
  block_partials<2, false>(v, rd);
}
struct OpLpCombine {
  double* S;
  int has_prev;
  __device__ void operator()(const double* t) const {
    S[S_ZR] = t[0];
    S[S_LP_ZQ] = t[1];
    const double prev = S[S_ZR_PREV];
    S[S_UBETA] = (has_prev && prev != 0.0) ? -S[S_TMP0] * t[1] / prev : 0.0;
  }
};

// training pair statistics: t = Qinv r; sums t^2, e, e^2
__global__ void k_ds_stats(const double* __restrict__ r, const double* __restrict__ Qinv,
                           const double* __restrict__ e, long long n, double* S, Reducer rd) {
  double v[3] = {0.0, 0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double t = Qinv[i] * r[i];
    v[0] += t * t;
    v[1] += e[i];
    v[2] += e[i] * e[i];
  }
  block_partials<3, false>(v, rd);
}
struct OpDsStats {
  double* S;
  __device__ void operator()(const double* t) const {
    S[S_DS_TT] = t[0];
    S[S_DS_E] = t[1];
    S[S_DS_EE] = t[2];
  }
};

// in2 = {t / rms, logeta}; target = (e - t) / rms; weight = rms^2 / |P e|^2
__global__ void k_ds_write(const double* __restrict__ r, const double* __restrict__ Qinv,
                           const double* __restrict__ e, const float* __restrict__ logeta, long long n,
                           const double* S, int proj, float* __restrict__ in2, float* __restrict__ tgt,
                           float* __restrict__ w, int* __restrict__ pj) {
  const double rms = sqrt(S[S_DS_TT] / (double)n);
  const double inv = rms > 0.0 ? 1.0 / rms : 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double t = Qinv[i] * r[i];
    in2[2 * i] = (float)(t * inv);
    in2[2 * i + 1] = logeta[i];
    tgt[i] = (float)((e[i] - t) * inv);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double se = S[S_DS_E], see = S[S_DS_EE];
    const double pe = proj ? see - se * se / (double)n : see;
    w[0] = (pe > 0.0 && rms > 0.0) ? (float)(rms * rms / pe) : 0.f;
    pj[0] = proj;
  }
}

// partial[b][j] = sum over this block's elements of z . qring[j]
__global__ void k_mdot_partial(const double* __restrict__ z, const double* __restrict__ qring, int c, long long n,
                               double* __restrict__ partial) {
  __shared__ double sh[kMaxWindow][kBlock / 64];
  double acc[kMaxWindow];
#pragma unroll
  for (int j = 0; j < kMaxWindow; j++) acc[j] = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double zi = z[i];
#pragma unroll
    for (int j = 0; j < kMaxWindow; j++)
      if (j < c) acc[j] += zi * qring[(long long)j * n + i];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kMaxWindow; j++) {
    if (j < c) {  // c is uniform across the block
      const double v = wave_sum(acc[j]);
      if (lane == 0) sh[j][wid] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < c) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[threadIdx.x][w];
    partial[(long long)blockIdx.x * kMaxWindow + threadIdx.x] = t;
  }
}

// beta[j] = (sum_b partial[b][j]) / sq[j]; one wavefront per j (fixed order:
// lane-strided partial sums, then the wavefront tree)
__global__ void k_mdot_final(const double* __restrict__ partial, int nb, int c, const double* __restrict__ sq,
                             double* __restrict__ beta) {
  const int j = blockIdx.x, lane = threadIdx.x;
  if (j >= c) return;
  double t = 0.0;
  for (int b = lane; b < nb; b += 64) t += partial[(long long)b * kMaxWindow + j];
  t = wave_sum(t);
  if (lane == 0) beta[j] = sq[j] != 0.0 ? t / sq[j] : 0.0;
}

// s_out = z - sum_j beta[j] s_j  (s_out may be one of the ring slots: each
// element is read and written by the same thread)
__global__ void k_orth(const double* __restrict__ z, double* sring, const double* __restrict__ beta, int c, int out,
                       long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double v = z[i];
    for (int j = 0; j < c; j++) v -= beta[j] * sring[(long long)j * n + i];
    sring[(long long)out * n + i] = v;
  }
}

__global__ void k_flex_alpha(const double* __restrict__ s, const double* __restrict__ r, const double* __restrict__ q,
                             long long n, double* S, double* sq_out, Reducer rd) {
  double v[2] = {0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    v[0] += s[i] * r[i];
    v[1] += s[i] * q[i];
  }
  block_partials<2, false>(v, rd);
}
struct OpFlexAlpha {
  double* S;
  double* sq_out;
  __device__ void operator()(const double* t) const {
    S[S_ZR] = t[0];
    S[S_SQ] = t[1];
    sq_out[0] = t[1];
  }
};

void launch_orthogonalise(const double* z, double* sring, const double* qring, const double* sq, int c, int out,
                          long long n, double* partials, double* beta, hipStream_t st) {
  if (c > 0) {
    int nb = reduce_grid(n);
    hipLaunchKernelGGL(k_mdot_partial, dim3(nb), dim3(kBlock), 0, st, z, qring, c, n, partials);
    hipLaunchKernelGGL(k_mdot_final, dim3(c), dim3(64), 0, st, partials, nb, c, sq, beta);
  }
  hipLaunchKernelGGL(k_orth, dim3(grid_for(n)), dim3(kBlock), 0, st, z, sring, beta, c, out, n);
}

void launch_flex_alpha(const double* s, const double* r, const double* q, long long n, double* S, double* sq_out,
                       Reducer rd, hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_flex_alpha, dim3(nb), dim3(kBlock), 0, st, s, r, q, n, S, sq_out, rd);
  reduce_finish<2, false>(rd, nb, OpFlexAlpha{S, sq_out}, st);
}

void launch_elem_logeta(const Geo& g, const double4* pk, const int* off, float* out, hipStream_t st) {
  double vol = 1.0;
  for (int d = 0; d < g.dim; d++) vol *= (g.hi[d] - g.lo[d]) / g.res[d];
  hipLaunchKernelGGL(k_elem_logeta, dim3(grid_for(g.E)), dim3(kBlock), 0, st, pk, off, g.E, vol, out);
}

static int ensure_infer(gs_lpc* L, long long V) {
  if (L->infer_V == V) return GS_OK;
  if (L->in2) (void)hipFree(L->in2);
  if (L->out) (void)hipFree(L->out);
  L->in2 = L->out = nullptr;
  L->infer_V = 0;
  LCK(hipMalloc(&L->in2, sizeof(float) * 2 * V));
  LCK(hipMalloc(&L->out, sizeof(float) * V));
  L->infer_V = V;
  return GS_OK;
}

int lpc_precond(gs_lpc* L, const Geo& g, const double* r, const double* Qinv, const float* logeta,
                const double* qprev, double* z, double* S, Reducer rd, hipStream_t st) {
  if (!L || !lpc_shape_ok(g) || !Qinv) return GS_E_INVALID;
  const long long n = g.E;
  int e = ensure_infer(L, n);
  if (e) return e;
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_lp_prep, dim3(nb), dim3(kBlock), 0, st, r, Qinv, z, n, S, rd);
  reduce_finish<1, false>(rd, nb, OpStore{S, S_LP_TT}, st);
  hipLaunchKernelGGL(k_lp_input, dim3(grid_for(n)), dim3(kBlock), 0, st, z, logeta, L->in2, n, S);
  L->net.st = st;
  const net::Shape sh{g.res[0], g.res[1], g.res[2], 1};
  if ((e = net::forward(&L->net, sh, L->in2, L->out))) return e;
  hipLaunchKernelGGL(k_lp_combine, dim3(nb), dim3(kBlock), 0, st, r, qprev, z, L->out, n, S, rd);
  reduce_finish<2, false>(rd, nb, OpLpCombine{S, qprev ? 1 : 0}, st);
  return GS_OK;
}

int lpc_record(gs_lpc_dataset* D, const Geo& g, const double* rlike, const double* elike, const double* Qinv,
               const float* logeta, int proj, double* S, Reducer rd, hipStream_t st) {
  if (!D || !Qinv) return GS_E_INVALID;
  if (D->nx != g.res[0] || D->ny != g.res[1] || D->nz != g.res[2] || g.dim != 3) return GS_E_INVALID;
  if (D->size >= D->cap) return GS_OK;
  const long long n = g.E;
  const long long idx = D->size;
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_ds_stats, dim3(nb), dim3(kBlock), 0, st, rlike, Qinv, elike, n, S, rd);
  reduce_finish<3, false>(rd, nb, OpDsStats{S}, st);
  hipLaunchKernelGGL(k_ds_write, dim3(grid_for(n)), dim3(kBlock), 0, st, rlike, Qinv, elike, logeta, n, S, proj,
                     D->in2 + idx * 2 * n, D->tgt + idx * n, D->w + idx, D->proj + idx);
  D->size++;
  return GS_OK;
}

}  // namespace gs

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int32_t gs_lpc_create(uint64_t seed, gs_lpc** out) {
  if (!out) return GS_E_INVALID;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return GS_E_NO_DEVICE;
  gs_lpc* L = new gs_lpc();
  LCK(hipStreamCreate(&L->st));
  int e = net::create(&L->net, seed, L->st);
  if (e) {
    net::destroy(&L->net);
    delete L;
    return e;
  }
  *out = L;
  return GS_OK;
}

static void free_train(gs_lpc* L) {
  float* fs[] = {L->bin2, L->btgt, L->bout, L->bgout, L->bw};
  for (float* p : fs)
    if (p) (void)hipFree(p);
  if (L->bproj) (void)hipFree(L->bproj);
  L->bin2 = L->btgt = L->bout = L->bgout = L->bw = nullptr;
  L->bproj = nullptr;
  L->train_n = 0;
  L->train_B = 0;
}

extern "C" void gs_lpc_destroy(gs_lpc* L) {
  if (!L) return;
  if (L->st) (void)hipStreamSynchronize(L->st);
  net::destroy(&L->net);
  free_train(L);
  if (L->in2) (void)hipFree(L->in2);
  if (L->out) (void)hipFree(L->out);
  if (L->st) (void)hipStreamDestroy(L->st);
  delete L;
}

extern "C" int64_t gs_lpc_num_params(void) { return net::num_params(); }

extern "C" int32_t gs_lpc_get_params(gs_lpc* L, float* params) {
  if (!L || !params) return GS_E_INVALID;
  LCK(hipStreamSynchronize(L->st));
  LCK(hipMemcpy(params, L->net.w, sizeof(float) * net::num_params(), hipMemcpyDeviceToHost));
  return GS_OK;
}

extern "C" int32_t gs_lpc_set_params(gs_lpc* L, const float* params) {
  if (!L || !params) return GS_E_INVALID;
  LCK(hipStreamSynchronize(L->st));
  LCK(hipMemcpy(L->net.w, params, sizeof(float) * net::num_params(), hipMemcpyHostToDevice));
  L->net.st = L->st;
  net::pack(&L->net);
  LCK(hipStreamSynchronize(L->st));
  return GS_OK;
}

static int shape_of(int32_t nx, int32_t ny, int32_t nz, int32_t batch, net::Shape* sh) {
  const int m = net::kGridMultiple;
  if (nx < m || ny < m || nz < m || nx % m || ny % m || nz % m || batch < 1) return GS_E_INVALID;
  *sh = net::Shape{nx, ny, nz, batch};
  return GS_OK;
}

extern "C" int32_t gs_lpc_forward(gs_lpc* L, int32_t nx, int32_t ny, int32_t nz, int32_t batch, const float* d_in,
                                  float* d_out) {
  net::Shape sh;
  if (!L || !d_in || !d_out || shape_of(nx, ny, nz, batch, &sh)) return GS_E_INVALID;
  L->net.st = L->st;
  int e = net::forward(&L->net, sh, d_in, d_out);
  if (e) return e;
  LCK(hipStreamSynchronize(L->st));
  return GS_OK;
}

extern "C" int32_t gs_lpc_backward(gs_lpc* L, int32_t nx, int32_t ny, int32_t nz, int32_t batch, const float* d_in,
                                   const float* d_gout, float* grads) {
  net::Shape sh;
  if (!L || !d_in || !d_gout || !grads || shape_of(nx, ny, nz, batch, &sh)) return GS_E_INVALID;
  L->net.st = L->st;
  int e = net::forward(&L->net, sh, d_in, nullptr);
  if (e) return e;
  if ((e = net::backward(&L->net, sh, d_gout))) return e;
  LCK(hipStreamSynchronize(L->st));
  LCK(hipMemcpy(grads, L->net.g, sizeof(float) * net::num_params(), hipMemcpyDeviceToHost));
  return GS_OK;
}

static int ensure_train(gs_lpc* L, int B, long long V) {
  if (L->train_B == B && L->train_n == V) return GS_OK;
  free_train(L);
  const long long n = (long long)B * V;
  LCK(hipMalloc(&L->bin2, sizeof(float) * 2 * n));
  LCK(hipMalloc(&L->btgt, sizeof(float) * n));
  LCK(hipMalloc(&L->bout, sizeof(float) * n));
  LCK(hipMalloc(&L->bgout, sizeof(float) * n));
  LCK(hipMalloc(&L->bw, sizeof(float) * B));
  LCK(hipMalloc(&L->bproj, sizeof(int) * B));
  L->train_B = B;
  L->train_n = V;
  return GS_OK;
}

extern "C" int32_t gs_lpc_loss(gs_lpc* L, int32_t nx, int32_t ny, int32_t nz, int32_t batch, const float* d_out,
                               const float* d_target, const float* weights, const int32_t* proj, float* d_gout,
                               double* loss) {
  net::Shape sh;
  if (!L || !d_out || !d_target || !weights || !proj || !d_gout || shape_of(nx, ny, nz, batch, &sh))
    return GS_E_INVALID;
  L->net.st = L->st;
  int e = net::ensure_work(&L->net, sh);
  if (e) return e;
  if ((e = ensure_train(L, batch, (long long)nx * ny * nz))) return e;
  LCK(hipMemcpyAsync(L->bw, weights, sizeof(float) * batch, hipMemcpyHostToDevice, L->st));
  LCK(hipMemcpyAsync(L->bproj, proj, sizeof(int) * batch, hipMemcpyHostToDevice, L->st));
  double l = 0.0;
  if ((e = net::loss_grad(&L->net, sh, d_out, d_target, L->bw, L->bproj, d_gout, &l))) return e;
  LCK(hipStreamSynchronize(L->st));
  if (loss) *loss = l;
  return GS_OK;
}

extern "C" int32_t gs_lpc_dataset_create(int32_t nx, int32_t ny, int32_t nz, int32_t capacity,
                                         gs_lpc_dataset** out) {
  if (!out || capacity < 1) return GS_E_INVALID;
  net::Shape sh;
  if (shape_of(nx, ny, nz, 1, &sh)) return GS_E_INVALID;
  gs_lpc_dataset* D = new gs_lpc_dataset();
  D->nx = nx;
  D->ny = ny;
  D->nz = nz;
  D->V = (long long)nx * ny * nz;
  D->cap = capacity;
  const long long n = (long long)capacity * D->V;
  if (hipMalloc(&D->in2, sizeof(float) * 2 * n) != hipSuccess || hipMalloc(&D->tgt, sizeof(float) * n) != hipSuccess ||
      hipMalloc(&D->w, sizeof(float) * capacity) != hipSuccess ||
      hipMalloc(&D->proj, sizeof(int) * capacity) != hipSuccess) {
    gs_lpc_dataset_destroy(D);
    return GS_E_HIP;
  }
  *out = D;
  return GS_OK;
}

extern "C" void gs_lpc_dataset_destroy(gs_lpc_dataset* D) {
  if (!D) return;
  (void)hipDeviceSynchronize();
  if (D->in2) (void)hipFree(D->in2);
  if (D->tgt) (void)hipFree(D->tgt);
  if (D->w) (void)hipFree(D->w);
  if (D->proj) (void)hipFree(D->proj);
  delete D;
}

extern "C" int32_t gs_lpc_dataset_size(const gs_lpc_dataset* D) { return D ? D->size : 0; }

extern "C" int32_t gs_lpc_dataset_get(gs_lpc_dataset* D, int32_t idx, float* in, float* target, float* weight,
                                      int32_t* proj) {
  if (!D || idx < 0 || idx >= D->size) return GS_E_INVALID;
  LCK(hipDeviceSynchronize());
  const long long V = D->V;
  if (in) LCK(hipMemcpy(in, D->in2 + (long long)idx * 2 * V, sizeof(float) * 2 * V, hipMemcpyDeviceToHost));
  if (target) LCK(hipMemcpy(target, D->tgt + (long long)idx * V, sizeof(float) * V, hipMemcpyDeviceToHost));
  if (weight) LCK(hipMemcpy(weight, D->w + idx, sizeof(float), hipMemcpyDeviceToHost));
  if (proj) LCK(hipMemcpy(proj, D->proj + idx, sizeof(int), hipMemcpyDeviceToHost));
  return GS_OK;
}

static uint32_t xorshift(uint32_t* s) {
  uint32_t x = *s;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  *s = x;
  return x;
}

static int gather_batch(gs_lpc* L, gs_lpc_dataset* D, const std::vector<int>& idx) {
  const long long V = D->V;
  for (size_t b = 0; b < idx.size(); b++) {
    const long long i = idx[b];
    LCK(hipMemcpyAsync(L->bin2 + b * 2 * V, D->in2 + i * 2 * V, sizeof(float) * 2 * V, hipMemcpyDeviceToDevice,
                       L->st));
    LCK(hipMemcpyAsync(L->btgt + b * V, D->tgt + i * V, sizeof(float) * V, hipMemcpyDeviceToDevice, L->st));
    LCK(hipMemcpyAsync(L->bw + b, D->w + i, sizeof(float), hipMemcpyDeviceToDevice, L->st));
    LCK(hipMemcpyAsync(L->bproj + b, D->proj + i, sizeof(int), hipMemcpyDeviceToDevice, L->st));
  }
  return GS_OK;
}

extern "C" int32_t gs_lpc_train(gs_lpc* L, gs_lpc_dataset* D, const gs_lpc_train_params* p, float* loss_history) {
  if (!L || !D || !p || p->steps < 0 || p->batch < 1) return GS_E_INVALID;
  const int first = p->first < 0 ? 0 : p->first;
  const int count = p->count > 0 ? p->count : D->size - first;
  if (count < 1 || first + count > D->size) return GS_E_INVALID;
  LCK(hipDeviceSynchronize());  // recording may have run on a solver stream
  const int B = p->batch;
  const net::Shape sh{D->nx, D->ny, D->nz, B};
  L->net.st = L->st;
  int e = net::ensure_work(&L->net, sh);
  if (e) return e;
  if ((e = ensure_train(L, B, D->V))) return e;
  uint32_t rs = p->seed ? p->seed : 0x9e3779b9u;
  std::vector<int> idx(B);
  const float b1 = p->beta1 > 0 ? p->beta1 : 0.9f, b2 = p->beta2 > 0 ? p->beta2 : 0.999f;
  const float eps = p->eps > 0 ? p->eps : 1e-8f;
  for (int step = 0; step < p->steps; step++) {
    for (int b = 0; b < B; b++) idx[b] = first + (int)(xorshift(&rs) % (uint32_t)count);
    if ((e = gather_batch(L, D, idx))) return e;
    float lr;
    if (p->warmup > 0 && step < p->warmup)
      lr = p->lr * (float)(step + 1) / (float)p->warmup;
    else {
      const int span = p->steps - p->warmup;
      const float t = span > 0 ? (float)(step - p->warmup) / (float)span : 1.f;
      lr = p->lr_final + 0.5f * (p->lr - p->lr_final) * (1.f + std::cos(3.14159265f * t));
    }
    if ((e = net::forward(&L->net, sh, L->bin2, L->bout))) return e;
    double loss = 0.0;
    if ((e = net::loss_grad(&L->net, sh, L->bout, L->btgt, L->bw, L->bproj, L->bgout, &loss))) return e;
    if ((e = net::backward(&L->net, sh, L->bgout))) return e;
    if ((e = net::adam_step(&L->net, lr, b1, b2, eps, p->clip, nullptr))) return e;
    if (loss_history) loss_history[step] = (float)loss;
    if (!std::isfinite(loss)) return GS_E_NAN;
  }
  LCK(hipStreamSynchronize(L->st));
  return GS_OK;
}

extern "C" int32_t gs_lpc_eval(gs_lpc* L, gs_lpc_dataset* D, int32_t first, int32_t count, double* loss) {
  if (!L || !D || !loss || first < 0 || count < 1 || first + count > D->size) return GS_E_INVALID;
  LCK(hipDeviceSynchronize());
  const int B = count < 8 ? count : 8;
  L->net.st = L->st;
  double tot = 0.0;
  int done = 0;
  while (done < count) {
    const int nb = (count - done) < B ? (count - done) : B;
    const net::Shape sh{D->nx, D->ny, D->nz, nb};
    int e = net::ensure_work(&L->net, sh);
    if (e) return e;
    if ((e = ensure_train(L, nb, D->V))) return e;
    std::vector<int> idx(nb);
    for (int b = 0; b < nb; b++) idx[b] = first + done + b;
    if ((e = gather_batch(L, D, idx))) return e;
    if ((e = net::forward(&L->net, sh, L->bin2, L->bout))) return e;
    double l = 0.0;
    if ((e = net::loss_grad(&L->net, sh, L->bout, L->btgt, L->bw, L->bproj, L->bgout, &l))) return e;
    tot += l * nb;
    done += nb;
  }
  *loss = tot / count;
  return GS_OK;
}

extern "C" int32_t gs_lpc_selftest(float* max_err) {
  if (!max_err) return GS_E_INVALID;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return GS_E_NO_DEVICE;
  return net::selftest_mfma(max_err);
}
