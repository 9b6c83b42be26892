// gs_batch.hip -- batched solve of many independent 2D Stokes instances in one
// persistent launch (include/gale_batch.h).
//
// Each workgroup takes instances from a queue (largest first) and runs the
// complete reference Uzawa loop for it with workgroup barriers only:
//   K apply     element tiles of TX x (BS/TX) staged in LDS, forces
//               accumulated in four conflict-free corner phases, the node box
//               added to y tile after tile (deterministic);
//   G, G^T      closed-form Q1-P0 gradient / divergence (node gather,
//               element loop), as the single-problem kernels;
//   inner solve FP64 flexible CG on K v = b, preconditioned by one FP32
//               geometric multigrid V-cycle per iteration (the instance's
//               Galerkin moment levels, Chebyshev-Jacobi smoothing, dense
//               coarsest solve), all FP32 K applies;
//   outer loop  Stokes_SLE_UzawaSolver.cxx:553-837 (diagonal Q, the same
//               residual, stopping rule and firewalls), FP64.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/gale_batch.h"
#include "gs_internal.h"
#include "gs_mg.h"
#include "gs_reduce.h"

namespace gs {
struct BatchLevelInfo {
  int nx, ny;
  double lmax;
};
int32_t batch_mg_info(gs_solver* s, int* nlev, BatchLevelInfo* lv, int max_levels, int* nfc);
int32_t batch_copy_fine(gs_solver* s, double* f, double* dinv, double* Qinv, double* M64, hipStream_t st);
int32_t batch_copy_level(gs_solver* s, int l, float* M32, float* dinv32, uint8_t* bcflag, hipStream_t st);
int32_t batch_copy_dense(gs_solver* s, float* Kinv32, int* freeidx, hipStream_t st);
}  // namespace gs

using namespace gs;

#define BCK(x)                                                                                               \
  do {                                                                                                       \
    hipError_t e_ = (x);                                                                                     \
    if (e_ != hipSuccess) {                                                                                  \
      std::fprintf(stderr, "gale_stokes(batch): HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                   __LINE__);                                                                                \
      return GS_E_HIP;                                                                                       \
    }                                                                                                        \
  } while (0)

namespace {

struct BInst {
  int nx, ny;
  long long node_off, elem_off;
  double hx, hy;
  int lev0, nlev, nfc;  // levels [lev0, lev0 + nlev) in the level table; dense coarsest size (<= 0: none)
  long long kinv_off, free_off;
};

struct BLevel {
  int nx, ny;
  long long node_off, elem_off;  // into the flat level arrays
  double hx, hy, lmax;
};

struct BArgs {
  const BInst* inst;
  const BLevel* lev;
  const float* LM;       // FP32 moments of every level
  const float* Ldinv;    // FP32 masked inverse diagonal of every level
  const uint8_t* Lbc;    // node flags of every level
  float *Lb, *Lx, *Lr, *Ld, *LAx;  // level work vectors
  const float* Kinv;     // dense coarsest inverses
  const int* freeidx;
  const int* order;  // queue order (largest first)
  int B;
  const double* M;
  const double* f;
  const double* dinv;
  const double* Qinv;
  const uint8_t* bc;
  double *u, *vstar, *ft, *cr, *cz, *cp, *cAp;
  double *p, *r, *s, *q, *z;
  int* status;
  unsigned long long* itime;  // per instance: device wall-clock ticks of its solve
  int* outer;
  long long* inner;
  double* relres;
  unsigned* queue;
  double *fs, *fq;   // relaxed loop: per-instance rings of window directions and their Schur images
  // parameters
  double tol, irt;
  int max_outer, max_inner, nu;
  double relax;      // > 0: relaxed inner tolerance with a truncated flexible outer loop (0: reference loop)
  int window;        // directions of the flexible outer loop
};
constexpr int kBatchMaxWin = 16;

template <int BS>
__device__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < BS / 64; i++) t += sh[i];
  return t;
}

// Nodal forces of one Q1 element from its 4 nodal velocities and 9 particle
// moments (moment form of the per-particle integration, see gs_tile.hip).
template <typename T>
__device__ __forceinline__ void forces2(const T (&u)[4][2], T ix, T iy, const T (&M)[9], T (&ye)[4][2]) {
  T A[2][2], Bm[2][2];
#pragma unroll
  for (int cc = 0; cc < 2; cc++) {
    const T u0 = u[0][cc], u1 = u[1][cc], u2 = u[2][cc], u3 = u[3][cc];
    const T t1 = (u1 - u0) + (u3 - u2), t2 = (u2 + u3) - (u0 + u1), t3 = (u3 - u2) - (u1 - u0);
    A[cc][0] = ix * t1;
    A[cc][1] = ix * t3;
    Bm[cc][0] = iy * t2;
    Bm[cc][1] = iy * t3;
  }
  const T mxx0 = 2 * (A[0][0] * M[0] + A[0][1] * M[3]);
  const T mxxs = 2 * (A[0][0] * M[3] + A[0][1] * M[6]);
  const T myy0 = 2 * (Bm[1][0] * M[0] + Bm[1][1] * M[1]);
  const T myyr = 2 * (Bm[1][0] * M[1] + Bm[1][1] * M[2]);
  const T txy0 = Bm[0][0] * M[0] + Bm[0][1] * M[1] + A[1][0] * M[0] + A[1][1] * M[3];
  const T txyr = Bm[0][0] * M[1] + Bm[0][1] * M[2] + A[1][0] * M[1] + A[1][1] * M[4];
  const T txys = Bm[0][0] * M[3] + Bm[0][1] * M[4] + A[1][0] * M[3] + A[1][1] * M[6];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const T ax = (a & 1) ? (T)1 : (T)-1, ay = (a & 2) ? (T)1 : (T)-1;
    ye[a][0] = ix * ax * (mxx0 + ay * mxxs) + iy * ay * (txy0 + ax * txyr);
    ye[a][1] = iy * ay * (myy0 + ax * myyr) + ix * ax * (txy0 + ay * txys);
  }
}

// y = K x (masked rows) on an nx x ny grid; sx, sy: two LDS node boxes of T.
template <typename T, int BS, int TX>
__device__ void bapply(int nx, int ny, double hx, double hy, const T* __restrict__ Mom, const T* __restrict__ x,
                       T* __restrict__ y, const uint8_t* __restrict__ bc, T* sx, T* sy) {
  constexpr int TY = BS / TX, NBX = TX + 1, NBY = TY + 1, NB = NBX * NBY;
  const int nnx = nx + 1;
  const long long N = (long long)(nx + 1) * (ny + 1), E = (long long)nx * ny;
  const T ix = (T)(0.5 / hx), iy = (T)(0.5 / hy);
  for (long long i = threadIdx.x; i < 2 * N; i += BS) y[i] = 0;
  __syncthreads();
  const int ex = threadIdx.x % TX, ey = threadIdx.x / TX;
  for (int y0 = 0; y0 < ny; y0 += TY)
    for (int x0 = 0; x0 < nx; x0 += TX) {
      const int nbx = min(TX, nx - x0) + 1, nby = min(TY, ny - y0) + 1;
      for (int q = threadIdx.x; q < NB * 2; q += BS) {
        const int c = q & 1, nd = q >> 1, lx = nd % NBX, ly = nd / NBX;
        T v = 0;
        if (lx < nbx && ly < nby) v = x[((long long)(x0 + lx) + (long long)nnx * (y0 + ly)) * 2 + c];
        sx[q] = v;
        sy[q] = 0;
      }
      __syncthreads();
      const bool act = (x0 + ex < nx) && (y0 + ey < ny);
      T ye[4][2];
      if (act) {
        T u[4][2];
#pragma unroll
        for (int a = 0; a < 4; a++) {
          const int ln = (ex + (a & 1)) + NBX * (ey + ((a >> 1) & 1));
          u[a][0] = sx[2 * ln];
          u[a][1] = sx[2 * ln + 1];
        }
        const long long e = (x0 + ex) + (long long)nx * (y0 + ey);
        T M[9];
#pragma unroll
        for (int k = 0; k < 9; k++) M[k] = Mom[k * E + e];
        forces2(u, ix, iy, M, ye);
      }
#pragma unroll
      for (int a = 0; a < 4; a++) {  // corner phases: element -> node e + a is injective
        if (act) {
          const int ln = (ex + (a & 1)) + NBX * (ey + ((a >> 1) & 1));
          sy[2 * ln] += ye[a][0];
          sy[2 * ln + 1] += ye[a][1];
        }
        __syncthreads();
      }
      for (int q = threadIdx.x; q < NB * 2; q += BS) {
        const int c = q & 1, nd = q >> 1, lx = nd % NBX, ly = nd / NBX;
        if (lx >= nbx || ly >= nby) continue;
        const long long gn = (long long)(x0 + lx) + (long long)nnx * (y0 + ly);
        if (!((bc[gn] >> c) & 1u)) y[gn * 2 + c] += sy[q];
      }
      __syncthreads();
    }
}

// y = G p (masked), node gather with the closed-form element gradient.
template <int BS>
__device__ void bapply_G(const BInst& I, const double* __restrict__ p, double* __restrict__ y,
                         const uint8_t* __restrict__ bc) {
  const int nx = I.nx, ny = I.ny, nnx = nx + 1;
  const long long N = (long long)(nx + 1) * (ny + 1);
  for (long long n = threadIdx.x; n < N; n += BS) {
    const int i = (int)(n % nnx), j = (int)(n / nnx);
    double a0 = 0.0, a1 = 0.0;
    for (int jj = j - 1; jj <= j; jj++) {
      if (jj < 0 || jj >= ny) continue;
      const double sgy = (jj == j - 1) ? 1.0 : -1.0;
      for (int ii = i - 1; ii <= i; ii++) {
        if (ii < 0 || ii >= nx) continue;
        const double sgx = (ii == i - 1) ? 1.0 : -1.0;
        const double pe = p[ii + (long long)nx * jj];
        a0 -= sgx * (I.hy * 0.5) * pe;
        a1 -= sgy * (I.hx * 0.5) * pe;
      }
    }
    y[2 * n] = (bc[n] & 1u) ? 0.0 : a0;
    y[2 * n + 1] = (bc[n] & 2u) ? 0.0 : a1;
  }
  __syncthreads();
}

// q = G^T x (divergence per element)
template <int BS>
__device__ void bapply_Gt(const BInst& I, const double* __restrict__ x, double* __restrict__ q) {
  const int nx = I.nx, nnx = nx + 1;
  const long long E = (long long)nx * I.ny;
  for (long long e = threadIdx.x; e < E; e += BS) {
    const long long i = e % nx, j = e / nx;
    const long long n0 = i + (long long)nnx * j;
    const double* u0 = x + n0 * 2;
    const double* u1 = x + (n0 + 1) * 2;
    const double* u2 = x + (n0 + nnx) * 2;
    const double* u3 = x + (n0 + nnx + 1) * 2;
    const double fx = (u1[0] - u0[0]) + (u3[0] - u2[0]);
    const double fy = (u2[1] - u0[1]) + (u3[1] - u1[1]);
    q[e] = -(I.hy * 0.5) * fx - (I.hx * 0.5) * fy;
  }
  __syncthreads();
}

template <int BS>
__device__ double bdot(const double* a, const double* b, long long n, double* sh) {
  double v = 0.0;
  for (long long i = threadIdx.x; i < n; i += BS) v += a[i] * b[i];
  return block_sum<BS>(v, sh);
}

struct Inst {
  BInst I;
  long long N2, E;
  const double *M, *f, *dinv, *Qinv;
  const uint8_t* bc;
  double *u, *vstar, *ft, *cr, *cz, *cp, *cAp, *p, *r, *s, *q, *z;
  double *fs, *fq;
};

// Chebyshev-Jacobi smoothing on one level (FP32), window [0.1, 1.1] lmax,
// as the single-problem smoother: zero = start from x = 0.
template <int BS, int TX>
__device__ void bsmooth(const BArgs& A, const BLevel& L, int nu, bool zero, float* sx, float* sy) {
  const long long n = 2LL * (L.nx + 1) * (L.ny + 1);
  const float* M = A.LM + 9 * L.elem_off;
  const float* dinv = A.Ldinv + 2 * L.node_off;
  const uint8_t* bc = A.Lbc + L.node_off;
  float *b = A.Lb + 2 * L.node_off, *x = A.Lx + 2 * L.node_off, *d = A.Ld + 2 * L.node_off,
        *Ax = A.LAx + 2 * L.node_off;
  const double emax = 1.1 * L.lmax, emin = 0.1 * L.lmax;
  const double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin), sigma = theta / delta;
  double rho_prev = 1.0 / sigma;
  const float b0 = (float)(1.0 / theta);
  if (!zero) bapply<float, BS, TX>(L.nx, L.ny, L.hx, L.hy, M, x, Ax, bc, sx, sy);
  for (long long i = threadIdx.x; i < n; i += BS) {
    const float r = zero ? b[i] : b[i] - Ax[i];
    const float di = b0 * dinv[i] * r;
    d[i] = di;
    x[i] = zero ? di : x[i] + di;
  }
  __syncthreads();
  for (int k = 1; k < nu; k++) {
    const double rho = 1.0 / (2.0 * sigma - rho_prev);
    const float c1 = (float)(rho * rho_prev), c2 = (float)(2.0 * rho / delta);
    bapply<float, BS, TX>(L.nx, L.ny, L.hx, L.hy, M, x, Ax, bc, sx, sy);
    for (long long i = threadIdx.x; i < n; i += BS) {
      const float di = c1 * d[i] + c2 * dinv[i] * (b[i] - Ax[i]);
      d[i] = di;
      x[i] += di;
    }
    __syncthreads();
    rho_prev = rho;
  }
}

// One V-cycle on the instance's levels: level-0 b in, level-0 x out (FP32).
template <int BS, int TX>
__device__ void bvcycle(const BArgs& A, const BInst& I, int nu, float* sx, float* sy) {
  const int Lc = I.nlev - 1;
  for (int l = 0; l < Lc; l++) {
    const BLevel F = A.lev[I.lev0 + l], C = A.lev[I.lev0 + l + 1];
    bsmooth<BS, TX>(A, F, nu, true, sx, sy);
    const long long nf = 2LL * (F.nx + 1) * (F.ny + 1);
    float *b = A.Lb + 2 * F.node_off, *x = A.Lx + 2 * F.node_off, *r = A.Lr + 2 * F.node_off,
          *Ax = A.LAx + 2 * F.node_off;
    bapply<float, BS, TX>(F.nx, F.ny, F.hx, F.hy, A.LM + 9 * F.elem_off, x, Ax, A.Lbc + F.node_off, sx, sy);
    for (long long i = threadIdx.x; i < nf; i += BS) r[i] = b[i] - Ax[i];
    __syncthreads();
    // full-weighting restriction to the coarse level (masked)
    const int fnx = F.nx + 1, fny = F.ny + 1, cnx = C.nx + 1;
    const long long NC = (long long)(C.nx + 1) * (C.ny + 1);
    float* bcv = A.Lb + 2 * C.node_off;
    const uint8_t* cbc = A.Lbc + C.node_off;
    for (long long n = threadIdx.x; n < NC; n += BS) {
      const int Ic = (int)(n % cnx), Jc = (int)(n / cnx);
      float a0 = 0.f, a1 = 0.f;
      for (int dy = -1; dy <= 1; dy++) {
        const int j = 2 * Jc + dy;
        if (j < 0 || j >= fny) continue;
        const float wy = dy == 0 ? 1.f : 0.5f;
        for (int dx = -1; dx <= 1; dx++) {
          const int i = 2 * Ic + dx;
          if (i < 0 || i >= fnx) continue;
          const float w = wy * (dx == 0 ? 1.f : 0.5f);
          const long long f = i + (long long)fnx * j;
          a0 += w * r[2 * f];
          a1 += w * r[2 * f + 1];
        }
      }
      bcv[2 * n] = (cbc[n] & 1u) ? 0.f : a0;
      bcv[2 * n + 1] = (cbc[n] & 2u) ? 0.f : a1;
    }
    __syncthreads();
  }
  {
    const BLevel C = A.lev[I.lev0 + Lc];
    if (I.nfc > 0) {
      const long long nc = 2LL * (C.nx + 1) * (C.ny + 1);
      float *b = A.Lb + 2 * C.node_off, *x = A.Lx + 2 * C.node_off;
      for (long long i = threadIdx.x; i < nc; i += BS) x[i] = 0.f;
      __syncthreads();
      const float* K = A.Kinv + I.kinv_off;
      const int* fi = A.freeidx + I.free_off;
      for (int row = threadIdx.x; row < I.nfc; row += BS) {
        float acc = 0.f;
        for (int j = 0; j < I.nfc; j++) acc += K[(long long)row * I.nfc + j] * b[fi[j]];
        x[fi[row]] = acc;
      }
      __syncthreads();
    } else {
      bsmooth<BS, TX>(A, C, 8, true, sx, sy);
    }
  }
  for (int l = Lc - 1; l >= 0; l--) {
    const BLevel F = A.lev[I.lev0 + l], C = A.lev[I.lev0 + l + 1];
    // bilinear prolongation added to the fine iterate (masked)
    const int fnx = F.nx + 1, cnx = C.nx + 1;
    const long long NF = (long long)(F.nx + 1) * (F.ny + 1);
    const float* xc = A.Lx + 2 * C.node_off;
    float* xf = A.Lx + 2 * F.node_off;
    const uint8_t* fbc = A.Lbc + F.node_off;
    for (long long n = threadIdx.x; n < NF; n += BS) {
      const int i = (int)(n % fnx), j = (int)(n / fnx);
      const int ny_ = (j & 1) ? 2 : 1, nx_ = (i & 1) ? 2 : 1;
      const float wy = ny_ == 2 ? 0.5f : 1.f, wx = nx_ == 2 ? 0.5f : 1.f;
      float a0 = 0.f, a1 = 0.f;
      for (int bb = 0; bb < ny_; bb++)
        for (int cc = 0; cc < nx_; cc++) {
          const long long cn = ((i >> 1) + cc) + (long long)cnx * ((j >> 1) + bb);
          a0 += wx * wy * xc[2 * cn];
          a1 += wx * wy * xc[2 * cn + 1];
        }
      if (!(fbc[n] & 1u)) xf[2 * n] += a0;
      if (!(fbc[n] & 2u)) xf[2 * n + 1] += a1;
    }
    __syncthreads();
    bsmooth<BS, TX>(A, F, nu, false, sx, sy);
  }
}

// z = V-cycle(r): FP64 -> FP32 level-0 right-hand side, FP32 -> FP64 result.
template <int BS, int TX>
__device__ void bprec(const BArgs& A, const Inst& W, const double* r, double* z, float* sx, float* sy) {
  const BLevel L0 = A.lev[W.I.lev0];
  float* b = A.Lb + 2 * L0.node_off;
  float* x = A.Lx + 2 * L0.node_off;
  for (long long i = threadIdx.x; i < W.N2; i += BS) b[i] = (float)r[i];
  __syncthreads();
  bvcycle<BS, TX>(A, W.I, A.nu, sx, sy);
  for (long long i = threadIdx.x; i < W.N2; i += BS) z[i] = (double)x[i];
  __syncthreads();
}

// x = K^-1 b (flexible CG, FP64), stopping at |r| <= rtol |b|.  Returns 0 or
// GS_E_INNER_MAXIT / GS_E_NAN.
template <int BS, int TX>
__device__ int inner_solve(const BArgs& A, const Inst& W, const double* b, double* x, double rtol, long long* its,
                           float* lds32, double* lds64, double* sh) {
  constexpr int NB2 = 2 * (TX + 1) * (BS / TX + 1);
  float* sx = lds32;
  float* sy = lds32 + NB2;
  const double bb = bdot<BS>(b, b, W.N2, sh);
  for (long long i = threadIdx.x; i < W.N2; i += BS) {
    x[i] = 0.0;
    W.cr[i] = b[i];
  }
  __syncthreads();
  if (!(bb > 0.0)) return isfinite(bb) ? 0 : GS_E_NAN;
  bprec<BS, TX>(A, W, W.cr, W.cz, sx, sy);
  for (long long i = threadIdx.x; i < W.N2; i += BS) W.cp[i] = W.cz[i];
  __syncthreads();
  double rz = bdot<BS>(W.cr, W.cz, W.N2, sh);
  const double stop = rtol * rtol * bb;
  for (int k = 0; k < A.max_inner; k++) {
    bapply<double, BS, TX>(W.I.nx, W.I.ny, W.I.hx, W.I.hy, W.M, W.cp, W.cAp, W.bc, lds64, lds64 + NB2);
    const double pap = bdot<BS>(W.cp, W.cAp, W.N2, sh);
    const double alpha = (pap != 0.0 && isfinite(pap)) ? rz / pap : 0.0;
    double rr = 0.0;
    for (long long i = threadIdx.x; i < W.N2; i += BS) {
      x[i] += alpha * W.cp[i];
      const double ri = W.cr[i] - alpha * W.cAp[i];
      W.cr[i] = ri;
      rr += ri * ri;
    }
    rr = block_sum<BS>(rr, sh);
    (*its)++;
    if (!isfinite(rr)) return GS_E_NAN;
    if (rr <= stop) return 0;
    bprec<BS, TX>(A, W, W.cr, W.cz, sx, sy);
    double zr = 0.0, zap = 0.0;
    for (long long i = threadIdx.x; i < W.N2; i += BS) {
      zr += W.cz[i] * W.cr[i];
      zap += W.cz[i] * W.cAp[i];
    }
    zr = block_sum<BS>(zr, sh);
    zap = block_sum<BS>(zap, sh);
    const double beta = rz != 0.0 ? -alpha * zap / rz : 0.0;
    rz = zr;
    for (long long i = threadIdx.x; i < W.N2; i += BS) W.cp[i] = W.cz[i] + beta * W.cp[i];
    __syncthreads();
  }
  return GS_E_INNER_MAXIT;
}

// Stokes_SLE_UzawaSolver.cxx:553-837 for one instance.
template <int BS, int TX>
__device__ int uzawa_one(const Inst& W, const BArgs& A, int* outer_out, long long* inner_out, double* relres_out,
                         float* lds32, double* lds64, double* sh) {
  const long long N2 = W.N2, E = W.E;
  long long inner = 0;
  const double qRecip = 1.0 / sqrt((double)E);
  // zero-force firewall (:476-493)
  double nfree = 0.0;
  for (long long i = threadIdx.x; i < N2; i += BS) nfree += W.dinv[i] != 0.0 ? 1.0 : 0.0;
  nfree = block_sum<BS>(nfree, sh);
  const double ff = bdot<BS>(W.f, W.f, N2, sh);
  *outer_out = 0;
  *inner_out = 0;
  *relres_out = 0.0;
  if (sqrt(ff) / sqrt(nfree > 0 ? nfree : 1.0) <= 1e-99) return GS_E_ZERO_FORCE;
  int st;
  // STEP 1: vstar = K^-1 f; z = G^T vstar - h (h = 0, homogeneous conditions)
  if ((st = inner_solve<BS, TX>(A, W, W.f, W.vstar, A.tol, &inner, lds32, lds64, sh))) return st;
  bapply_Gt<BS>(W.I, W.vstar, W.z);
  const double scale = sqrt(bdot<BS>(W.z, W.z, E, sh)) * qRecip;
  if (!isfinite(scale)) return GS_E_NAN;
  // STEP 3: u = K^-1 (f - G p), p = 0; r = G^T u - h
  for (long long e = threadIdx.x; e < E; e += BS) W.p[e] = 0.0;
  __syncthreads();
  if ((st = inner_solve<BS, TX>(A, W, W.f, W.u, A.irt, &inner, lds32, lds64, sh))) return st;
  bapply_Gt<BS>(W.I, W.u, W.r);
  double zr_prev = 1.0, relRes = 0.0;
  int it = 0;
  int err = 0;
  do {
    double zr = 0.0;
    for (long long e = threadIdx.x; e < E; e += BS) {
      const double zi = W.Qinv[e] * W.r[e];
      W.z[e] = zi;
      zr += zi * W.r[e];
    }
    zr = block_sum<BS>(zr, sh);
    const double beta = it == 0 ? 0.0 : (zr_prev != 0.0 ? zr / zr_prev : 0.0);
    for (long long e = threadIdx.x; e < E; e += BS) W.s[e] = it == 0 ? W.z[e] : W.z[e] + beta * W.s[e];
    __syncthreads();
    bapply_G<BS>(W.I, W.s, W.ft, W.bc);
    if ((err = inner_solve<BS, TX>(A, W, W.ft, W.vstar, A.irt, &inner, lds32, lds64, sh))) break;
    bapply_Gt<BS>(W.I, W.vstar, W.q);
    const double sq = bdot<BS>(W.s, W.q, E, sh);
    const double alpha = zr / sq;
    if (!isfinite(zr) || !isfinite(sq) || !isfinite(alpha)) {
      err = GS_E_NAN;
      break;
    }
    double rr = 0.0;
    for (long long e = threadIdx.x; e < E; e += BS) {
      W.p[e] += alpha * W.s[e];
      const double ri = W.r[e] - alpha * W.q[e];
      W.r[e] = ri;
      rr += ri * ri;
    }
    for (long long i = threadIdx.x; i < N2; i += BS) W.u[i] -= alpha * W.vstar[i];
    rr = block_sum<BS>(rr, sh);
    zr_prev = zr;
    relRes = sqrt(rr) * qRecip / scale;
    if (!isfinite(relRes)) {
      err = GS_E_NAN;
      break;
    }
    if (!(it < A.max_outer)) {
      err = GS_E_MAXIT;
      break;
    }
    it++;
  } while (relRes > A.tol);
  *outer_out = it;
  *inner_out = inner;
  *relres_out = relRes;
  return err;
}

// The same solve with inexact inner solves: the inner tolerance is relaxed as
// the outer residual falls, rtol_k = min(relax, max(irt, irt |r_0| / |r_k|))
// (the inexact-Krylov relaxation rule), and the outer iteration is flexible CG
// on S = G^T K^-1 G with each new direction S-orthogonalised against the last
// `window` ones, which keeps it converging with inexact applications of S.
// When the recursive residual reaches the tolerance, the true residual is
// recomputed from an accurate inner solve u = K^-1 (f - G p) (the reference's
// STEP 3) and the loop resumes from it if needed; a round that stagnates
// (less than 5 % residual drop over 50 iterations) also ends in that check.
// Each new round relaxes ten times less and the fifth not at all (then the
// loop is flexible CG with the reference inner tolerance).  The reported
// residual is the true one.  Same firewalls as uzawa_one.
template <int BS, int TX>
__device__ int uzawa_relaxed(const Inst& W, const BArgs& A, int* outer_out, long long* inner_out, double* relres_out,
                             float* lds32, double* lds64, double* sh, double* sqv) {
  const long long N2 = W.N2, E = W.E;
  long long inner = 0;
  const double qRecip = 1.0 / sqrt((double)E);
  double nfree = 0.0;
  for (long long i = threadIdx.x; i < N2; i += BS) nfree += W.dinv[i] != 0.0 ? 1.0 : 0.0;
  nfree = block_sum<BS>(nfree, sh);
  const double ff = bdot<BS>(W.f, W.f, N2, sh);
  *outer_out = 0;
  *inner_out = 0;
  *relres_out = 0.0;
  if (sqrt(ff) / sqrt(nfree > 0 ? nfree : 1.0) <= 1e-99) return GS_E_ZERO_FORCE;
  int st;
  if ((st = inner_solve<BS, TX>(A, W, W.f, W.vstar, A.tol, &inner, lds32, lds64, sh))) return st;
  bapply_Gt<BS>(W.I, W.vstar, W.z);
  const double scale = sqrt(bdot<BS>(W.z, W.z, E, sh)) * qRecip;
  if (!isfinite(scale)) return GS_E_NAN;
  for (long long e = threadIdx.x; e < E; e += BS) W.p[e] = 0.0;
  __syncthreads();
  if ((st = inner_solve<BS, TX>(A, W, W.f, W.u, A.irt, &inner, lds32, lds64, sh))) return st;
  bapply_Gt<BS>(W.I, W.u, W.r);
  double relRes = sqrt(bdot<BS>(W.r, W.r, E, sh)) * qRecip / scale;
  const int m = A.window;
  int it = 0, err = 0;
  constexpr int kRounds = 5, kStagWin = 50;
  for (int round = 0; round < kRounds; round++) {
    // each round relaxes ten times less; the last one not at all
    const double relax = round == kRounds - 1 ? 0.0 : A.relax * pow(0.1, (double)round);
    const double relax_ref = relRes;
    const int base = it;
    double stag_ref = relRes;  // residual at the start of the current stagnation window
    int stag_it = it;
    bool conv = relRes <= A.tol && it >= 1;
    while (!conv) {
      for (long long e = threadIdx.x; e < E; e += BS) W.z[e] = W.Qinv[e] * W.r[e];
      __syncthreads();
      const int k = it - base, c = k < m ? k : m, slot = k % m;
      double* sk = W.fs + (long long)slot * E;
      double* qk = W.fq + (long long)slot * E;
      double beta[kBatchMaxWin];
#pragma unroll
      for (int j = 0; j < kBatchMaxWin; j++) {
        beta[j] = 0.0;
        if (j < c) beta[j] = bdot<BS>(W.z, W.fq + (long long)j * E, E, sh) / sqv[j];
      }
      for (long long e = threadIdx.x; e < E; e += BS) {
        double v = W.z[e];
#pragma unroll
        for (int j = 0; j < kBatchMaxWin; j++)
          if (j < c) v -= beta[j] * W.fs[(long long)j * E + e];
        sk[e] = v;
      }
      __syncthreads();
      bapply_G<BS>(W.I, sk, W.ft, W.bc);
      double irt_k = A.irt;
      if (relax > 0 && relRes > 0) irt_k = fmin(relax, fmax(A.irt, A.irt * relax_ref / relRes));
      if ((err = inner_solve<BS, TX>(A, W, W.ft, W.vstar, irt_k, &inner, lds32, lds64, sh))) break;
      bapply_Gt<BS>(W.I, W.vstar, qk);
      const double sq = bdot<BS>(sk, qk, E, sh);
      const double sr = bdot<BS>(sk, W.r, E, sh);
      if (threadIdx.x == 0) sqv[slot] = sq;
      const double alpha = sr / sq;
      if (!isfinite(sq) || !isfinite(alpha)) {
        err = GS_E_NAN;
        break;
      }
      double rr = 0.0;
      for (long long e = threadIdx.x; e < E; e += BS) {
        W.p[e] += alpha * sk[e];
        const double ri = W.r[e] - alpha * qk[e];
        W.r[e] = ri;
        rr += ri * ri;
      }
      for (long long i = threadIdx.x; i < N2; i += BS) W.u[i] -= alpha * W.vstar[i];
      rr = block_sum<BS>(rr, sh);
      relRes = sqrt(rr) * qRecip / scale;
      if (!isfinite(relRes)) {
        err = GS_E_NAN;
        break;
      }
      if (!(it < A.max_outer)) {
        err = GS_E_MAXIT;
        break;
      }
      it++;
      conv = relRes <= A.tol;
      // stagnation (less than 5 % drop over kStagWin iterations) while still
      // relaxing: end the round, re-check the true residual, relax less
      if (!conv && relax > 0 && it - stag_it >= kStagWin) {
        if (relRes > 0.95 * stag_ref) break;
        stag_ref = relRes;
        stag_it = it;
      }
    }
    if (err) break;
    // true residual: u = K^-1 (f - G p) with the reference inner tolerance
    bapply_G<BS>(W.I, W.p, W.ft, W.bc);
    for (long long i = threadIdx.x; i < N2; i += BS) W.ft[i] = W.f[i] - W.ft[i];
    __syncthreads();
    if ((err = inner_solve<BS, TX>(A, W, W.ft, W.u, A.irt, &inner, lds32, lds64, sh))) break;
    bapply_Gt<BS>(W.I, W.u, W.r);
    relRes = sqrt(bdot<BS>(W.r, W.r, E, sh)) * qRecip / scale;
    if (!isfinite(relRes)) {
      err = GS_E_NAN;
      break;
    }
    if (relRes <= A.tol) break;
    if (round == kRounds - 1) err = GS_E_MAXIT;
  }
  *outer_out = it;
  *inner_out = inner;
  *relres_out = relRes;
  return err;
}

template <int BS, int TX>
__global__ __launch_bounds__(BS) void k_batch_uzawa(BArgs A) {
  constexpr int TY = BS / TX;
  constexpr int NB2 = 2 * (TX + 1) * (TY + 1);
  __shared__ double lds64[2 * NB2];
  __shared__ float lds32[2 * NB2];
  __shared__ double sh[BS / 64];
  __shared__ double sqv[kBatchMaxWin];
  __shared__ int cur;
  for (;;) {
    if (threadIdx.x == 0) {
      const unsigned k = atomicAdd(A.queue, 1u);
      cur = k < (unsigned)A.B ? A.order[k] : -1;
    }
    __syncthreads();
    const int i = cur;
    __syncthreads();
    if (i < 0) return;
    Inst W;
    W.I = A.inst[i];
    const long long no = W.I.node_off, eo = W.I.elem_off;
    W.N2 = 2LL * (W.I.nx + 1) * (W.I.ny + 1);
    W.E = (long long)W.I.nx * W.I.ny;
    W.M = A.M + 9 * eo;
    W.f = A.f + 2 * no;
    W.dinv = A.dinv + 2 * no;
    W.Qinv = A.Qinv + eo;
    W.bc = A.bc + no;
    W.u = A.u + 2 * no;
    W.vstar = A.vstar + 2 * no;
    W.ft = A.ft + 2 * no;
    W.cr = A.cr + 2 * no;
    W.cz = A.cz + 2 * no;
    W.cp = A.cp + 2 * no;
    W.cAp = A.cAp + 2 * no;
    W.p = A.p + eo;
    W.r = A.r + eo;
    W.s = A.s + eo;
    W.q = A.q + eo;
    W.z = A.z + eo;
    W.fs = A.fs ? A.fs + (long long)A.window * eo : nullptr;
    W.fq = A.fq ? A.fq + (long long)A.window * eo : nullptr;
    int outer = 0;
    long long inner = 0;
    double rel = 0.0;
    const unsigned long long t0 = wall_clock64();
    const int st = A.relax > 0 ? uzawa_relaxed<BS, TX>(W, A, &outer, &inner, &rel, lds32, lds64, sh, sqv)
                               : uzawa_one<BS, TX>(W, A, &outer, &inner, &rel, lds32, lds64, sh);
    if (threadIdx.x == 0) {
      A.itime[i] = wall_clock64() - t0;
      A.status[i] = st;
      A.outer[i] = outer;
      A.inner[i] = inner;
      A.relres[i] = rel;
    }
    __syncthreads();
  }
}

__global__ void k_zero_queue(unsigned* q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) q[0] = 0u;
}

struct Tile {
  int bs, tx;
};
const Tile kTiles[6] = {{256, 16}, {256, 32}, {512, 32}, {512, 64}, {1024, 32}, {1024, 64}};

}  // namespace

struct gs_batch {
  int B = 0;
  hipStream_t st = nullptr;
  std::vector<BInst> hinst;
  std::vector<BLevel> hlev;
  BInst* inst = nullptr;
  BLevel* lev = nullptr;
  int* order = nullptr;
  long long nodes = 0, elems = 0, lnodes = 0, lelems = 0, kinv_n = 0, free_n = 0;
  double *M = nullptr, *f = nullptr, *dinv = nullptr, *Qinv = nullptr;
  uint8_t* bc = nullptr;
  float *LM = nullptr, *Ldinv = nullptr, *Lb = nullptr, *Lx = nullptr, *Lr = nullptr, *Ld = nullptr, *LAx = nullptr;
  uint8_t* Lbc = nullptr;
  float* Kinv = nullptr;
  int* freeidx = nullptr;
  double *u = nullptr, *vstar = nullptr, *ft = nullptr, *cr = nullptr, *cz = nullptr, *cp = nullptr, *cAp = nullptr;
  double *p = nullptr, *r = nullptr, *s = nullptr, *q = nullptr, *z = nullptr;
  int *status = nullptr, *outer = nullptr;
  unsigned long long* itime = nullptr;
  double clock_khz = 100000.0;  // wall_clock64 rate
  long long* inner = nullptr;
  double* relres = nullptr;
  double *fs = nullptr, *fq = nullptr;  // relaxed loop direction rings (window x elements)
  int fwin = 0;
  unsigned* queue = nullptr;
  int tile = 4;  // index into kTiles
  int ncu = 256;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  gs_batch_params gparams{};
  int gtile = -1;
};

template <typename T>
static hipError_t balloc(T** p, long long n) {
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n > 0 ? n : 1));
}

static void free_graph(gs_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->gtile = -1;
}

extern "C" void gs_batch_destroy(gs_batch* b) {
  if (!b) return;
  if (b->st) (void)hipStreamSynchronize(b->st);
  free_graph(b);
  void* ptrs[] = {b->inst, b->lev, b->order, b->M, b->f, b->dinv, b->Qinv, b->bc, b->LM, b->Ldinv, b->Lb, b->Lx,
                  b->Lr, b->Ld, b->LAx, b->Lbc, b->Kinv, b->freeidx, b->u, b->vstar, b->ft, b->cr, b->cz, b->cp,
                  b->cAp, b->p, b->r, b->s, b->q, b->z, b->status, b->outer, b->inner, b->relres, b->queue,
                  b->fs, b->fq, b->itime};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

extern "C" int32_t gs_batch_create(int32_t count, const gs_problem_desc* descs, gs_batch** out) {
  if (!out || count < 1 || !descs) return GS_E_INVALID;
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return GS_E_NO_DEVICE;
  for (int i = 0; i < count; i++)
    if (descs[i].dim != 2 || descs[i].res[0] < 1 || descs[i].res[1] < 1) return GS_E_INVALID;
  gs_batch* b = new gs_batch();
  b->B = count;
  std::vector<gs_problem*> pbs(count, nullptr);
  std::vector<gs_solver*> svs(count, nullptr);
  auto fail = [&](int32_t s) {
    for (int i = 0; i < count; i++) {
      if (svs[i]) gs_solver_destroy(svs[i]);
      if (pbs[i]) gs_problem_destroy(pbs[i]);
    }
    gs_batch_destroy(b);
    return s;
  };
  if (hipStreamCreate(&b->st) != hipSuccess || hipEventCreate(&b->e0) != hipSuccess ||
      hipEventCreate(&b->e1) != hipSuccess)
    return fail(GS_E_HIP);
  {
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
      b->ncu = prop.multiProcessorCount;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
      b->clock_khz = khz;
  }
  // per-instance setup through the single-problem path: particle location,
  // constitutive pass, right-hand side, multigrid levels (Galerkin moments,
  // Chebyshev bounds, dense coarsest inverse)
  gs_uzawa_params up{};
  up.tolerance = 1e-6;
  up.max_iterations = 1;
  up.preconditioner = GS_PC_DIAGONAL;
  up.precision = GS_PREC_MIXED;
  up.inner_preconditioner = GS_INNER_MULTIGRID;
  constexpr int kMaxLev = 16;
  std::vector<std::vector<BatchLevelInfo>> info(count, std::vector<BatchLevelInfo>(kMaxLev));
  b->hinst.resize(count);
  long long no = 0, eo = 0, lno = 0, leo = 0, ko = 0, fo = 0;
  for (int i = 0; i < count; i++) {
    const gs_problem_desc& d = descs[i];
    int32_t s = gs_problem_create(&d, &pbs[i]);
    if (s != GS_OK) return fail(s);
    if ((s = gs_solver_create(pbs[i], &up, &svs[i])) != GS_OK) return fail(s);
    if ((s = gs_solver_setup(svs[i], nullptr, nullptr, 0)) != GS_OK) return fail(s);
    int nl = 0, nfc = 0;
    if ((s = batch_mg_info(svs[i], &nl, info[i].data(), kMaxLev, &nfc)) != GS_OK) return fail(s);
    BInst& I = b->hinst[i];
    I.nx = d.res[0];
    I.ny = d.res[1];
    I.node_off = no;
    I.elem_off = eo;
    I.hx = (d.hi[0] - d.lo[0]) / d.res[0];
    I.hy = (d.hi[1] - d.lo[1]) / d.res[1];
    I.lev0 = (int)b->hlev.size();
    I.nlev = nl;
    I.nfc = nfc;
    I.kinv_off = ko;
    I.free_off = fo;
    if (nfc > 0) {
      ko += (long long)nfc * nfc;
      fo += nfc;
    }
    no += (long long)(I.nx + 1) * (I.ny + 1);
    eo += (long long)I.nx * I.ny;
    for (int l = 0; l < nl; l++) {
      BLevel L{};
      L.nx = info[i][l].nx;
      L.ny = info[i][l].ny;
      L.node_off = lno;
      L.elem_off = leo;
      L.hx = (d.hi[0] - d.lo[0]) / L.nx;
      L.hy = (d.hi[1] - d.lo[1]) / L.ny;
      L.lmax = info[i][l].lmax;
      b->hlev.push_back(L);
      lno += (long long)(L.nx + 1) * (L.ny + 1);
      leo += (long long)L.nx * L.ny;
    }
  }
  b->nodes = no;
  b->elems = eo;
  b->lnodes = lno;
  b->lelems = leo;
  b->kinv_n = ko;
  b->free_n = fo;
  const long long N2 = 2 * no, L2 = 2 * lno;
  if (balloc(&b->inst, count) || balloc(&b->lev, (long long)b->hlev.size()) || balloc(&b->order, count) ||
      balloc(&b->M, 9 * eo) || balloc(&b->f, N2) || balloc(&b->dinv, N2) || balloc(&b->Qinv, eo) ||
      balloc(&b->bc, no) || balloc(&b->LM, 9 * leo) || balloc(&b->Ldinv, L2) || balloc(&b->Lb, L2) ||
      balloc(&b->Lx, L2) || balloc(&b->Lr, L2) || balloc(&b->Ld, L2) || balloc(&b->LAx, L2) ||
      balloc(&b->Lbc, lno) || balloc(&b->Kinv, ko) || balloc(&b->freeidx, fo) || balloc(&b->u, N2) ||
      balloc(&b->vstar, N2) || balloc(&b->ft, N2) || balloc(&b->cr, N2) || balloc(&b->cz, N2) ||
      balloc(&b->cp, N2) || balloc(&b->cAp, N2) || balloc(&b->p, eo) || balloc(&b->r, eo) || balloc(&b->s, eo) ||
      balloc(&b->q, eo) || balloc(&b->z, eo) || balloc(&b->status, count) || balloc(&b->outer, count) ||
      balloc(&b->inner, count) || balloc(&b->relres, count) || balloc(&b->queue, 1) || balloc(&b->itime, count))
    return fail(GS_E_HIP);
  for (int i = 0; i < count; i++) {
    const BInst& I = b->hinst[i];
    int32_t s = batch_copy_fine(svs[i], b->f + 2 * I.node_off, b->dinv + 2 * I.node_off, b->Qinv + I.elem_off,
                                b->M + 9 * I.elem_off, b->st);
    for (int l = 0; l < I.nlev && s == GS_OK; l++) {
      const BLevel& L = b->hlev[I.lev0 + l];
      s = batch_copy_level(svs[i], l, b->LM + 9 * L.elem_off, b->Ldinv + 2 * L.node_off, b->Lbc + L.node_off,
                           b->st);
    }
    if (s == GS_OK && I.nfc > 0) s = batch_copy_dense(svs[i], b->Kinv + I.kinv_off, b->freeidx + I.free_off, b->st);
    if (s == GS_OK && hipStreamSynchronize(b->st) != hipSuccess) s = GS_E_HIP;
    if (s != GS_OK) return fail(s);
    // the fine-level node flags equal level 0's
    const BLevel& L0 = b->hlev[I.lev0];
    if (hipMemcpyAsync(b->bc + I.node_off, b->Lbc + L0.node_off, (size_t)(I.nx + 1) * (I.ny + 1),
                       hipMemcpyDeviceToDevice, b->st) != hipSuccess)
      return fail(GS_E_HIP);
    gs_solver_destroy(svs[i]);
    svs[i] = nullptr;
    gs_problem_destroy(pbs[i]);
    pbs[i] = nullptr;
  }
  std::vector<int> ord(count);
  for (int i = 0; i < count; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) {
    return (long long)b->hinst[a].nx * b->hinst[a].ny > (long long)b->hinst[c].nx * b->hinst[c].ny;
  });
  if (hipMemcpyAsync(b->inst, b->hinst.data(), sizeof(BInst) * count, hipMemcpyHostToDevice, b->st) ||
      hipMemcpyAsync(b->lev, b->hlev.data(), sizeof(BLevel) * b->hlev.size(), hipMemcpyHostToDevice, b->st) ||
      hipMemcpyAsync(b->order, ord.data(), sizeof(int) * count, hipMemcpyHostToDevice, b->st) ||
      hipStreamSynchronize(b->st))
    return fail(GS_E_HIP);
  *out = b;
  return GS_OK;
}

extern "C" int32_t gs_batch_set_tile(gs_batch* b, int32_t block, int32_t tx) {
  if (!b) return GS_E_INVALID;
  for (int k = 0; k < 6; k++)
    if (kTiles[k].bs == block && kTiles[k].tx == tx) {
      if (b->tile != k) free_graph(b);
      b->tile = k;
      return GS_OK;
    }
  return GS_E_INVALID;
}

static BArgs make_args(gs_batch* b, const gs_batch_params* p) {
  BArgs A{};
  A.inst = b->inst;
  A.lev = b->lev;
  A.LM = b->LM;
  A.Ldinv = b->Ldinv;
  A.Lbc = b->Lbc;
  A.Lb = b->Lb;
  A.Lx = b->Lx;
  A.Lr = b->Lr;
  A.Ld = b->Ld;
  A.LAx = b->LAx;
  A.Kinv = b->Kinv;
  A.freeidx = b->freeidx;
  A.order = b->order;
  A.B = b->B;
  A.M = b->M;
  A.f = b->f;
  A.dinv = b->dinv;
  A.Qinv = b->Qinv;
  A.bc = b->bc;
  A.u = b->u;
  A.vstar = b->vstar;
  A.ft = b->ft;
  A.cr = b->cr;
  A.cz = b->cz;
  A.cp = b->cp;
  A.cAp = b->cAp;
  A.p = b->p;
  A.r = b->r;
  A.s = b->s;
  A.q = b->q;
  A.z = b->z;
  A.status = b->status;
  A.itime = b->itime;
  A.outer = b->outer;
  A.inner = b->inner;
  A.relres = b->relres;
  A.queue = b->queue;
  A.tol = p->tolerance;
  A.irt = (p->inner_rtol_factor > 0 ? p->inner_rtol_factor : 0.1) * p->tolerance;
  A.max_outer = p->max_iterations > 0 ? p->max_iterations : 1000;
  A.max_inner = p->max_inner > 0 ? p->max_inner : 20000;
  A.nu = p->mg_smooth > 0 ? p->mg_smooth : 2;
  A.relax = p->inner_relax_max > 0 ? p->inner_relax_max : 0.0;
  A.window = p->flexible_window > 0 ? std::min(p->flexible_window, kBatchMaxWin) : 8;
  A.fs = b->fs;
  A.fq = b->fq;
  return A;
}

// Direction rings of the relaxed loop (allocated on first use, before any
// graph capture).
static int32_t ensure_rings(gs_batch* b, const gs_batch_params* p) {
  if (!(p->inner_relax_max > 0)) return GS_OK;
  const int m = p->flexible_window > 0 ? std::min(p->flexible_window, kBatchMaxWin) : 8;
  if (b->fs && b->fwin == m) return GS_OK;
  if (b->fs) (void)hipFree(b->fs);
  if (b->fq) (void)hipFree(b->fq);
  b->fs = b->fq = nullptr;
  b->fwin = 0;
  BCK(balloc(&b->fs, (long long)m * b->elems));
  BCK(balloc(&b->fq, (long long)m * b->elems));
  b->fwin = m;
  return GS_OK;
}

static void launch_batch(gs_batch* b, const BArgs& A, int tile, hipStream_t st) {
  // one workgroup per instance, up to 2 per CU; the queue hands out instances
  const unsigned grid = (unsigned)std::min(b->B, 2 * b->ncu);
  hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(64), 0, st, b->queue);
  switch (tile) {
#define CASE(k, BS_, TX_)                                                                   \
  case k:                                                                                   \
    hipLaunchKernelGGL((k_batch_uzawa<BS_, TX_>), dim3(grid), dim3(BS_), 0, st, A);         \
    break;
    CASE(0, 256, 16)
    CASE(1, 256, 32)
    CASE(2, 512, 32)
    CASE(3, 512, 64)
    CASE(4, 1024, 32)
    default:
      hipLaunchKernelGGL((k_batch_uzawa<1024, 64>), dim3(grid), dim3(1024), 0, st, A);
      break;
#undef CASE
  }
}

static bool same_params(const gs_batch_params& a, const gs_batch_params& c) {
  return a.tolerance == c.tolerance && a.max_iterations == c.max_iterations && a.max_inner == c.max_inner &&
         a.inner_rtol_factor == c.inner_rtol_factor && a.mg_smooth == c.mg_smooth &&
         a.inner_relax_max == c.inner_relax_max && a.flexible_window == c.flexible_window;
}

extern "C" int32_t gs_batch_solve(gs_batch* b, const gs_batch_params* p, gs_batch_result* results, double* seconds) {
  if (!b || !p || !(p->tolerance > 0)) return GS_E_INVALID;
  if (p->flexible_window < 0 || p->flexible_window > kBatchMaxWin) return GS_E_INVALID;
  {
    const int32_t s = ensure_rings(b, p);
    if (s != GS_OK) return s;
  }
  const BArgs A = make_args(b, p);
  if (p->use_graph) {
    if (!b->exec || b->gtile != b->tile || !same_params(b->gparams, *p)) {
      free_graph(b);
      BCK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      launch_batch(b, A, b->tile, b->st);
      BCK(hipStreamEndCapture(b->st, &b->graph));
      BCK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->gparams = *p;
      b->gtile = b->tile;
    }
    BCK(hipEventRecord(b->e0, b->st));
    BCK(hipGraphLaunch(b->exec, b->st));
    BCK(hipEventRecord(b->e1, b->st));
  } else {
    BCK(hipEventRecord(b->e0, b->st));
    launch_batch(b, A, b->tile, b->st);
    BCK(hipGetLastError());
    BCK(hipEventRecord(b->e1, b->st));
  }
  BCK(hipEventSynchronize(b->e1));
  float ms = 0.f;
  BCK(hipEventElapsedTime(&ms, b->e0, b->e1));
  if (seconds) *seconds = ms * 1e-3;
  std::vector<int> st(b->B), ou(b->B);
  std::vector<long long> in(b->B);
  std::vector<double> rr(b->B);
  std::vector<unsigned long long> tk(b->B);
  BCK(hipMemcpy(st.data(), b->status, sizeof(int) * b->B, hipMemcpyDeviceToHost));
  BCK(hipMemcpy(tk.data(), b->itime, sizeof(unsigned long long) * b->B, hipMemcpyDeviceToHost));
  BCK(hipMemcpy(ou.data(), b->outer, sizeof(int) * b->B, hipMemcpyDeviceToHost));
  BCK(hipMemcpy(in.data(), b->inner, sizeof(long long) * b->B, hipMemcpyDeviceToHost));
  BCK(hipMemcpy(rr.data(), b->relres, sizeof(double) * b->B, hipMemcpyDeviceToHost));
  int32_t worst = GS_OK;
  for (int i = 0; i < b->B; i++) {
    if (results) results[i] = gs_batch_result{st[i], ou[i], (int64_t)in[i], rr[i], (double)tk[i] / (b->clock_khz * 1e3)};
    if (st[i] != GS_OK && worst == GS_OK) worst = st[i];
  }
  return worst;
}

extern "C" int32_t gs_batch_autotune(gs_batch* b, const gs_batch_params* p, int32_t* block, int32_t* tx, float* ms) {
  if (!b || !p) return GS_E_INVALID;
  if (p->flexible_window < 0 || p->flexible_window > kBatchMaxWin) return GS_E_INVALID;
  {
    const int32_t s = ensure_rings(b, p);
    if (s != GS_OK) return s;
  }
  gs_batch_params q = *p;
  q.use_graph = 0;
  if (q.max_iterations <= 0 || q.max_iterations > 3) q.max_iterations = 3;  // a few Uzawa iterations per instance
  const BArgs A = make_args(b, &q);
  int best = b->tile;
  float best_ms = 1e30f;
  for (int k = 0; k < 6; k++) {
    BCK(hipEventRecord(b->e0, b->st));
    launch_batch(b, A, k, b->st);
    BCK(hipGetLastError());
    BCK(hipEventRecord(b->e1, b->st));
    BCK(hipEventSynchronize(b->e1));
    float t = 0.f;
    BCK(hipEventElapsedTime(&t, b->e0, b->e1));
    if (ms) ms[k] = t;
    if (t < best_ms) {
      best_ms = t;
      best = k;
    }
  }
  if (b->tile != best) free_graph(b);
  b->tile = best;
  if (block) *block = kTiles[best].bs;
  if (tx) *tx = kTiles[best].tx;
  return GS_OK;
}

extern "C" int32_t gs_batch_solution(gs_batch* b, int32_t i, double* u, double* p) {
  if (!b || i < 0 || i >= b->B) return GS_E_INVALID;
  const BInst& I = b->hinst[i];
  BCK(hipStreamSynchronize(b->st));
  if (u)
    BCK(hipMemcpy(u, b->u + 2 * I.node_off, sizeof(double) * 2 * (I.nx + 1) * (I.ny + 1), hipMemcpyDeviceToHost));
  if (p) BCK(hipMemcpy(p, b->p + I.elem_off, sizeof(double) * I.nx * I.ny, hipMemcpyDeviceToHost));
  return GS_OK;
}
