// gs_mg.hip -- geometric multigrid for the inner velocity solve.
//
// Because the Q1 spaces of a regular mesh are nested and the coarse basis
// functions are exactly representable on the fine mesh, the Galerkin coarse
// operator P^T K P equals the coarse Q1 operator integrated with the SAME
// material points and weights omega_p = w_p detJ eta_p.  Its element matrix
// depends on the points only through 27 (3D) / 9 (2D) weighted moments
// M[a,b,c] = sum_p omega_p r^a s^b t^c (a,b,c <= 2) in element-local
// coordinates, since every entry of B^T D B is a polynomial of degree <= 2 in
// each local coordinate.  The fine moments come from one pass over the
// particles; each coarser level's moments are an exact polynomial change of
// variables of its children's.  Every level is applied matrix-free from its
// moments (216 B per 3D element instead of 32 B per particle), smoothed by
// Chebyshev-Jacobi, and the coarsest level is solved with a dense inverse.
#include "gs_internal.h"
#include "gs_mg.h"
#include "gs_tile.h"

#include <math.h>

namespace gs {

__device__ __forceinline__ long long xcd_block_mg(long long b, long long nb) {
  const long long q = nb / 8, rem = nb % 8;
  const long long x = b % 8, i = b / 8;
  return x * q + (x < rem ? x : rem) + i;
}

__device__ __forceinline__ double vcoord_mg(const Geo& g, int d, long long idx) {
#pragma clang fp contract(off)
  if (idx <= 0) return g.lo[d];
  if (idx >= g.res[d]) return g.hi[d];
  return g.lo[d] + ((double)idx / (double)g.res[d]) * (g.hi[d] - g.lo[d]);
}

__device__ __forceinline__ void colour_elem_mg(const Geo& g, int c, long long t, long long* ijk) {
  const long long mx = (g.res[0] - (c & 1) + 1) / 2;
  const long long my = g.dim > 1 ? (g.res[1] - ((c >> 1) & 1) + 1) / 2 : 1;
  const long long a = t % mx, rem = t / mx;
  const long long b = rem % my, cc = rem / my;
  ijk[0] = 2 * a + (c & 1);
  ijk[1] = 2 * b + ((c >> 1) & 1);
  ijk[2] = 2 * cc + ((c >> 2) & 1);
}

static inline unsigned gridn(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// moments
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void k_moments_fine(Geo g, const double4* __restrict__ pk, const int* __restrict__ off,
                               double* __restrict__ M) {
  constexpr int NM = DIM == 2 ? 9 : 27;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= g.E) return;
  double m[NM];
#pragma unroll
  for (int k = 0; k < NM; k++) m[k] = 0.0;
  for (int p = off[e]; p < off[e + 1]; p++) {
    const double4 q = pk[p];
    const double r[3] = {1.0, q.x, q.x * q.x};
    const double s[3] = {1.0, q.y, q.y * q.y};
    if (DIM == 2) {
#pragma unroll
      for (int b = 0; b < 3; b++)
#pragma unroll
        for (int a = 0; a < 3; a++) m[a + 3 * b] += q.w * r[a] * s[b];
    } else {
      const double t[3] = {1.0, q.z, q.z * q.z};
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
          const double ws = q.w * s[b] * t[c];
#pragma unroll
          for (int a = 0; a < 3; a++) m[a + 3 * b + 9 * c] += ws * r[a];
        }
    }
  }
#pragma unroll
  for (int k = 0; k < NM; k++) M[mom_index<double>(g, k, e)] = m[k];
}

// Parent moments from children: child local r_f maps to parent r_c = (r_f + sg)/2,
// sg = -1 for the lower child, +1 for the upper.  Per axis r_c^a = T[a][k] r_f^k.
template <int DIM>
__global__ void k_moments_coarsen(Geo gc, Geo gf, const double* __restrict__ Mf, double* __restrict__ Mc) {
  constexpr int NM = DIM == 2 ? 9 : 27;
  const long long E = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (E >= gc.E) return;
  const long long I = E % gc.res[0], J = (E / gc.res[0]) % gc.res[1], K = DIM == 3 ? E / ((long long)gc.res[0] * gc.res[1]) : 0;
  double acc[NM];
#pragma unroll
  for (int k = 0; k < NM; k++) acc[k] = 0.0;
  for (int ch = 0; ch < (DIM == 2 ? 4 : 8); ch++) {
    const int ox = ch & 1, oy = (ch >> 1) & 1, oz = (ch >> 2) & 1;
    const long long e = (2 * I + ox) + (long long)gf.res[0] * ((2 * J + oy) + (long long)gf.res[1] * (2 * K + oz));
    double T[3][3][3];  // T[axis][a][k]
    const double sg[3] = {ox ? 1.0 : -1.0, oy ? 1.0 : -1.0, oz ? 1.0 : -1.0};
    for (int d = 0; d < 3; d++) {
      const double s = sg[d];
      T[d][0][0] = 1.0; T[d][0][1] = 0.0; T[d][0][2] = 0.0;
      T[d][1][0] = 0.5 * s; T[d][1][1] = 0.5; T[d][1][2] = 0.0;
      T[d][2][0] = 0.25; T[d][2][1] = 0.5 * s; T[d][2][2] = 0.25;  // (r+s)^2/4 = (r^2 + 2 s r + 1)/4
    }
    double m[NM];
    for (int k = 0; k < NM; k++) m[k] = Mf[mom_index<double>(gf, k, e)];
    if (DIM == 2) {
      for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) {
          double v = 0.0;
          for (int kb = 0; kb <= b; kb++)
            for (int ka = 0; ka <= a; ka++) v += T[0][a][ka] * T[1][b][kb] * m[ka + 3 * kb];
          acc[a + 3 * b] += v;
        }
    } else {
      for (int c = 0; c < 3; c++)
        for (int b = 0; b < 3; b++)
          for (int a = 0; a < 3; a++) {
            double v = 0.0;
            for (int kc = 0; kc <= c; kc++)
              for (int kb = 0; kb <= b; kb++)
                for (int ka = 0; ka <= a; ka++)
                  v += T[0][a][ka] * T[1][b][kb] * T[2][c][kc] * m[ka + 3 * kb + 9 * kc];
            acc[a + 3 * b + 9 * c] += v;
          }
    }
  }
#pragma unroll
  for (int k = 0; k < NM; k++) Mc[mom_index<double>(gc, k, E)] = acc[k];
}

void launch_moments_fine(const Geo& g, const double4* pk, const int* off, double* M, hipStream_t st) {
  if (g.dim == 3)
    hipLaunchKernelGGL(k_moments_fine<3>, dim3(gridn(g.E)), dim3(kBlock), 0, st, g, pk, off, M);
  else
    hipLaunchKernelGGL(k_moments_fine<2>, dim3(gridn(g.E)), dim3(kBlock), 0, st, g, pk, off, M);
}
void launch_moments_coarsen(const Geo& gc, const Geo& gf, const double* Mf, double* Mc, hipStream_t st) {
  if (gc.dim == 3)
    hipLaunchKernelGGL(k_moments_coarsen<3>, dim3(gridn(gc.E)), dim3(kBlock), 0, st, gc, gf, Mf, Mc);
  else
    hipLaunchKernelGGL(k_moments_coarsen<2>, dim3(gridn(gc.E)), dim3(kBlock), 0, st, gc, gf, Mf, Mc);
}

template <typename S, typename T>
__global__ void k_convert(const S* __restrict__ a, T* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = (T)a[i];
}
// FP32 copy of a level's moments (the two precisions group the moments
// differently in 3D, so the copy goes through the element and moment index).
template <int DIM>
__global__ void k_mom_d2f(Geo g, const double* __restrict__ a, float* __restrict__ b) {
  constexpr int NM = DIM == 2 ? 9 : 27;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= g.E) return;
#pragma unroll
  for (int k = 0; k < NM; k++) b[mom_index<float>(g, k, e)] = (float)a[mom_index<double>(g, k, e)];
}
void launch_mom_d2f(const Geo& g, const double* a, float* b, hipStream_t st) {
  if (g.dim == 3)
    hipLaunchKernelGGL(k_mom_d2f<3>, dim3(gridn(g.E)), dim3(kBlock), 0, st, g, a, b);
  else
    hipLaunchKernelGGL(k_mom_d2f<2>, dim3(gridn(g.E)), dim3(kBlock), 0, st, g, a, b);
}

void launch_d2f(const double* a, float* b, long long n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL((k_convert<double, float>), dim3(gridn(n) < 8192 ? gridn(n) : 8192), dim3(kBlock), 0, st, a, b, n);
}
void launch_f2d(const float* a, double* b, long long n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL((k_convert<float, double>), dim3(gridn(n) < 8192 ? gridn(n) : 8192), dim3(kBlock), 0, st, a, b, n);
}

template <typename T>
__global__ void k_fill_mg(T* __restrict__ y, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = v;
}

template <>
void launch_mom_apply<double>(const Geo& g, const double* x, double* y, const double* Mom, const uint8_t* bcflag,
                              double* halo, hipStream_t st) {
  tile_apply_moments(g, x, y, halo, Mom, bcflag, st);
}
template <>
void launch_mom_apply<float>(const Geo& g, const float* x, float* y, const float* Mom, const uint8_t* bcflag,
                             float* halo, hipStream_t st) {
  tile_apply_moments32(g, x, y, halo, Mom, bcflag, st);
}

// diag(K) from moments: node a, component d: sum_p omega (|grad N_a|^2 + (dN_a/dx_d)^2).
template <int DIM>
__global__ void k_mom_diag(Geo g, int c, long long mcount, const double* __restrict__ Mom, double* __restrict__ diag) {
  constexpr int NEN = DIM == 2 ? 4 : 8;
  constexpr int NM = DIM == 2 ? 9 : 27;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= mcount) return;
  long long ijk[3];
  colour_elem_mg(g, c, t, ijk);
  const long long e = ijk[0] + (long long)g.res[0] * (ijk[1] + (long long)g.res[1] * ijk[2]);
  double inv[3];
  for (int d = 0; d < DIM; d++) inv[d] = 2.0 / (vcoord_mg(g, d, ijk[d] + 1) - vcoord_mg(g, d, ijk[d]));
  const double cN = DIM == 2 ? 0.25 : 0.125;
  double M[NM];
  for (int k = 0; k < NM; k++) M[k] = Mom[mom_index<double>(g, k, e)];
  const long long sy = g.nn[0], sz = g.nn[0] * g.nn[1];
  const long long n0 = ijk[0] + sy * ijk[1] + sz * ijk[2];
  for (int a = 0; a < NEN; a++) {
    // (dN_a/dx_d)^2 = (cN inv_d)^2 prod_{d' != d} (1 + s_d' x_d')^2 ; (1 + s x)^2 = 1 + 2 s x + x^2
    double sq[3];
    for (int d = 0; d < DIM; d++) {
      double v = 0.0;
      // polynomial coefficients per axis: axis d contributes only degree 0 (coefficient 1)
      double cf[3][3];
      for (int d2 = 0; d2 < 3; d2++) {
        const double s = ((a >> d2) & 1) ? 1.0 : -1.0;
        if (d2 >= DIM || d2 == d) {
          cf[d2][0] = 1.0; cf[d2][1] = 0.0; cf[d2][2] = 0.0;
        } else {
          cf[d2][0] = 1.0; cf[d2][1] = 2.0 * s; cf[d2][2] = 1.0;
        }
      }
      if (DIM == 2) {
        for (int b = 0; b < 3; b++)
          for (int aa = 0; aa < 3; aa++) v += cf[0][aa] * cf[1][b] * M[aa + 3 * b];
      } else {
        for (int cc = 0; cc < 3; cc++)
          for (int b = 0; b < 3; b++)
            for (int aa = 0; aa < 3; aa++) v += cf[0][aa] * cf[1][b] * cf[2][cc] * M[aa + 3 * b + 9 * cc];
      }
      sq[d] = (cN * inv[d]) * (cN * inv[d]) * v;
    }
    double tot = 0.0;
    for (int d = 0; d < DIM; d++) tot += sq[d];
    const long long nd = n0 + (a & 1) + ((a >> 1) & 1) * sy + ((a >> 2) & 1) * sz;
    for (int d = 0; d < DIM; d++) diag[nd * DIM + d] += tot + sq[d];
  }
}

void launch_mom_diag(const Geo& g, const double* Mom, double* diag, hipStream_t st) {
  const long long n = g.N * g.dim;
  unsigned gb = gridn(n);
  if (gb > 8192) gb = 8192;
  hipLaunchKernelGGL(k_fill_mg<double>, dim3(gb), dim3(kBlock), 0, st, diag, n, 0.0);
  const int ncol = g.dim == 2 ? 4 : 8;
  for (int c = 0; c < ncol; c++) {
    const long long m = colour_count(g, c);
    if (m <= 0) continue;
    if (g.dim == 3)
      hipLaunchKernelGGL(k_mom_diag<3>, dim3(gridn(m)), dim3(kBlock), 0, st, g, c, m, Mom, diag);
    else
      hipLaunchKernelGGL(k_mom_diag<2>, dim3(gridn(m)), dim3(kBlock), 0, st, g, c, m, Mom, diag);
  }
}

// ---------------------------------------------------------------------------
// transfers: P = Q1 interpolation coarse -> fine, R = P^T (full weighting)
// ---------------------------------------------------------------------------
// Lattice indices of node n (x fastest); 32-bit division when the node count
// allows it (64-bit division is a long software sequence on the GPU).
__device__ __forceinline__ void node_ijk(const Geo& g, long long n, long long& i, long long& j, long long& k) {
  if (g.N < (1LL << 31)) {
    const unsigned un = (unsigned)n, nx = (unsigned)g.nn[0], ny = (unsigned)g.nn[1];
    const unsigned q = un / nx;
    i = un - q * nx;
    j = q % ny;
    k = q / ny;
  } else {
    i = n % g.nn[0];
    j = (n / g.nn[0]) % g.nn[1];
    k = n / (g.nn[0] * g.nn[1]);
  }
}
// Full-weighting restriction b_C = R r_F on the nodes of the coarse level
// (3 x 3 (x 3) stencil, weights 1 and 1/2 per axis, in stencil order).
template <typename T, int DIM>
__global__ void k_restrict(Geo gc, Geo gf, const T* __restrict__ rf, T* __restrict__ bc, const uint8_t* __restrict__ bcflag_c) {
  const long long n = xcd_block_mg(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= gc.N) return;
  long long I, J, K;
  node_ijk(gc, n, I, J, K);
  T acc[3] = {0, 0, 0};
  const int kz0 = DIM == 3 ? -1 : 0, kz1 = DIM == 3 ? 1 : 0;
  for (int dz = kz0; dz <= kz1; dz++) {
    const long long k = 2 * K + dz;
    if (k < 0 || k >= gf.nn[2]) continue;
    const T wz = dz == 0 ? (T)1 : (T)0.5;
    for (int dy = -1; dy <= 1; dy++) {
      const long long j = 2 * J + dy;
      if (j < 0 || j >= gf.nn[1]) continue;
      const T wy = dy == 0 ? (T)1 : (T)0.5;
      for (int dx = -1; dx <= 1; dx++) {
        const long long i = 2 * I + dx;
        if (i < 0 || i >= gf.nn[0]) continue;
        const T w = wz * wy * (dx == 0 ? (T)1 : (T)0.5);
        const long long f = i + gf.nn[0] * (j + gf.nn[1] * k);
        for (int d = 0; d < DIM; d++) acc[d] += w * rf[f * DIM + d];
      }
    }
  }
  const unsigned bf = bcflag_c ? bcflag_c[n] : 0u;
  for (int d = 0; d < DIM; d++) bc[n * DIM + d] = ((bf >> d) & 1u) ? (T)0 : acc[d];
}

// x_F += P x_C (Q1 interpolation) on the nodes of the fine level: the up to
// 2^dim coarse parents and the fine value are loaded at once.
template <typename T, int DIM>
__global__ void k_prolong_add(Geo gc, Geo gf, const T* __restrict__ xc, T* __restrict__ xf, const uint8_t* __restrict__ bcflag_f) {
  const long long n = xcd_block_mg(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= gf.N) return;
  long long i, j, k;
  node_ijk(gf, n, i, j, k);
  const int nz = (DIM == 3 && (k & 1)) ? 2 : 1, ny = (j & 1) ? 2 : 1, nx = (i & 1) ? 2 : 1;
  const T wz = nz == 2 ? (T)0.5 : (T)1, wy = ny == 2 ? (T)0.5 : (T)1, wx = nx == 2 ? (T)0.5 : (T)1;
  constexpr int NA = DIM == 3 ? 2 : 1;
  T v[NA][2][2][DIM];
  T xv[DIM];
  const unsigned bf = bcflag_f ? bcflag_f[n] : 0u;
#pragma unroll
  for (int d = 0; d < DIM; d++) xv[d] = xf[n * DIM + d];
#pragma unroll
  for (int a = 0; a < NA; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int cc = 0; cc < 2; cc++) {
        const bool in = a < nz && b < ny && cc < nx;
        const long long K = DIM == 3 ? (k >> 1) + a : 0, J = (j >> 1) + b, I = (i >> 1) + cc;
        const long long cn = in ? I + gc.nn[0] * (J + gc.nn[1] * K) : 0;
#pragma unroll
        for (int d = 0; d < DIM; d++) v[a][b][cc][d] = in ? xc[cn * DIM + d] : (T)0;
      }
  T acc[DIM];
#pragma unroll
  for (int d = 0; d < DIM; d++) acc[d] = 0;
#pragma unroll
  for (int a = 0; a < NA; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int cc = 0; cc < 2; cc++)
#pragma unroll
        for (int d = 0; d < DIM; d++) acc[d] += wx * wy * wz * v[a][b][cc][d];
#pragma unroll
  for (int d = 0; d < DIM; d++)
    if (!((bf >> d) & 1u)) xf[n * DIM + d] = xv[d] + acc[d];
}

template <typename T>
void launch_restrict(const Geo& gc, const Geo& gf, const T* rf, T* bc, const uint8_t* bcflag_c, hipStream_t st) {
  if (gc.dim == 3)
    hipLaunchKernelGGL((k_restrict<T, 3>), dim3(gridn(gc.N)), dim3(kBlock), 0, st, gc, gf, rf, bc, bcflag_c);
  else
    hipLaunchKernelGGL((k_restrict<T, 2>), dim3(gridn(gc.N)), dim3(kBlock), 0, st, gc, gf, rf, bc, bcflag_c);
}
template <typename T>
void launch_prolong_add(const Geo& gc, const Geo& gf, const T* xc, T* xf, const uint8_t* bcflag_f, hipStream_t st) {
  if (gf.dim == 3)
    hipLaunchKernelGGL((k_prolong_add<T, 3>), dim3(gridn(gf.N)), dim3(kBlock), 0, st, gc, gf, xc, xf, bcflag_f);
  else
    hipLaunchKernelGGL((k_prolong_add<T, 2>), dim3(gridn(gf.N)), dim3(kBlock), 0, st, gc, gf, xc, xf, bcflag_f);
}
template void launch_restrict<double>(const Geo&, const Geo&, const double*, double*, const uint8_t*, hipStream_t);
template void launch_restrict<float>(const Geo&, const Geo&, const float*, float*, const uint8_t*, hipStream_t);
template void launch_prolong_add<double>(const Geo&, const Geo&, const double*, double*, const uint8_t*, hipStream_t);
template void launch_prolong_add<float>(const Geo&, const Geo&, const float*, float*, const uint8_t*, hipStream_t);

// ---------------------------------------------------------------------------
// Chebyshev-Jacobi step: r = b - Ax ; d = alpha d + beta Dinv r ; x += d
// (with x == 0 on entry, "first" uses Ax = 0 and skips reading x / Ax).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_cheb(const T* __restrict__ b, const T* __restrict__ Ax, const T* __restrict__ dinv, T* __restrict__ d,
                       T* __restrict__ x, long long n, T alpha, T beta, int first) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const T r = first ? b[i] : b[i] - Ax[i];
    const T di = (first ? (T)0 : alpha * d[i]) + beta * dinv[i] * r;
    d[i] = di;
    x[i] = first ? di : x[i] + di;
  }
}
template <typename T>
void launch_cheb(const T* b, const T* Ax, const T* dinv, T* d, T* x, long long n, T alpha, T beta, int first, hipStream_t st) {
  unsigned gb = gridn(n);
  if (gb > 8192) gb = 8192;
  hipLaunchKernelGGL(k_cheb<T>, dim3(gb), dim3(kBlock), 0, st, b, Ax, dinv, d, x, n, alpha, beta, first);
}
template void launch_cheb<double>(const double*, const double*, const double*, double*, double*, long long, double, double, int, hipStream_t);
template void launch_cheb<float>(const float*, const float*, const float*, float*, float*, long long, float, float, int, hipStream_t);

// r = b - Ax
template <typename T>
__global__ void k_resid(const T* __restrict__ b, const T* __restrict__ Ax, T* __restrict__ r, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    r[i] = b[i] - Ax[i];
}
template <typename T>
void launch_resid(const T* b, const T* Ax, T* r, long long n, hipStream_t st) {
  unsigned gb = gridn(n);
  if (gb > 8192) gb = 8192;
  hipLaunchKernelGGL(k_resid<T>, dim3(gb), dim3(kBlock), 0, st, b, Ax, r, n);
}
template void launch_resid<double>(const double*, const double*, double*, long long, hipStream_t);
template void launch_resid<float>(const float*, const float*, float*, long long, hipStream_t);

// Dense coarse solve: x[free[i]] = sum_j Kinv[i][j] b[free[j]], one wavefront per row.
template <typename T>
__global__ void k_dense_solve(const T* __restrict__ Kinv, const int* __restrict__ freeidx, int nf, const T* __restrict__ b,
                              T* __restrict__ x) {
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nf) return;
  T s = 0;
  for (int j = lane; j < nf; j += 64) s += Kinv[(long long)row * nf + j] * b[freeidx[j]];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (lane == 0) x[freeidx[row]] = s;
}
// Wide form: 16 rows per 1024-thread workgroup, one wavefront per row; the
// gathered right-hand side b[free[j]] is read once per workgroup into LDS,
// and each row is read with 16-byte loads (rows padded to ld, a multiple of
// 4 values, so every row starts 16-byte aligned).
template <typename T>
__global__ __launch_bounds__(1024) void k_dense_solve_wide(const T* __restrict__ Kinv, int ld,
                                                           const int* __restrict__ freeidx, int nf,
                                                           const T* __restrict__ b, T* __restrict__ x) {
  constexpr int E = 16 / (int)sizeof(T);
  typedef T vec_t __attribute__((ext_vector_type(E)));
  __shared__ __align__(16) T sb[kDenseWideMax];
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const bool live = row < nf;
  const vec_t* r = reinterpret_cast<const vec_t*>(Kinv + (long long)(live ? row : 0) * ld);
  // the row's first segment is requested before the right-hand side is
  // staged, and each later one while the previous one is being used (the
  // products and their order are unchanged)
  vec_t vn = {};
  if (live && lane * E < nf) vn = r[lane];
  for (int j = threadIdx.x; j < nf; j += blockDim.x) sb[j] = b[freeidx[j]];
  __syncthreads();
  if (!live) return;
  T s = 0;
  for (int j0 = lane * E; j0 < nf; j0 += 64 * E) {
    const vec_t v = vn;
    if (j0 + 64 * E < nf) vn = r[(j0 + 64 * E) / E];
#pragma unroll
    for (int e = 0; e < E; e++)
      if (j0 + e < nf) s += v[e] * sb[j0 + e];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (lane == 0) x[freeidx[row]] = s;
}

template <typename T>
void launch_dense_solve(const T* Kinv, const int* freeidx, int nf, int ld, const T* b, T* x, long long n,
                        hipStream_t st) {
  unsigned gb = gridn(n);
  if (gb > 8192) gb = 8192;
  hipLaunchKernelGGL(k_fill_mg<T>, dim3(gb), dim3(kBlock), 0, st, x, n, (T)0);
  if (nf <= 0) return;
  if (ld > 0)
    hipLaunchKernelGGL(k_dense_solve_wide<T>, dim3((nf + 15) / 16), dim3(1024), 0, st, Kinv, ld, freeidx, nf, b, x);
  else
    hipLaunchKernelGGL(k_dense_solve<T>, dim3((nf + 3) / 4), dim3(kBlock), 0, st, Kinv, freeidx, nf, b, x);
}
template void launch_dense_solve<double>(const double*, const int*, int, int, const double*, double*, long long,
                                         hipStream_t);
template void launch_dense_solve<float>(const float*, const int*, int, int, const float*, float*, long long,
                                        hipStream_t);

__global__ void k_symmetrize_lower(double* __restrict__ A, int n, int ld) {
  const long long nn = (long long)n * n;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nn; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % n), j = (int)(t / n);
    if (i < j) A[i + (long long)j * ld] = A[j + (long long)i * ld];
  }
}
void launch_symmetrize_lower(double* A, int n, int ld, hipStream_t st) {
  unsigned gb = gridn((long long)n * n);
  if (gb > 65536) gb = 65536;
  hipLaunchKernelGGL(k_symmetrize_lower, dim3(gb), dim3(kBlock), 0, st, A, n, ld);
}

}  // namespace gs
