// gs_kernels.hip -- CDNA4 (gfx950) kernels of the Q1-P0 PIC Stokes hot path.
//
// The constitutive setup pass runs one thread per element in 2^dim parity
// colours so that no two threads of a launch share a node: its nodal scatter
// is a plain read-modify-write, deterministic and free of atomics.  Within a
// colour the block index is remapped so that each of the eight XCDs (blocks
// are dealt round-robin to XCDs) works on one contiguous range of elements.
//
// K apply (reference: K_e = sum_p w_p detJ B_p^T D_p B_p assembled by
// ConstitutiveMatrixCartesian.cxx:363-485 and applied by PETSc MatMult) is
// evaluated matrix-free in gs_tile.hip: the element's nodal velocities are
// turned into the coefficients of the (bi/tri)linear velocity gradient by a
// Walsh-Hadamard butterfly, each material point adds its stress times the
// monomials {1, r, s, t, rs, rt, st} to 30 (3D) / 7 (2D) element moments, and
// the nodal forces are read back from the moments.
#include "gs_internal.h"
#include "gs_reduce.h"
#include "gs_tile.h"

#include <math.h>

namespace gs {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ long long xcd_block(long long b, long long nb) {
  const long long q = nb / 8, rem = nb % 8;
  const long long x = b % 8, i = b / 8;
  return x * q + (x < rem ? x : rem) + i;
}

// CartesianGenerator.cxx:2174-2215 vertex coordinate (contactDepth 0).
__device__ __forceinline__ double vcoord(const Geo& g, int d, long long idx) {
#pragma clang fp contract(off)
  if (idx <= 0) return g.lo[d];
  if (idx >= g.res[d]) return g.hi[d];
  return g.lo[d] + ((double)idx / (double)g.res[d]) * (g.hi[d] - g.lo[d]);
}

__device__ __forceinline__ int num_approx0(double v) { return v >= -1e-12 && v <= 1e-12; }

// Mesh_RegularAlgorithms.cxx:191-224, bit for bit: per axis
// index = (unsigned)modf((x-min)/sep); a point within 1e-12 of a face goes to
// the lower element; within 1e-12 of max goes to the last element.
__device__ __forceinline__ int search_element(const Geo& g, const double* x, long long* el) {
#pragma clang fp contract(off)
  long long ind[3] = {0, 0, 0};
  for (int d = 0; d < g.dim; d++) {
    const double sep = (g.hi[d] - g.lo[d]) / (double)g.res[d];
    if (num_approx0(x[d] - g.hi[d])) {
      ind[d] = g.res[d] - 1;
    } else if (x[d] < g.lo[d] || x[d] > g.hi[d]) {
      return 0;
    } else {
      double integer;
      const double out = (x[d] - g.lo[d]) / sep;
      const double frac = modf(out, &integer);
      ind[d] = (long long)(unsigned)integer;
      if (ind[d] > 0 && num_approx0(frac)) ind[d]--;
    }
  }
  *el = ind[0] + (long long)g.res[0] * (ind[1] + (long long)g.res[1] * ind[2]);
  return 1;
}



// Element (i,j,k) of parity colour c from thread t.
__device__ __forceinline__ void colour_elem(const Geo& g, int c, long long t, long long* ijk) {
  const long long mx = (g.res[0] - (c & 1) + 1) / 2;
  const long long my = g.dim > 1 ? (g.res[1] - ((c >> 1) & 1) + 1) / 2 : 1;
  const long long a = t % mx, rem = t / mx;
  const long long b = rem % my, cc = rem / my;
  ijk[0] = 2 * a + (c & 1);
  ijk[1] = 2 * b + ((c >> 1) & 1);
  ijk[2] = 2 * cc + ((c >> 2) & 1);
}

// ---------------------------------------------------------------------------
// fills and vector kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* __restrict__ y, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = v;
}

void launch_fill(double* y, long long n, double v, hipStream_t st) {
  if (n <= 0) return;
  unsigned gb = grid_for(n);
  if (gb > 8192) gb = 8192;
  hipLaunchKernelGGL(k_fill<double>, dim3(gb), dim3(kBlock), 0, st, y, n, v);
}
void launch_fill_f(float* y, long long n, float v, hipStream_t st) {
  if (n <= 0) return;
  unsigned gb = grid_for(n);
  if (gb > 8192) gb = 8192;
  hipLaunchKernelGGL(k_fill<float>, dim3(gb), dim3(kBlock), 0, st, y, n, v);
}

__global__ void k_masked_copy(const double* __restrict__ src, double* __restrict__ dst,
                              const uint8_t* __restrict__ bc, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = bc[i] ? 0.0 : src[i];
}
void launch_masked_copy(const double* src, double* dst, const uint8_t* bcdof, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_masked_copy, dim3(grid_for(n)), dim3(kBlock), 0, st, src, dst, bcdof, n);
}

// f <- mask(f - K u_bc)   (StiffnessMatrix.cxx:2092-2117 lifting)
__global__ void k_lift(double* __restrict__ f, const double* __restrict__ Ku, const uint8_t* __restrict__ bc,
                       long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    f[i] = bc[i] ? 0.0 : f[i] - Ku[i];
}
void launch_lift(double* f, const double* Kubc, const uint8_t* bcdof, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_lift, dim3(grid_for(n)), dim3(kBlock), 0, st, f, Kubc, bcdof, n);
}

__global__ void k_make_dinv(const double* __restrict__ d, const uint8_t* __restrict__ bc, double* __restrict__ di,
                            long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    di[i] = bc[i] ? 0.0 : 1.0 / d[i];
}
void launch_make_dinv(const double* diag, const uint8_t* bcdof, double* dinv, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_make_dinv, dim3(grid_for(n)), dim3(kBlock), 0, st, diag, bcdof, dinv, n);
}

__global__ void k_recip(const double* __restrict__ a, double* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = 1.0 / a[i];
}
void launch_recip(const double* a, double* b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_recip, dim3(grid_for(n)), dim3(kBlock), 0, st, a, b, n);
}

__global__ void k_scale(double* __restrict__ y, const double* __restrict__ x, double a, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}
void launch_scale(double* y, const double* x, double a, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kBlock), 0, st, y, x, a, n);
}

__global__ void k_axpby(double* __restrict__ y, const double* __restrict__ x, double a, double b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
void launch_axpby(double* y, const double* x, double a, double b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0, st, y, x, a, b, n);
}

__global__ void k_add(double* __restrict__ y, const double* __restrict__ a, const double* __restrict__ b,
                      long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = a[i] + b[i];
}
void launch_add(double* y, const double* a, const double* b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_add, dim3(grid_for(n)), dim3(kBlock), 0, st, y, a, b, n);
}

__global__ void k_mul(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    c[i] = a[i] * b[i];
}
void launch_mul(const double* a, const double* b, double* c, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(kBlock), 0, st, a, b, c, n);
}

__global__ void k_dot(const double* __restrict__ a, const double* __restrict__ b, long long n, double* S, int slot,
                      Reducer rd) {
  double v[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    v[0] += a[i] * b[i];
  block_partials<1, false>(v, rd);
}
void launch_dot(const double* a, const double* b, long long n, double* S, int slot, Reducer r, hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kBlock), 0, st, a, b, n, S, slot, r);
  reduce_finish<1, false>(r, nb, OpStore{S, slot}, st);
}

__global__ void k_vmax(const double* __restrict__ u, long long nnode, int dim, double* S, int slot, Reducer rd) {
  double v[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nnode;
       i += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int d = 0; d < dim; d++) s += u[i * dim + d] * u[i * dim + d];
    v[0] = fmax(v[0], sqrt(s));
  }
  block_partials<1, true>(v, rd);
}
void launch_vmax(const double* u, long long nnode, int dim, double* S, int slot, Reducer rd, hipStream_t st) {
  const int nb = reduce_grid(nnode);
  hipLaunchKernelGGL(k_vmax, dim3(nb), dim3(kBlock), 0, st, u, nnode, dim, S, slot, rd);
  reduce_finish<1, true>(rd, nb, OpStore{S, slot}, st);
}

// ---------------------------------------------------------------------------
// Inner velocity CG with Jacobi preconditioner.  Scalars stay on the device.
// ---------------------------------------------------------------------------
__global__ void k_cg_init(const double* __restrict__ b, const double* __restrict__ dinv, double* __restrict__ x,
                          double* __restrict__ r, double* __restrict__ p, long long n, double* S, Reducer rd) {
  double v[2] = {0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double bi = b[i], zi = dinv[i] * bi;
    x[i] = 0.0;
    r[i] = bi;
    p[i] = zi;
    v[0] += bi * zi;
    v[1] += bi * bi;
  }
  block_partials<2, false>(v, rd);
}
struct OpCgInit {
  double* S;
  __device__ void operator()(const double* t) const {
    S[S_RZ] = t[0];
    S[S_BB] = t[1];
    S[S_RR] = t[1];
  }
};
void launch_cg_init(const double* b, const double* dinv, double* x, double* r, double* p, long long n, double* S,
                    Reducer rd, hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_cg_init, dim3(nb), dim3(kBlock), 0, st, b, dinv, x, r, p, n, S, rd);
  reduce_finish<2, false>(rd, nb, OpCgInit{S}, st);
}

__global__ void k_cg_update(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                            const double* __restrict__ Ap, const double* __restrict__ dinv, long long n, double* S,
                            Reducer rd) {
  const double rz = S[S_RZ], pap = S[S_PAP];
  const double alpha = (pap != 0.0 && isfinite(pap)) ? rz / pap : 0.0;
  double v[2] = {0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    v[0] += ri * (dinv[i] * ri);
    v[1] += ri * ri;
  }
  block_partials<2, false>(v, rd);
}
struct OpCgUpdate {
  double* S;
  __device__ void operator()(const double* t) const {
    const double rz = S[S_RZ];
    S[S_BETA] = rz != 0.0 ? t[0] / rz : 0.0;
    S[S_RZ] = t[0];
    S[S_RR] = t[1];
  }
};
void launch_cg_update(double* x, double* r, const double* p, const double* Ap, const double* dinv, long long n,
                      double* S, Reducer rd, hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(kBlock), 0, st, x, r, p, Ap, dinv, n, S, rd);
  reduce_finish<2, false>(rd, nb, OpCgUpdate{S}, st);
}

__global__ void k_cg_pupdate(double* __restrict__ p, const double* __restrict__ r, const double* __restrict__ dinv,
                             long long n, const double* S) {
  const double beta = S[S_BETA];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = dinv[i] * r[i] + beta * p[i];
}
void launch_cg_pupdate(double* p, const double* r, const double* dinv, long long n, const double* S,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_cg_pupdate, dim3(grid_for(n)), dim3(kBlock), 0, st, p, r, dinv, n, S);
}

// Flexible CG: alpha = rz / pAp; x += alpha p; r -= alpha Ap.
__global__ void k_fcg_update(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                             const double* __restrict__ Ap, long long n, double* S, Reducer rd) {
  const double rz = S[S_RZ], pap = S[S_PAP];
  const double alpha = (pap != 0.0 && isfinite(pap)) ? rz / pap : 0.0;
  double v[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    v[0] += ri * ri;
  }
  block_partials<1, false>(v, rd);
}
struct OpFcgUpdate {
  double* S;
  __device__ void operator()(const double* t) const {
    const double rz = S[S_RZ], pap = S[S_PAP];
    S[S_RR] = t[0];
    S[S_ALPHA] = (pap != 0.0 && isfinite(pap)) ? rz / pap : 0.0;
  }
};
void launch_fcg_update(double* x, double* r, const double* p, const double* Ap, long long n, double* S, Reducer rd,
                       hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_fcg_update, dim3(nb), dim3(kBlock), 0, st, x, r, p, Ap, n, S, rd);
  reduce_finish<1, false>(rd, nb, OpFcgUpdate{S}, st);
}

// z.r and z.Ap; beta = (z.(r_new - r_old)) / rz_old = -alpha z.Ap / rz_old (Polak-Ribiere).
__global__ void k_fcg_dots(const double* __restrict__ z, const double* __restrict__ r, const double* __restrict__ Ap,
                           long long n, double* S, Reducer rd) {
  double v[2] = {0.0, 0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double zi = z[i];
    v[0] += zi * r[i];
    v[1] += zi * Ap[i];
  }
  block_partials<2, false>(v, rd);
}
struct OpFcgDots {
  double* S;
  __device__ void operator()(const double* t) const {
    const double rz = S[S_RZ];
    S[S_BETA] = rz != 0.0 ? -S[S_ALPHA] * t[1] / rz : 0.0;
    S[S_RZ] = t[0];
  }
};
void launch_fcg_dots(const double* z, const double* r, const double* Ap, long long n, double* S, Reducer rd,
                     hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_fcg_dots, dim3(nb), dim3(kBlock), 0, st, z, r, Ap, n, S, rd);
  reduce_finish<2, false>(rd, nb, OpFcgDots{S}, st);
}

__global__ void k_fcg_p(double* __restrict__ p, const double* __restrict__ z, long long n, const double* S) {
  const double beta = S[S_BETA];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = z[i] + beta * p[i];
}
void launch_fcg_p(double* p, const double* z, long long n, const double* S, hipStream_t st) {
  hipLaunchKernelGGL(k_fcg_p, dim3(grid_for(n)), dim3(kBlock), 0, st, p, z, n, S);
}

// ---------------------------------------------------------------------------
// Uzawa outer loop vector steps (Stokes_SLE_UzawaSolver.cxx:664-837).
// ---------------------------------------------------------------------------
__global__ void k_uz_precond(const double* __restrict__ r, const double* __restrict__ Qinv, double* __restrict__ z,
                             long long n, double* S, Reducer rd) {
  double v[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double zi = Qinv ? r[i] * Qinv[i] : r[i];
    z[i] = zi;
    v[0] += zi * r[i];
  }
  block_partials<1, false>(v, rd);
}
struct OpUzPrecond {
  double* S;
  __device__ void operator()(const double* t) const {
    const double prev = S[S_ZR_PREV];
    S[S_UBETA] = prev != 0.0 ? t[0] / prev : 0.0;
    S[S_ZR] = t[0];
  }
};
void launch_uz_precond(const double* r, const double* Qinv, double* z, long long n, double* S, Reducer rd,
                       hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_uz_precond, dim3(nb), dim3(kBlock), 0, st, r, Qinv, z, n, S, rd);
  reduce_finish<1, false>(rd, nb, OpUzPrecond{S}, st);
}

// launch_uz_precond in place of a network output that is not a descent
// direction: the blocks and the finishing op test S[S_ZR] (z.r of the
// network pass) the same way, so either both do exactly the work and the
// scalar updates of launch_uz_precond, counting the case in S[S_NBAD], or
// neither changes z or the scalars.
__device__ __forceinline__ bool uz_descent(double zr) { return zr > 0.0 && isfinite(zr); }
__global__ void k_uz_precond_if_bad(const double* __restrict__ r, const double* __restrict__ Qinv,
                                    double* __restrict__ z, long long n, double* S, Reducer rd) {
  double v[1] = {0.0};
  if (!uz_descent(S[S_ZR])) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
      const double zi = Qinv ? r[i] * Qinv[i] : r[i];
      z[i] = zi;
      v[0] += zi * r[i];
    }
  }
  block_partials<1, false>(v, rd);
}
struct OpUzPrecondIfBad {
  double* S;
  __device__ void operator()(const double* t) const {
    if (uz_descent(S[S_ZR])) return;
    OpUzPrecond{S}(t);
    S[S_NBAD] += 1.0;
  }
};
void launch_uz_precond_if_bad(const double* r, const double* Qinv, double* z, long long n, double* S, Reducer rd,
                              hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_uz_precond_if_bad, dim3(nb), dim3(kBlock), 0, st, r, Qinv, z, n, S, rd);
  reduce_finish<1, false>(rd, nb, OpUzPrecondIfBad{S}, st);
}

__global__ void k_uz_sdir(double* __restrict__ s, const double* __restrict__ z, long long n, const double* S,
                          int first) {
  const double beta = first ? 0.0 : S[S_UBETA];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s[i] = first ? z[i] : z[i] + beta * s[i];
}
void launch_uz_sdir(double* s, const double* z, long long n, const double* S, int first, hipStream_t st) {
  hipLaunchKernelGGL(k_uz_sdir, dim3(grid_for(n)), dim3(kBlock), 0, st, s, z, n, S, first);
}

__global__ void k_uz_update(double* __restrict__ p, const double* __restrict__ s, double* __restrict__ r,
                            const double* __restrict__ q, long long n, double* S, Reducer rd) {
  const double alpha = S[S_ZR] / S[S_SQ];
  double v[1] = {0.0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    p[i] += alpha * s[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    v[0] += ri * ri;
  }
  block_partials<1, false>(v, rd);
}
struct OpUzUpdate {
  double* S;
  __device__ void operator()(const double* t) const {
    S[S_ORR] = t[0];
    S[S_TMP0] = S[S_ZR] / S[S_SQ];
    S[S_ZR_PREV] = S[S_ZR];
  }
};
void launch_uz_update(double* p, const double* s, double* r, const double* q, long long n, double* S, Reducer rd,
                      hipStream_t st) {
  const int nb = reduce_grid(n);
  hipLaunchKernelGGL(k_uz_update, dim3(nb), dim3(kBlock), 0, st, p, s, r, q, n, S, rd);
  reduce_finish<1, false>(rd, nb, OpUzUpdate{S}, st);
}

__global__ void k_uz_uupdate(double* __restrict__ u, const double* __restrict__ v, long long n, const double* S) {
  const double alpha = S[S_TMP0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    u[i] -= alpha * v[i];
}
void launch_uz_uupdate(double* u, const double* v, long long n, const double* S, hipStream_t st) {
  hipLaunchKernelGGL(k_uz_uupdate, dim3(grid_for(n)), dim3(kBlock), 0, st, u, v, n, S);
}

// ---------------------------------------------------------------------------
// K apply: the tiled kernels of gs_tile.hip.
// ---------------------------------------------------------------------------
void launch_apply_K(const Geo& g, const double* x, double* y, const double4* pk, const int* off,
                    const uint8_t* bcflag, double* halo, hipStream_t st) {
  tile_apply_particles(g, x, y, halo, pk, off, bcflag, st);
}
void launch_apply_K32(const Geo& g, const float* x, float* y, const float4* pk, const int* off,
                      const uint8_t* bcflag, float* halo, hipStream_t st) {
  tile_apply_particles32(g, x, y, halo, pk, off, bcflag, st);
}

__global__ void k_pack32(const double4* __restrict__ a, float4* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double4 v = a[i];
    b[i] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
  }
}
void launch_pack32(const double4* pk, float4* pk32, long long P, hipStream_t st) {
  if (P > 0) hipLaunchKernelGGL(k_pack32, dim3(grid_for(P)), dim3(kBlock), 0, st, pk, pk32, P);
}

// StrainWeakening as a TimeIntegrand (StrainWeakening.cxx:432-441: the time
// derivative is the increment of the last constitutive pass, so the RK2 step
// is w + dt * inc) followed by _StrainWeakening_MakeValuesPositive (:443-468).
// The strain-limited shape test uses the same arithmetic as the host's
// Sphere / Box containment (Sphere.cxx:189-208, Box.cxx:209-221).
__device__ bool strain_limit_inside(const StrainLimit& L, const double* x) {
  if (L.type == 0) return true;
  if (L.type == 1) {
    double v = 0.0;
    for (int d = 0; d < L.dim; d++) {
      const double t = x[d] - L.p[d];
      v += t * t;
    }
    return v <= L.p[3] * L.p[3];
  }
  for (int d = 0; d < L.dim; d++) {
    const double c = 0.5 * (L.p[d] + L.p[3 + d]);
    const double wd = L.p[3 + d] - L.p[d];
    if (fabs(x[d] - c) > 0.5 * wd) return false;
  }
  return true;
}

__global__ void k_weakening_step(double* __restrict__ w, const double* __restrict__ inc,
                                 const double* __restrict__ coord, long long P, double dt, StrainLimit lim) {
#pragma clang fp contract(off)
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= P) return;
  double v = w[p] + dt * inc[p];
  if (lim.type >= 0 && strain_limit_inside(lim, coord + 3 * p)) v = lim.limit;
  w[p] = v < 0.0 ? 0.0 : v;
}

void launch_weakening_step(double* w, const double* inc, const double* coord, long long P, double dt,
                           const StrainLimit& lim, hipStream_t st) {
  if (P > 0)
    hipLaunchKernelGGL(k_weakening_step, dim3(grid_for(P)), dim3(kBlock), 0, st, w, inc, coord, P, dt, lim);
}

// ---------------------------------------------------------------------------
// G p (node-centric gather) and G^T u (element-centric).  G_e[a,d] is the
// closed form of GradientStiffnessMatrixTerm.cxx:225-245 on an affine
// element: -int_e dN_a/dx_d = -s_ad * prod_{d'!=d} h_d' / 2^(dim-1).
// ---------------------------------------------------------------------------
// y = G p on the nodes: each node gathers the pressures of its (up to 2^dim)
// elements, all loaded at once (elements outside the mesh read as 0 and add
// exact zeros); element spacings are taken once per axis; 32-bit lattice
// indices when the mesh allows.  Sum order: element layers, rows, columns.
template <int DIM>
__global__ void k_apply_G(Geo g, const double* __restrict__ p, double* __restrict__ y,
                          const uint8_t* __restrict__ bcflag) {
  // blocks remapped so each XCD gathers from a contiguous slab (its L2 serves
  // the pressures shared by neighbouring nodes)
  const long long n = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  long long i, j, k;
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
  const unsigned bf = bcflag ? bcflag[n] : 0u;
  constexpr int NZ = DIM == 3 ? 2 : 1;
  double pe[NZ][2][2];
#pragma unroll
  for (int az = 0; az < NZ; az++) {
    const long long kk = DIM == 3 ? k - 1 + az : 0;
#pragma unroll
    for (int ay = 0; ay < 2; ay++) {
      const long long jj = j - 1 + ay;
#pragma unroll
      for (int ax = 0; ax < 2; ax++) {
        const long long ii = i - 1 + ax;
        const bool in = kk >= 0 && kk < g.res[2] && jj >= 0 && jj < g.res[1] && ii >= 0 && ii < g.res[0];
        pe[az][ay][ax] = in ? p[ii + (long long)g.res[0] * (jj + (long long)g.res[1] * kk)] : 0.0;
      }
    }
  }
  // element spacings of the columns / rows / layers i-1 and i
  double h[3][2];
  for (int d = 0; d < DIM; d++) {
    const long long c = d == 0 ? i : (d == 1 ? j : k);
    const double v0 = vcoord(g, d, c - 1), v1 = vcoord(g, d, c), v2 = vcoord(g, d, c + 1);
    h[d][0] = v1 - v0;
    h[d][1] = v2 - v1;
  }
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int az = 0; az < NZ; az++) {
    const double hz = DIM == 3 ? h[2][az] : 1.0;
    const double sz = az == 0 ? 1.0 : -1.0;
#pragma unroll
    for (int ay = 0; ay < 2; ay++) {
      const double hy = h[1][ay];
      const double sy = ay == 0 ? 1.0 : -1.0;
#pragma unroll
      for (int ax = 0; ax < 2; ax++) {
        const double hx = h[0][ax];
        const double sx = ax == 0 ? 1.0 : -1.0;
        const double v = pe[az][ay][ax];
        if (DIM == 2) {
          acc[0] -= sx * (hy * 0.5) * v;
          acc[1] -= sy * (hx * 0.5) * v;
        } else {
          acc[0] -= sx * (hy * hz * 0.25) * v;
          acc[1] -= sy * (hx * hz * 0.25) * v;
          acc[2] -= sz * (hx * hy * 0.25) * v;
        }
      }
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) y[n * DIM + d] = (bf >> d) & 1u ? 0.0 : acc[d];
}
void launch_apply_G(const Geo& g, const double* p, double* y, const uint8_t* bcflag, hipStream_t st) {
  if (g.dim == 3)
    hipLaunchKernelGGL(k_apply_G<3>, dim3(grid_for(g.N)), dim3(kBlock), 0, st, g, p, y, bcflag);
  else
    hipLaunchKernelGGL(k_apply_G<2>, dim3(grid_for(g.N)), dim3(kBlock), 0, st, g, p, y, bcflag);
}

__global__ void k_apply_Gt(Geo g, const double* __restrict__ x, double* __restrict__ q) {
  // blocks remapped so each XCD gathers from a contiguous slab of nodes
  const long long e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= g.E) return;
  long long i, j, k;
  if (g.E < (1LL << 31)) {  // 32-bit lattice indices (64-bit division is a long software sequence)
    const unsigned ue = (unsigned)e, rx = (unsigned)g.res[0], ry = (unsigned)g.res[1];
    const unsigned qq = ue / rx;
    i = ue - qq * rx;
    j = qq % ry;
    k = g.dim == 3 ? qq / ry : 0;
  } else {
    i = e % g.res[0];
    j = (e / g.res[0]) % g.res[1];
    k = g.dim == 3 ? e / ((long long)g.res[0] * g.res[1]) : 0;
  }
  const double hx = vcoord(g, 0, i + 1) - vcoord(g, 0, i);
  const double hy = vcoord(g, 1, j + 1) - vcoord(g, 1, j);
  const long long n0 = i + g.nn[0] * (j + g.nn[1] * k);
  const int dim = g.dim;
  if (dim == 2) {
    const long long sy = g.nn[0];
    const double* u0 = x + n0 * 2;
    const double* u1 = x + (n0 + 1) * 2;
    const double* u2 = x + (n0 + sy) * 2;
    const double* u3 = x + (n0 + sy + 1) * 2;
    const double fx = (u1[0] - u0[0]) + (u3[0] - u2[0]);
    const double fy = (u2[1] - u0[1]) + (u3[1] - u1[1]);
    q[e] = -(hy * 0.5) * fx - (hx * 0.5) * fy;
  } else {
    const double hz = vcoord(g, 2, k + 1) - vcoord(g, 2, k);
    const long long sy = g.nn[0], sz = g.nn[0] * g.nn[1];
    double fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const long long nd = n0 + (a & 1) + ((a >> 1) & 1) * sy + ((a >> 2) & 1) * sz;
      const double* u = x + nd * 3;
      fx += (a & 1) ? u[0] : -u[0];
      fy += (a & 2) ? u[1] : -u[1];
      fz += (a & 4) ? u[2] : -u[2];
    }
    q[e] = -(hy * hz * 0.25) * fx - (hx * hz * 0.25) * fy - (hx * hy * 0.25) * fz;
  }
}
void launch_apply_Gt(const Geo& g, const double* x, double* q, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_Gt, dim3(grid_for(g.E)), dim3(kBlock), 0, st, g, x, q);
}

// ---------------------------------------------------------------------------
// Fused constitutive pass (one per linear solve): effective viscosity at each
// material point (MaterialViscosity -> YieldRheology -> VonMises /
// DruckerPrager), omega_p = w_p detJ eta_p written into the particle stream,
// element diag(K_e), Q_e = sum G_e^2 / K_e,vv (UzawaPreconditionerTerm.cxx
// :195-215) and the buoyancy force (BuoyancyForceTerm.cxx:474-533).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sym_inv(int dim, const double* t) {
  double v;
  if (dim == 2)
    v = 0.5 * (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2];
  else
    v = 0.5 * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) + t[3] * t[3] + t[4] * t[4] + t[5] * t[5];
  return sqrt(v);
}

// Yield-limited viscosity (YieldRheology.cxx:203-231, VonMises.cxx:207-291,
// DruckerPrager.cxx:296-445).  L = velocity gradient at the point.
// StrainWeakening_CalcRatio (StrainWeakening.cxx:497-519): softening ratio of
// accumulated plastic strain w, clipped to [0, 1].
__device__ inline double weakening_ratio(const MatDev& m, double w) {
  if (w < 0.0) w = 0.0;
  double r = (w - m.sw_eps0) / (m.sw_eps1 - m.sw_eps0);
  if (r > 1.0) r = 1.0;
  if (r < 0.0) r = 0.0;
  return r;
}

// Viscosity after the yield rheology of material m.  With strain weakening
// (ratio r) the cohesion and friction are interpolated towards their
// after-softening values (VonMises.cxx:196-199, DruckerPrager.cxx:448-489),
// and *crit / *ind return the yield criterion and indicator the
// StrainWeakening increment uses.
__device__ double yield_viscosity(int dim, const MatDev& m, double eta, const double (&L)[3][3], double pres,
                                  double r, int* yielded, double* crit_out, double* ind_out) {
  double sr[6], stv[6];
  const int nst = dim == 2 ? 3 : 6;
  if (dim == 2) {
    sr[0] = L[0][0];
    sr[1] = L[1][1];
    sr[2] = 0.5 * (L[0][1] + L[1][0]);
  } else {
    sr[0] = L[0][0];
    sr[1] = L[1][1];
    sr[2] = L[2][2];
    sr[3] = 0.5 * (L[0][1] + L[1][0]);
    sr[4] = 0.5 * (L[0][2] + L[2][0]);
    sr[5] = 0.5 * (L[1][2] + L[2][1]);
  }
  for (int k = 0; k < nst; k++) stv[k] = (k < dim ? 2.0 * eta : eta * 2.0) * sr[k];
  double trs = 0.0, trr = 0.0;
  for (int k = 0; k < dim; k++) {
    trs += stv[k];
    trr += sr[k];
  }
  for (int k = 0; k < dim; k++) {
    sr[k] -= trr / dim;
    stv[k] -= trs / dim;
  }
  const double srII = sym_inv(dim, sr), tauII = sym_inv(dim, stv);
  *yielded = 0;
  *ind_out = tauII;
  const double C = m.sw_on ? m.cohesion * (1.0 - r) + m.cohesion_after * r : m.cohesion;
  if (m.yield_type == 1) {
    const double mu = m.sw_on ? m.friction * (1.0 - r) + m.friction_after * r : m.friction;
    double factor, fs;
    if (dim == 2) {
      factor = 1 / sqrt(1 + mu * mu);
      fs = mu * pres * factor + C * factor;
    } else {
      const double cphi = 1 / sqrt(1 + mu * mu);
      const double sphi = mu * cphi;
      factor = 2 * cphi / (sqrt(3.0) * (3 - sphi));
      fs = mu * factor * pres + C * 3 * factor;
    }
    double minys = m.min_yield_stress;
    if (minys == 0.0) minys = C;
    if (fs < minys * factor) fs = minys * factor;
    *crit_out = fs;
    if (fs < tauII) {
      double v = (m.max_strain_rate > 0 && srII > m.max_strain_rate) ? fs / (2 * m.max_strain_rate)
                                                                      : fs / (2 * srII);
      if (v < m.min_viscosity) v = m.min_viscosity;
      *yielded = 1;
      return v;
    }
  } else if (m.yield_type == 2) {
    const double crit = C;
    *crit_out = crit;
    if (crit < tauII) {
      const double beta = 1.0 - crit / tauII;
      double corr = -eta * beta;
      if (eta + corr < m.min_viscosity) corr = m.min_viscosity - eta;
      *yielded = 1;
      return eta + corr;
    }
  }
  return eta;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_setup(Geo g, int c, long long mcount, SetupArgs a, double4* __restrict__ pk,
                                                  const int* __restrict__ off, double* __restrict__ diag,
                                                  double* __restrict__ f, double* __restrict__ Q,
                                                  double* __restrict__ eta_out, uint8_t* __restrict__ yielded) {
  constexpr int NEN = DIM == 2 ? 4 : 8;
  const long long t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= mcount) return;
  long long ijk[3];
  colour_elem(g, c, t, ijk);
  const long long e = ijk[0] + (long long)g.res[0] * (ijk[1] + (long long)g.res[1] * ijk[2]);
  double h[3] = {1.0, 1.0, 1.0};
  for (int d = 0; d < DIM; d++) h[d] = vcoord(g, d, ijk[d] + 1) - vcoord(g, d, ijk[d]);
  const double detJ = DIM == 2 ? (h[0] * 0.5) * (h[1] * 0.5) : (h[0] * 0.5) * (h[1] * 0.5) * (h[2] * 0.5);
  double gi[3];
  for (int d = 0; d < DIM; d++) gi[d] = 2.0 / h[d];
  const long long sy = g.nn[0], sz = g.nn[0] * g.nn[1];
  const long long n0 = ijk[0] + sy * ijk[1] + sz * ijk[2];
  long long nid[NEN];
  for (int q = 0; q < NEN; q++) nid[q] = n0 + (q & 1) + ((q >> 1) & 1) * sy + ((q >> 2) & 1) * sz;
  double ue[NEN][3];
  const bool need_u = a.yield_on && a.u_prev != nullptr;
  if (need_u)
    for (int q = 0; q < NEN; q++)
      for (int d = 0; d < DIM; d++) ue[q][d] = a.u_prev[nid[q] * DIM + d];
  const double pres = a.p_prev ? a.p_prev[e] : 0.0;
  double dg[NEN][3], fe[NEN][3];
  for (int q = 0; q < NEN; q++)
    for (int d = 0; d < 3; d++) dg[q][d] = fe[q][d] = 0.0;
  const int pb = off[e], pe = off[e + 1];
  const double wc = (DIM == 2 ? 4.0 : 8.0) / (double)(pe - pb);
  for (int p = pb; p < pe; p++) {
    const double w = a.wp ? a.wp[p] : wc;
    double4 q4 = pk[p];
    const double xi[3] = {q4.x, q4.y, q4.z};
    const int mi = a.mat[p];
    const MatDev m = a.mats[mi];
    double eta = a.eta0p ? a.eta0p[p] : m.eta0;
    // shape functions and global derivatives (affine element)
    double N[NEN], dN[3][NEN];
    for (int q = 0; q < NEN; q++) {
      double fct[3], sg[3];
      for (int d = 0; d < DIM; d++) {
        sg[d] = ((q >> d) & 1) ? 1.0 : -1.0;
        fct[d] = 1.0 + sg[d] * xi[d];
      }
      const double cN = DIM == 2 ? 0.25 : 0.125;
      double nv = cN;
      for (int d = 0; d < DIM; d++) nv *= fct[d];
      N[q] = nv;
      for (int d = 0; d < DIM; d++) {
        double v = cN * sg[d] * gi[d];
        for (int d2 = 0; d2 < DIM; d2++)
          if (d2 != d) v *= fct[d2];
        dN[d][q] = v;
      }
    }
    int yl = 0;
    double inc = 0.0;
    if (need_u && m.yield_type != 0) {
      double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int q = 0; q < NEN; q++)
        for (int i = 0; i < DIM; i++)
          for (int j = 0; j < DIM; j++) L[i][j] += dN[j][q] * ue[q][i];
      const bool sw = m.sw_on && a.pstr;
      const double w = sw ? (a.pstr[p] < 0.0 ? 0.0 : a.pstr[p]) : 0.0;
      double crit = 0.0, ind = 0.0;
      const double eta_in = eta;
      eta = yield_viscosity(DIM, m, eta, L, pres, sw ? weakening_ratio(m, w) : 0.0, &yl, &crit, &ind);
      if (sw) {
        // _StrainWeakening_CalcIncrementIsotropic (StrainWeakening.cxx:470-495),
        // with the viscosity before yielding: growth from the plastic part of
        // the strain rate, decay by healing
        double beta = 1.0 - crit / ind;
        if (beta < 0.0) beta = 0.0;
        inc = (crit / eta_in) * (beta / (1.0 - beta) - m.sw_healing * w);
      }
    }
    if (a.pinc) a.pinc[p] = inc;
    if (a.clamp_min > 0 && eta < a.clamp_min) eta = a.clamp_min;
    if (a.clamp_max > 0 && eta > a.clamp_max) eta = a.clamp_max;
    const double cw = w * detJ;
    const double om = cw * eta;
    q4.w = om;
    pk[p] = q4;
    if (eta_out) eta_out[p] = eta;
    if (yielded) yielded[p] = (uint8_t)yl;
    const double fr = cw * m.density;
    for (int q = 0; q < NEN; q++) {
      double sq = 0.0, d2[3];
      for (int d = 0; d < DIM; d++) {
        d2[d] = dN[d][q] * dN[d][q];
        sq += d2[d];
      }
      for (int d = 0; d < DIM; d++) {
        dg[q][d] += om * (sq + d2[d]);
        fe[q][d] += a.g[d] * fr * N[q];
      }
    }
  }
  // closed-form element G and Q_e
  double qe = 0.0;
  for (int q = 0; q < NEN; q++)
    for (int d = 0; d < DIM; d++) {
      double area = 1.0;
      for (int d2 = 0; d2 < DIM; d2++)
        if (d2 != d) area *= h[d2];
      const double s = ((q >> d) & 1) ? 1.0 : -1.0;
      const double ge = -s * area * (DIM == 2 ? 0.5 : 0.25);
      qe += ge * (ge / dg[q][d]);
    }
  Q[e] = qe;
  for (int q = 0; q < NEN; q++)
    for (int d = 0; d < DIM; d++) {
      diag[nid[q] * DIM + d] += dg[q][d];
      f[nid[q] * DIM + d] += fe[q][d];
    }
}

// Output fields of the material points after a solve: strain rate of the
// velocity u at the integration point (the symmetric part of the velocity
// gradient, StrainRateField), its second invariant (TensorMath.cxx:431-459),
// and the stress of the isotropic constitutive matrix with the point's
// viscosity from the last constitutive pass (StoreStress.cxx / StoreViscosity
// .cxx: ConstitutiveMatrix_CalculateStress, ConstitutiveMatrixCartesian
// .cxx:727-775).  Sorted particle order; escaped particles get zeros.
template <int DIM>
__global__ void k_particle_fields(Geo g, const double4* __restrict__ pk, const unsigned* __restrict__ keys,
                                  long long P, const double* __restrict__ u, const double* __restrict__ eta,
                                  double* __restrict__ sr_out, double* __restrict__ inv_out,
                                  double* __restrict__ st_out) {
  constexpr int NEN = DIM == 2 ? 4 : 8, NST = DIM == 2 ? 3 : 6;
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= P) return;
  const long long e = keys[p];
  double sr[6] = {0, 0, 0, 0, 0, 0}, st[6] = {0, 0, 0, 0, 0, 0}, inv = 0.0;
  if (e < g.E) {
    const long long ijk[3] = {e % g.res[0], (e / g.res[0]) % g.res[1],
                              DIM == 3 ? e / ((long long)g.res[0] * g.res[1]) : 0};
    double gi[3] = {1, 1, 1};
    for (int d = 0; d < DIM; d++) gi[d] = 2.0 / (vcoord(g, d, ijk[d] + 1) - vcoord(g, d, ijk[d]));
    const double4 q4 = pk[p];
    const double xi[3] = {q4.x, q4.y, q4.z};
    const long long sy = g.nn[0], sz = g.nn[0] * g.nn[1];
    const long long n0 = ijk[0] + sy * ijk[1] + sz * ijk[2];
    double L[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int q = 0; q < NEN; q++) {
      double fct[3], sg[3];
      for (int d = 0; d < DIM; d++) {
        sg[d] = ((q >> d) & 1) ? 1.0 : -1.0;
        fct[d] = 1.0 + sg[d] * xi[d];
      }
      const double cN = DIM == 2 ? 0.25 : 0.125;
      double dN[3];
      for (int d = 0; d < DIM; d++) {
        double v = cN * sg[d] * gi[d];
        for (int d2 = 0; d2 < DIM; d2++)
          if (d2 != d) v *= fct[d2];
        dN[d] = v;
      }
      const long long nd = n0 + (q & 1) + ((q >> 1) & 1) * sy + ((q >> 2) & 1) * sz;
      for (int i = 0; i < DIM; i++)
        for (int j = 0; j < DIM; j++) L[i][j] += dN[j] * u[nd * DIM + i];
    }
    if (DIM == 2) {
      sr[0] = L[0][0];
      sr[1] = L[1][1];
      sr[2] = 0.5 * (L[0][1] + L[1][0]);
    } else {
      sr[0] = L[0][0];
      sr[1] = L[1][1];
      sr[2] = L[2][2];
      sr[3] = 0.5 * (L[0][1] + L[1][0]);
      sr[4] = 0.5 * (L[0][2] + L[2][0]);
      sr[5] = 0.5 * (L[1][2] + L[2][1]);
    }
    inv = sym_inv(DIM, sr);
    const double et = eta[p];
    for (int k = 0; k < NST; k++) st[k] = (k < DIM ? 2.0 * et : et * 2.0) * sr[k];
  }
  for (int k = 0; k < NST; k++) {
    sr_out[p * NST + k] = sr[k];
    st_out[p * NST + k] = st[k];
  }
  inv_out[p] = inv;
}

void launch_particle_fields(const Geo& g, const double4* pk, const unsigned* keys, long long P, const double* u,
                            const double* eta, double* sr, double* inv, double* st, hipStream_t st_) {
  if (P <= 0) return;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_particle_fields<3>, dim3(grid_for(P)), dim3(kBlock), 0, st_, g, pk, keys, P, u, eta, sr, inv,
                       st);
  else
    hipLaunchKernelGGL(k_particle_fields<2>, dim3(grid_for(P)), dim3(kBlock), 0, st_, g, pk, keys, P, u, eta, sr, inv,
                       st);
}

void launch_setup(const Geo& g, const SetupArgs& a, double4* pk, const int* off, double* diag, double* f, double* Q,
                  double* eta_out, uint8_t* yielded, hipStream_t st) {
  launch_fill(diag, g.N * g.dim, 0.0, st);
  launch_fill(f, g.N * g.dim, 0.0, st);
  const int ncol = g.dim == 2 ? 4 : 8;
  for (int c = 0; c < ncol; c++) {
    const long long m = colour_count(g, c);
    if (m <= 0) continue;
    if (g.dim == 3)
      hipLaunchKernelGGL(k_setup<3>, dim3(grid_for(m)), dim3(kBlock), 0, st, g, c, m, a, pk, off, diag, f, Q,
                         eta_out, yielded);
    else
      hipLaunchKernelGGL(k_setup<2>, dim3(grid_for(m)), dim3(kBlock), 0, st, g, c, m, a, pk, off, diag, f, Q,
                         eta_out, yielded);
  }
}

// ---------------------------------------------------------------------------
// Particles: location (K9), offsets, remap (K10), RK2 advection (K8).
// ---------------------------------------------------------------------------
__global__ void k_locate(Geo g, const double* __restrict__ coord, long long P, int* __restrict__ keys,
                         int* __restrict__ vals, unsigned* esc) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const double x[3] = {coord[3 * i], coord[3 * i + 1], coord[3 * i + 2]};
  long long el;
  if (search_element(g, x, &el)) {
    keys[i] = (int)el;
  } else {
    keys[i] = (int)g.E;
    atomicAdd(esc, 1u);
  }
  vals[i] = (int)i;
}
void launch_locate(const Geo& g, const double* coord, long long P, int* keys, int* vals, double* S,
                   hipStream_t st) {
  if (P <= 0) return;
  hipLaunchKernelGGL(k_locate, dim3(grid_for(P)), dim3(kBlock), 0, st, g, coord, P, keys, vals,
                     reinterpret_cast<unsigned*>(S + S_ESC));
}

__global__ void k_offsets(const int* __restrict__ keys, long long P, long long E, int* __restrict__ off) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i > P) return;
  const long long kprev = i == 0 ? -1 : keys[i - 1];
  long long kcur = i < P ? keys[i] : E;
  if (kcur > E) kcur = E;
  for (long long e = kprev + 1; e <= kcur; e++) off[e] = (int)i;
}
void launch_offsets(const int* keys_sorted, long long P, long long E, int* off, hipStream_t st) {
  hipLaunchKernelGGL(k_offsets, dim3(grid_for(P + 1)), dim3(kBlock), 0, st, keys_sorted, P, E, off);
}

// CoincidentMapper global -> local: for an affine element the Newton
// iteration of ElementType.cxx:221-333 converges to xi = (x - centre) * 2/h.
// Local coordinates of a global point in element ijk by the reference's
// Newton-Raphson iteration (ElementType.cxx:221-333): from the element centre,
// J dxi = x - sum_n N_n(xi) x_n with J = sum_n dN_n/dxi x_n, solved by
// Cramer's rule, until max |dxi| < 1e-4 (at most 100 iterations).  The
// operations follow the oracle's restatement one for one (nodes in order, no
// contraction), so both give the same bits; particles placed exactly on the
// sub-grid of the Voronoi weights (PCDVC's Inflow) make ties that depend on
// them.  Node n of the box element sits at (lo or hi) per axis by bit d of n.
__device__ void newton_local(const Geo& g, const long long* ijk, const double* x, double* xi) {
#pragma clang fp contract(off)
  const int dim = g.dim, nen = dim == 2 ? 4 : 8;
  const double c = dim == 2 ? 0.25 : 0.125;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < dim; d++) {
    lo[d] = vcoord(g, d, ijk[d]);
    hi[d] = vcoord(g, d, ijk[d] + 1);
  }
  for (int d = 0; d < 3; d++) xi[d] = 0.0;
  for (int it = 0; it < 100; it++) {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, rhs[3] = {0, 0, 0}, inc[3] = {0, 0, 0};
    for (int a = 0; a < nen; a++) {
      double s[3], f[3], xn[3];
      for (int d = 0; d < dim; d++) {
        s[d] = ((a >> d) & 1) ? 1.0 : -1.0;
        f[d] = 1.0 + s[d] * xi[d];
        xn[d] = ((a >> d) & 1) ? hi[d] : lo[d];
      }
      double N = c;
      for (int d = 0; d < dim; d++) N *= f[d];
      double gn[3];
      for (int d = 0; d < dim; d++) {
        double v = c * s[d];
        for (int d2 = 0; d2 < dim; d2++)
          if (d2 != d) v *= f[d2];
        gn[d] = v;
      }
      for (int i = 0; i < dim; i++) {
        for (int j = 0; j < dim; j++) J[i][j] += gn[j] * xn[i];
        rhs[i] -= N * xn[i];
      }
    }
    for (int i = 0; i < dim; i++) rhs[i] += x[i];
    if (dim == 2) {
      const double D = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      inc[0] = (rhs[0] * J[1][1] - J[0][1] * rhs[1]) / D;
      inc[1] = (J[0][0] * rhs[1] - rhs[0] * J[1][0]) / D;
    } else {
      const double D = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      for (int cc = 0; cc < 3; cc++) {
        double M[3][3];
        for (int r = 0; r < 3; r++)
          for (int q = 0; q < 3; q++) M[r][q] = q == cc ? rhs[r] : J[r][q];
        inc[cc] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                   M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                   M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) / D;
      }
    }
    double mr = 0.0;
    for (int d = 0; d < dim; d++) {
      xi[d] += inc[d];
      if (fabs(inc[d]) > mr) mr = fabs(inc[d]);
    }
    if (mr < 1e-4) break;
  }
}

// One thread per (sorted) particle; particles outside every element (owner
// E) keep their previous values.
__global__ void k_remap(Geo g, const double* __restrict__ coord, const unsigned* __restrict__ owner,
                        double4* __restrict__ pk, long long P) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= P) return;
  const long long e = owner[p];
  if (e >= g.E) return;
  const long long ijk[3] = {e % g.res[0], (e / g.res[0]) % g.res[1],
                            g.dim == 3 ? e / ((long long)g.res[0] * g.res[1]) : 0};
  const double x[3] = {coord[3 * p], coord[3 * p + 1], coord[3 * p + 2]};
  double xi[3];
  newton_local(g, ijk, x, xi);
  pk[p] = make_double4(xi[0], xi[1], g.dim == 3 ? xi[2] : 0.0, 0.0);
}
void launch_remap(const Geo& g, const double* coord, const unsigned* owner, double4* pk, long long P, hipStream_t st) {
  if (P <= 0) return;
  hipLaunchKernelGGL(k_remap, dim3(grid_for(P)), dim3(kBlock), 0, st, g, coord, owner, pk, P);
}

// Velocity at a global point (FeVariable_InterpolateValueAt): 0 if outside.
__device__ int interp_velocity(const Geo& g, const double* __restrict__ u, const double* x, double* v) {
  long long e;
  if (!search_element(g, x, &e)) return 0;
  const long long ijk[3] = {e % g.res[0], (e / g.res[0]) % g.res[1],
                            g.dim == 3 ? e / ((long long)g.res[0] * g.res[1]) : 0};
  double xi[3] = {0, 0, 0};
  for (int d = 0; d < g.dim; d++) {
    const double a = vcoord(g, d, ijk[d]), b = vcoord(g, d, ijk[d] + 1);
    xi[d] = (x[d] - 0.5 * (a + b)) * (2.0 / (b - a));
  }
  const long long sy = g.nn[0], sz = g.nn[0] * g.nn[1];
  const long long n0 = ijk[0] + sy * ijk[1] + sz * ijk[2];
  const int nen = g.dim == 2 ? 4 : 8;
  for (int d = 0; d < g.dim; d++) v[d] = 0.0;
  for (int q = 0; q < nen; q++) {
    double nv = g.dim == 2 ? 0.25 : 0.125;
    for (int d = 0; d < g.dim; d++) nv *= 1.0 + (((q >> d) & 1) ? 1.0 : -1.0) * xi[d];
    const long long nd = n0 + (q & 1) + ((q >> 1) & 1) * sy + ((q >> 2) & 1) * sz;
    for (int d = 0; d < g.dim; d++) v[d] += nv * u[nd * g.dim + d];
  }
  for (int d = 0; d < g.dim; d++)
    if (isinf(v[d])) return 0;
  return 1;
}

// TimeIntegrand.cxx:265-335: x_half = x0 + 0.5 dt v(x0); x1 = x0 + dt v(x_half),
// falling back to x0 + dt v(x0) when the midpoint lookup fails.  Position
// updates are not contracted into FMAs so that they round as the reference's
// separate multiply and add.
__global__ void k_advect(Geo g, double* __restrict__ coord, long long P, const double* __restrict__ u, double dt,
                         unsigned* esc) {
#pragma clang fp contract(off)
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P) return;
  double x0[3] = {coord[3 * i], coord[3 * i + 1], coord[3 * i + 2]};
  double v1[3], v2[3], xh[3] = {x0[0], x0[1], x0[2]};
  if (!interp_velocity(g, u, x0, v1)) {
    atomicAdd(esc, 1u);
    return;
  }
  for (int d = 0; d < g.dim; d++) xh[d] = x0[d] + 0.5 * dt * v1[d];
  const int ok = interp_velocity(g, u, xh, v2);
  for (int d = 0; d < g.dim; d++) coord[3 * i + d] = x0[d] + dt * (ok ? v2[d] : v1[d]);
}
void launch_advect_rk2(const Geo& g, double* coord, long long P, const double* u, double dt, double* S,
                       hipStream_t st) {
  if (P <= 0) return;
  hipLaunchKernelGGL(k_advect, dim3(grid_for(P)), dim3(kBlock), 0, st, g, coord, P, u, dt,
                     reinterpret_cast<unsigned*>(S + S_ESC));
}

}  // namespace gs
