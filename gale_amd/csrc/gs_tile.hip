// gs_tile.hip -- tiled matrix-free K apply (particle form and moment form).
//
// One workgroup owns a tile of TX x TY (x TZ) elements.  It stages the
// tile's nodal velocities in LDS with coalesced row loads, evaluates each
// element (one thread per element), accumulates the element forces into an
// LDS node array in 2^dim corner phases (deterministic, no atomics), and then
// stores every node it owns straight to y.  Nodes on a tile's upper faces are
// owned by the next tile: their partial sums go to a per-tile halo buffer and
// a second, node-parallel kernel adds them in a fixed order.  So one apply is
// two launches, every y entry is written once (plus one read-modify-write on
// tile faces), and the result is bitwise reproducible.
#include "gs_internal.h"
#include "gs_tile.h"

#include <math.h>

namespace gs {

namespace {

__device__ __forceinline__ long long xcd_remap(long long b, long long nb) {
  const long long q = nb / 8, rem = nb % 8;
  const long long x = b % 8, i = b / 8;
  return x * q + (x < rem ? x : rem) + i;
}

__device__ __forceinline__ double vcoord_t(const Geo& g, int d, long long idx) {
#pragma clang fp contract(off)
  if (idx <= 0) return g.lo[d];
  if (idx >= g.res[d]) return g.hi[d];
  return g.lo[d] + ((double)idx / (double)g.res[d]) * (g.hi[d] - g.lo[d]);
}

// Lattice indices of entry n of an nx x ny x nz lattice (x fastest); 32-bit
// division when the lattice is small enough (64-bit division is a long
// software sequence on the GPU).
__device__ __forceinline__ void lattice3(long long n, long long nx, long long ny, long long total, long long& i,
                                         long long& j, long long& k) {
  if (total < (1LL << 31)) {
    const unsigned un = (unsigned)n, ux = (unsigned)nx, uy = (unsigned)ny;
    const unsigned q = un / ux;
    i = un - q * ux;
    j = q % uy;
    k = q / uy;
  } else {
    i = n % nx;
    j = (n / nx) % ny;
    k = n / (nx * ny);
  }
}

template <typename T>
struct V4;
template <>
struct V4<double> {
  using type = double4;
};
template <>
struct V4<float> {
  using type = float4;
};

#define MI(a, b, c) ((a) + 3 * (b) + 9 * (c))

__device__ __forceinline__ double wave_sum_bot(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Completes dof i with ax = (K x)_i (see TileEpi); xv = x_i.
template <typename T, int EPI>
__device__ __forceinline__ void epi_store(const TileEpi<T>& ep, T* __restrict__ y, long long i, T ax, T xv) {
  if (EPI == 0) {
    y[i] = ax;
  } else if (EPI == 1) {
    const T rr = ep.b[i] - ax;
    const T di = ep.alpha * ep.d[i] + ep.beta * ep.dinv[i] * rr;
    ep.d[i] = di;
    ep.xout[i] = xv + di;
  } else {
    ep.r[i] = ep.b[i] - ax;
  }
}

// ---- 3D element kernels: gradient coefficients from the 8 nodal vectors ----
template <typename T>
__device__ __forceinline__ void grad_coeffs3(const T (&u)[8][3], T ix, T iy, T iz, T (&A)[3][4], T (&B)[3][4],
                                             T (&C)[3][4]) {
#pragma unroll
  for (int cc = 0; cc < 3; cc++) {
    T v[8];
#pragma unroll
    for (int a = 0; a < 8; a++) v[a] = u[a][cc];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int bit = 1 << d;
#pragma unroll
      for (int a = 0; a < 8; a++)
        if (!(a & bit)) {
          const T lo = v[a], hi = v[a | bit];
          v[a] = lo + hi;
          v[a | bit] = hi - lo;
        }
    }
    A[cc][0] = ix * v[1]; A[cc][1] = ix * v[3]; A[cc][2] = ix * v[5]; A[cc][3] = ix * v[7];
    B[cc][0] = iy * v[2]; B[cc][1] = iy * v[3]; B[cc][2] = iy * v[6]; B[cc][3] = iy * v[7];
    C[cc][0] = iz * v[4]; C[cc][1] = iz * v[5]; C[cc][2] = iz * v[6]; C[cc][3] = iz * v[7];
  }
}

// Nodal forces from the 30 stress moments (see gs_kernels.hip for the layout):
//   ye[a][0] = ix ax P(ay,az) + iy ay Q(ax,az) + iz az R(ax,ay), ...
// where every P, Q, R is a two-sign combination p0 + s p1 + t p2 + s t p3,
// so each is evaluated for its four sign pairs with one 2x2 Hadamard step.
template <typename T>
__device__ __forceinline__ void had4(T p0, T p1, T p2, T p3, T sc, T (&h)[2][2]) {
  // h[(s+1)/2][(t+1)/2] = sc (p0 + s p1 + t p2 + s t p3)
  p0 *= sc; p1 *= sc; p2 *= sc; p3 *= sc;
  const T am = p0 - p1, ap = p0 + p1, bm = p2 - p3, bp = p2 + p3;
  h[0][0] = am - bm; h[0][1] = am + bm;
  h[1][0] = ap - bp; h[1][1] = ap + bp;
}

template <typename T>
__device__ __forceinline__ void epilogue3(const T (&m)[30], T ix, T iy, T iz, T (&ye)[8][3]) {
  T P[2][2], Q[2][2], R[2][2];
  // component x: P(ay,az) from m0..3, Q(ax,az) from m12,13,15,16, R(ax,ay) from m18,19,20,22
  had4(m[0], m[1], m[2], m[3], ix, P);
  had4(m[12], m[13], m[15], m[16], iy, Q);
  had4(m[18], m[19], m[20], m[22], iz, R);
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const int sx = a & 1, sy = (a >> 1) & 1, sz = (a >> 2) & 1;
    const T p = sx ? P[sy][sz] : -P[sy][sz], q = sy ? Q[sx][sz] : -Q[sx][sz], r = sz ? R[sx][sy] : -R[sx][sy];
    ye[a][0] = p + q + r;
  }
  // component y: P(ax,az) from m4..7 (iy), Q(ay,az) from m12,14,15,17 (ix), R(ax,ay) from m24,25,26,28 (iz)
  had4(m[4], m[5], m[6], m[7], iy, P);
  had4(m[12], m[14], m[15], m[17], ix, Q);
  had4(m[24], m[25], m[26], m[28], iz, R);
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const int sx = a & 1, sy = (a >> 1) & 1, sz = (a >> 2) & 1;
    const T p = sy ? P[sx][sz] : -P[sx][sz], q = sx ? Q[sy][sz] : -Q[sy][sz], r = sz ? R[sx][sy] : -R[sx][sy];
    ye[a][1] = p + q + r;
  }
  // component z: P(ax,ay) from m8..11 (iz), Q(ay,az) from m18,20,21,23 (ix), R(ax,az) from m24,25,27,29 (iy)
  had4(m[8], m[9], m[10], m[11], iz, P);
  had4(m[18], m[20], m[21], m[23], ix, Q);
  had4(m[24], m[25], m[27], m[29], iy, R);
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const int sx = a & 1, sy = (a >> 1) & 1, sz = (a >> 2) & 1;
    const T p = sz ? P[sx][sy] : -P[sx][sy], q = sx ? Q[sy][sz] : -Q[sy][sz], r = sy ? R[sx][sz] : -R[sx][sz];
    ye[a][2] = p + q + r;
  }
}

template <typename T>
__device__ __forceinline__ void element_particles3(const T (&A)[3][4], const T (&B)[3][4], const T (&C)[3][4],
                                                   const typename V4<T>::type* __restrict__ pk, int pb, int pe,
                                                   T (&m)[30]) {
#pragma unroll
  for (int k = 0; k < 30; k++) m[k] = 0;
  // particles in batches of four whose loads are issued together (one
  // memory round trip per batch); a slot past the element's last particle
  // carries weight 0 and adds exact zeros
  constexpr int PB = 4;
  for (int p0 = pb; p0 < pe; p0 += PB) {
    typename V4<T>::type qb[PB];
#pragma unroll
    for (int k = 0; k < PB; k++) {
      if (p0 + k < pe) {
        qb[k] = pk[p0 + k];
      } else {
        qb[k].x = 0; qb[k].y = 0; qb[k].z = 0; qb[k].w = 0;
      }
    }
#pragma unroll
    for (int k = 0; k < PB; k++) {
    const typename V4<T>::type q = qb[k];
    const T r = q.x, s = q.y, tt = q.z, w = q.w;
    const T st = s * tt, rt = r * tt, rs = r * s;
    const T uxx = A[0][0] + A[0][1] * s + A[0][2] * tt + A[0][3] * st;
    const T uxy = B[0][0] + B[0][1] * r + B[0][2] * tt + B[0][3] * rt;
    const T uxz = C[0][0] + C[0][1] * r + C[0][2] * s + C[0][3] * rs;
    const T uyx = A[1][0] + A[1][1] * s + A[1][2] * tt + A[1][3] * st;
    const T uyy = B[1][0] + B[1][1] * r + B[1][2] * tt + B[1][3] * rt;
    const T uyz = C[1][0] + C[1][1] * r + C[1][2] * s + C[1][3] * rs;
    const T uzx = A[2][0] + A[2][1] * s + A[2][2] * tt + A[2][3] * st;
    const T uzy = B[2][0] + B[2][1] * r + B[2][2] * tt + B[2][3] * rt;
    const T uzz = C[2][0] + C[2][1] * r + C[2][2] * s + C[2][3] * rs;
    const T w2 = w + w;
    const T sxx = w2 * uxx, syy = w2 * uyy, szz = w2 * uzz;
    const T sxy = w * (uxy + uyx), sxz = w * (uxz + uzx), syz = w * (uyz + uzy);
    m[0] += sxx; m[1] += sxx * s; m[2] += sxx * tt; m[3] += sxx * st;
    m[4] += syy; m[5] += syy * r; m[6] += syy * tt; m[7] += syy * rt;
    m[8] += szz; m[9] += szz * r; m[10] += szz * s; m[11] += szz * rs;
    m[12] += sxy; m[13] += sxy * r; m[14] += sxy * s; m[15] += sxy * tt; m[16] += sxy * rt; m[17] += sxy * st;
    m[18] += sxz; m[19] += sxz * r; m[20] += sxz * s; m[21] += sxz * tt; m[22] += sxz * rs; m[23] += sxz * st;
    m[24] += syz; m[25] += syz * r; m[26] += syz * s; m[27] += syz * tt; m[28] += syz * rs; m[29] += syz * rt;
    }
  }
}

// The 27 moments of the element at place `local` of apply tile `tile`
// (tile-major layout, gs_tile.h): one 16-byte load per group of
// G = 16 / sizeof(T) moments, consecutive threads reading consecutive
// 16-byte slots.  NT marks the loads non-temporal: a stream larger than the
// last-level cache is read once per apply and should not evict the node
// vectors; a small level's moments are reused by the next applies of the
// cycle and are loaded normally.
typedef double dvec2_t __attribute__((ext_vector_type(2)));
typedef float fvec4_t __attribute__((ext_vector_type(4)));
template <bool NT, typename V>
__device__ __forceinline__ V ld_mom(const V* p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT>
__device__ __forceinline__ void load_tile_moments3(const double* __restrict__ Mom, long long tile, int local,
                                                   double (&M)[27]) {
  const double* base = Mom + tile * 27 * kTileThreads;
  const dvec2_t* b2 = reinterpret_cast<const dvec2_t*>(base) + local;
#pragma unroll
  for (int j = 0; j < 13; j++) {
    const dvec2_t v = ld_mom<NT>(b2 + j * kTileThreads);
    M[2 * j] = v.x;
    M[2 * j + 1] = v.y;
  }
  M[26] = ld_mom<NT>(base + 26 * kTileThreads + local);
}
template <bool NT>
__device__ __forceinline__ void load_tile_moments3(const float* __restrict__ Mom, long long tile, int local,
                                                   float (&M)[27]) {
  const float* base = Mom + tile * 27 * kTileThreads;
  const fvec4_t* b4 = reinterpret_cast<const fvec4_t*>(base) + local;
#pragma unroll
  for (int j = 0; j < 6; j++) {
    const fvec4_t v = ld_mom<NT>(b4 + j * kTileThreads);
    M[4 * j] = v.x;
    M[4 * j + 1] = v.y;
    M[4 * j + 2] = v.z;
    M[4 * j + 3] = v.w;
  }
#pragma unroll
  for (int k = 24; k < 27; k++) M[k] = ld_mom<NT>(base + k * kTileThreads + local);
}

// The 27 moments of the element at lattice place (gi, gj, gk).
template <typename T>
__device__ __forceinline__ void load_moments3(const Geo& g, const T* __restrict__ Mom, int gi, int gj, int gk,
                                              T (&M)[27]) {
  const int tnx = (g.res[0] + kTile3X - 1) / kTile3X, tny = (g.res[1] + kTile3Y - 1) / kTile3Y;
  const long long tile = gi / kTile3X + (long long)tnx * (gj / kTile3Y + (long long)tny * (gk / kTile3Z));
  const int local = gi % kTile3X + kTile3X * (gj % kTile3Y + kTile3Y * (gk % kTile3Z));
  load_tile_moments3<false>(Mom, tile, local, M);
}

template <typename T>
__device__ __forceinline__ void element_moments3(const T (&A)[3][4], const T (&B)[3][4], const T (&C)[3][4],
                                                 const T (&M)[27], T (&m)[30]) {
  constexpr int XE[4][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}};
  constexpr int YE[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1}};
  constexpr int ZE[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}};
  constexpr int PXY[6][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}};
  constexpr int PXZ[6][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}};
  constexpr int PYZ[6][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    T vx = 0, vy = 0, vz = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      vx += A[0][k] * M[MI(XE[k][0] + XE[j][0], XE[k][1] + XE[j][1], XE[k][2] + XE[j][2])];
      vy += B[1][k] * M[MI(YE[k][0] + YE[j][0], YE[k][1] + YE[j][1], YE[k][2] + YE[j][2])];
      vz += C[2][k] * M[MI(ZE[k][0] + ZE[j][0], ZE[k][1] + ZE[j][1], ZE[k][2] + ZE[j][2])];
    }
    m[j] = (T)2 * vx;
    m[4 + j] = (T)2 * vy;
    m[8 + j] = (T)2 * vz;
  }
#pragma unroll
  for (int j = 0; j < 6; j++) {
    T vxy = 0, vxz = 0, vyz = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      vxy += B[0][k] * M[MI(YE[k][0] + PXY[j][0], YE[k][1] + PXY[j][1], YE[k][2] + PXY[j][2])] +
             A[1][k] * M[MI(XE[k][0] + PXY[j][0], XE[k][1] + PXY[j][1], XE[k][2] + PXY[j][2])];
      vxz += C[0][k] * M[MI(ZE[k][0] + PXZ[j][0], ZE[k][1] + PXZ[j][1], ZE[k][2] + PXZ[j][2])] +
             A[2][k] * M[MI(XE[k][0] + PXZ[j][0], XE[k][1] + PXZ[j][1], XE[k][2] + PXZ[j][2])];
      vyz += C[1][k] * M[MI(ZE[k][0] + PYZ[j][0], ZE[k][1] + PYZ[j][1], ZE[k][2] + PYZ[j][2])] +
             B[2][k] * M[MI(YE[k][0] + PYZ[j][0], YE[k][1] + PYZ[j][1], YE[k][2] + PYZ[j][2])];
    }
    m[12 + j] = vxy;
    m[18 + j] = vxz;
    m[24 + j] = vyz;
  }
}

// ---- 2D element kernels ----
template <typename T>
__device__ __forceinline__ void element2(const T (&u)[4][2], T ix, T iy, int form, const T* __restrict__ Mom,
                                         long long E, long long e, const typename V4<T>::type* __restrict__ pk,
                                         int pb, int pe, T (&ye)[4][2]) {
  T A[2][2], B[2][2];
#pragma unroll
  for (int cc = 0; cc < 2; cc++) {
    const T u0 = u[0][cc], u1 = u[1][cc], u2 = u[2][cc], u3 = u[3][cc];
    const T t1 = (u1 - u0) + (u3 - u2), t2 = (u2 + u3) - (u0 + u1), t3 = (u3 - u2) - (u1 - u0);
    A[cc][0] = ix * t1; A[cc][1] = ix * t3;
    B[cc][0] = iy * t2; B[cc][1] = iy * t3;
  }
  T mxx0 = 0, mxxs = 0, myy0 = 0, myyr = 0, txy0 = 0, txyr = 0, txys = 0;
  if (form == 0) {
    T M[9];
#pragma unroll
    for (int k = 0; k < 9; k++) M[k] = Mom[k * E + e];
    mxx0 = 2 * (A[0][0] * M[0] + A[0][1] * M[3]);
    mxxs = 2 * (A[0][0] * M[3] + A[0][1] * M[6]);
    myy0 = 2 * (B[1][0] * M[0] + B[1][1] * M[1]);
    myyr = 2 * (B[1][0] * M[1] + B[1][1] * M[2]);
    txy0 = B[0][0] * M[0] + B[0][1] * M[1] + A[1][0] * M[0] + A[1][1] * M[3];
    txyr = B[0][0] * M[1] + B[0][1] * M[2] + A[1][0] * M[1] + A[1][1] * M[4];
    txys = B[0][0] * M[3] + B[0][1] * M[4] + A[1][0] * M[3] + A[1][1] * M[6];
  } else {
    for (int p = pb; p < pe; p++) {
      const typename V4<T>::type q = pk[p];
      const T r = q.x, s = q.y, w = q.w;
      const T uxx = A[0][0] + A[0][1] * s, uxy = B[0][0] + B[0][1] * r;
      const T uyx = A[1][0] + A[1][1] * s, uyy = B[1][0] + B[1][1] * r;
      const T w2 = w + w;
      const T sxx = w2 * uxx, syy = w2 * uyy, sxy = w * (uxy + uyx);
      mxx0 += sxx; mxxs += sxx * s;
      myy0 += syy; myyr += syy * r;
      txy0 += sxy; txyr += sxy * r; txys += sxy * s;
    }
  }
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const T ax = (a & 1) ? (T)1 : (T)-1, ay = (a & 2) ? (T)1 : (T)-1;
    ye[a][0] = ix * ax * (mxx0 + ay * mxxs) + iy * ay * (txy0 + ax * txyr);
    ye[a][1] = iy * ay * (myy0 + ax * myyr) + ix * ax * (txy0 + ay * txys);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// 3D tile kernel
// ---------------------------------------------------------------------------
// The FP32 moment form is held to 5 waves per SIMD (86 VGPRs, no spills;
// 118 and 4 waves unconstrained); the other forms keep the compiler's choice.
template <typename T, int FORM>
__global__ __launch_bounds__(kTileThreads)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && FORM == 0 ? 5 : 1))) void
k_tile_apply3(Geo g, TileGrid tg, const T* __restrict__ x,
                                                              T* __restrict__ y, T* __restrict__ halo,
                                                              const T* __restrict__ Mom,
                                                              const typename V4<T>::type* __restrict__ pk,
                                                              const int* __restrict__ off,
                                                              const uint8_t* __restrict__ bcflag, int mom_nt) {
  constexpr int TX = kTile3X, TY = kTile3Y, TZ = kTile3Z;
  constexpr int NBX = TX + 1, NBY = TY + 1, NBZ = TZ + 1, NB = NBX * NBY * NBZ;
  // node box, one array per component (consecutive elements hit consecutive
  // LDS words in the corner phases)
  __shared__ T sx[3][NB];
  __shared__ T sy[3][NB];
  __shared__ T sih[TX + TY + TZ];  // 1 / (4 h) of the tile's columns, rows, layers
  const long long tile = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned tq = (unsigned)tile / (unsigned)tg.n[0];  // the grid has < 2^32 tiles
  const int tx = (int)((unsigned)tile - tq * (unsigned)tg.n[0]), ty = (int)(tq % (unsigned)tg.n[1]),
            tz = (int)(tq / (unsigned)tg.n[1]);
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int nbx = min(TX, g.res[0] - x0) + 1, nby = min(TY, g.res[1] - y0) + 1, nbz = min(TZ, g.res[2] - z0) + 1;
  const long long sy_ = g.nn[0], sz_ = g.nn[0] * g.nn[1];
  const int ex = threadIdx.x % TX, ey = (threadIdx.x / TX) % TY, ez = threadIdx.x / (TX * TY);
  const bool active = (x0 + ex < g.res[0]) && (y0 + ey < g.res[1]) && (z0 + ez < g.res[2]);
  const long long e = (x0 + ex) + (long long)g.res[0] * ((y0 + ey) + (long long)g.res[1] * (z0 + ez));
  // the element's moments are requested before the node box, so both
  // memory round trips overlap
  T M[27];
  if (FORM == 0 && active) {
    if (mom_nt)
      load_tile_moments3<true>(Mom, tile, (int)threadIdx.x, M);
    else
      load_tile_moments3<false>(Mom, tile, (int)threadIdx.x, M);
  }
  // Node box: every thread's (at most NQ) values and their Dirichlet flags
  // are requested at once, so the box costs one memory round trip; the
  // flags are kept for the stores at the end (same value-to-thread map).
  constexpr int NQ = (NB * 3 + kTileThreads - 1) / kTileThreads;
  T vq[NQ];
  unsigned mq = 0;  // bit it: value it is a Dirichlet dof
#pragma unroll
  for (int it = 0; it < NQ; it++) {
    const int q = threadIdx.x + it * kTileThreads;
    vq[it] = 0;
    if (q < NB * 3) {
      const int c = q % 3, nd = q / 3;
      const int lx = nd % NBX, ly = (nd / NBX) % NBY, lz = nd / (NBX * NBY);
      if (lx < nbx && ly < nby && lz < nbz) {
        const long long gn = (x0 + lx) + sy_ * (y0 + ly) + sz_ * (z0 + lz);
        vq[it] = x[gn * 3 + c];
        if (bcflag) mq |= ((bcflag[gn] >> c) & 1u) << it;
      }
    }
  }
#pragma unroll
  for (int it = 0; it < NQ; it++) {
    const int q = threadIdx.x + it * kTileThreads;
    if (q < NB * 3) {
      const int c = q % 3, nd = q / 3;
      sx[c][nd] = vq[it];
      sy[c][nd] = 0;
    }
  }
  if (threadIdx.x < TX + TY + TZ) {
    const int t = threadIdx.x;
    const int d = t < TX ? 0 : (t < TX + TY ? 1 : 2);
    const long long i = d == 0 ? x0 + t : (d == 1 ? y0 + t - TX : z0 + t - TX - TY);
    sih[t] = i < g.res[d] ? (T)(0.25 / (vcoord_t(g, d, i + 1) - vcoord_t(g, d, i))) : (T)0;
  }
  __syncthreads();
  T ye[8][3];
  if (active) {
    const T ix = sih[ex], iy = sih[TX + ey], iz = sih[TX + TY + ez];
    T u[8][3];
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int ln = (ex + (a & 1)) + NBX * ((ey + ((a >> 1) & 1)) + NBY * (ez + ((a >> 2) & 1)));
#pragma unroll
      for (int c = 0; c < 3; c++) u[a][c] = sx[c][ln];
    }
    T A[3][4], B[3][4], C[3][4];
    grad_coeffs3(u, ix, iy, iz, A, B, C);
    T m[30];
    if (FORM == 0)
      element_moments3(A, B, C, M, m);
    else
      element_particles3(A, B, C, pk, off[e], off[e + 1], m);
    epilogue3(m, ix, iy, iz, ye);
  }
  // One phase per element corner a: for fixed a, element -> node (e + a) is
  // injective, so all threads add without conflicts; each node receives its
  // (up to 8) contributions in corner order, so the sum is reproducible.
#pragma unroll
  for (int a = 0; a < 8; a++) {
    if (active) {
      const int ln = (ex + (a & 1)) + NBX * ((ey + ((a >> 1) & 1)) + NBY * (ez + ((a >> 2) & 1)));
#pragma unroll
      for (int c = 0; c < 3; c++) sy[c][ln] += ye[a][c];
    }
    __syncthreads();
  }
  const bool lastx = tx == tg.n[0] - 1, lasty = ty == tg.n[1] - 1, lastz = tz == tg.n[2] - 1;
#pragma unroll
  for (int it = 0; it < NQ; it++) {
    const int q = threadIdx.x + it * kTileThreads;
    if (q >= NB * 3) continue;
    const int c = q % 3, nd = q / 3;
    const int lx = nd % NBX, ly = (nd / NBX) % NBY, lz = nd / (NBX * NBY);
    if (lx >= nbx || ly >= nby || lz >= nbz) continue;
    const bool hal = (lx == nbx - 1 && !lastx) || (ly == nby - 1 && !lasty) || (lz == nbz - 1 && !lastz);
    if (hal) {
      halo[(tile * NB + nd) * 3 + c] = sy[c][nd];
    } else {
      const long long gn = (x0 + lx) + sy_ * (y0 + ly) + sz_ * (z0 + lz);
      y[gn * 3 + c] = ((mq >> it) & 1u) ? (T)0 : sy[c][nd];
    }
  }
}

// Completes the tile-face nodes of a plain apply (no fused update): one
// thread per node, only nodes on a lower tile face (about 40 %) do any work,
// adding the halo partial sums of up to seven lower neighbour tiles.
template <typename T>
__global__ void k_tile_face3(Geo g, TileGrid tg, T* __restrict__ y, const T* __restrict__ halo,
                             const uint8_t* __restrict__ bcflag) {
  constexpr int TX = kTile3X, TY = kTile3Y, TZ = kTile3Z;
  constexpr int NBX = TX + 1, NBY = TY + 1, NBZ = TZ + 1, NB = NBX * NBY * NBZ;
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  int i, j, k;
  if (g.N < (1LL << 31)) {
    const unsigned un = (unsigned)n, nx = (unsigned)g.nn[0], ny = (unsigned)g.nn[1];
    const unsigned q = un / nx;
    i = (int)(un - q * nx);
    j = (int)(q % ny);
    k = (int)(q / ny);
  } else {
    i = (int)(n % g.nn[0]);
    j = (int)((n / g.nn[0]) % g.nn[1]);
    k = (int)(n / (g.nn[0] * g.nn[1]));
  }
  const int tx = min(i / TX, tg.n[0] - 1), ty = min(j / TY, tg.n[1] - 1), tz = min(k / TZ, tg.n[2] - 1);
  const int onx = (i == tx * TX && tx > 0), ony = (j == ty * TY && ty > 0), onz = (k == tz * TZ && tz > 0);
  if (!(onx | ony | onz)) return;
  T acc[3] = {0, 0, 0};
  for (int az = 0; az <= onz; az++)
    for (int ay = 0; ay <= ony; ay++)
      for (int ax = 0; ax <= onx; ax++) {
        if (!(ax | ay | az)) continue;
        const int ux = tx - ax, uy = ty - ay, uz = tz - az;
        const long long t2 = ux + (long long)tg.n[0] * (uy + (long long)tg.n[1] * uz);
        const int ln = (i - ux * TX) + NBX * ((j - uy * TY) + NBY * (k - uz * TZ));
        const T* h = halo + (t2 * NB + ln) * 3;
        acc[0] += h[0];
        acc[1] += h[1];
        acc[2] += h[2];
      }
  const unsigned bf = bcflag ? bcflag[n] : 0u;
  for (int c = 0; c < 3; c++)
    if (!((bf >> c) & 1u)) y[n * 3 + c] += acc[c];
}

// Completes an apply with a fused node update (EPI != 0; EPI == 0 is
// k_tile_face3): every dof passes through here, so the update streams all
// vectors once after the apply, adding on tile-face nodes the halo partial
// sums of up to seven lower neighbour tiles.  One thread per dof (node-major:
// dof i = 3 n + c), so every vector access of a wavefront is one contiguous
// run (a thread per node with three strided accesses moved about twice the
// bytes); blocks are remapped so each XCD works on a contiguous range.
template <typename T, int EPI>
__global__ void k_tile_halo3(Geo g, TileGrid tg, const T* __restrict__ x, T* __restrict__ y,
                             const T* __restrict__ halo, const uint8_t* __restrict__ bcflag, TileEpi<T> ep) {
  constexpr int TX = kTile3X, TY = kTile3Y, TZ = kTile3Z;
  constexpr int NBX = TX + 1, NBY = TY + 1, NBZ = TZ + 1, NB = NBX * NBY * NBZ;
  const long long id = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (id >= g.N * 3) return;
  long long n;
  int c, i, j, k;
  if (g.N * 3 < (1LL << 31)) {  // 32-bit lattice indices (64-bit division is a long software sequence)
    const unsigned uid = (unsigned)id;
    const unsigned un = uid / 3u;
    c = (int)(uid - 3u * un);
    n = un;
    const unsigned nx = (unsigned)g.nn[0], ny = (unsigned)g.nn[1];
    const unsigned q = un / nx;
    i = (int)(un - q * nx);
    j = (int)(q % ny);
    k = (int)(q / ny);
  } else {
    n = id / 3;
    c = (int)(id - 3 * n);
    i = (int)(n % g.nn[0]);
    j = (int)((n / g.nn[0]) % g.nn[1]);
    k = (int)(n / (g.nn[0] * g.nn[1]));
  }
  const int tx = min(i / TX, tg.n[0] - 1), ty = min(j / TY, tg.n[1] - 1), tz = min(k / TZ, tg.n[2] - 1);
  const int onx = (i == tx * TX && tx > 0), ony = (j == ty * TY && ty > 0), onz = (k == tz * TZ && tz > 0);
  if (EPI == 0 && !(onx | ony | onz)) return;
  T acc = 0;
  for (int az = 0; az <= onz; az++)
    for (int ay = 0; ay <= ony; ay++)
      for (int ax = 0; ax <= onx; ax++) {
        if (!(ax | ay | az)) continue;
        const int ux = tx - ax, uy = ty - ay, uz = tz - az;
        const long long t2 = ux + (long long)tg.n[0] * (uy + (long long)tg.n[1] * uz);
        const int ln = (i - ux * TX) + NBX * ((j - uy * TY) + NBY * (k - uz * TZ));
        acc += halo[(t2 * NB + ln) * 3 + c];
      }
  const bool fixed = bcflag ? ((bcflag[n] >> c) & 1u) : false;
  if (EPI == 0) {
    if (!fixed) y[id] += acc;
  } else {
    const T ax = fixed ? (T)0 : y[id] + acc;
    epi_store<T, EPI>(ep, y, id, ax, x[id]);
  }
}

// ---------------------------------------------------------------------------
// 2D tile kernel
// ---------------------------------------------------------------------------
template <typename T, int FORM>
__global__ __launch_bounds__(kTileThreads) void k_tile_apply2(Geo g, TileGrid tg, const T* __restrict__ x,
                                                              T* __restrict__ y, T* __restrict__ halo,
                                                              const T* __restrict__ Mom,
                                                              const typename V4<T>::type* __restrict__ pk,
                                                              const int* __restrict__ off,
                                                              const uint8_t* __restrict__ bcflag) {
  constexpr int TX = kTile2X, TY = kTile2Y;
  constexpr int NBX = TX + 1, NBY = TY + 1, NB = NBX * NBY;
  __shared__ T sx[NB * 2];
  __shared__ T sy[NB * 2];
  const long long tile = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(tile % tg.n[0]), ty = (int)(tile / tg.n[0]);
  const int x0 = tx * TX, y0 = ty * TY;
  const int nbx = min(TX, g.res[0] - x0) + 1, nby = min(TY, g.res[1] - y0) + 1;
  const long long sy_ = g.nn[0];
  for (int q = threadIdx.x; q < NB * 2; q += kTileThreads) {
    const int c = q % 2, nd = q / 2;
    const int lx = nd % NBX, ly = nd / NBX;
    T v = 0;
    if (lx < nbx && ly < nby) v = x[((x0 + lx) + sy_ * (y0 + ly)) * 2 + c];
    sx[q] = v;
    sy[q] = 0;
  }
  __syncthreads();
  const int ex = threadIdx.x % TX, ey = threadIdx.x / TX;
  const bool active = (x0 + ex < g.res[0]) && (y0 + ey < g.res[1]);
  T ye[4][2];
  if (active) {
    const long long gi = x0 + ex, gj = y0 + ey;
    const T ix = (T)(0.5 / (vcoord_t(g, 0, gi + 1) - vcoord_t(g, 0, gi)));
    const T iy = (T)(0.5 / (vcoord_t(g, 1, gj + 1) - vcoord_t(g, 1, gj)));
    T u[4][2];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int ln = (ex + (a & 1)) + NBX * (ey + ((a >> 1) & 1));
      u[a][0] = sx[ln * 2];
      u[a][1] = sx[ln * 2 + 1];
    }
    const long long e = gi + (long long)g.res[0] * gj;
    element2(u, ix, iy, FORM, Mom, g.E, e, pk, FORM ? off[e] : 0, FORM ? off[e + 1] : 0, ye);
  }
#pragma unroll
  for (int a = 0; a < 4; a++) {  // one conflict-free phase per corner (see the 3D kernel)
    if (active) {
      const int ln = (ex + (a & 1)) + NBX * (ey + ((a >> 1) & 1));
      sy[ln * 2] += ye[a][0];
      sy[ln * 2 + 1] += ye[a][1];
    }
    __syncthreads();
  }
  const bool lastx = tx == tg.n[0] - 1, lasty = ty == tg.n[1] - 1;
  for (int q = threadIdx.x; q < NB * 2; q += kTileThreads) {
    const int c = q % 2, nd = q / 2;
    const int lx = nd % NBX, ly = nd / NBX;
    if (lx >= nbx || ly >= nby) continue;
    const bool hal = (lx == nbx - 1 && !lastx) || (ly == nby - 1 && !lasty);
    if (hal) {
      halo[(tile * NB + nd) * 2 + c] = sy[q];
    } else {
      const long long gn = (x0 + lx) + sy_ * (y0 + ly);
      const unsigned bf = bcflag ? bcflag[gn] : 0u;
      y[gn * 2 + c] = ((bf >> c) & 1u) ? (T)0 : sy[q];
    }
  }
}

template <typename T, int EPI>
__global__ void k_tile_halo2(Geo g, TileGrid tg, const T* __restrict__ x, T* __restrict__ y,
                             const T* __restrict__ halo, const uint8_t* __restrict__ bcflag, TileEpi<T> ep) {
  constexpr int TX = kTile2X, TY = kTile2Y;
  constexpr int NBX = TX + 1, NBY = TY + 1, NB = NBX * NBY;
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= g.N) return;
  const int i = (int)(n % g.nn[0]), j = (int)(n / g.nn[0]);
  const int tx = min(i / TX, tg.n[0] - 1), ty = min(j / TY, tg.n[1] - 1);
  const int onx = (i == tx * TX && tx > 0), ony = (j == ty * TY && ty > 0);
  if (EPI == 0 && !(onx | ony)) return;
  T acc[2] = {0, 0};
  for (int ay = 0; ay <= ony; ay++)
    for (int ax = 0; ax <= onx; ax++) {
      if (!(ax | ay)) continue;
      const int ux = tx - ax, uy = ty - ay;
      const long long t2 = ux + (long long)tg.n[0] * uy;
      const int ln = (i - ux * TX) + NBX * (j - uy * TY);
      acc[0] += halo[(t2 * NB + ln) * 2];
      acc[1] += halo[(t2 * NB + ln) * 2 + 1];
    }
  const unsigned bf = bcflag ? bcflag[n] : 0u;
  if (EPI == 0) {
    for (int c = 0; c < 2; c++)
      if (!((bf >> c) & 1u)) y[n * 2 + c] += acc[c];
  } else {
    for (int c = 0; c < 2; c++) {
      const T ax = ((bf >> c) & 1u) ? (T)0 : y[n * 2 + c] + acc[c];
      epi_store<T, EPI>(ep, y, n * 2 + c, ax, x[n * 2 + c]);
    }
  }
}

// ---------------------------------------------------------------------------
TileGrid tile_grid(const Geo& g) {
  TileGrid tg{};
  if (g.dim == 3) {
    tg.n[0] = (g.res[0] + kTile3X - 1) / kTile3X;
    tg.n[1] = (g.res[1] + kTile3Y - 1) / kTile3Y;
    tg.n[2] = (g.res[2] + kTile3Z - 1) / kTile3Z;
    tg.nb = (kTile3X + 1) * (kTile3Y + 1) * (kTile3Z + 1);
  } else {
    tg.n[0] = (g.res[0] + kTile2X - 1) / kTile2X;
    tg.n[1] = (g.res[1] + kTile2Y - 1) / kTile2Y;
    tg.n[2] = 1;
    tg.nb = (kTile2X + 1) * (kTile2Y + 1);
  }
  tg.ntiles = (long long)tg.n[0] * tg.n[1] * tg.n[2];
  return tg;
}

long long tile_halo_size(const Geo& g) {
  const TileGrid tg = tile_grid(g);
  return tg.ntiles * tg.nb * g.dim;
}

template <typename T, int FORM, int EPI>
static void tile_launch(const Geo& g, const T* x, T* y, T* halo, const T* Mom, const typename V4<T>::type* pk,
                        const int* off, const uint8_t* bcflag, const TileEpi<T>& ep, hipStream_t st) {
  const TileGrid tg = tile_grid(g);
  const unsigned nbn = (unsigned)((g.N + kBlock - 1) / kBlock);
  if (g.dim == 3) {
    // non-temporal moment loads when the moments exceed a quarter of the
    // 256 MB last-level cache (they are then streamed, not reused)
    const int nt = FORM == 0 && (double)mom_count<T>(g) * sizeof(T) > 64.0e6 ? 1 : 0;
    hipLaunchKernelGGL((k_tile_apply3<T, FORM>), dim3((unsigned)tg.ntiles), dim3(kTileThreads), 0, st, g, tg, x,
                       y, halo, Mom, pk, off, bcflag, nt);
    if (EPI == 0) {
      hipLaunchKernelGGL((k_tile_face3<T>), dim3(nbn), dim3(kBlock), 0, st, g, tg, y, halo, bcflag);
    } else {
      const unsigned nbd = (unsigned)((g.N * 3 + kBlock - 1) / kBlock);
      hipLaunchKernelGGL((k_tile_halo3<T, EPI>), dim3(nbd), dim3(kBlock), 0, st, g, tg, x, y, halo, bcflag, ep);
    }
  } else {
    hipLaunchKernelGGL((k_tile_apply2<T, FORM>), dim3((unsigned)tg.ntiles), dim3(kTileThreads), 0, st, g, tg, x,
                       y, halo, Mom, pk, off, bcflag);
    hipLaunchKernelGGL((k_tile_halo2<T, EPI>), dim3(nbn), dim3(kBlock), 0, st, g, tg, x, y, halo, bcflag, ep);
  }
}

template <typename T, int FORM>
static void tile_apply(const Geo& g, const T* x, T* y, T* halo, const T* Mom, const typename V4<T>::type* pk,
                       const int* off, const uint8_t* bcflag, hipStream_t st) {
  tile_launch<T, FORM, 0>(g, x, y, halo, Mom, pk, off, bcflag, TileEpi<T>{}, st);
}

template <typename T>
static void tile_apply_epi(const Geo& g, const T* x, T* y, T* halo, const T* Mom, const uint8_t* bcflag,
                           const TileEpi<T>& ep, hipStream_t st) {
  switch (ep.kind) {
    case 1: tile_launch<T, 0, 1>(g, x, y, halo, Mom, nullptr, nullptr, bcflag, ep, st); break;
    case 2: tile_launch<T, 0, 2>(g, x, y, halo, Mom, nullptr, nullptr, bcflag, ep, st); break;
    default: tile_launch<T, 0, 0>(g, x, y, halo, Mom, nullptr, nullptr, bcflag, ep, st); break;
  }
}

void tile_apply_particles(const Geo& g, const double* x, double* y, double* halo, const double4* pk, const int* off,
                          const uint8_t* bcflag, hipStream_t st) {
  tile_apply<double, 1>(g, x, y, halo, nullptr, pk, off, bcflag, st);
}
void tile_apply_particles32(const Geo& g, const float* x, float* y, float* halo, const float4* pk, const int* off,
                            const uint8_t* bcflag, hipStream_t st) {
  tile_apply<float, 1>(g, x, y, halo, nullptr, pk, off, bcflag, st);
}
void tile_apply_moments(const Geo& g, const double* x, double* y, double* halo, const double* Mom,
                        const uint8_t* bcflag, hipStream_t st) {
  tile_apply<double, 0>(g, x, y, halo, Mom, nullptr, nullptr, bcflag, st);
}
void tile_apply_moments32(const Geo& g, const float* x, float* y, float* halo, const float* Mom,
                          const uint8_t* bcflag, hipStream_t st) {
  tile_apply<float, 0>(g, x, y, halo, Mom, nullptr, nullptr, bcflag, st);
}

void tile_apply_moments_epi(const Geo& g, const double* x, double* y, double* halo, const double* Mom,
                            const uint8_t* bcflag, const TileEpi<double>& ep, hipStream_t st) {
  tile_apply_epi<double>(g, x, y, halo, Mom, bcflag, ep, st);
}
void tile_apply_moments_epi(const Geo& g, const float* x, float* y, float* halo, const float* Mom,
                            const uint8_t* bcflag, const TileEpi<float>& ep, hipStream_t st) {
  tile_apply_epi<float>(g, x, y, halo, Mom, bcflag, ep, st);
}

// ---------------------------------------------------------------------------
// Single-workgroup bottom of the 3D V-cycle
// ---------------------------------------------------------------------------
namespace {

// sy[0, 3N) += K x (unmasked), elements in chunks of one per thread, each
// chunk's forces added in eight conflict-free corner phases.
template <typename T, int BS>
__device__ void bot_kx(const BotLevel<T>& L, const T* __restrict__ x, T* sy) {
  const Geo& g = L.g;
  const long long E = g.E, sy_ = g.nn[0], sz_ = g.nn[0] * g.nn[1];
  for (long long e0 = 0; e0 < E; e0 += BS) {
    const long long e = e0 + threadIdx.x;
    const bool act = e < E;
    T ye[8][3];
    long long n0 = 0;
    if (act) {
      long long gi, gj, gk;
      lattice3(e, g.res[0], g.res[1], E, gi, gj, gk);
      n0 = gi + sy_ * gj + sz_ * gk;
      const T ix = (T)(0.25 / (vcoord_t(g, 0, gi + 1) - vcoord_t(g, 0, gi)));
      const T iy = (T)(0.25 / (vcoord_t(g, 1, gj + 1) - vcoord_t(g, 1, gj)));
      const T iz = (T)(0.25 / (vcoord_t(g, 2, gk + 1) - vcoord_t(g, 2, gk)));
      T u[8][3];
#pragma unroll
      for (int a = 0; a < 8; a++) {
        const long long nd = n0 + (a & 1) + sy_ * ((a >> 1) & 1) + sz_ * ((a >> 2) & 1);
#pragma unroll
        for (int c = 0; c < 3; c++) u[a][c] = x[nd * 3 + c];
      }
      T A[3][4], B[3][4], C[3][4];
      grad_coeffs3(u, ix, iy, iz, A, B, C);
      T M[27], m[30];
      load_moments3(g, L.M, (int)gi, (int)gj, (int)gk, M);
      element_moments3(A, B, C, M, m);
      epilogue3(m, ix, iy, iz, ye);
    }
#pragma unroll
    for (int a = 0; a < 8; a++) {
      if (act) {
        const long long nd = n0 + (a & 1) + sy_ * ((a >> 1) & 1) + sz_ * ((a >> 2) & 1);
#pragma unroll
        for (int c = 0; c < 3; c++) sy[nd * 3 + c] += ye[a][c];
      }
      __syncthreads();
    }
  }
}

// (K x)_i with the Dirichlet mask; clears the accumulator entry.
template <typename T>
__device__ __forceinline__ T bot_take(const BotLevel<T>& L, T* sy, long long i) {
  const unsigned bf = L.bc ? L.bc[i / 3] : 0u;
  const T v = sy[i];
  sy[i] = 0;
  return ((bf >> (i % 3)) & 1u) ? (T)0 : v;
}

// Chebyshev-Jacobi smoothing (the multi-launch cheb_smooth, step for step);
// x == nullptr starts from zero.  Returns the buffer holding x.
template <typename T, int BS>
__device__ T* bot_cheb(const BotLevel<T>& L, int nu, T* x, T* sy) {
  const long long n = L.g.N * 3;
  const double emax = 1.1 * L.lmax, emin = 0.1 * L.lmax;
  const double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin), sigma = theta / delta;
  double rho_prev = 1.0 / sigma;
  int k0 = 0;
  if (!x) {
    const T beta = (T)(1.0 / theta);
    for (long long i = threadIdx.x; i < n; i += BS) {
      const T di = beta * L.dinv[i] * L.b[i];
      L.d[i] = di;
      L.x[i] = di;
    }
    __syncthreads();
    x = L.x;
    k0 = 1;
  }
  for (int k = k0; k < nu; k++) {
    T alpha = 0, beta = (T)(1.0 / theta);
    if (k > 0) {
      const double rho = 1.0 / (2.0 * sigma - rho_prev);
      alpha = (T)(rho * rho_prev);
      beta = (T)(2.0 * rho / delta);
      rho_prev = rho;
    }
    bot_kx<T, BS>(L, x, sy);
    T* xo = x == L.x ? L.x2 : L.x;
    for (long long i = threadIdx.x; i < n; i += BS) {
      const T rr = L.b[i] - bot_take(L, sy, i);
      const T di = alpha * L.d[i] + beta * L.dinv[i] * rr;
      L.d[i] = di;
      xo[i] = x[i] + di;
    }
    __syncthreads();
    x = xo;
  }
  return x;
}

// C.b = R r_F (full weighting, the multi-launch k_restrict).
template <typename T, int BS>
__device__ void bot_restrict(const BotLevel<T>& C, const BotLevel<T>& F) {
  const Geo &gc = C.g, &gf = F.g;
  for (long long n = threadIdx.x; n < gc.N; n += BS) {
    long long I, J, K;
    lattice3(n, gc.nn[0], gc.nn[1], gc.N, I, J, K);
    T acc[3] = {0, 0, 0};
    for (int dz = -1; dz <= 1; dz++) {
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
          for (int d = 0; d < 3; d++) acc[d] += w * F.r[f * 3 + d];
        }
      }
    }
    const unsigned bf = C.bc ? C.bc[n] : 0u;
    for (int d = 0; d < 3; d++) C.b[n * 3 + d] = ((bf >> d) & 1u) ? (T)0 : acc[d];
  }
  __syncthreads();
}

// xf += P xc (Q1 interpolation, the multi-launch k_prolong_add).
template <typename T, int BS>
__device__ void bot_prolong(const BotLevel<T>& C, const T* __restrict__ xc, const BotLevel<T>& F, T* xf) {
  const Geo &gc = C.g, &gf = F.g;
  for (long long n = threadIdx.x; n < gf.N; n += BS) {
    long long i, j, k;
    lattice3(n, gf.nn[0], gf.nn[1], gf.N, i, j, k);
    T acc[3] = {0, 0, 0};
    const int nz = (k & 1) ? 2 : 1, ny = (j & 1) ? 2 : 1, nx = (i & 1) ? 2 : 1;
    const T wz = nz == 2 ? (T)0.5 : (T)1, wy = ny == 2 ? (T)0.5 : (T)1, wx = nx == 2 ? (T)0.5 : (T)1;
    for (int a = 0; a < nz; a++)
      for (int b = 0; b < ny; b++)
        for (int cc = 0; cc < nx; cc++) {
          const long long cn = ((i >> 1) + cc) + gc.nn[0] * (((j >> 1) + b) + gc.nn[1] * ((k >> 1) + a));
          for (int d = 0; d < 3; d++) acc[d] += wx * wy * wz * xc[cn * 3 + d];
        }
    const unsigned bf = F.bc ? F.bc[n] : 0u;
    for (int d = 0; d < 3; d++)
      if (!((bf >> d) & 1u)) xf[n * 3 + d] += acc[d];
  }
  __syncthreads();
}

template <typename T, int BS>
__global__ __launch_bounds__(BS) void k_mg_bottom3(BotArgs<T> a) {
  extern __shared__ __align__(16) unsigned char bot_smem[];
  T* sy = reinterpret_cast<T*>(bot_smem);
  for (long long i = threadIdx.x; i < a.lv[0].g.N * 3; i += BS) sy[i] = 0;
  __syncthreads();
  const int Lc = a.nlev - 1;
  unsigned in2 = 0;  // bit l: level l's x is in x2
  for (int l = 0; l < Lc; l++) {
    const BotLevel<T>& L = a.lv[l];
    T* x = bot_cheb<T, BS>(L, a.nu, (T*)nullptr, sy);
    in2 |= (x == L.x2 ? 1u : 0u) << l;
    bot_kx<T, BS>(L, x, sy);
    for (long long i = threadIdx.x; i < L.g.N * 3; i += BS) L.r[i] = L.b[i] - bot_take(L, sy, i);
    __syncthreads();
    bot_restrict<T, BS>(a.lv[l + 1], L);
  }
  {
    const BotLevel<T>& L = a.lv[Lc];
    if (a.nfc > 0) {
      for (long long i = threadIdx.x; i < L.g.N * 3; i += BS) L.x[i] = 0;
      __syncthreads();
      const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      for (int row = w; row < a.nfc; row += BS / 64) {
        double s = 0;
        for (int j = lane; j < a.nfc; j += 64) s += (double)(a.Kinv[(long long)row * a.nfc + j] * L.b[a.freeidx[j]]);
        s = wave_sum_bot(s);
        if (lane == 0) L.x[a.freeidx[row]] = (T)s;
      }
      __syncthreads();
    } else {
      T* x = bot_cheb<T, BS>(L, 8, (T*)nullptr, sy);
      in2 |= (x == L.x2 ? 1u : 0u) << Lc;
    }
  }
  for (int l = Lc - 1; l >= 0; l--) {
    const BotLevel<T>& L = a.lv[l];
    const BotLevel<T>& C = a.lv[l + 1];
    T* x = ((in2 >> l) & 1u) ? L.x2 : L.x;
    bot_prolong<T, BS>(C, ((in2 >> (l + 1)) & 1u) ? C.x2 : C.x, L, x);
    x = bot_cheb<T, BS>(L, a.nu, x, sy);
    in2 = (in2 & ~(1u << l)) | ((x == L.x2 ? 1u : 0u) << l);
  }
  if (in2 & 1u) {
    const BotLevel<T>& L = a.lv[0];
    for (long long i = threadIdx.x; i < L.g.N * 3; i += BS) L.x[i] = L.x2[i];
  }
}

// FP32: 1024 threads (the element evaluation fits 128 VGPRs); FP64: 512.
template <typename T>
void bottom_launch(const BotArgs<T>& a, hipStream_t st) {
  constexpr int BS = sizeof(T) == 4 ? 1024 : 512;
  const size_t lds = (size_t)a.lv[0].g.N * 3 * sizeof(T);
  hipLaunchKernelGGL((k_mg_bottom3<T, BS>), dim3(1), dim3(BS), lds, st, a);
}

}  // namespace

void mg_bottom3(const BotArgs<float>& a, hipStream_t st) { bottom_launch<float>(a, st); }
void mg_bottom3(const BotArgs<double>& a, hipStream_t st) { bottom_launch<double>(a, st); }

}  // namespace gs
