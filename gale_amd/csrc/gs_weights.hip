// gs_weights.hip -- discrete Voronoi cell (DVC) integration weights of the
// material points (include/gale_weights.h).
//
// One wavefront per FE cell grows the discrete Voronoi regions of the cell's
// particles on a fixed sub-grid of the local box [-1,1]^dim exactly as
// DVCWeights.cxx:693-786 does, then moves every integration point to the
// centroid of its region and gives it the region's volume as weight
// (DVCWeights.cxx:875-951).  The growth is sequential over particles by
// construction (particle i contests what particles 0..i-1 just took), but
// within one particle's step every boundary sub-cell is decided on its own,
// so the 64 lanes decide a step's sub-cells in parallel: the boundary of each
// particle is a bit set over the sub-cells, the owner of every sub-cell an
// LDS array.  The result is the reference's, sub-cell for sub-cell.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/gale_weights.h"
#include "gs_internal.h"
#include "gs_weights.h"

namespace gs {

namespace {

constexpr int kDvcLanes = 64;

// Vertex coordinate of a regular mesh (CartesianGenerator): lo + (i / n)(hi - lo),
// the end vertices exactly lo and hi.
__device__ __forceinline__ double vcoord_w(const Geo& g, int d, long long idx) {
#pragma clang fp contract(off)
  if (idx <= 0) return g.lo[d];
  if (idx >= g.res[d]) return g.hi[d];
  return g.lo[d] + ((double)idx / (double)g.res[d]) * (g.hi[d] - g.lo[d]);
}

// Closer-claimant test of ClaimCells (DVCWeights.cxx:629-634): positive when
// q is strictly closer to the sub-cell centre c than the owner o.
__device__ __forceinline__ double dvc_contest(int dim, const double* c, const double* o, const double* q) {
#pragma clang fp contract(off)
  double s = 0.0;
  for (int d = 0; d < dim; d++) s += (o[d] + q[d] - c[d] - c[d]) * (o[d] - q[d]);
  return s;
}

__device__ __forceinline__ int wave_sum_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sub-cell coordinates packed as ix | iy << 10 | iz << 20 (LDS table, so
// the hot loops need no integer division).
__device__ __forceinline__ int dvc_ix(unsigned v) { return (int)(v & 1023u); }
__device__ __forceinline__ int dvc_iy(unsigned v) { return (int)((v >> 10) & 1023u); }
__device__ __forceinline__ int dvc_iz(unsigned v) { return (int)(v >> 20); }

// Adds to particle i's boundary the face neighbours of sub-cell c that it
// does not own (UpdateBchain, DVCWeights.cxx:420-523); returns how many bits
// it newly set.
__device__ __forceinline__ int dvc_rebound(const DvcGrid& G, const short* owner, const unsigned* cxyz, unsigned* bi,
                                           int i, int c) {
  const unsigned v = cxyz[c];
  const int cx = dvc_ix(v), cy = dvc_iy(v), cz = dvc_iz(v);
  int nb[6], nn = 0;
  if (cy > 0) nb[nn++] = c - G.r[0];
  if (cy < G.r[1] - 1) nb[nn++] = c + G.r[0];
  if (cx < G.r[0] - 1) nb[nn++] = c + 1;
  if (cx > 0) nb[nn++] = c - 1;
  if (G.dim == 3 && cz < G.r[2] - 1) nb[nn++] = c + G.r[0] * G.r[1];
  if (G.dim == 3 && cz > 0) nb[nn++] = c - G.r[0] * G.r[1];
  int added = 0;
  for (int k = 0; k < nn; k++) {
    const int m = nb[k];
    if (owner[m] != i) {
      const unsigned bit = 1u << (m & 31);
      if (!(atomicOr(&bi[m >> 5], bit) & bit)) added++;
    }
  }
  return added;
}

// LDS of one cell's growth (bytes): centres, particle coordinates, a table
// of particle slots (order), boundary sizes, sub-cell coordinates, owners,
// the current step's claims, and one boundary bit set per particle.
__host__ __device__ inline size_t dvc_lds_bytes(const DvcGrid& G, int maxp) {
  size_t b = sizeof(double) * (G.r[0] + G.r[1] + G.r[2]);
  b += sizeof(double) * 3 * maxp;
  b += sizeof(int) * 2 * maxp;
  b += sizeof(unsigned) * G.nc;
  b += (sizeof(short) * G.nc + 7) / 8 * 8;
  b += sizeof(unsigned) * (G.nwords + 1);
  b += sizeof(unsigned) * (size_t)G.nwords * maxp;
  return b;
}

struct DvcLds {
  double* ctr;      // sub-cell centres per axis
  double* xs;       // 3 per particle, in table order
  int* ord;         // table of the cell (its meaning is the caller's)
  int* bsize;       // boundary size of each particle
  unsigned* cxyz;   // packed sub-cell coordinates
  short* owner;     // owning particle of each sub-cell
  unsigned* took;   // the current step's claims
  unsigned* bound;  // one boundary bit set per particle
};

__device__ __forceinline__ DvcLds dvc_carve(const DvcGrid& G, int maxp, unsigned char* base) {
  DvcLds L;
  L.ctr = reinterpret_cast<double*>(base);
  L.xs = L.ctr + (G.r[0] + G.r[1] + G.r[2]);
  L.ord = reinterpret_cast<int*>(L.xs + 3 * maxp);
  L.bsize = L.ord + maxp;
  L.cxyz = reinterpret_cast<unsigned*>(L.bsize + maxp);
  L.owner = reinterpret_cast<short*>(L.cxyz + G.nc);
  L.took = reinterpret_cast<unsigned*>(reinterpret_cast<unsigned char*>(L.owner) + (sizeof(short) * G.nc + 7) / 8 * 8);
  L.bound = L.took + G.nwords + 1;
  return L;
}

// Sub-cell centres, accumulated one step at a time (DVCWeights.cxx:525-619),
// and the packed sub-cell coordinates.  The caller synchronises.
__device__ void dvc_setup(const DvcGrid& G, const DvcLds& L) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x;
  if (lane < G.dim) {
    const int d = lane;
    double* cd = L.ctr + (d == 0 ? 0 : (d == 1 ? G.r[0] : G.r[0] + G.r[1]));
    double x = -1.0 - G.h[d] / 2.0;
    for (int i = 0; i < G.r[d]; i++) {
      x = x + G.h[d];
      cd[i] = x;
    }
  }
  for (int c = lane; c < G.nc; c += kDvcLanes) {
    const int ix = c % G.r[0], iy = (c / G.r[0]) % G.r[1], iz = c / (G.r[0] * G.r[1]);
    L.cxyz[c] = (unsigned)ix | ((unsigned)iy << 10) | ((unsigned)iz << 20);
  }
}

__device__ __forceinline__ void dvc_centre(const DvcGrid& G, const DvcLds& L, int c, double* x) {
  const unsigned v = L.cxyz[c];
  x[0] = L.ctr[dvc_ix(v)];
  x[1] = L.ctr[G.r[0] + dvc_iy(v)];
  x[2] = G.dim == 3 ? L.ctr[G.r[0] + G.r[1] + dvc_iz(v)] : 0.0;
}

// Grows the discrete Voronoi regions of particles 0..n-1 (coordinates in
// L.xs, table order) as DVCWeights.cxx:693-786 does; L.owner holds the result.
__device__ void dvc_grow(const DvcGrid& G, const DvcLds& L, int n) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x;
  for (int c = lane; c < G.nc; c += kDvcLanes) L.owner[c] = -1;
  for (int q = lane; q < n * G.nwords; q += kDvcLanes) L.bound[q] = 0;
  __syncthreads();
  // seeds, in order (a later particle in the same sub-cell takes it over)
  if (lane == 0) {
    for (int i = 0; i < n; i++) {
      int idx[3] = {0, 0, 0};
      for (int d = 0; d < G.dim; d++) {
        int k = (int)((L.xs[3 * i + d] - (-1.0)) / G.h[d]);
        if (k == G.r[d]) k--;
        idx[d] = k;
      }
      const int c = idx[0] + G.r[0] * (idx[1] + G.r[1] * idx[2]);
      L.owner[c] = (short)i;
      L.bsize[i] = dvc_rebound(G, L.owner, L.cxyz, L.bound + (size_t)i * G.nwords, i, c);
    }
  }
  __syncthreads();
  const double* cx = L.ctr;
  const double* cy = L.ctr + G.r[0];
  const double* cz = L.ctr + G.r[0] + G.r[1];
  // Rounds of claims until one claims nothing.  Lanes take the boundary's
  // 32-bit words; a particle with an empty boundary (it claimed nothing last
  // time) is skipped, as ClaimCells over an empty list claims nothing.
  for (;;) {
    int total = 0;
    for (int i = 0; i < n; i++) {
      if (L.bsize[i] == 0) continue;  // uniform across the wave (LDS value)
      unsigned* bi = L.bound + (size_t)i * G.nwords;
      const double* xi_ = L.xs + 3 * i;
      int cnt = 0;
      for (int wd = lane; wd < G.nwords; wd += kDvcLanes) {
        unsigned bits = bi[wd], tk = 0;
        while (bits) {
          const int t = __ffs(bits) - 1;
          bits &= bits - 1u;
          const int c = wd * 32 + t;
          const int o = L.owner[c];
          bool take = o < 0;
          if (!take && o != i) {
            const unsigned v = L.cxyz[c];
            const double c3[3] = {cx[dvc_ix(v)], cy[dvc_iy(v)], G.dim == 3 ? cz[dvc_iz(v)] : 0.0};
            take = dvc_contest(G.dim, c3, L.xs + 3 * o, xi_) > 0.0;
          }
          if (take) {
            L.owner[c] = (short)i;
            tk |= 1u << t;
            cnt++;
          }
        }
        L.took[wd] = tk;
        bi[wd] = 0u;
      }
      __syncthreads();
      cnt = wave_sum_i(cnt);
      total += cnt;
      int added = 0;
      if (cnt > 0)
        for (int wd = lane; wd < G.nwords; wd += kDvcLanes) {
          unsigned tk = L.took[wd];
          while (tk) {
            const int t = __ffs(tk) - 1;
            tk &= tk - 1u;
            added += dvc_rebound(G, L.owner, L.cxyz, bi, i, wd * 32 + t);
          }
        }
      added = wave_sum_i(added);
      if (lane == 0) L.bsize[i] = added;
      __syncthreads();
    }
    if (total == 0) break;
  }
}

// Centroid and volume of every region (DVCWeights.cxx:202-258): each particle
// sums the centres of its sub-cells in index order, walking a bit set of the
// sub-cells it owns (the boundary storage, no longer needed).  out(i, c0,
// c1, c2, cnt) runs on the lane that handles particle i.
template <class F>
__device__ void dvc_regions(const DvcGrid& G, const DvcLds& L, int n, F out) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x;
  for (int q = lane; q < n * G.nwords; q += kDvcLanes) L.bound[q] = 0;
  __syncthreads();
  for (int c = lane; c < G.nc; c += kDvcLanes) {
    const int o = L.owner[c];
    if (o >= 0) atomicOr(&L.bound[(size_t)o * G.nwords + (c >> 5)], 1u << (c & 31));
  }
  __syncthreads();
  const double* cx = L.ctr;
  const double* cy = L.ctr + G.r[0];
  const double* cz = L.ctr + G.r[0] + G.r[1];
  for (int i = lane; i < n; i += kDvcLanes) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    int cnt = 0;
    const unsigned* own = L.bound + (size_t)i * G.nwords;
    for (int wd = 0; wd < G.nwords; wd++) {
      unsigned bits = own[wd];
      while (bits) {
        const int t = __ffs(bits) - 1;
        bits &= bits - 1u;
        const unsigned v = L.cxyz[wd * 32 + t];
        s0 += cx[dvc_ix(v)];
        s1 += cy[dvc_iy(v)];
        if (G.dim == 3) s2 += cz[dvc_iz(v)];
        cnt++;
      }
    }
    out(i, cnt ? s0 / cnt : 0.0, cnt ? s1 / cnt : 0.0, cnt ? s2 / cnt : 0.0, cnt);
  }
}

// One FE cell per workgroup.  Particles of cell e are pk[off[e] .. off[e+1])
// and take part in the growth in ascending pid order (the order in which the
// reference's cell table lists them).  pk.xyz becomes the centroid, w the
// weight.
__global__ __launch_bounds__(kDvcLanes) void k_dvc(DvcGrid G, const int* __restrict__ off,
                                                   const int* __restrict__ pid, double4* __restrict__ pk,
                                                   double* __restrict__ w, int maxp) {
  extern __shared__ __align__(16) unsigned char dvc_smem[];
  const long long e = blockIdx.x;
  const int b = off[e], n = off[e + 1] - b;
  if (n <= 0 || n > maxp) return;
  const int lane = threadIdx.x;
  const DvcLds L = dvc_carve(G, maxp, dvc_smem);
  dvc_setup(G, L);
  // rank of each particle by pid
  for (int j = lane; j < n; j += kDvcLanes) {
    const int pj = pid[b + j];
    int rk = 0;
    for (int k = 0; k < n; k++) rk += pid[b + k] < pj ? 1 : 0;
    L.ord[rk] = j;
  }
  __syncthreads();
  for (int i = lane; i < n; i += kDvcLanes) {
    const double4 q = pk[b + L.ord[i]];
    L.xs[3 * i] = q.x;
    L.xs[3 * i + 1] = q.y;
    L.xs[3 * i + 2] = G.dim == 3 ? q.z : 0.0;
  }
  dvc_grow(G, L, n);
  dvc_regions(G, L, n, [&](int i, double c0, double c1, double c2, int cnt) {
    const int j = b + L.ord[i];
    double4 q = pk[j];
    q.x = c0;
    q.y = c1;
    if (G.dim == 3) q.z = c2;
    pk[j] = q;
    w[j] = cnt * G.vol;
  });
}

}  // namespace

namespace {

bool make_grid(int dim, const int* res, DvcGrid* G) {
  if (dim != 2 && dim != 3) return false;
  G->dim = dim;
  G->nc = 1;
  for (int d = 0; d < 3; d++) {
    G->r[d] = d < dim ? res[d] : 1;
    if (G->r[d] < 1 || G->r[d] > kDvcMaxRes) return false;
    G->h[d] = d < dim ? 2.0 / G->r[d] : 0.0;
    G->nc *= G->r[d];
  }
  G->nwords = (G->nc + 31) / 32;
  G->vol = dim == 2 ? G->h[0] * G->h[1] : G->h[0] * G->h[1] * G->h[2];
  return true;
}

}  // namespace

int dvc_weights(int dim, const int* res, long long ncell, const int* off, const int* pid, double4* pk, double* w,
                int maxp, hipStream_t st) {
  DvcGrid G;
  if (!make_grid(dim, res, &G)) return GS_E_INVALID;
  if (ncell <= 0 || maxp <= 0) return GS_OK;
  if (maxp > 32767) return GS_E_INVALID;
  const size_t lds = dvc_lds_bytes(G, maxp);
  if (lds > kDvcMaxLds) return GS_E_INVALID;
  hipLaunchKernelGGL(k_dvc, dim3((unsigned)ncell), dim3(kDvcLanes), lds, st, G, off, pid, pk, w, maxp);
  return hipGetLastError() == hipSuccess ? GS_OK : GS_E_HIP;
}

long long dvc_max_particles(int dim, const int* res) {
  DvcGrid G;
  if (!res) return 0;
  if (!make_grid(dim, res, &G)) return 0;
  long long lo = 0, hi = 32767;
  while (lo < hi) {
    const long long mid = (lo + hi + 1) / 2;
    if (dvc_lds_bytes(G, (int)mid) <= kDvcMaxLds)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

}  // namespace gs

namespace gs {

namespace {

// PCDVC, one wavefront per FE cell (PCDVC.cxx:656-1090 in 3D, 1092-1512 in
// 2D).  The cell's table is an LDS list of particle references: r >= 0 is
// an existing particle (its sorted index), r < 0 a particle the pass created,
// record -1 - r of the staging area (its original source particle and its
// exact local coordinate).  The table starts as the cell's particles in pid
// order (CoincidentMapper.cxx:207-240); a split appends to it, a deletion
// moves its last entry into the freed slot (SwarmClass.cxx:976-990).
__device__ __forceinline__ int pc_src(const PcData& D, int ref) { return ref >= 0 ? ref : D.rsrc[-1 - ref]; }

__device__ __forceinline__ void pc_load(const DvcGrid& G, const DvcLds& L, const PcData& D, int n) {
  for (int i = threadIdx.x; i < n; i += kDvcLanes) {
    const int ref = L.ord[i];
    if (ref >= 0) {
      const double4 q = D.pk[ref];
      L.xs[3 * i] = q.x;
      L.xs[3 * i + 1] = q.y;
      L.xs[3 * i + 2] = G.dim == 3 ? q.z : 0.0;
    } else {
      const double* x = D.rxi + 3LL * (-1 - ref);
      L.xs[3 * i] = x[0];
      L.xs[3 * i + 1] = x[1];
      L.xs[3 * i + 2] = G.dim == 3 ? x[2] : 0.0;
    }
  }
}

__host__ __device__ inline size_t pc_lds_extra(int maxcap) { return sizeof(double) * 4 * maxcap + sizeof(int) * 8; }

// The pass over one cell, up to the deletions that need the swarm's global
// order: pass-1 growth, the Inflow re-population (triggered cells), the
// splits and the deletion candidates.  In probe mode a triggered cell stops
// after the trigger test.  rand() values come from rs at roff (per listed
// cell).
__global__ __launch_bounds__(kDvcLanes) void k_pc_cell(PcParams A, PcData D, const int* __restrict__ cells,
                                                       const long long* __restrict__ roff) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char pc_smem[];
  const long long e = cells ? cells[blockIdx.x] : (long long)blockIdx.x;
  const long long o = roff ? roff[blockIdx.x] : 0;
  const DvcGrid& G = A.G;
  const int lane = threadIdx.x;
  const DvcLds L = dvc_carve(G, A.maxcap, pc_smem);
  double* cent = reinterpret_cast<double*>(pc_smem + dvc_lds_bytes(G, A.maxcap));
  double* wgt = cent + 3 * A.maxcap;
  int* sh = reinterpret_cast<int*>(wgt + A.maxcap);
  int* tb = L.ord;
  const int b = D.off[e], n0 = D.off[e + 1] - b;
  const long long so = D.soff[e];
  if (n0 == 0) {  // an empty cell is left alone
    if (lane == 0) {
      D.tcnt[e] = 0;
      D.nins[e] = D.nspl[e] = D.ninfdel[e] = D.ncand[e] = D.cntdel[e] = D.ndraw[e] = 0;
      D.flags[e] = 0;
    }
    return;
  }
  dvc_setup(G, L);
  for (int j = lane; j < n0; j += kDvcLanes) {
    const int pj = D.pid[b + j];
    int rk = 0;
    for (int k = 0; k < n0; k++) rk += D.pid[b + k] < pj ? 1 : 0;
    tb[rk] = b + j;
  }
  __syncthreads();
  pc_load(G, L, D, n0);
  auto regions = [&](int i, double c0, double c1, double c2, int cnt) {
    cent[3 * i] = c0;
    cent[3 * i + 1] = c1;
    cent[3 * i + 2] = c2;
    wgt[i] = cnt * G.vol;
  };
  dvc_grow(G, L, n0);
  dvc_regions(G, L, n0, regions);
  __syncthreads();
  int n = n0, nd = 0, nrec = 0, ninf = 0, flags = 0, nins = 0;
  bool err = false;
  auto draw = [&]() -> double {
    const long long i = o + nd++;
    if (i >= D.rs_n) {
      err = true;
      return 0.0;
    }
    return (double)D.rs[i] / 2147483648.0;
  };
  if (A.inflow) {
    double dist = 0.0;
    for (int d = 0; d < G.dim; d++) {
      const double t = cent[3 * (n0 - 1) + d] - L.xs[3 * (n0 - 1) + d];
      dist += t * t;
    }
    if ((1.0 * n0) / A.ppc < A.thresh || dist > A.cpr * 2.0) {
      flags |= kPcTrig;
      if (A.probe) {
        if (lane == 0) {
          D.ndraw[e] = A.ppc + (n0 > 5 ? n0 : 0);
          D.flags[e] = flags;
        }
        return;
      }
      if (lane == 0) {
        for (int i = 0; i < A.ppc; i++) {  // PCDVC.cxx:824-832
          const int j = (int)(G.nc * draw());
          const long long r = so + nrec++;
          D.rsrc[r] = pc_src(D, tb[L.owner[j]]);
          dvc_centre(G, L, j, D.rxi + 3 * r);
          tb[n++] = (int)(-1 - r);
        }
        nins = A.ppc;
        if (n - A.ppc > 5) {  // PCDVC.cxx:838-846
          const int del_num = n - A.ppc;
          for (int i = 0; i < del_num; i++) {
            const int j = (int)(n * draw());
            D.infdel[so + ninf++] = tb[j];
            tb[j] = tb[n - 1];
            n--;
          }
        }
        sh[0] = n;
        sh[1] = nrec;
        sh[2] = nd;
        sh[3] = ninf;
        sh[4] = err ? 1 : 0;
      }
      __syncthreads();
      n = sh[0];
      nrec = sh[1];
      nd = sh[2];
      ninf = sh[3];
      err = sh[4] != 0;
      nins = A.ppc;
      if (A.ppc + ninf > 0) {  // redo the diagram (PCDVC.cxx:876-906)
        if (n < 3) err = true;  // the reference's under-resolved-cell firewall
        if (!err) {
          pc_load(G, L, D, n);
          dvc_grow(G, L, n);
          dvc_regions(G, L, n, regions);
          __syncthreads();
        }
      }
    }
  }
  if (lane == 0) {
    int mS = A.max_splits, mD = A.max_deletions;
    const int m0 = D.mat[pc_src(D, tb[0])];
    for (int i = 0; i < n; i++)
      if (D.mat[pc_src(D, tb[i])] != m0) {
        if (!A.delete_if) mD = 0;
        if (!A.split_if) mS = 0;
        break;
      }
    const int npre = n;
    int ns = 0;
    for (int q = 0; q < npre && ns < mS && !err; q++) {
      if (!(wgt[q] > A.max_w)) continue;
      const long long r = so + nrec++;
      double* x = D.rxi + 3 * r;
      if (!A.inflow) {
        x[0] = cent[3 * q];
        x[1] = cent[3 * q + 1];
        x[2] = G.dim == 3 ? cent[3 * q + 2] : 0.0;
      } else {
        // a random entry of q's owned sub-cell list, which holds its first
        // (int)(w / da' + 0.5) sub-cells in descending order (PCDVC.cxx:898-918)
        const int ls = (int)(wgt[q] * A.one_oda + 0.5);
        int k = ls - 1 - (int)(ls * draw());
        int sub = -1;
        for (int c = 0; c < G.nc && sub < 0; c++)
          if (L.owner[c] == q) {
            if (k == 0) sub = c;
            k--;
          }
        if (sub < 0) err = true;
        else dvc_centre(G, L, sub, x);
      }
      D.rsrc[r] = pc_src(D, tb[q]);
      tb[n++] = (int)(-1 - r);
      ns++;
    }
    int ncd = 0;
    for (int q = 0; q < npre; q++)
      if (wgt[q] < A.min_w) D.cand[so + ncd++] = tb[q];
    if (mD > npre - 4) mD = (G.dim == 2 && A.inflow) ? npre - 4 : npre / 2;
    int cd = mD < ncd ? mD : ncd;
    if (cd < 0) cd = 0;
    for (int i = 0; i < n; i++) D.tbl[so + i] = tb[i];
    D.tcnt[e] = n;
    D.nins[e] = nins;
    D.nspl[e] = ns;
    D.ninfdel[e] = ninf;
    D.ncand[e] = ncd;
    D.cntdel[e] = cd;
    D.ndraw[e] = nd;
    if (D.ndraw_list) D.ndraw_list[blockIdx.x] = nd;
    if (ns > 0 || cd > 0) flags |= kPcFinal;
    if (A.probe && A.inflow && ns > 0) flags |= kPcRedo;
    if (err) flags |= kPcErr;
    D.flags[e] = flags;
    sh[5] = flags;
  }
  __syncthreads();
  if (!(sh[5] & kPcFinal))
    for (int i = lane; i < n; i += kDvcLanes) {
      D.scent[3 * (so + i)] = cent[3 * i];
      D.scent[3 * (so + i) + 1] = cent[3 * i + 1];
      D.scent[3 * (so + i) + 2] = cent[3 * i + 2];
      D.sw[so + i] = wgt[i];
    }
}

// The deletions in the order the swarm's global indices give (decided on the
// host), then the cell's final diagram (PCDVC.cxx:1015-1061).
__global__ __launch_bounds__(kDvcLanes) void k_pc_final(PcParams A, PcData D, const int* __restrict__ cells,
                                                        const int* __restrict__ dorder,
                                                        const long long* __restrict__ doff) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char pc_smem[];
  const long long e = cells[blockIdx.x];
  const DvcGrid& G = A.G;
  const int lane = threadIdx.x;
  const DvcLds L = dvc_carve(G, A.maxcap, pc_smem);
  int* sh = reinterpret_cast<int*>(pc_smem + dvc_lds_bytes(G, A.maxcap) + sizeof(double) * 4 * A.maxcap);
  int* tb = L.ord;
  const long long so = D.soff[e];
  int n = D.tcnt[e];
  dvc_setup(G, L);
  for (int i = lane; i < n; i += kDvcLanes) tb[i] = D.tbl[so + i];
  __syncthreads();
  if (lane == 0) {
    const long long d0 = doff[blockIdx.x];
    bool err = false;
    for (int k = 0; k < D.cntdel[e]; k++) {
      const int ref = dorder[d0 + k];
      int s = 0;
      while (s < n && tb[s] != ref) s++;
      if (s == n) {
        err = true;
        break;
      }
      tb[s] = tb[n - 1];
      n--;
    }
    if (n < 3) err = true;
    if (err) D.flags[e] |= kPcErr;
    sh[0] = n;
    sh[1] = err ? 1 : 0;
  }
  __syncthreads();
  n = sh[0];
  if (sh[1]) return;
  pc_load(G, L, D, n);
  dvc_grow(G, L, n);
  dvc_regions(G, L, n, [&](int i, double c0, double c1, double c2, int cnt) {
    D.scent[3 * (so + i)] = c0;
    D.scent[3 * (so + i) + 1] = c1;
    D.scent[3 * (so + i) + 2] = c2;
    D.sw[so + i] = cnt * G.vol;
  });
  for (int i = lane; i < n; i += kDvcLanes) D.tbl[so + i] = tb[i];
  if (lane == 0) D.tcnt[e] = n;
}

// The deletion candidates and the Inflow deletions of the listed cells, as
// caller indices (pid) for existing particles and records (-1 - r) for new
// ones, packed for the host.
__global__ void k_pc_gather(long long ncells, const int* __restrict__ cells, const long long* __restrict__ goff,
                            PcData D, int* __restrict__ out) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= ncells) return;
  const long long e = cells[k], so = D.soff[e];
  long long w = goff[k];
  for (int i = 0; i < D.ninfdel[e]; i++) {
    const int ref = D.infdel[so + i];
    out[w++] = ref;
    out[w++] = ref >= 0 ? D.pid[ref] : ref;
  }
  for (int i = 0; i < D.ncand[e]; i++) {
    const int ref = D.cand[so + i];
    out[w++] = ref;
    out[w++] = ref >= 0 ? D.pid[ref] : ref;
  }
}

// Vertex coordinate of a regular mesh as vcoord_w; a new particle's global
// position is 0.5 ((1 - xi) a + (1 + xi) b) per axis.
__global__ void k_pc_rebuild(Geo g, PcData D, const int* __restrict__ off2, const int* __restrict__ fidx,
                             const int* __restrict__ nbase, const int* __restrict__ fidx_new,
                             const double* __restrict__ coord, const int* __restrict__ mat, PartDoubles x,
                             double* __restrict__ coordN, int* __restrict__ matN, int* __restrict__ pidN,
                             PartDoubles xN, double4* __restrict__ pkN, double* __restrict__ wN,
                             unsigned* __restrict__ keysN) {
#pragma clang fp contract(off)
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= g.E) return;
  const long long so = D.soff[e];
  const long long ijk[3] = {e % g.res[0], (e / g.res[0]) % g.res[1],
                            g.dim == 3 ? e / ((long long)g.res[0] * g.res[1]) : 0};
  for (int s = 0; s < D.tcnt[e]; s++) {
    const int ref = D.tbl[so + s];
    const long long dst = off2[e] + s;
    const int src = pc_src(D, ref);
    if (ref >= 0) {
      for (int d = 0; d < 3; d++) coordN[3 * dst + d] = coord[3LL * ref + d];
      pidN[dst] = fidx[ref];
    } else {
      const long long r = -1 - ref;
      const double* xi = D.rxi + 3 * r;
      for (int d = 0; d < 3; d++) {
        double v = 0.0;
        if (d < g.dim) {
          const double lo = vcoord_w(g, d, ijk[d]), hi = vcoord_w(g, d, ijk[d] + 1);
          v = 0.5 * ((1.0 - xi[d]) * lo + (1.0 + xi[d]) * hi);
        }
        coordN[3 * dst + d] = v;
      }
      pidN[dst] = fidx_new[nbase[e] + (r - so)];
    }
    matN[dst] = mat[src];
    for (int q = 0; q < kPartDoubles; q++)
      if (xN.a[q]) xN.a[q][dst] = x.a[q][src];
    pkN[dst] = make_double4(D.scent[3 * (so + s)], D.scent[3 * (so + s) + 1], D.scent[3 * (so + s) + 2], 0.0);
    wN[dst] = D.sw[so + s];
    keysN[dst] = (unsigned)e;
  }
}

}  // namespace

size_t pc_lds_bytes(const PcParams& A) { return dvc_lds_bytes(A.G, A.maxcap) + pc_lds_extra(A.maxcap); }

bool pc_grid(int dim, const int* res, PcParams* A) { return make_grid(dim, res, &A->G); }

long long pc_max_capacity(const PcParams& A) {
  long long lo = 0, hi = 32767;
  while (lo < hi) {
    const long long mid = (lo + hi + 1) / 2;
    if (dvc_lds_bytes(A.G, (int)mid) + pc_lds_extra((int)mid) <= kDvcMaxLds)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

void pc_cells(const PcParams& A, const PcData& D, long long ncells, const int* cells, const long long* roff,
              hipStream_t st) {
  if (ncells <= 0) return;
  hipLaunchKernelGGL(k_pc_cell, dim3((unsigned)ncells), dim3(kDvcLanes), pc_lds_bytes(A), st, A, D, cells, roff);
}

void pc_final(const PcParams& A, const PcData& D, long long ncells, const int* cells, const int* dorder,
              const long long* doff, hipStream_t st) {
  if (ncells <= 0) return;
  hipLaunchKernelGGL(k_pc_final, dim3((unsigned)ncells), dim3(kDvcLanes), pc_lds_bytes(A), st, A, D, cells, dorder,
                     doff);
}

void pc_gather(const PcData& D, long long ncells, const int* cells, const long long* goff, int* out,
               hipStream_t st) {
  if (ncells <= 0) return;
  hipLaunchKernelGGL(k_pc_gather, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, st, ncells, cells, goff, D,
                     out);
}

void pc_rebuild(const Geo& g, const PcData& D, const int* off2, const int* fidx, const int* nbase,
                const int* fidx_new, const double* coord, const int* mat, PartDoubles x, double* coordN, int* matN,
                int* pidN, PartDoubles xN, double4* pkN, double* wN, unsigned* keysN, hipStream_t st) {
  if (g.E <= 0) return;
  hipLaunchKernelGGL(k_pc_rebuild, dim3((unsigned)((g.E + 255) / 256)), dim3(256), 0, st, g, D, off2, fidx, nbase,
                     fidx_new, coord, mat, x, coordN, matN, pidN, xN, pkN, wN, keysN);
}

}  // namespace gs

using namespace gs;

#define WCK(x)                                \
  do {                                        \
    if ((x) != hipSuccess) return GS_E_HIP;   \
  } while (0)

extern "C" int32_t gs_dvc_cells(int32_t dim, const int32_t* res, int64_t ncell, const int32_t* off,
                                const double* xi, double* centroid, double* weight) {
  if ((dim != 2 && dim != 3) || !res || ncell < 0 || (ncell > 0 && (!off || !xi || !centroid || !weight)))
    return GS_E_INVALID;
  if (ncell == 0) return GS_OK;
  if (ncell > 0x7fffffff) return GS_E_INVALID;
  const long long P = off[ncell];
  int maxp = 0;
  for (long long e = 0; e < ncell; e++) {
    const int c = off[e + 1] - off[e];
    if (c < 0 || off[e] < 0) return GS_E_INVALID;
    maxp = c > maxp ? c : maxp;
  }
  if (maxp > dvc_max_particles(dim, res)) return GS_E_INVALID;
  std::vector<double4> pk(P > 0 ? P : 1);
  std::vector<int> pid(P > 0 ? P : 1);
  for (long long p = 0; p < P; p++) {
    pk[p] = make_double4(xi[p * dim], xi[p * dim + 1], dim == 3 ? xi[p * dim + 2] : 0.0, 0.0);
    pid[p] = (int)(p - off[0]);
  }
  int *d_off = nullptr, *d_pid = nullptr;
  double4* d_pk = nullptr;
  double* d_w = nullptr;
  int32_t st = GS_OK;
  hipStream_t s = nullptr;
  if (hipMalloc(&d_off, sizeof(int) * (ncell + 1)) != hipSuccess || hipMalloc(&d_pid, sizeof(int) * pid.size()) ||
      hipMalloc(&d_pk, sizeof(double4) * pk.size()) || hipMalloc(&d_w, sizeof(double) * pk.size())) {
    st = GS_E_HIP;
  } else if (hipMemcpy(d_off, off, sizeof(int) * (ncell + 1), hipMemcpyHostToDevice) ||
             hipMemcpy(d_pid, pid.data(), sizeof(int) * pid.size(), hipMemcpyHostToDevice) ||
             hipMemcpy(d_pk, pk.data(), sizeof(double4) * pk.size(), hipMemcpyHostToDevice)) {
    st = GS_E_HIP;
  } else {
    st = dvc_weights(dim, res, ncell, d_off, d_pid, d_pk, d_w, maxp, s);
    if (st == GS_OK) {
      std::vector<double> wv(pk.size());
      if (hipMemcpy(pk.data(), d_pk, sizeof(double4) * pk.size(), hipMemcpyDeviceToHost) ||
          hipMemcpy(wv.data(), d_w, sizeof(double) * wv.size(), hipMemcpyDeviceToHost)) {
        st = GS_E_HIP;
      } else {
        for (long long p = off[0]; p < P; p++) {
          centroid[p * dim] = pk[p].x;
          centroid[p * dim + 1] = pk[p].y;
          if (dim == 3) centroid[p * dim + 2] = pk[p].z;
          weight[p] = wv[p];
        }
      }
    }
  }
  if (d_off) (void)hipFree(d_off);
  if (d_pid) (void)hipFree(d_pid);
  if (d_pk) (void)hipFree(d_pk);
  if (d_w) (void)hipFree(d_w);
  return st;
}

extern "C" int64_t gs_dvc_max_particles(int32_t dim, const int32_t* res) { return dvc_max_particles(dim, res); }
