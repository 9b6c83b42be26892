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

struct DvcGrid {
  int dim;
  int r[3];
  int nc, nwords;
  double h[3];
  double vol;
};

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

// LDS layout of one cell (bytes): centres, particle coordinates, order,
// boundary sizes, sub-cell coordinates, owners, the current step's claims,
// and one boundary bit set per particle.
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
  double* ctr = reinterpret_cast<double*>(dvc_smem);
  double* xs = ctr + (G.r[0] + G.r[1] + G.r[2]);
  int* ord = reinterpret_cast<int*>(xs + 3 * maxp);
  int* bsize = ord + maxp;  // boundary size of each particle
  unsigned* cxyz = reinterpret_cast<unsigned*>(bsize + maxp);
  short* owner = reinterpret_cast<short*>(cxyz + G.nc);
  unsigned* took = reinterpret_cast<unsigned*>(reinterpret_cast<unsigned char*>(owner) +
                                               (sizeof(short) * G.nc + 7) / 8 * 8);
  unsigned* bound = took + G.nwords + 1;
  // sub-cell centres, accumulated one step at a time (DVCWeights.cxx:525-619)
  if (lane < G.dim) {
    const int d = lane;
    double* cd = ctr + (d == 0 ? 0 : (d == 1 ? G.r[0] : G.r[0] + G.r[1]));
    double x = -1.0 - G.h[d] / 2.0;
    for (int i = 0; i < G.r[d]; i++) {
      x = x + G.h[d];
      cd[i] = x;
    }
  }
  // rank of each particle by pid
  for (int j = lane; j < n; j += kDvcLanes) {
    const int pj = pid[b + j];
    int rk = 0;
    for (int k = 0; k < n; k++) rk += pid[b + k] < pj ? 1 : 0;
    ord[rk] = j;
  }
  for (int c = lane; c < G.nc; c += kDvcLanes) {
    owner[c] = -1;
    const int ix = c % G.r[0], iy = (c / G.r[0]) % G.r[1], iz = c / (G.r[0] * G.r[1]);
    cxyz[c] = (unsigned)ix | ((unsigned)iy << 10) | ((unsigned)iz << 20);
  }
  for (int q = lane; q < n * G.nwords; q += kDvcLanes) bound[q] = 0;
  __syncthreads();
  for (int i = lane; i < n; i += kDvcLanes) {
    const double4 q = pk[b + ord[i]];
    xs[3 * i] = q.x;
    xs[3 * i + 1] = q.y;
    xs[3 * i + 2] = G.dim == 3 ? q.z : 0.0;
  }
  __syncthreads();
  // seeds, in order (a later particle in the same sub-cell takes it over)
  if (lane == 0) {
    for (int i = 0; i < n; i++) {
      int idx[3] = {0, 0, 0};
      for (int d = 0; d < G.dim; d++) {
        int k = (int)((xs[3 * i + d] - (-1.0)) / G.h[d]);
        if (k == G.r[d]) k--;
        idx[d] = k;
      }
      const int c = idx[0] + G.r[0] * (idx[1] + G.r[1] * idx[2]);
      owner[c] = (short)i;
      bsize[i] = dvc_rebound(G, owner, cxyz, bound + (size_t)i * G.nwords, i, c);
    }
  }
  __syncthreads();
  const double* cx = ctr;
  const double* cy = ctr + G.r[0];
  const double* cz = ctr + G.r[0] + G.r[1];
  // Rounds of claims until one claims nothing.  Lanes take the boundary's
  // 32-bit words; a particle with an empty boundary (it claimed nothing last
  // time) is skipped, as ClaimCells over an empty list claims nothing.
  for (;;) {
    int total = 0;
    for (int i = 0; i < n; i++) {
      if (bsize[i] == 0) continue;  // uniform across the wave (LDS value)
      unsigned* bi = bound + (size_t)i * G.nwords;
      const double* xi_ = xs + 3 * i;
      int cnt = 0;
      for (int wd = lane; wd < G.nwords; wd += kDvcLanes) {
        unsigned bits = bi[wd], tk = 0;
        while (bits) {
          const int t = __ffs(bits) - 1;
          bits &= bits - 1u;
          const int c = wd * 32 + t;
          const int o = owner[c];
          bool take = o < 0;
          if (!take && o != i) {
            const unsigned v = cxyz[c];
            const double c3[3] = {cx[dvc_ix(v)], cy[dvc_iy(v)], G.dim == 3 ? cz[dvc_iz(v)] : 0.0};
            take = dvc_contest(G.dim, c3, xs + 3 * o, xi_) > 0.0;
          }
          if (take) {
            owner[c] = (short)i;
            tk |= 1u << t;
            cnt++;
          }
        }
        took[wd] = tk;
        bi[wd] = 0u;
      }
      __syncthreads();
      cnt = wave_sum_i(cnt);
      total += cnt;
      int added = 0;
      if (cnt > 0)
        for (int wd = lane; wd < G.nwords; wd += kDvcLanes) {
          unsigned tk = took[wd];
          while (tk) {
            const int t = __ffs(tk) - 1;
            tk &= tk - 1u;
            added += dvc_rebound(G, owner, cxyz, bi, i, wd * 32 + t);
          }
        }
      added = wave_sum_i(added);
      if (lane == 0) bsize[i] = added;
      __syncthreads();
    }
    if (total == 0) break;
  }
  // centroids and volumes (DVCWeights.cxx:202-258): each particle sums the
  // centres of its sub-cells in index order, walking a bit set of the
  // sub-cells it owns (the boundary storage, no longer needed)
  for (int q = lane; q < n * G.nwords; q += kDvcLanes) bound[q] = 0;
  __syncthreads();
  for (int c = lane; c < G.nc; c += kDvcLanes) {
    const int o = owner[c];
    if (o >= 0) atomicOr(&bound[(size_t)o * G.nwords + (c >> 5)], 1u << (c & 31));
  }
  __syncthreads();
  for (int i = lane; i < n; i += kDvcLanes) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    int cnt = 0;
    const unsigned* own = bound + (size_t)i * G.nwords;
    for (int wd = 0; wd < G.nwords; wd++) {
      unsigned bits = own[wd];
      while (bits) {
        const int t = __ffs(bits) - 1;
        bits &= bits - 1u;
        const unsigned v = cxyz[wd * 32 + t];
        s0 += cx[dvc_ix(v)];
        s1 += cy[dvc_iy(v)];
        if (G.dim == 3) s2 += cz[dvc_iz(v)];
        cnt++;
      }
    }
    const int j = b + ord[i];
    double4 q = pk[j];
    q.x = cnt ? s0 / cnt : 0.0;
    q.y = cnt ? s1 / cnt : 0.0;
    if (G.dim == 3) q.z = cnt ? s2 / cnt : 0.0;
    pk[j] = q;
    w[j] = cnt * G.vol;
  }
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

__global__ void k_pcdvc_decide(Geo g, const int* __restrict__ off, const int* __restrict__ pid,
                               const int* __restrict__ mat, const double4* __restrict__ pk,
                               const double* __restrict__ w, PcdvcArgs a, unsigned char* __restrict__ del,
                               int* __restrict__ nsplit, double* __restrict__ scoord, int* __restrict__ sparent) {
#pragma clang fp contract(off)
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= g.E) return;
  const int b = off[e], n = off[e + 1] - b;
  for (int j = 0; j < n; j++) del[b + j] = 0;
  nsplit[e] = 0;
  if (n == 0) return;
  // first particle in pid order and whether the cell holds several materials
  int first = 0;
  for (int j = 1; j < n; j++)
    if (pid[b + j] < pid[b + first]) first = j;
  bool interface = false;
  for (int j = 0; j < n; j++) interface = interface || mat[b + j] != mat[b + first];
  int mS = a.max_splits, mD = a.max_deletions;
  if (interface) {
    if (!a.delete_if) mD = 0;
    if (!a.split_if) mS = 0;
  }
  if (mD > n - 4) mD = n / 2;
  const long long ijk[3] = {e % g.res[0], (e / g.res[0]) % g.res[1],
                            g.dim == 3 ? e / ((long long)g.res[0] * g.res[1]) : 0};
  int ns = 0, last = -1;
  for (int t = 0; t < n; t++) {
    // next particle in pid order
    int j = -1;
    for (int k = 0; k < n; k++) {
      const int pk_ = pid[b + k];
      if (pk_ > last && (j < 0 || pk_ < pid[b + j])) j = k;
    }
    last = pid[b + j];
    const double wj = w[b + j];
    if (wj > a.max_w && ns < mS) {
      const double4 c = pk[b + j];
      const double xi[3] = {c.x, c.y, c.z};
      double* out = scoord + ((long long)e * a.slots + ns) * 3;
      for (int d = 0; d < 3; d++) {
        double x = 0.0;
        if (d < g.dim) {
          const double lo = vcoord_w(g, d, ijk[d]), hi = vcoord_w(g, d, ijk[d] + 1);
          x = 0.5 * ((1.0 - xi[d]) * lo + (1.0 + xi[d]) * hi);
        }
        out[d] = x;
      }
      sparent[(long long)e * a.slots + ns] = b + j;
      ns++;
    }
  }
  // deletions: of the particles below minW, the reference sorts the list by
  // swarm index in descending order and deletes its first maxDeletions
  // (compare_indexOnCPU, PCDVC.cxx:611-620, :974, :1015-1021), i.e. the
  // candidates with the highest indices
  int ncand = 0;
  for (int j = 0; j < n; j++) ncand += w[b + j] < a.min_w ? 1 : 0;
  const int nD = mD < ncand ? mD : ncand;
  if (nD > 0)
    for (int j = 0; j < n; j++) {
      if (!(w[b + j] < a.min_w)) continue;
      int higher = 0;
      for (int k = 0; k < n; k++) higher += (w[b + k] < a.min_w && pid[b + k] > pid[b + j]) ? 1 : 0;
      if (higher < nD) del[b + j] = 1;
    }
  nsplit[e] = ns;
}

__global__ void k_pcdvc_compact(long long P, const int* __restrict__ map, const double* __restrict__ coord,
                                const int* __restrict__ mat, PartDoubles x, double* __restrict__ coordN,
                                int* __restrict__ matN, PartDoubles xN) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int m = map[i];
  if (m < 0) return;
  for (int d = 0; d < 3; d++) coordN[3LL * m + d] = coord[3 * i + d];
  matN[m] = mat[i];
  for (int k = 0; k < kPartDoubles; k++)
    if (xN.a[k]) xN.a[k][m] = x.a[k][i];
}

// A split copies its parent's particle state (Swarm_CopyParticleWithinSwarm
// in PCDVC.cxx): material, per-particle eta0 and the StrainWeakening values.
__global__ void k_pcdvc_append(long long E, const int* __restrict__ nsplit, const int* __restrict__ base, int slots,
                               const double* __restrict__ scoord, const int* __restrict__ sparent,
                               const int* __restrict__ mat, PartDoubles x, long long n_surv,
                               double* __restrict__ coordN, int* __restrict__ matN, PartDoubles xN) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= E) return;
  for (int k = 0; k < nsplit[e]; k++) {
    const long long np = n_surv + base[e] + k;
    const long long s = e * slots + k;
    for (int d = 0; d < 3; d++) coordN[3 * np + d] = scoord[3 * s + d];
    const int par = sparent[s];
    matN[np] = mat[par];
    for (int q = 0; q < kPartDoubles; q++)
      if (xN.a[q]) xN.a[q][np] = x.a[q][par];
  }
}

}  // namespace

void pcdvc_decide(const Geo& g, const int* off, const int* pid, const int* mat, const double4* pk, const double* w,
                  const PcdvcArgs& a, unsigned char* del, int* nsplit, double* scoord, int* sparent, hipStream_t st) {
  if (g.E <= 0) return;
  hipLaunchKernelGGL(k_pcdvc_decide, dim3((unsigned)((g.E + 255) / 256)), dim3(256), 0, st, g, off, pid, mat, pk, w,
                     a, del, nsplit, scoord, sparent);
}

void pcdvc_rebuild(long long P, long long E, const int* map, const double* coord, const int* mat, PartDoubles x,
                   const int* nsplit, const int* base, int slots, const double* scoord, const int* sparent,
                   long long n_surv, double* coordN, int* matN, PartDoubles xN, hipStream_t st) {
  if (P > 0)
    hipLaunchKernelGGL(k_pcdvc_compact, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, P, map, coord, mat, x,
                       coordN, matN, xN);
  if (E > 0)
    hipLaunchKernelGGL(k_pcdvc_append, dim3((unsigned)((E + 255) / 256)), dim3(256), 0, st, E, nsplit, base, slots,
                       scoord, sparent, mat, x, n_surv, coordN, matN, xN);
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
