// gs_api.hip -- host driver and C ABI (include/gale_stokes.h).
//
// The host side owns device buffers and drives the Uzawa / CG / Picard loops
// with all scalars kept on the device; it waits for the device only at
// convergence checks.  Reference behaviour each function follows is cited in
// include/gale_stokes.h.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocsolver/rocsolver.h>

#include "../../include/gale_stokes.h"
#include "gs_crand.h"
#include "gs_internal.h"
#include "gs_learned.h"
#include "gs_mg.h"
#include "gs_tile.h"
#include "gs_weights.h"
#include "../../include/gale_weights.h"

using namespace gs;

#define CK(x)                                                                 \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::fprintf(stderr, "gale_stokes: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return GS_E_HIP;                                                        \
    }                                                                         \
  } while (0)

#define CKS(x)              \
  do {                      \
    int32_t s_ = (x);       \
    if (s_ != GS_OK) return s_; \
  } while (0)

template <typename T>
static hipError_t dalloc(T** p, long long n) {
  *p = nullptr;
  if (n <= 0) n = 1;
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (size_t)n);
}
template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

struct gs_problem {
  Geo g;
  int dim = 0, nen = 0;
  hipStream_t st = nullptr;
  std::vector<gs_material> mats_h;
  MatDev* mats_d = nullptr;
  std::vector<MatDev> mats_dh;  // host copy of mats_d
  double gravity = 0, ghat[3] = {0, 0, 0}, clamp_min = 0, clamp_max = 0;
  uint8_t* bcdof = nullptr;   // N*dim
  uint8_t* bcflag = nullptr;  // N (bit d = dof d is Dirichlet)
  double* ubc = nullptr;      // N*dim, Dirichlet values, zeros elsewhere
  bool has_bc = false;
  bool closed = false;  // every boundary node has its normal velocity fixed
  long long n_free = 0;
  long long P = 0, escaped = 0;
  double4* pk = nullptr;
  float4* pk32 = nullptr;
  bool pk32_valid = false;
  double* coord = nullptr;
  int* mat = nullptr;
  int* pid = nullptr;
  double* eta0p = nullptr;
  double* pstr = nullptr;   // StrainWeakening: accumulated plastic strain (sorted), null when off
  double* pinc = nullptr;   // its rate from the last constitutive pass (sorted)
  StrainLimit slim{-1, 0, {0, 0, 0, 0, 0, 0}, 0.0};  // strain-limited region (type -1: none)
  int* off = nullptr;
  unsigned* keys = nullptr;  // sorted owner of each sorted particle (E = outside)
  unsigned* keys2 = nullptr;
  int* vals = nullptr;
  int* vals2 = nullptr;
  double* coord2 = nullptr;
  double* eta0p2 = nullptr;
  double* pstr2 = nullptr;
  double* pinc2 = nullptr;
  int* mat2 = nullptr;
  int* pid2 = nullptr;
  void* sort_tmp = nullptr;
  size_t sort_bytes = 0;
  double* S = nullptr;
  double* hS = nullptr;  // pinned
  Reducer rd{};
  bool min_one_particle = true;
  double* halo64 = nullptr;  // tiled K apply scratch (fine level)
  float* halo32 = nullptr;
  int wkind = GS_WEIGHTS_CONSTANT;  // integration weights of the material points
  int wres[3] = {10, 10, 10};       // DVC sub-grid
  double* wp = nullptr;             // DVC weight of each (sorted) particle
  unsigned long long gen = 0;       // bumped by every relocation (solvers re-run their setup)
  int device = 0;                   // HIP device current at creation (the problem's buffers live there)
  // the C library's rand() stream of PCDVC's Inflow branch, positioned after
  // rng_pos calls since srand(rng_seed) (rng_seed < 0: not started)
  CRand rng;
  long long rng_seed = -1, rng_pos = 0;
  void* scratch = nullptr;          // that device's address of the no-allocation scratch array
};


// One multigrid level: moments (FP64 + FP32 copies), Dirichlet flags,
// Jacobi diagonal, Chebyshev bound and work vectors in both precisions.
struct MGLevel {
  Geo g{};
  long long n = 0;
  double* M64 = nullptr;
  float* M32 = nullptr;
  uint8_t* bcflag = nullptr;
  uint8_t* bcdof = nullptr;
  bool own_bc = false;
  double* diag = nullptr;
  double* dinv64 = nullptr;
  float* dinv32 = nullptr;
  double lmax = 2.0;
  double *b = nullptr, *x = nullptr, *r = nullptr, *d = nullptr, *Ax = nullptr, *x2 = nullptr;
  float *bf = nullptr, *xf = nullptr, *rf = nullptr, *df = nullptr, *Axf = nullptr, *xf2 = nullptr;
  double* h64 = nullptr;  // tiled-apply halo scratch
  float* h32 = nullptr;
};

// Free dofs up to which the 3D hierarchy stops for a dense coarsest solve.
// 96^3 coarsens 96-48-24-12 and solves the 12^3 level (5577 free dofs with
// free-slip walls) densely: one matrix-vector product replaces the 12^3 and
// 6^3 levels' smoothing, transfer and solve launches of every V-cycle.
// GALE_MG_DENSE_MAX overrides (A/B timing); at most kDenseWideMax.
static constexpr int kDenseHostMax = 2000;  // larger inverses are formed on the device
static int dense_max_free() {
  const char* e = std::getenv("GALE_MG_DENSE_MAX");
  int v = e ? std::atoi(e) : 6000;
  if (v < 0) v = 0;
  return std::min(v, kDenseWideMax);
}

struct MGData {
  bool structured = false;
  bool bottom = false;  // single-workgroup bottom of the cycle (3D) where levels fit
  int L = 0;  // index of the coarsest level
  std::vector<MGLevel> lv;
  int nfc = 0;
  int ldk = 0;  // 0: dense inverse packed nfc x nfc; > 0: rows padded to ldk (wide solve)
  int* freeidx = nullptr;
  double* Kinv64 = nullptr;
  float* Kinv32 = nullptr;
};

static void mg_free(MGData& mg);

struct gs_solver {
  gs_problem* pb = nullptr;
  gs_uzawa_params up{};
  long long n = 0, np = 0;
  double *f = nullptr, *diag = nullptr, *dinv = nullptr, *h = nullptr, *Q = nullptr, *Qinv = nullptr;
  double *u = nullptr, *ufull = nullptr, *p = nullptr, *r = nullptr, *s = nullptr, *z = nullptr, *q = nullptr;
  double *fT = nullptr, *vstar = nullptr, *cg_r = nullptr, *cg_p = nullptr, *cg_Ap = nullptr, *tmpN = nullptr;
  double *tmpE = nullptr, *eta = nullptr, *uprev = nullptr;
  uint8_t* yielded = nullptr;
  float *x32 = nullptr, *y32 = nullptr;
  double* S = nullptr;
  double* hS = nullptr;
  Reducer rd{};
  double* cg_z = nullptr;
  long long P_alloc = 0;  // particle count of eta / yielded
  MGData mg;
  bool setup_done = false;
  unsigned long long setup_gen = 0;  // problem generation the setup was made for
  bool has_solution = false;
  // previousSolutionExists of the reference's ConstitutiveMatrix: set after
  // the first solve of a nonlinear execute (sle->hasExecuted,
  // SystemLinearEquations.cxx:794; ConstitutiveMatrixCartesian.cxx:270-273)
  // or by a restart (ConstitutiveMatrix.cxx:127-130), and kept for later
  // time steps, whose first solve then yields from the stored solution
  int prev_solution = 0;
  long long k_applies = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // learned Schur preconditioner and training-pair recording
  gs_lpc* lpc = nullptr;
  gs_lpc_dataset* rec = nullptr;
  int rec_every = 0;
  float* logeta = nullptr;
  double* rec_r0 = nullptr;
  double* rec_p0 = nullptr;
  // truncated flexible CG directions (learned preconditioner)
  int fl_m = 0;
  double* fl_s = nullptr;
  double* fl_q = nullptr;
  double* fl_aux = nullptr;
};
// The solver's arrays describe the problem's particles as of its last setup.
static bool needs_setup(const gs_solver* s) { return !s->setup_done || s->setup_gen != s->pb->gen; }


// ---------------------------------------------------------------------------
extern "C" const char* gs_strerror(int32_t s) {
  switch (s) {
    case GS_OK: return "ok";
    case GS_E_NAN: return "NaN or Inf in the Uzawa iteration";
    case GS_E_MAXIT: return "Uzawa reached maxIterations without converging";
    case GS_E_ZERO_FORCE: return "the momentum force vector is zero";
    case GS_E_STAGNATION: return "residual stagnated";
    case GS_E_HIP: return "HIP runtime error";
    case GS_E_INVALID: return "invalid argument";
    case GS_E_ESCAPED: return "a particle left the domain";
    case GS_E_INNER_MAXIT: return "inner velocity solve reached its iteration cap";
    case GS_E_PICARD: return "Picard iteration did not converge";
    case GS_E_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

extern "C" int32_t gs_version(void) { return 1; }

extern "C" int32_t gs_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

static int32_t alloc_reducer(Reducer* rd) {
  CK(dalloc(&rd->partials, (long long)kMaxReduceBlocks * 4));
  return GS_OK;
}

// Carry the particle arrays along the sort permutation.
__global__ void k_gather_particles(const int* __restrict__ perm, long long P, const double* __restrict__ c,
                                   double* __restrict__ c2, const int* __restrict__ m, int* __restrict__ m2,
                                   const int* __restrict__ id, int* __restrict__ id2, PartDoubles x, PartDoubles x2) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const long long j = perm[i];
  c2[3 * i] = c[3 * j];
  c2[3 * i + 1] = c[3 * j + 1];
  c2[3 * i + 2] = c[3 * j + 2];
  m2[i] = m[j];
  id2[i] = id[j];
  for (int k = 0; k < kPartDoubles; k++)
    if (x.a[k]) x2.a[k][i] = x.a[k][j];
}

// Locate every particle (bit-exact owner rule), sort by owner (stable radix
// sort: within an element particles keep their previous relative order),
// carry all particle arrays along, rebuild the element offsets and remap xi.
static int32_t relocate(gs_problem* pb) {
  pb->gen++;
  const long long P = pb->P;
  hipStream_t st = pb->st;
  launch_fill(pb->S + S_ESC, 1, 0.0, st);
  if (P > 0) {
    launch_locate(pb->g, pb->coord, P, reinterpret_cast<int*>(pb->keys2), pb->vals2, pb->S, st);
    unsigned end_bit = 1;
    while ((1ull << end_bit) <= (unsigned long long)pb->g.E) end_bit++;
    size_t bytes = 0;
    CK(rocprim::radix_sort_pairs(nullptr, bytes, pb->keys2, pb->keys, pb->vals2, pb->vals, (size_t)P, 0, end_bit, st));
    if (bytes > pb->sort_bytes) {
      if (pb->sort_tmp) CK(hipFree(pb->sort_tmp));
      CK(hipMalloc(&pb->sort_tmp, bytes));
      pb->sort_bytes = bytes;
    }
    CK(rocprim::radix_sort_pairs(pb->sort_tmp, bytes, pb->keys2, pb->keys, pb->vals2, pb->vals, (size_t)P, 0, end_bit,
                                 st));
    hipLaunchKernelGGL(k_gather_particles, dim3((unsigned)((P + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, pb->vals, P,
                       pb->coord, pb->coord2, pb->mat, pb->mat2, pb->pid, pb->pid2,
                       PartDoubles{{pb->eta0p, pb->pstr, pb->pinc}}, PartDoubles{{pb->eta0p2, pb->pstr2, pb->pinc2}});
    std::swap(pb->coord, pb->coord2);
    std::swap(pb->mat, pb->mat2);
    std::swap(pb->pid, pb->pid2);
    std::swap(pb->eta0p, pb->eta0p2);
    std::swap(pb->pstr, pb->pstr2);
    std::swap(pb->pinc, pb->pinc2);
  }
  launch_offsets(reinterpret_cast<const int*>(pb->keys), P, pb->g.E, pb->off, st);
  launch_remap(pb->g, pb->coord, pb->keys, pb->pk, P, st);
  pb->pk32_valid = false;
  CK(hipMemcpyAsync(pb->hS, pb->S, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  unsigned esc;
  std::memcpy(&esc, pb->hS + S_ESC, sizeof(unsigned));
  pb->escaped = esc;
  // every element must hold at least one material point (ConstantWeights divides by the count)
  std::vector<int> off(pb->g.E + 1);
  CK(hipMemcpy(off.data(), pb->off, sizeof(int) * (pb->g.E + 1), hipMemcpyDeviceToHost));
  pb->min_one_particle = true;
  for (long long e = 0; e < pb->g.E; e++)
    if (off[e + 1] == off[e]) {
      pb->min_one_particle = false;
      break;
    }
  if (pb->wkind == GS_WEIGHTS_DVC && P > 0) {
    // DVCWeights: integration point at the centroid of its discrete Voronoi
    // region, weight = the region's volume (gs_weights.hip)
    int maxp = 0;
    for (long long e = 0; e < pb->g.E; e++) maxp = std::max(maxp, off[e + 1] - off[e]);
    if (maxp > dvc_max_particles(pb->dim, pb->wres)) return GS_E_INVALID;
    const int32_t r = dvc_weights(pb->dim, pb->wres, pb->g.E, pb->off, pb->pid, pb->pk, pb->wp, maxp, st);
    if (r != GS_OK) return r;
    CK(hipStreamSynchronize(st));  // relocation returns with the weights in place
  }
  return esc ? GS_E_ESCAPED : GS_OK;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation K apply (gs_apply_K_noalloc): the tile-face
// partial sums of one apply live in this file-scope device array instead of
// a buffer the library allocates.  Each device holds its own copy until the
// process ends.  In host code `scratch` only stands for the symbol: its
// device address is resolved once per device with hipGetSymbolAddress (with
// that device current, at problem creation) and only that address is handed
// to kernels.  Every call on a device shares that device's array, so calls
// that overlap on one device race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device id

static int32_t scratch_address(int dev, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0) return GS_E_INVALID;
  if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      std::fprintf(stderr, "gale_stokes: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return GS_E_HIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return GS_OK;
}

extern "C" int32_t gs_problem_create(const gs_problem_desc* d, gs_problem** out) {
  if (!d || !out) return GS_E_INVALID;
  *out = nullptr;
  if (d->dim != 2 && d->dim != 3) return GS_E_INVALID;
  if (d->n_particles < 0 || (d->n_particles > 0 && (!d->coord || !d->material))) return GS_E_INVALID;
  if (d->n_materials <= 0 || !d->materials) return GS_E_INVALID;
  if (gs_device_count() <= 0) return GS_E_NO_DEVICE;
  for (int i = 0; i < d->dim; i++)
    if (d->res[i] <= 0 || !(d->hi[i] > d->lo[i])) return GS_E_INVALID;
  auto* pb = new gs_problem();
  Geo& g = pb->g;
  g.dim = d->dim;
  g.N = 1;
  g.E = 1;
  for (int i = 0; i < 3; i++) {
    g.res[i] = i < d->dim ? d->res[i] : 1;
    g.nn[i] = i < d->dim ? d->res[i] + 1 : 1;
    g.lo[i] = i < d->dim ? d->lo[i] : 0.0;
    g.hi[i] = i < d->dim ? d->hi[i] : 1.0;
    g.N *= g.nn[i];
    g.E *= g.res[i];
  }
  if (g.E >= (1ll << 31) - 1 || d->n_particles >= (1ll << 31) - 1) {
    delete pb;
    return GS_E_INVALID;
  }
  pb->dim = d->dim;
  pb->nen = d->dim == 2 ? 4 : 8;
  pb->gravity = d->gravity;
  for (int i = 0; i < 3; i++) pb->ghat[i] = d->ghat[i];
  pb->clamp_min = d->eta_clamp_min;
  pb->clamp_max = d->eta_clamp_max;
  pb->mats_h.assign(d->materials, d->materials + d->n_materials);
  for (long long i = 0; i < d->n_particles; i++)
    if (d->material[i] < 0 || d->material[i] >= d->n_materials) {
      delete pb;
      return GS_E_INVALID;
    }
  auto fail = [&](int32_t s) {
    gs_problem_destroy(pb);
    return s;
  };
  if (hipGetDevice(&pb->device) != hipSuccess) return fail(GS_E_HIP);
  if (scratch_address(pb->device, &pb->scratch) != GS_OK) return fail(GS_E_HIP);
  if (hipStreamCreate(&pb->st) != hipSuccess) return fail(GS_E_HIP);
  const long long ND = g.N * g.dim;
  std::vector<MatDev> md(d->n_materials);
  for (int i = 0; i < d->n_materials; i++) {
    const gs_material& m = d->materials[i];
    md[i] = MatDev{m.eta0, m.density, m.yield_type, 0, m.cohesion, m.friction, m.min_yield_stress, m.min_viscosity,
                   m.max_strain_rate, 0.0, 0.0, 0.0, 0.0, 0.0};
  }
  pb->mats_dh = md;
  std::vector<uint8_t> bcdof(ND, 0), bcflag(g.N, 0);
  std::vector<double> ubc(ND, 0.0);
  long long nfree = 0;
  for (long long k = 0; k < ND; k++) {
    const uint8_t m = d->bc_mask ? (d->bc_mask[k] ? 1 : 0) : 0;
    bcdof[k] = m;
    if (m) {
      bcflag[k / g.dim] |= (uint8_t)(1u << (k % g.dim));
      ubc[k] = d->bc_value ? d->bc_value[k] : 0.0;
      pb->has_bc = true;
    } else {
      nfree++;
    }
  }
  pb->n_free = nfree;
  {
    bool closed = true;
    for (long long nd = 0; nd < g.N && closed; nd++) {
      long long idx[3] = {nd % g.nn[0], (nd / g.nn[0]) % g.nn[1], nd / (g.nn[0] * g.nn[1])};
      for (int dd = 0; dd < g.dim; dd++)
        if ((idx[dd] == 0 || idx[dd] == g.res[dd]) && !(bcflag[nd] & (1u << dd))) closed = false;
    }
    pb->closed = closed;
  }
  const long long P = d->n_particles;
  pb->P = P;
  if (dalloc(&pb->mats_d, d->n_materials) || dalloc(&pb->bcdof, ND) || dalloc(&pb->bcflag, g.N) ||
      dalloc(&pb->ubc, ND) || dalloc(&pb->pk, P) || dalloc(&pb->coord, 3 * P) || dalloc(&pb->coord2, 3 * P) ||
      dalloc(&pb->mat, P) || dalloc(&pb->mat2, P) || dalloc(&pb->pid, P) || dalloc(&pb->pid2, P) ||
      dalloc(&pb->off, g.E + 1) || dalloc(&pb->keys, P) || dalloc(&pb->keys2, P) || dalloc(&pb->vals, P) ||
      dalloc(&pb->vals2, P) || dalloc(&pb->S, S_COUNT) || dalloc(&pb->halo64, tile_halo_size(g)) ||
      dalloc(&pb->halo32, tile_halo_size(g)))
    return fail(GS_E_HIP);
  if (d->eta0_particle && (dalloc(&pb->eta0p, P) || dalloc(&pb->eta0p2, P))) return fail(GS_E_HIP);
  if (hipHostMalloc(reinterpret_cast<void**>(&pb->hS), sizeof(double) * S_COUNT) != hipSuccess) return fail(GS_E_HIP);
  if (alloc_reducer(&pb->rd)) return fail(GS_E_HIP);
  std::vector<int> ids(P > 0 ? P : 1);
  for (long long i = 0; i < P; i++) ids[i] = (int)i;
  if (hipMemcpy(pb->mats_d, md.data(), sizeof(MatDev) * md.size(), hipMemcpyHostToDevice) ||
      hipMemcpy(pb->bcdof, bcdof.data(), ND, hipMemcpyHostToDevice) ||
      hipMemcpy(pb->bcflag, bcflag.data(), g.N, hipMemcpyHostToDevice) ||
      hipMemcpy(pb->ubc, ubc.data(), sizeof(double) * ND, hipMemcpyHostToDevice) ||
      (P > 0 && (hipMemcpy(pb->coord, d->coord, sizeof(double) * 3 * P, hipMemcpyHostToDevice) ||
                 hipMemcpy(pb->mat, d->material, sizeof(int) * P, hipMemcpyHostToDevice) ||
                 hipMemcpy(pb->pid, ids.data(), sizeof(int) * P, hipMemcpyHostToDevice))) ||
      (d->eta0_particle && hipMemcpy(pb->eta0p, d->eta0_particle, sizeof(double) * P, hipMemcpyHostToDevice)))
    return fail(GS_E_HIP);
  launch_fill(pb->S, S_COUNT, 0.0, pb->st);
  const int32_t rs = relocate(pb);
  if (rs == GS_E_HIP) return fail(rs);
  *out = pb;
  return rs;  // GS_E_ESCAPED leaves a usable problem with the outside particles dropped from every cell
}

extern "C" void gs_problem_destroy(gs_problem* pb) {
  if (!pb) return;
  if (pb->st) (void)hipStreamSynchronize(pb->st);
  dfree(pb->mats_d); dfree(pb->bcdof); dfree(pb->bcflag); dfree(pb->ubc); dfree(pb->pk); dfree(pb->pk32);
  dfree(pb->coord); dfree(pb->coord2); dfree(pb->mat); dfree(pb->mat2); dfree(pb->pid); dfree(pb->pid2);
  dfree(pb->eta0p); dfree(pb->eta0p2); dfree(pb->off); dfree(pb->keys); dfree(pb->keys2); dfree(pb->vals);
  dfree(pb->wp); dfree(pb->pstr); dfree(pb->pstr2); dfree(pb->pinc); dfree(pb->pinc2);
  dfree(pb->vals2); dfree(pb->S); dfree(pb->rd.partials); dfree(pb->halo64);
  dfree(pb->halo32);
  if (pb->sort_tmp) (void)hipFree(pb->sort_tmp);
  if (pb->hS) (void)hipHostFree(pb->hS);
  if (pb->st) (void)hipStreamDestroy(pb->st);
  delete pb;
}

extern "C" int32_t gs_problem_info(const gs_problem* pb, int64_t* out5) {
  if (!pb || !out5) return GS_E_INVALID;
  out5[0] = pb->g.N;
  out5[1] = pb->g.E;
  out5[2] = pb->P;
  out5[3] = pb->dim;
  out5[4] = pb->escaped;
  return GS_OK;
}

extern "C" int32_t gs_problem_particles(gs_problem* pb, double* coord, int32_t* owner, double* xi, double* weight) {
  if (!pb) return GS_E_INVALID;
  const long long P = pb->P, E = pb->g.E;
  const int dim = pb->dim;
  CK(hipStreamSynchronize(pb->st));
  std::vector<int> pid(P > 0 ? P : 1);
  std::vector<unsigned> keys(P > 0 ? P : 1);
  CK(hipMemcpy(pid.data(), pb->pid, sizeof(int) * P, hipMemcpyDeviceToHost));
  CK(hipMemcpy(keys.data(), pb->keys, sizeof(unsigned) * P, hipMemcpyDeviceToHost));
  if (coord) {
    std::vector<double> c(3 * P > 0 ? 3 * P : 1);
    CK(hipMemcpy(c.data(), pb->coord, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
    for (long long i = 0; i < P; i++)
      for (int d = 0; d < 3; d++) coord[3 * (long long)pid[i] + d] = c[3 * i + d];
  }
  if (owner)
    for (long long i = 0; i < P; i++) owner[pid[i]] = keys[i] >= (unsigned)E ? -1 : (int32_t)keys[i];
  if (xi || weight) {
    std::vector<double4> pk(P > 0 ? P : 1);
    std::vector<int> off(E + 1);
    CK(hipMemcpy(pk.data(), pb->pk, sizeof(double4) * P, hipMemcpyDeviceToHost));
    CK(hipMemcpy(off.data(), pb->off, sizeof(int) * (E + 1), hipMemcpyDeviceToHost));
    std::vector<double> wv;
    if (pb->wkind == GS_WEIGHTS_DVC && P > 0) {
      wv.resize(P);
      CK(hipMemcpy(wv.data(), pb->wp, sizeof(double) * P, hipMemcpyDeviceToHost));
    }
    for (long long e = 0; e < E; e++)
      for (int i = off[e]; i < off[e + 1]; i++) {
        const long long o = pid[i];
        if (xi) {
          xi[o * dim] = pk[i].x;
          xi[o * dim + 1] = pk[i].y;
          if (dim == 3) xi[o * dim + 2] = pk[i].z;
        }
        if (weight) weight[o] = wv.empty() ? (dim == 2 ? 4.0 : 8.0) / (double)(off[e + 1] - off[e]) : wv[i];
      }
    for (long long i = off[E]; i < P; i++) {
      const long long o = pid[i];
      if (xi)
        for (int d = 0; d < dim; d++) xi[o * dim + d] = 0.0;
      if (weight) weight[o] = 0.0;
    }
  }
  return GS_OK;
}

extern "C" int32_t gs_relocate(gs_problem* pb, int32_t* owner_out) {
  if (!pb) return GS_E_INVALID;
  const int32_t st = relocate(pb);
  if (st == GS_E_HIP) return st;
  if (owner_out) CKS(gs_problem_particles(pb, nullptr, owner_out, nullptr, nullptr));
  return st;
}

extern "C" int32_t gs_problem_set_strain_weakening(gs_problem* pb, int32_t material, const gs_strain_weakening* sw) {
  if (!pb || material < 0 || material >= (int32_t)pb->mats_dh.size()) return GS_E_INVALID;
  MatDev& m = pb->mats_dh[material];
  if (sw) {
    // StrainWeakening belongs to a yield rheology (YieldRheology.cxx:141-147)
    if (m.yield_type == GS_YIELD_NONE) return GS_E_INVALID;
    if (!(sw->final_softening_strain > sw->initial_softening_strain) || sw->healing_rate < 0.0) return GS_E_INVALID;
    m.sw_on = 1;
    m.sw_healing = sw->healing_rate;
    m.sw_eps0 = sw->initial_softening_strain;
    m.sw_eps1 = sw->final_softening_strain;
    m.cohesion_after = sw->cohesion_after;
    m.friction_after = sw->friction_after;
  } else {
    m.sw_on = 0;
  }
  CK(hipStreamSynchronize(pb->st));
  if (sw && !pb->pstr) {
    // _StrainWeakening_Initialise: no initial damage, zero increments
    const long long PA = pb->P > 0 ? pb->P : 1;
    CK(dalloc(&pb->pstr, PA));
    CK(dalloc(&pb->pstr2, PA));
    CK(dalloc(&pb->pinc, PA));
    CK(dalloc(&pb->pinc2, PA));
    CK(hipMemset(pb->pstr, 0, sizeof(double) * PA));
    CK(hipMemset(pb->pinc, 0, sizeof(double) * PA));
  }
  CK(hipMemcpy(pb->mats_d, pb->mats_dh.data(), sizeof(MatDev) * pb->mats_dh.size(), hipMemcpyHostToDevice));
  pb->gen++;  // solvers re-run their constitutive pass
  return GS_OK;
}

extern "C" int32_t gs_problem_set_strain_limit(gs_problem* pb, int32_t shape_type, const double* shape_params,
                                               double limit) {
  if (!pb || shape_type < -1 || shape_type > 2 || (shape_type > 0 && !shape_params)) return GS_E_INVALID;
  pb->slim.type = shape_type;
  pb->slim.dim = pb->dim;
  for (int i = 0; i < 6; i++) pb->slim.p[i] = shape_type > 0 ? shape_params[i] : 0.0;
  pb->slim.limit = limit;
  return GS_OK;
}

extern "C" int32_t gs_problem_set_plastic_strain(gs_problem* pb, const double* w) {
  if (!pb || !pb->pstr || (!w && pb->P > 0)) return GS_E_INVALID;
  const long long P = pb->P;
  if (P == 0) return GS_OK;
  CK(hipStreamSynchronize(pb->st));
  std::vector<int> pid(P);
  std::vector<double> ws(P);
  CK(hipMemcpy(pid.data(), pb->pid, sizeof(int) * P, hipMemcpyDeviceToHost));
  for (long long i = 0; i < P; i++) ws[i] = w[pid[i]];
  CK(hipMemcpy(pb->pstr, ws.data(), sizeof(double) * P, hipMemcpyHostToDevice));
  pb->gen++;
  return GS_OK;
}

extern "C" int32_t gs_problem_plastic_strain(gs_problem* pb, double* w, double* inc) {
  if (!pb || !pb->pstr) return GS_E_INVALID;
  const long long P = pb->P;
  if (P == 0) return GS_OK;
  CK(hipStreamSynchronize(pb->st));
  std::vector<int> pid(P);
  std::vector<double> v(P);
  CK(hipMemcpy(pid.data(), pb->pid, sizeof(int) * P, hipMemcpyDeviceToHost));
  if (w) {
    CK(hipMemcpy(v.data(), pb->pstr, sizeof(double) * P, hipMemcpyDeviceToHost));
    for (long long i = 0; i < P; i++) w[pid[i]] = v[i];
  }
  if (inc) {
    CK(hipMemcpy(v.data(), pb->pinc, sizeof(double) * P, hipMemcpyDeviceToHost));
    for (long long i = 0; i < P; i++) inc[pid[i]] = v[i];
  }
  return GS_OK;
}

extern "C" int32_t gs_problem_set_weights(gs_problem* pb, int32_t kind, const int32_t* res) {
  if (!pb || (kind != GS_WEIGHTS_CONSTANT && kind != GS_WEIGHTS_DVC)) return GS_E_INVALID;
  if (kind == GS_WEIGHTS_DVC) {
    if (!res) return GS_E_INVALID;
    int r[3] = {res[0], res[1], pb->dim == 3 ? res[2] : 1};
    if (dvc_max_particles(pb->dim, r) <= 0) return GS_E_INVALID;
    if (!pb->wp && pb->P > 0) CK(dalloc(&pb->wp, pb->P));
    for (int d = 0; d < 3; d++) pb->wres[d] = r[d];
  }
  pb->wkind = kind;
  CK(hipStreamSynchronize(pb->st));
  return relocate(pb);
}

// One PCDVC pass (PCDVC.cxx:656-1090 / 1092-1512).  The cells run on the
// device, one workgroup each, up to the deletions whose order depends on the
// swarm's global indices; the host replays the swarm's array operations in
// cell order (appends, last-particle moves; an integer walk over the cells
// that changed) to fix that order and the new caller order; the device then
// finishes the changed cells and rebuilds the particle arrays.
//
// With Inflow the cells share one rand() stream in cell order, and a
// re-populated cell's split draws depend on the values it drew, so the
// offset of a cell is known only once every earlier re-populated cell is
// done.  The re-populated cells run in windows, each window assuming its
// cells make no split draws; the first cell that makes some is still right
// (its offset was), the window's later cells are run again from the
// corrected offsets.
extern "C" int32_t gs_problem_population_control(gs_problem* pb, const gs_pcdvc_params* pp, int64_t* counts) {
  if (!pb || !pp || pp->max_splits < 0 || pp->max_deletions < 0) return GS_E_INVALID;
  if (pb->escaped) return GS_E_ESCAPED;  // every particle must sit in a cell table
  const int inflow = pp->inflow != 0;
  if (inflow && (pp->particles_per_cell < 0 || pp->rand_seed < 0 || pp->rand_seed > 0xffffffffLL || pp->rand_skip < 0))
    return GS_E_INVALID;
  int r[3] = {pp->res[0], pp->res[1], pb->dim == 3 ? pp->res[2] : 1};
  PcParams A{};
  if (!pc_grid(pb->dim, r, &A)) return GS_E_INVALID;
  double upT = pp->upper_t, lowT = pp->lower_t;
  if (upT < lowT) {  // PCDVC.cxx:265-270 resets inconsistent thresholds
    lowT = 0.6;
    upT = 25.0;
  }
  const double V = pb->dim == 2 ? 4.0 : 8.0;
  A.max_w = upT * V / 100.0;
  A.min_w = lowT * V / 100.0;
  A.thresh = pp->threshold;
  A.cpr = pp->cent_pos_ratio;
  A.max_splits = pp->max_splits;
  A.max_deletions = pp->max_deletions;
  A.split_if = pp->split_in_interface_cells;
  A.delete_if = pp->delete_in_interface_cells;
  A.inflow = inflow;
  A.ppc = inflow ? pp->particles_per_cell : 0;
  const double da = A.G.vol;
  A.one_oda = (int)(1.0 / da + 0.5);
  if (inflow)  // a region list longer than the region would hold entries the reference never wrote
    for (int k = 1; k <= A.G.nc; k++)
      if ((int)(k * da * A.one_oda + 0.5) > k) return GS_E_INVALID;
  hipStream_t st = pb->st;
  const long long P = pb->P, E = pb->g.E;
  CK(hipStreamSynchronize(st));
  std::vector<int> off(E + 1);
  CK(hipMemcpy(off.data(), pb->off, sizeof(int) * (E + 1), hipMemcpyDeviceToHost));
  // staging: one slot per table entry a cell can reach
  std::vector<long long> soff(E + 1, 0);
  long long maxcap = 1, bound_draws = 0;
  for (long long e = 0; e < E; e++) {
    const long long n = off[e + 1] - off[e];
    const long long cap = n > 0 ? n + A.ppc + A.max_splits : 0;
    soff[e + 1] = soff[e] + cap;
    maxcap = std::max(maxcap, cap);
    if (n > 0 && inflow) bound_draws += A.ppc + n + A.max_splits;
  }
  const long long S = soff[E];
  if (maxcap > pc_max_capacity(A) || S >= 0x7fffffffLL) return GS_E_INVALID;
  A.maxcap = (int)maxcap;
  launch_remap(pb->g, pb->coord, pb->keys, pb->pk, P, st);  // local coordinates of the particles

  // device buffers
  long long* d_soff = nullptr;
  int *tbl = nullptr, *tcnt = nullptr, *rsrc = nullptr, *cnts = nullptr, *infdel = nullptr, *cand = nullptr;
  double *rxi = nullptr, *scent = nullptr, *sw = nullptr;
  int *rs = nullptr, *lcells = nullptr, *dlist = nullptr, *gout = nullptr, *fdel = nullptr;
  long long *lroff = nullptr, *goff = nullptr, *fdoff = nullptr;
  int *off2 = nullptr, *fidx = nullptr, *nbase_d = nullptr, *fidx_new = nullptr;
  double *coordN = nullptr, *eta0N = nullptr, *pstrN = nullptr, *pincN = nullptr, *wN = nullptr;
  int *matN = nullptr, *pidN = nullptr;
  double4* pkN = nullptr;
  unsigned* keysN = nullptr;
  auto cleanup = [&]() {
    dfree(d_soff); dfree(tbl); dfree(tcnt); dfree(rsrc); dfree(cnts); dfree(infdel); dfree(cand); dfree(rxi);
    dfree(scent); dfree(sw); dfree(rs); dfree(lcells); dfree(dlist); dfree(gout); dfree(fdel); dfree(lroff);
    dfree(goff); dfree(fdoff); dfree(off2); dfree(fidx); dfree(nbase_d); dfree(fidx_new); dfree(coordN);
    dfree(eta0N); dfree(pstrN); dfree(pincN); dfree(wN); dfree(matN); dfree(pidN); dfree(pkN); dfree(keysN);
  };
  // a refused pass leaves the problem as it was (its weights recomputed)
  auto refuse = [&]() -> int32_t {
    const int32_t rr = relocate(pb);
    return rr == GS_OK ? GS_E_INVALID : rr;
  };
#define PCK(x)                                                                                      \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) {                                                                         \
      std::fprintf(stderr, "gale_stokes: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      cleanup();                                                                                    \
      return GS_E_HIP;                                                                              \
    }                                                                                               \
  } while (0)
  PCK(dalloc(&d_soff, E + 1));
  PCK(dalloc(&tbl, S));
  PCK(dalloc(&tcnt, E));
  PCK(dalloc(&rsrc, S));
  PCK(dalloc(&rxi, 3 * S));
  PCK(dalloc(&cnts, 7 * E));
  PCK(dalloc(&infdel, S));
  PCK(dalloc(&cand, S));
  PCK(dalloc(&scent, 3 * S));
  PCK(dalloc(&sw, S));
  PCK(hipMemcpyAsync(d_soff, soff.data(), sizeof(long long) * (E + 1), hipMemcpyHostToDevice, st));
  // the rand() values this pass can use, from the stream position rand_skip
  if (inflow) {
    if (!(pb->rng_seed == pp->rand_seed && pb->rng_pos == pp->rand_skip)) {
      pb->rng = CRand((unsigned)pp->rand_seed);
      for (long long i = 0; i < pp->rand_skip; i++) (void)pb->rng.next();
      pb->rng_seed = pp->rand_seed;
      pb->rng_pos = pp->rand_skip;
    }
    std::vector<int> hv(bound_draws > 0 ? bound_draws : 1);
    CRand g = pb->rng;
    for (long long i = 0; i < bound_draws; i++) hv[i] = g.next();
    PCK(dalloc(&rs, (long long)hv.size()));
    PCK(hipMemcpyAsync(rs, hv.data(), sizeof(int) * hv.size(), hipMemcpyHostToDevice, st));
    PCK(hipStreamSynchronize(st));
  }
  PcData D{};
  D.off = pb->off;
  D.pid = pb->pid;
  D.mat = pb->mat;
  D.pk = pb->pk;
  D.soff = d_soff;
  D.tbl = tbl;
  D.tcnt = tcnt;
  D.rsrc = rsrc;
  D.rxi = rxi;
  D.nins = cnts;
  D.nspl = cnts + E;
  D.ninfdel = cnts + 2 * E;
  D.ncand = cnts + 3 * E;
  D.cntdel = cnts + 4 * E;
  D.ndraw = cnts + 5 * E;
  D.flags = cnts + 6 * E;
  D.infdel = infdel;
  D.cand = cand;
  D.scent = scent;
  D.sw = sw;
  D.rs = rs;
  D.rs_n = inflow ? bound_draws : 0;
  D.ndraw_list = nullptr;
  std::vector<int> hc(7 * E);
  long long n_repop = 0, n_windows = 0;
  auto fetch_counts = [&]() -> hipError_t {
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipMemcpy(hc.data(), cnts, sizeof(int) * 7 * E, hipMemcpyDeviceToHost);
    return e;
  };
  if (!inflow) {
    A.probe = 0;
    pc_cells(A, D, E, nullptr, nullptr, st);
    PCK(fetch_counts());
  } else {
    // probe: triggers, and the draw counts of the cells that are not re-populated
    A.probe = 1;
    pc_cells(A, D, E, nullptr, nullptr, st);
    PCK(fetch_counts());
    A.probe = 0;
    const int* hdraw = hc.data() + 5 * E;
    const int* hflag = hc.data() + 6 * E;
    std::vector<long long> pre(E + 1, 0);  // offsets with no split draws in re-populated cells
    std::vector<int> trig;
    for (long long e = 0; e < E; e++) {
      pre[e + 1] = pre[e] + hdraw[e];
      if (hflag[e] & kPcTrig) trig.push_back((int)e);
    }
    std::vector<int> extra(trig.size(), 0);  // split draws of each re-populated cell
    const long long T = (long long)trig.size();
    if (T > 0) {
      PCK(dalloc(&lcells, T));
      PCK(dalloc(&lroff, T));
      PCK(dalloc(&dlist, T));
      PCK(hipMemcpyAsync(lcells, trig.data(), sizeof(int) * T, hipMemcpyHostToDevice, st));
    }
    long long pos = 0, delta = 0, win = 256;
    n_repop = T;
    std::vector<long long> ro;
    std::vector<int> got;
    while (pos < T) {
      const long long w = std::min(win, T - pos);
      ro.resize(w);
      for (long long i = 0; i < w; i++) ro[i] = pre[trig[pos + i]] + delta;
      PCK(hipMemcpyAsync(lroff, ro.data(), sizeof(long long) * w, hipMemcpyHostToDevice, st));
      D.ndraw_list = dlist;
      pc_cells(A, D, w, lcells + pos, lroff, st);
      D.ndraw_list = nullptr;
      n_windows++;
      got.resize(w);
      PCK(hipMemcpyAsync(got.data(), dlist, sizeof(int) * w, hipMemcpyDeviceToHost, st));
      PCK(hipStreamSynchronize(st));
      long long k = 0;
      for (; k < w; k++) {
        const int n0 = off[trig[pos + k] + 1] - off[trig[pos + k]];
        const int a = got[k] - (A.ppc + (n0 > 5 ? n0 : 0));
        extra[pos + k] = a;
        if (a != 0) break;  // later cells of the window drew from a wrong offset
      }
      if (k < w) {
        delta += extra[pos + k];
        pos += k + 1;
        win = std::max<long long>(64, 2 * (k + 1));
      } else {
        pos += w;
        win = std::min<long long>(16384, 2 * win);
      }
    }
    // cells that split without re-population draw from their exact offsets now
    std::vector<int> redo;
    std::vector<long long> redo_off;
    long long acc = 0, t = 0;
    for (long long e = 0; e < E; e++) {
      if (t < T && trig[t] == e) {
        acc += extra[t];
        t++;
        continue;
      }
      if (hflag[e] & kPcRedo) {
        redo.push_back((int)e);
        redo_off.push_back(pre[e] + acc);
      }
    }
    if (!redo.empty()) {
      const long long R = (long long)redo.size();
      dfree(lcells);
      dfree(lroff);
      PCK(dalloc(&lcells, R));
      PCK(dalloc(&lroff, R));
      PCK(hipMemcpyAsync(lcells, redo.data(), sizeof(int) * R, hipMemcpyHostToDevice, st));
      PCK(hipMemcpyAsync(lroff, redo_off.data(), sizeof(long long) * R, hipMemcpyHostToDevice, st));
      pc_cells(A, D, R, lcells, lroff, st);
    }
    PCK(fetch_counts());
  }
  const int* h_nins = hc.data();
  const int* h_nspl = hc.data() + E;
  const int* h_ninf = hc.data() + 2 * E;
  const int* h_ncand = hc.data() + 3 * E;
  const int* h_cntdel = hc.data() + 4 * E;
  const int* h_draw = hc.data() + 5 * E;
  const int* h_flag = hc.data() + 6 * E;
  long long draws = 0, added = 0, nnew = 0;
  for (long long e = 0; e < E; e++) {
    if (h_flag[e] & kPcErr) {
      cleanup();
      return refuse();
    }
    draws += h_draw[e];
    nnew += h_nins[e] + h_nspl[e];
  }
  added = nnew;
  // the lists the host needs: Inflow deletions and deletion candidates
  std::vector<int> gcells;
  std::vector<long long> gofs;
  long long gtot = 0;
  for (long long e = 0; e < E; e++)
    if (h_ninf[e] + h_ncand[e] > 0) {
      gcells.push_back((int)e);
      gofs.push_back(gtot);
      gtot += 2LL * (h_ninf[e] + h_ncand[e]);
    }
  std::vector<int> hg(gtot > 0 ? gtot : 1);
  if (!gcells.empty()) {
    const long long G_ = (long long)gcells.size();
    dfree(lcells);
    PCK(dalloc(&lcells, G_));
    PCK(dalloc(&goff, G_));
    PCK(dalloc(&gout, gtot));
    PCK(hipMemcpyAsync(lcells, gcells.data(), sizeof(int) * G_, hipMemcpyHostToDevice, st));
    PCK(hipMemcpyAsync(goff, gofs.data(), sizeof(long long) * G_, hipMemcpyHostToDevice, st));
    pc_gather(D, G_, lcells, goff, gout, st);
    PCK(hipMemcpyAsync(hg.data(), gout, sizeof(int) * gtot, hipMemcpyDeviceToHost, st));
    PCK(hipStreamSynchronize(st));
  }
  // the swarm's global array, cell by cell (PCDVC.cxx:339-610): ids < P are
  // the caller's particles, P + k the k-th particle the pass created
  if (P + nnew > 0x7fffffffLL) {
    cleanup();
    return refuse();
  }
  std::vector<int> arr(P), idx(P + nnew), nbase(E, 0);
  for (long long i = 0; i < P; i++) arr[i] = idx[i] = (int)i;
  auto del = [&](int id) {
    const int k = idx[id], last = arr.back();
    arr[k] = last;
    idx[last] = k;
    arr.pop_back();
    idx[id] = -1;
  };
  std::vector<int> fcells, fdorder;
  std::vector<long long> fdoffs;
  std::vector<std::pair<int, int>> cl;  // (id, reference)
  long long next = P, gi = 0, deleted = 0;
  for (long long e = 0; e < E; e++) {
    if (h_nins[e] + h_nspl[e] + h_ninf[e] + h_ncand[e] == 0) continue;
    nbase[e] = (int)(next - P);
    auto id_of = [&](int code) -> int {
      return code >= 0 ? code : (int)(P + nbase[e] + ((-1 - (long long)code) - soff[e]));
    };
    const long long g0 = gi < (long long)gcells.size() && gcells[gi] == e ? gofs[gi] : -1;
    if (g0 >= 0) gi++;
    for (int k = 0; k < h_nins[e]; k++) {
      idx[next] = (int)arr.size();
      arr.push_back((int)next);
      next++;
    }
    for (int k = 0; k < h_ninf[e]; k++) {
      del(id_of(hg[g0 + 2 * k + 1]));
      deleted++;
    }
    for (int k = 0; k < h_nspl[e]; k++) {
      idx[next] = (int)arr.size();
      arr.push_back((int)next);
      next++;
    }
    if (h_flag[e] & kPcFinal) {
      fcells.push_back((int)e);
      fdoffs.push_back((long long)fdorder.size());
    }
    if (h_cntdel[e] > 0) {
      // candidates sorted by their current index, highest first (compare_indexOnCPU)
      cl.clear();
      for (int k = 0; k < h_ncand[e]; k++) {
        const long long q = g0 + 2LL * (h_ninf[e] + k);
        cl.emplace_back(id_of(hg[q + 1]), hg[q]);
      }
      std::sort(cl.begin(), cl.end(), [&](const std::pair<int, int>& x, const std::pair<int, int>& y) {
        return idx[x.first] > idx[y.first];
      });
      for (int k = 0; k < h_cntdel[e]; k++) {
        del(cl[k].first);
        fdorder.push_back(cl[k].second);
        deleted++;
      }
    }
  }
  const long long P2 = (long long)arr.size();
  if (!fcells.empty()) {
    const long long F = (long long)fcells.size();
    dfree(lcells);
    PCK(dalloc(&lcells, F));
    PCK(dalloc(&fdoff, F));
    PCK(dalloc(&fdel, (long long)fdorder.size()));
    PCK(hipMemcpyAsync(lcells, fcells.data(), sizeof(int) * F, hipMemcpyHostToDevice, st));
    PCK(hipMemcpyAsync(fdoff, fdoffs.data(), sizeof(long long) * F, hipMemcpyHostToDevice, st));
    if (!fdorder.empty())
      PCK(hipMemcpyAsync(fdel, fdorder.data(), sizeof(int) * fdorder.size(), hipMemcpyHostToDevice, st));
    pc_final(A, D, F, lcells, fdel, fdoff, st);
    PCK(hipStreamSynchronize(st));
    PCK(hipMemcpy(hc.data() + 6 * E, D.flags, sizeof(int) * E, hipMemcpyDeviceToHost));
    for (long long e = 0; e < E; e++)
      if (h_flag[e] & kPcErr) {
        cleanup();
        return refuse();
      }
  }
  // new sorted order: cells in order, each in its final table order
  std::vector<int> htc(E), hoff2(E + 1, 0);
  PCK(hipMemcpy(htc.data(), tcnt, sizeof(int) * E, hipMemcpyDeviceToHost));
  for (long long e = 0; e < E; e++) hoff2[e + 1] = hoff2[e] + htc[e];
  if (hoff2[E] != P2) {  // the tables and the array must hold the same particles
    cleanup();
    return refuse();
  }
  std::vector<int> hpid(P > 0 ? P : 1), hf(P > 0 ? P : 1), hfn(nnew > 0 ? nnew : 1);
  PCK(hipMemcpy(hpid.data(), pb->pid, sizeof(int) * P, hipMemcpyDeviceToHost));
  for (long long i = 0; i < P; i++) hf[i] = idx[hpid[i]];
  for (long long k = 0; k < nnew; k++) hfn[k] = idx[P + k];
  const long long PA = P2 > 0 ? P2 : 1;
  PCK(dalloc(&off2, E + 1));
  PCK(dalloc(&fidx, P));
  PCK(dalloc(&nbase_d, E));
  PCK(dalloc(&fidx_new, nnew));
  PCK(hipMemcpyAsync(off2, hoff2.data(), sizeof(int) * (E + 1), hipMemcpyHostToDevice, st));
  PCK(hipMemcpyAsync(fidx, hf.data(), sizeof(int) * hf.size(), hipMemcpyHostToDevice, st));
  PCK(hipMemcpyAsync(nbase_d, nbase.data(), sizeof(int) * E, hipMemcpyHostToDevice, st));
  PCK(hipMemcpyAsync(fidx_new, hfn.data(), sizeof(int) * hfn.size(), hipMemcpyHostToDevice, st));
  PCK(dalloc(&coordN, 3 * PA));
  PCK(dalloc(&matN, PA));
  PCK(dalloc(&pidN, PA));
  PCK(dalloc(&pkN, PA));
  PCK(dalloc(&wN, PA));
  PCK(dalloc(&keysN, PA));
  if (pb->eta0p) PCK(dalloc(&eta0N, PA));
  if (pb->pstr) {
    PCK(dalloc(&pstrN, PA));
    PCK(dalloc(&pincN, PA));
  }
  pc_rebuild(pb->g, D, off2, fidx, nbase_d, fidx_new, pb->coord, pb->mat, PartDoubles{{pb->eta0p, pb->pstr, pb->pinc}},
             coordN, matN, pidN, PartDoubles{{eta0N, pstrN, pincN}}, pkN, wN, keysN, st);
  PCK(hipMemcpyAsync(pb->off, off2, sizeof(int) * (E + 1), hipMemcpyDeviceToDevice, st));
  PCK(hipStreamSynchronize(st));
  // the problem takes the new particle arrays
  dfree(pb->coord); dfree(pb->coord2); dfree(pb->mat); dfree(pb->mat2); dfree(pb->pid); dfree(pb->pid2);
  dfree(pb->eta0p); dfree(pb->eta0p2); dfree(pb->pk); dfree(pb->pk32); dfree(pb->keys); dfree(pb->keys2);
  dfree(pb->vals); dfree(pb->vals2); dfree(pb->wp);
  dfree(pb->pstr); dfree(pb->pstr2); dfree(pb->pinc); dfree(pb->pinc2);
  pb->coord = coordN; coordN = nullptr;
  pb->mat = matN; matN = nullptr;
  pb->pid = pidN; pidN = nullptr;
  pb->pk = pkN; pkN = nullptr;
  pb->wp = wN; wN = nullptr;
  pb->keys = keysN; keysN = nullptr;
  pb->eta0p = eta0N; eta0N = nullptr;
  pb->pstr = pstrN; pstrN = nullptr;
  pb->pinc = pincN; pincN = nullptr;
  cleanup();
#undef PCK
  pb->P = P2;
  CK(dalloc(&pb->coord2, 3 * PA));
  CK(dalloc(&pb->mat2, PA));
  CK(dalloc(&pb->pid2, PA));
  CK(dalloc(&pb->keys2, PA));
  CK(dalloc(&pb->vals, PA));
  CK(dalloc(&pb->vals2, PA));
  if (pb->eta0p) CK(dalloc(&pb->eta0p2, PA));
  if (pb->pstr) {
    CK(dalloc(&pb->pstr2, PA));
    CK(dalloc(&pb->pinc2, PA));
  }
  pb->pk32_valid = false;
  pb->wkind = GS_WEIGHTS_DVC;
  for (int d = 0; d < 3; d++) pb->wres[d] = r[d];
  pb->escaped = 0;
  pb->min_one_particle = true;
  for (long long e = 0; e < E; e++)
    if (htc[e] == 0) {
      pb->min_one_particle = false;
      break;
    }
  pb->gen++;  // solvers re-run their setup on the new particles
  if (inflow) {
    for (long long i = 0; i < draws; i++) (void)pb->rng.next();
    pb->rng_pos += draws;
  }
  if (counts) {
    counts[0] = added;
    counts[1] = deleted;
    counts[2] = draws;
    counts[3] = n_repop;
    counts[4] = n_windows;
  }
  return GS_OK;
}

extern "C" int32_t gs_problem_set_coords(gs_problem* pb, const double* coord) {
  if (!pb || (!coord && pb->P > 0)) return GS_E_INVALID;
  const long long P = pb->P;
  CK(hipStreamSynchronize(pb->st));
  std::vector<int> ids(P > 0 ? P : 1);
  for (long long i = 0; i < P; i++) ids[i] = (int)i;
  // restore caller order for material / eta0 first
  std::vector<int> pid(P > 0 ? P : 1), mat(P > 0 ? P : 1), mat_o(P > 0 ? P : 1);
  CK(hipMemcpy(pid.data(), pb->pid, sizeof(int) * P, hipMemcpyDeviceToHost));
  CK(hipMemcpy(mat.data(), pb->mat, sizeof(int) * P, hipMemcpyDeviceToHost));
  for (long long i = 0; i < P; i++) mat_o[pid[i]] = mat[i];
  for (double* arr : {pb->eta0p, pb->pstr, pb->pinc}) {
    if (!arr) continue;
    std::vector<double> e0(P), e0o(P);
    CK(hipMemcpy(e0.data(), arr, sizeof(double) * P, hipMemcpyDeviceToHost));
    for (long long i = 0; i < P; i++) e0o[pid[i]] = e0[i];
    CK(hipMemcpy(arr, e0o.data(), sizeof(double) * P, hipMemcpyHostToDevice));
  }
  CK(hipMemcpy(pb->mat, mat_o.data(), sizeof(int) * P, hipMemcpyHostToDevice));
  CK(hipMemcpy(pb->pid, ids.data(), sizeof(int) * P, hipMemcpyHostToDevice));
  CK(hipMemcpy(pb->coord, coord, sizeof(double) * 3 * P, hipMemcpyHostToDevice));
  return relocate(pb);
}

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------
static gs_uzawa_params default_params() {
  gs_uzawa_params p{};
  p.tolerance = 1e-5;
  p.max_iterations = 1000;
  p.min_iterations = 1;
  p.use_absolute_tolerance = 0;
  p.preconditioner = GS_PC_DIAGONAL;
  p.precision = GS_PREC_FP64;
  p.inner_max_iterations = 0;
  p.inner_rtol_factor = 0.1;
  p.inner_rtol_override = 0.0;
  p.monitor = 0;
  p.check_every = 4;
  p.inner_preconditioner = GS_INNER_MULTIGRID;
  p.mg_smooth = 2;
  return p;
}

extern "C" int32_t gs_solver_set_params(gs_solver* s, const gs_uzawa_params* up) {
  if (!s) return GS_E_INVALID;
  s->up = up ? *up : default_params();
  if (s->up.check_every <= 0) s->up.check_every = 4;
  if (s->up.inner_max_iterations <= 0) s->up.inner_max_iterations = 50000;
  if (s->up.inner_rtol_factor <= 0) s->up.inner_rtol_factor = 0.1;
  if (s->up.mg_smooth <= 0) s->up.mg_smooth = 2;
  // Below the finest level, one smoothing step when the inner solves are
  // relaxed (two or three V-cycles each): over ten 96^3 fields the learned
  // solve took 8 % less time with one step than with two, faster on every
  // field, while the plain solve (tight inner solves) took 15 % more.
  if (s->up.mg_smooth_coarse <= 0) s->up.mg_smooth_coarse = s->up.inner_relax_max > 0 ? 1 : s->up.mg_smooth;
  if (s->up.inner_preconditioner != GS_INNER_JACOBI && s->up.inner_preconditioner != GS_INNER_MULTIGRID)
    return GS_E_INVALID;
  if (s->up.preconditioner != GS_PC_NONE && s->up.preconditioner != GS_PC_DIAGONAL &&
      s->up.preconditioner != GS_PC_LEARNED)
    return GS_E_INVALID;
  s->setup_done = false;
  if (!(s->up.tolerance > 0)) return GS_E_INVALID;
  return GS_OK;
}

extern "C" int32_t gs_solver_create(gs_problem* pb, const gs_uzawa_params* up, gs_solver** out) {
  if (!pb || !out) return GS_E_INVALID;
  *out = nullptr;
  auto* s = new gs_solver();
  s->pb = pb;
  if (gs_solver_set_params(s, up) != GS_OK) {
    delete s;
    return GS_E_INVALID;
  }
  const long long n = pb->g.N * pb->dim, np = pb->g.E, P = pb->P;
  s->n = n;
  s->np = np;
  auto fail = [&](int32_t st) {
    gs_solver_destroy(s);
    return st;
  };
  if (dalloc(&s->f, n) || dalloc(&s->diag, n) || dalloc(&s->dinv, n) || dalloc(&s->h, np) || dalloc(&s->Q, np) ||
      dalloc(&s->Qinv, np) || dalloc(&s->u, n) || dalloc(&s->ufull, n) || dalloc(&s->p, np) || dalloc(&s->r, np) ||
      dalloc(&s->s, np) || dalloc(&s->z, np) || dalloc(&s->q, np) || dalloc(&s->fT, n) || dalloc(&s->vstar, n) ||
      dalloc(&s->cg_r, n) || dalloc(&s->cg_p, n) || dalloc(&s->cg_z, n) || dalloc(&s->cg_Ap, n) || dalloc(&s->tmpN, n) ||
      dalloc(&s->tmpE, np) || dalloc(&s->eta, P) || dalloc(&s->uprev, n) || dalloc(&s->yielded, P) ||
      dalloc(&s->S, S_COUNT))
    return fail(GS_E_HIP);
  s->P_alloc = P;
  if (hipHostMalloc(reinterpret_cast<void**>(&s->hS), sizeof(double) * S_COUNT) != hipSuccess) return fail(GS_E_HIP);
  if (alloc_reducer(&s->rd)) return fail(GS_E_HIP);
  if (hipEventCreate(&s->e0) != hipSuccess || hipEventCreate(&s->e1) != hipSuccess) return fail(GS_E_HIP);
  hipStream_t st = pb->st;
  launch_fill(s->S, S_COUNT, 0.0, st);
  launch_fill(s->p, np, 0.0, st);
  launch_fill(s->u, n, 0.0, st);
  launch_fill(s->ufull, n, 0.0, st);
  if (hipStreamSynchronize(st) != hipSuccess) return fail(GS_E_HIP);
  *out = s;
  return GS_OK;
}

extern "C" void gs_solver_destroy(gs_solver* s) {
  if (!s) return;
  if (s->pb && s->pb->st) (void)hipStreamSynchronize(s->pb->st);
  dfree(s->f); dfree(s->diag); dfree(s->dinv); dfree(s->h); dfree(s->Q); dfree(s->Qinv); dfree(s->u);
  dfree(s->ufull); dfree(s->p); dfree(s->r); dfree(s->s); dfree(s->z); dfree(s->q); dfree(s->fT); dfree(s->vstar);
  mg_free(s->mg);
  dfree(s->cg_z);
  dfree(s->cg_r); dfree(s->cg_p); dfree(s->cg_Ap); dfree(s->tmpN); dfree(s->tmpE); dfree(s->eta); dfree(s->uprev);
  dfree(s->yielded); dfree(s->x32); dfree(s->y32); dfree(s->S); dfree(s->rd.partials);
  dfree(s->logeta); dfree(s->rec_r0); dfree(s->rec_p0); dfree(s->fl_s); dfree(s->fl_q); dfree(s->fl_aux);
  if (s->hS) (void)hipHostFree(s->hS);
  if (s->e0) (void)hipEventDestroy(s->e0);
  if (s->e1) (void)hipEventDestroy(s->e1);
  delete s;
}

static int32_t read_scalars(gs_solver* s) {
  CK(hipMemcpyAsync(s->hS, s->S, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s->pb->st));
  CK(hipStreamSynchronize(s->pb->st));
  return GS_OK;
}

static void apply_K_masked(gs_solver* s, const double* x, double* y) {
  gs_problem* pb = s->pb;
  launch_apply_K(pb->g, x, y, pb->pk, pb->off, pb->bcflag, pb->halo64, pb->st);
  s->k_applies++;
}

// ---------------------------------------------------------------------------
// Geometric multigrid (levels from the particle moments, gs_mg.hip).
// ---------------------------------------------------------------------------
static void mg_free(MGData& mg) {
  for (auto& L : mg.lv) {
    dfree(L.M64); dfree(L.M32); dfree(L.diag); dfree(L.dinv64); dfree(L.dinv32);
    dfree(L.b); dfree(L.x); dfree(L.r); dfree(L.d); dfree(L.Ax);
    dfree(L.bf); dfree(L.xf); dfree(L.rf); dfree(L.df); dfree(L.Axf); dfree(L.h64); dfree(L.h32);
    dfree(L.x2); dfree(L.xf2);
    if (L.own_bc) {
      dfree(L.bcflag);
      dfree(L.bcdof);
    }
  }
  mg.lv.clear();
  dfree(mg.freeidx); dfree(mg.Kinv64); dfree(mg.Kinv32);
  mg.structured = false;
  mg.L = 0;
  mg.nfc = 0;
  mg.ldk = 0;
}

static int32_t mg_alloc_level(MGLevel& L) {
  L.n = L.g.N * L.g.dim;
  CK(dalloc(&L.M64, mom_count<double>(L.g))); CK(dalloc(&L.M32, mom_count<float>(L.g)));
  CK(dalloc(&L.diag, L.n)); CK(dalloc(&L.dinv64, L.n)); CK(dalloc(&L.dinv32, L.n));
  CK(dalloc(&L.b, L.n)); CK(dalloc(&L.x, L.n)); CK(dalloc(&L.r, L.n)); CK(dalloc(&L.d, L.n)); CK(dalloc(&L.Ax, L.n));
  CK(dalloc(&L.bf, L.n)); CK(dalloc(&L.xf, L.n)); CK(dalloc(&L.rf, L.n)); CK(dalloc(&L.df, L.n));
  CK(dalloc(&L.Axf, L.n)); CK(dalloc(&L.x2, L.n)); CK(dalloc(&L.xf2, L.n));
  CK(dalloc(&L.h64, tile_halo_size(L.g))); CK(dalloc(&L.h32, tile_halo_size(L.g)));
  return GS_OK;
}

// Level geometry and Dirichlet flags.  A coarse node takes the flags of the
// fine node it coincides with; coarsening stops where that would not make P
// map coarse free dofs onto fine free dofs exactly (then the Galerkin
// identity P^T K P = K_coarse would not hold for the constrained operator).
static int32_t mg_structure(gs_solver* s) {
  gs_problem* pb = s->pb;
  MGData& mg = s->mg;
  mg_free(mg);
  const int dim = pb->dim;
  MGLevel L0;
  L0.g = pb->g;
  L0.bcflag = pb->bcflag;
  L0.bcdof = pb->bcdof;
  mg.lv.push_back(L0);
  CKS(mg_alloc_level(mg.lv.back()));
  std::vector<uint8_t> flag(pb->g.N);
  CK(hipMemcpy(flag.data(), pb->bcflag, pb->g.N, hipMemcpyDeviceToHost));
  // 3D: coarsening stops at the first level small enough for the dense
  // coarsest solve (dense_max_free(), 6000 free dofs): an exact solve there
  // needs fewer inner iterations than smoothing further down, and its
  // matrix-vector product is one wide launch where the levels below would
  // take a dozen small ones per V-cycle (96^3: 96-48-24-12, 12^3 dense).
  // 2D levels coarsen as far as the mesh allows (the batched solver runs each
  // instance's dense coarsest solve inside one workgroup, where a small one
  // is cheaper).  GALE_MG_DENSE_STOP=0 / 1 overrides (A/B timing).
  const char* dstop = std::getenv("GALE_MG_DENSE_STOP");
  const bool dense_stop = dstop ? std::atoi(dstop) != 0 : dim == 3;
  const int dense_max = dense_max_free();
  for (int lev = 0; lev < 16; lev++) {
    const Geo& gf = mg.lv.back().g;
    if (dense_stop && lev > 0) {
      long long nfree = 0;
      for (long long n = 0; n < gf.N; n++)
        for (int d = 0; d < dim; d++) nfree += ((flag[n] >> d) & 1u) ? 0 : 1;
      if (nfree <= dense_max) break;
    }
    bool ok = true;
    for (int d = 0; d < dim; d++) ok = ok && gf.res[d] % 2 == 0 && gf.res[d] / 2 >= 2;
    if (!ok) break;
    Geo gc = gf;
    gc.N = 1;
    gc.E = 1;
    for (int d = 0; d < 3; d++) {
      gc.res[d] = d < dim ? gf.res[d] / 2 : 1;
      gc.nn[d] = d < dim ? gc.res[d] + 1 : 1;
      gc.N *= gc.nn[d];
      gc.E *= gc.res[d];
    }
    std::vector<uint8_t> cflag(gc.N);
    for (long long K = 0; K < gc.nn[2]; K++)
      for (long long J = 0; J < gc.nn[1]; J++)
        for (long long I = 0; I < gc.nn[0]; I++)
          cflag[I + gc.nn[0] * (J + gc.nn[1] * K)] = flag[2 * I + gf.nn[0] * (2 * J + gf.nn[1] * (dim == 3 ? 2 * K : 0))];
    // consistency: a fine dof is constrained iff all its interpolation parents are
    for (long long k = 0; k < gf.nn[2] && ok; k++)
      for (long long j = 0; j < gf.nn[1] && ok; j++)
        for (long long i = 0; i < gf.nn[0] && ok; i++) {
          unsigned all = 7u;
          for (int a = 0; a < ((dim == 3 && (k & 1)) ? 2 : 1); a++)
            for (int b = 0; b < ((j & 1) ? 2 : 1); b++)
              for (int c = 0; c < ((i & 1) ? 2 : 1); c++) {
                const long long I = (i >> 1) + c, J = (j >> 1) + b, K = dim == 3 ? (k >> 1) + a : 0;
                all &= cflag[I + gc.nn[0] * (J + gc.nn[1] * K)];
              }
          const unsigned want = all & ((1u << dim) - 1u);
          if (want != flag[i + gf.nn[0] * (j + gf.nn[1] * k)]) ok = false;
        }
    if (!ok) break;
    MGLevel Lc;
    Lc.g = gc;
    Lc.own_bc = true;
    CK(dalloc(&Lc.bcflag, gc.N));
    CK(dalloc(&Lc.bcdof, gc.N * dim));
    std::vector<uint8_t> cdof(gc.N * dim);
    for (long long n = 0; n < gc.N; n++)
      for (int d = 0; d < dim; d++) cdof[n * dim + d] = (cflag[n] >> d) & 1u;
    CK(hipMemcpy(Lc.bcflag, cflag.data(), gc.N, hipMemcpyHostToDevice));
    CK(hipMemcpy(Lc.bcdof, cdof.data(), gc.N * dim, hipMemcpyHostToDevice));
    mg.lv.push_back(Lc);
    CKS(mg_alloc_level(mg.lv.back()));
    flag.swap(cflag);
  }
  mg.L = (int)mg.lv.size() - 1;
  // coarsest free dofs
  const MGLevel& C = mg.lv[mg.L];
  std::vector<int> fidx;
  for (long long n = 0; n < C.g.N; n++)
    for (int d = 0; d < dim; d++)
      if (!((flag[n] >> d) & 1u)) fidx.push_back((int)(n * dim + d));
  mg.nfc = (int)fidx.size();
  if (mg.nfc <= dense_max) {
    CK(dalloc(&mg.freeidx, mg.nfc));
    CK(hipMemcpy(mg.freeidx, fidx.data(), sizeof(int) * mg.nfc, hipMemcpyHostToDevice));
    // up to kDenseHostMax free dofs the inverse is formed on the host and
    // stored packed; larger ones are inverted on the device and stored with
    // padded rows for the wide solve
    mg.ldk = mg.nfc > kDenseHostMax ? (mg.nfc + 3) & ~3 : 0;
    const long long cols = mg.ldk > 0 ? mg.ldk : mg.nfc;
    CK(dalloc(&mg.Kinv64, (long long)mg.nfc * cols));
    CK(dalloc(&mg.Kinv32, (long long)mg.nfc * cols));
  } else {
    mg.nfc = -1;  // no dense coarse solve: Chebyshev smoothing on the coarsest level
  }
  {
    // The single-workgroup bottom kernel is off by default: with the dense
    // solve at the first level that allows it, the levels above run faster as
    // wide launches than in one workgroup (96^3: 1.03 vs 1.78 ms per inner
    // iteration).  GALE_MG_BOTTOM=1 turns it on (A/B timing).
    const char* ev = std::getenv("GALE_MG_BOTTOM");
    mg.bottom = ev && std::atoi(ev) != 0;
  }
  mg.structured = true;
  return GS_OK;
}

// In-place inverse of an SPD matrix by Cholesky (host, coarsest level only).
// Inverse of the SPD n x n matrix held column-major in A (leading dimension
// ld) on the device: Cholesky factor and inverse from it (rocSOLVER potrf,
// potri on the lower triangle), then the upper triangle filled from the
// lower.  GS_E_NAN when the matrix is not positive definite.
static int32_t spd_inverse_device(double* A, int n, int ld, hipStream_t st) {
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return GS_E_HIP;
  int32_t ret = GS_OK;
  rocblas_int* dinfo = nullptr;
  rocblas_int info = 0;
  if (rocblas_set_stream(h, st) != rocblas_status_success || hipMalloc(&dinfo, sizeof(rocblas_int)) != hipSuccess) {
    ret = GS_E_HIP;
  } else if (rocsolver_dpotrf(h, rocblas_fill_lower, n, A, ld, dinfo) != rocblas_status_success ||
             hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess) {
    ret = GS_E_HIP;
  } else if (info != 0) {
    ret = GS_E_NAN;
  } else if (rocsolver_dpotri(h, rocblas_fill_lower, n, A, ld, dinfo) != rocblas_status_success ||
             hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess) {
    ret = GS_E_HIP;
  } else if (info != 0) {
    ret = GS_E_NAN;
  } else {
    launch_symmetrize_lower(A, n, ld, st);
    if (hipStreamSynchronize(st) != hipSuccess) ret = GS_E_HIP;
  }
  if (dinfo) (void)hipFree(dinfo);
  rocblas_destroy_handle(h);
  return ret;
}

static bool spd_inverse(std::vector<double>& A, int n) {
  std::vector<double> Lm(A);
  for (int j = 0; j < n; j++) {
    double d = Lm[(size_t)j * n + j];
    for (int k = 0; k < j; k++) d -= Lm[(size_t)j * n + k] * Lm[(size_t)j * n + k];
    if (!(d > 0)) return false;
    d = std::sqrt(d);
    Lm[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; i++) {
      double v = Lm[(size_t)i * n + j];
      for (int k = 0; k < j; k++) v -= Lm[(size_t)i * n + k] * Lm[(size_t)j * n + k];
      Lm[(size_t)i * n + j] = v / d;
    }
  }
  // inverse of L, then A^-1 = L^-T L^-1
  std::vector<double> Li((size_t)n * n, 0.0);
  for (int c = 0; c < n; c++) {
    for (int i = c; i < n; i++) {
      double v = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; k++) v -= Lm[(size_t)i * n + k] * Li[(size_t)k * n + c];
      Li[(size_t)i * n + c] = v / Lm[(size_t)i * n + i];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= i; j++) {
      double v = 0.0;
      for (int k = i; k < n; k++) v += Li[(size_t)k * n + i] * Li[(size_t)k * n + j];
      A[(size_t)i * n + j] = v;
      A[(size_t)j * n + i] = v;
    }
  return true;
}

// Per-setup numeric part: moments of every level, Jacobi diagonals, the
// Chebyshev bounds (power iteration on D^-1 K) and the dense coarse inverse.
static int32_t mg_numeric(gs_solver* s) {
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  MGData& mg = s->mg;
  if (!mg.structured) CKS(mg_structure(s));
  const int dim = pb->dim;
  launch_moments_fine(mg.lv[0].g, pb->pk, pb->off, mg.lv[0].M64, st);
  for (int l = 1; l <= mg.L; l++) launch_moments_coarsen(mg.lv[l].g, mg.lv[l - 1].g, mg.lv[l - 1].M64, mg.lv[l].M64, st);
  for (int l = 0; l <= mg.L; l++) {
    MGLevel& L = mg.lv[l];
    launch_mom_d2f(L.g, L.M64, L.M32, st);
    launch_mom_diag(L.g, L.M64, L.diag, st);
    launch_make_dinv(L.diag, L.bcdof, L.dinv64, L.n, st);
    launch_d2f(L.dinv64, L.dinv32, L.n, st);
  }
  // Chebyshev upper bound: power iteration on D^-1 K
  for (int l = 0; l < mg.L || (l == mg.L && mg.nfc < 0); l++) {
    MGLevel& L = mg.lv[l];
    launch_fill(L.x, L.n, 1.0, st);
    launch_masked_copy(L.x, L.x, L.bcdof, L.n, st);
    double lam = 1.0;
    for (int it = 0; it < 12; it++) {
      launch_mom_apply<double>(L.g, L.x, L.Ax, L.M64, L.bcflag, L.h64, st);
      // x <- D^-1 K x / |.|
      launch_mul(L.dinv64, L.Ax, L.d, L.n, st);
      launch_dot(L.d, L.d, L.n, s->S, S_TMP1, s->rd, st);
      launch_dot(L.x, L.x, L.n, s->S, S_TMP2, s->rd, st);
      CKS(read_scalars(s));
      const double dn = std::sqrt(s->hS[S_TMP1]), xn = std::sqrt(s->hS[S_TMP2]);
      if (!(xn > 0) || !std::isfinite(dn)) break;
      lam = dn / xn;
      launch_scale(L.x, L.d, 1.0 / dn, L.n, st);
    }
    L.lmax = lam;
  }
  // dense coarse operator by colour probing (nodes 3 apart do not interact)
  if (mg.nfc > 0) {
    MGLevel& C = mg.lv[mg.L];
    const long long N = C.g.N;
    const int nf = mg.nfc;
    std::vector<int> fidx(nf);
    CK(hipMemcpy(fidx.data(), mg.freeidx, sizeof(int) * nf, hipMemcpyDeviceToHost));
    std::vector<int> pos(C.n, -1);
    for (int i = 0; i < nf; i++) pos[fidx[i]] = i;
    std::vector<double> Kd((size_t)nf * nf, 0.0), v(C.n), y(C.n);
    const int ncz = dim == 3 ? 3 : 1;
    for (int cz = 0; cz < ncz; cz++)
      for (int cy = 0; cy < 3; cy++)
        for (int cx = 0; cx < 3; cx++)
          for (int d = 0; d < dim; d++) {
            std::fill(v.begin(), v.end(), 0.0);
            bool any = false;
            for (long long n = 0; n < N; n++) {
              const long long i = n % C.g.nn[0], j = (n / C.g.nn[0]) % C.g.nn[1], k = n / (C.g.nn[0] * C.g.nn[1]);
              if (i % 3 == cx && j % 3 == cy && k % 3 == cz && pos[n * dim + d] >= 0) {
                v[n * dim + d] = 1.0;
                any = true;
              }
            }
            if (!any) continue;
            CK(hipMemcpyAsync(C.x, v.data(), sizeof(double) * C.n, hipMemcpyHostToDevice, st));
            launch_mom_apply<double>(C.g, C.x, C.Ax, C.M64, C.bcflag, C.h64, st);
            CK(hipMemcpyAsync(y.data(), C.Ax, sizeof(double) * C.n, hipMemcpyDeviceToHost, st));
            CK(hipStreamSynchronize(st));
            for (long long n = 0; n < N; n++) {
              if (v[n * dim + d] == 0.0) continue;
              const int col = pos[n * dim + d];
              const long long i = n % C.g.nn[0], j = (n / C.g.nn[0]) % C.g.nn[1], k = n / (C.g.nn[0] * C.g.nn[1]);
              for (long long kk = k - 1; kk <= k + 1; kk++)
                for (long long jj = j - 1; jj <= j + 1; jj++)
                  for (long long ii = i - 1; ii <= i + 1; ii++) {
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= C.g.nn[0] || jj >= C.g.nn[1] || kk >= C.g.nn[2]) continue;
                    const long long m = ii + C.g.nn[0] * (jj + C.g.nn[1] * kk);
                    for (int d2 = 0; d2 < dim; d2++) {
                      const int row = pos[m * dim + d2];
                      if (row >= 0) Kd[(size_t)row * nf + col] = y[m * dim + d2];
                    }
                  }
            }
          }
    if (mg.ldk == 0) {
      if (!spd_inverse(Kd, nf)) return GS_E_NAN;
      std::vector<float> Kf(Kd.begin(), Kd.end());
      CK(hipMemcpyAsync(mg.Kinv64, Kd.data(), sizeof(double) * Kd.size(), hipMemcpyHostToDevice, st));
      CK(hipMemcpyAsync(mg.Kinv32, Kf.data(), sizeof(float) * Kf.size(), hipMemcpyHostToDevice, st));
      CK(hipStreamSynchronize(st));
    } else {
      // Kd is symmetric, so its row-major layout is also the column-major
      // matrix; padded rows (ld = mg.ldk), padding zero
      const int ld = mg.ldk;
      launch_fill(mg.Kinv64, (long long)nf * ld, 0.0, st);
      CK(hipMemcpy2DAsync(mg.Kinv64, sizeof(double) * ld, Kd.data(), sizeof(double) * nf, sizeof(double) * nf, nf,
                          hipMemcpyHostToDevice, st));
      CKS(spd_inverse_device(mg.Kinv64, nf, ld, st));
      launch_d2f(mg.Kinv64, mg.Kinv32, (long long)nf * ld, st);
      CK(hipStreamSynchronize(st));
    }
  }
  return GS_OK;
}

template <typename T>
struct LvPtr {
  T *b, *x, *x2, *r, *d, *Ax;
  const T *M, *dinv;
  T* h;
};
static LvPtr<double> lvp(MGLevel& L, double*) {
  return {L.b, L.x, L.x2, L.r, L.d, L.Ax, L.M64, L.dinv64, L.h64};
}
static LvPtr<float> lvp(MGLevel& L, float*) {
  return {L.bf, L.xf, L.xf2, L.rf, L.df, L.Axf, L.M32, L.dinv32, L.h32};
}

// Chebyshev-Jacobi smoothing of K x = b on [0.1, 1.1] lambda_max(D^-1 K).
// Each step after the first is one K apply whose node completion performs
// the Chebyshev update (TileEpi kind 1), writing x + d into the other of the
// level's two x buffers.  x == nullptr starts from zero (first step without
// an apply).  Returns the buffer that holds the smoothed x.
template <typename T>
static T* cheb_smooth(MGLevel& L, int nu, T* x, hipStream_t st) {
  LvPtr<T> P = lvp(L, (T*)nullptr);
  const double emax = 1.1 * L.lmax, emin = 0.1 * L.lmax;
  const double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin), sigma = theta / delta;
  double rho_prev = 1.0 / sigma;
  int k0 = 0;
  if (!x) {
    launch_cheb<T>(P.b, nullptr, P.dinv, P.d, P.x, L.n, (T)0, (T)(1.0 / theta), 1, st);
    x = P.x;
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
    T* xo = x == P.x ? P.x2 : P.x;
    TileEpi<T> ep{1, P.b, P.dinv, P.d, xo, nullptr, alpha, beta};
    tile_apply_moments_epi(L.g, x, P.Ax, P.h, P.M, L.bcflag, ep, st);
    x = xo;
  }
  return x;
}

// Levels l..L in one single-workgroup launch when they fit (gs_tile.hip).
template <typename T>
static bool vcycle_bottom(gs_solver* s, int l) {
  MGData& mg = s->mg;
  if (!mg.bottom || !bot_fits<T>(mg.lv[l].g) || mg.L - l + 1 > kBotMaxLevels || mg.nfc > kBotMaxDense) return false;
  BotArgs<T> a{};
  for (int k = l; k <= mg.L; k++) {
    MGLevel& Lk = mg.lv[k];
    LvPtr<T> P = lvp(Lk, (T*)nullptr);
    a.lv[k - l] = BotLevel<T>{Lk.g, P.M, P.dinv, Lk.bcflag, P.b, P.x, P.x2, P.r, P.d, Lk.lmax};
  }
  a.nlev = mg.L - l + 1;
  a.nu = l == 0 ? s->up.mg_smooth : s->up.mg_smooth_coarse;
  a.Kinv = mg.nfc > 0 ? (sizeof(T) == 8 ? (const T*)(void*)mg.Kinv64 : (const T*)(void*)mg.Kinv32) : nullptr;
  a.freeidx = mg.freeidx;
  a.nfc = mg.nfc > 0 ? mg.nfc : 0;
  mg_bottom3(a, s->pb->st);
  return true;
}

// One V-cycle for K x = b at level l (b in the level's b buffer); returns the
// buffer holding x.
template <typename T>
static T* vcycle(gs_solver* s, int l) {
  MGData& mg = s->mg;
  MGLevel& L = mg.lv[l];
  hipStream_t st = s->pb->st;
  const int nu = l == 0 ? s->up.mg_smooth : s->up.mg_smooth_coarse;
  LvPtr<T> P = lvp(L, (T*)nullptr);
  if (vcycle_bottom<T>(s, l)) return P.x;
  if (l == mg.L) {
    if (mg.nfc > 0) {
      launch_dense_solve<T>(sizeof(T) == 8 ? (const T*)(void*)mg.Kinv64 : (const T*)(void*)mg.Kinv32, mg.freeidx,
                            mg.nfc, mg.ldk, P.b, P.x, L.n, st);
      return P.x;
    }
    return cheb_smooth<T>(L, 8, (T*)nullptr, st);
  }
  MGLevel& C = mg.lv[l + 1];
  LvPtr<T> Q = lvp(C, (T*)nullptr);
  T* x = cheb_smooth<T>(L, nu, (T*)nullptr, st);
  TileEpi<T> ep{2, P.b, nullptr, nullptr, nullptr, P.r, 0, 0};
  tile_apply_moments_epi(L.g, x, P.Ax, P.h, P.M, L.bcflag, ep, st);  // r = b - K x
  launch_restrict<T>(C.g, L.g, P.r, Q.b, C.bcflag, st);
  T* xc = vcycle<T>(s, l + 1);
  launch_prolong_add<T>(C.g, L.g, xc, x, L.bcflag, st);
  return cheb_smooth<T>(L, nu, x, st);
}

// z = M^-1 r by one V-cycle (FP64, or FP32 in the mixed-precision mode).
static void mg_precond(gs_solver* s, const double* r, double* z) {
  MGLevel& L = s->mg.lv[0];
  hipStream_t st = s->pb->st;
  if (s->up.precision == GS_PREC_MIXED) {
    launch_d2f(r, L.bf, L.n, st);
    const float* x = vcycle<float>(s, 0);
    launch_f2d(x, z, L.n, st);
  } else {
    (void)hipMemcpyAsync(L.b, r, sizeof(double) * L.n, hipMemcpyDeviceToDevice, st);
    const double* x = vcycle<double>(s, 0);
    (void)hipMemcpyAsync(z, x, sizeof(double) * L.n, hipMemcpyDeviceToDevice, st);
  }
}

// The fused constitutive pass + diagonal operators + right-hand sides.
static int32_t setup_dev(gs_solver* s, const double* u_prev, const double* p_prev, int yield_on) {
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  if (s->P_alloc != pb->P) {  // the particle count changed (population control)
    dfree(s->eta);
    dfree(s->yielded);
    const long long PA = pb->P > 0 ? pb->P : 1;
    CK(dalloc(&s->eta, PA));
    CK(dalloc(&s->yielded, PA));
    s->P_alloc = pb->P;
  }
  if (!pb->min_one_particle) return GS_E_INVALID;
  SetupArgs a{};
  a.u_prev = u_prev;
  a.p_prev = p_prev;
  a.yield_on = yield_on;
  a.mats = pb->mats_d;
  a.mat = pb->mat;
  a.eta0p = pb->eta0p;
  a.pstr = pb->pstr;
  a.pinc = pb->pinc;
  a.wp = pb->wkind == GS_WEIGHTS_DVC ? pb->wp : nullptr;
  a.clamp_min = pb->clamp_min;
  a.clamp_max = pb->clamp_max;
  for (int d = 0; d < 3; d++) a.g[d] = pb->ghat[d] * pb->gravity;
  launch_setup(pb->g, a, pb->pk, pb->off, s->diag, s->f, s->Q, s->eta, s->yielded, st);
  pb->pk32_valid = false;
  launch_make_dinv(s->diag, pb->bcdof, s->dinv, s->n, st);
  launch_recip(s->Q, s->Qinv, s->np, st);
  if (pb->has_bc) {
    launch_apply_K(pb->g, pb->ubc, s->tmpN, pb->pk, pb->off, nullptr, pb->halo64, st);
    s->k_applies++;
    launch_lift(s->f, s->tmpN, pb->bcdof, s->n, st);
    launch_apply_Gt(pb->g, pb->ubc, s->h, st);
    launch_scale(s->h, s->h, -1.0, s->np, st);
  } else {
    launch_fill(s->h, s->np, 0.0, st);
  }
  if (s->up.inner_preconditioner == GS_INNER_MULTIGRID || s->up.precision == GS_PREC_MIXED) CKS(mg_numeric(s));
  if (s->lpc || s->rec) {
    if (!s->logeta) CK(dalloc(&s->logeta, s->np));
    launch_elem_logeta(pb->g, pb->pk, pb->off, s->logeta, st);
  }
  s->setup_done = true;
  s->setup_gen = pb->gen;
  return GS_OK;
}

// ---------------------------------------------------------------------------
// Export of a set-up solver's data for the batched solver (gs_batch.hip).
// The solver must use the multigrid inner preconditioner in mixed precision
// (so FP32 level data exist) and have completed gs_solver_setup.
// ---------------------------------------------------------------------------
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

int32_t gs::batch_mg_info(gs_solver* s, int* nlev, BatchLevelInfo* lv, int max_levels, int* nfc) {
  if (!s || !s->setup_done || !s->mg.structured || s->pb->dim != 2) return GS_E_INVALID;
  const int n = s->mg.L + 1;
  if (n > max_levels) return GS_E_INVALID;
  for (int l = 0; l < n; l++) lv[l] = BatchLevelInfo{s->mg.lv[l].g.res[0], s->mg.lv[l].g.res[1], s->mg.lv[l].lmax};
  *nlev = n;
  *nfc = s->mg.nfc;
  return GS_OK;
}

int32_t gs::batch_copy_fine(gs_solver* s, double* f, double* dinv, double* Qinv, double* M64, hipStream_t st) {
  const MGLevel& L = s->mg.lv[0];
  CK(hipMemcpyAsync(f, s->f, sizeof(double) * s->n, hipMemcpyDeviceToDevice, st));
  CK(hipMemcpyAsync(dinv, s->dinv, sizeof(double) * s->n, hipMemcpyDeviceToDevice, st));
  CK(hipMemcpyAsync(Qinv, s->Qinv, sizeof(double) * s->np, hipMemcpyDeviceToDevice, st));
  CK(hipMemcpyAsync(M64, L.M64, sizeof(double) * L.g.E * n_moments(L.g.dim), hipMemcpyDeviceToDevice, st));
  return GS_OK;
}

int32_t gs::batch_copy_level(gs_solver* s, int l, float* M32, float* dinv32, uint8_t* bcflag, hipStream_t st) {
  if (l < 0 || l > s->mg.L) return GS_E_INVALID;
  const MGLevel& L = s->mg.lv[l];
  CK(hipMemcpyAsync(M32, L.M32, sizeof(float) * L.g.E * n_moments(L.g.dim), hipMemcpyDeviceToDevice, st));
  CK(hipMemcpyAsync(dinv32, L.dinv32, sizeof(float) * L.n, hipMemcpyDeviceToDevice, st));
  CK(hipMemcpyAsync(bcflag, L.bcflag, L.g.N, hipMemcpyDeviceToDevice, st));
  return GS_OK;
}

int32_t gs::batch_copy_dense(gs_solver* s, float* Kinv32, int* freeidx, hipStream_t st) {
  const int nf = s->mg.nfc;
  if (s->mg.ldk != 0) return GS_E_INVALID;  // the batched solver takes packed inverses only
  if (nf <= 0) return GS_OK;
  CK(hipMemcpyAsync(Kinv32, s->mg.Kinv32, sizeof(float) * nf * nf, hipMemcpyDeviceToDevice, st));
  CK(hipMemcpyAsync(freeidx, s->mg.freeidx, sizeof(int) * nf, hipMemcpyDeviceToDevice, st));
  return GS_OK;
}

extern "C" int32_t gs_solver_setup(gs_solver* s, const double* u_prev, const double* p_prev, int32_t yield_on) {
  if (!s) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  const double* ud = nullptr;
  const double* pd = nullptr;
  if (u_prev) {
    CK(hipMemcpyAsync(s->uprev, u_prev, sizeof(double) * s->n, hipMemcpyHostToDevice, st));
    ud = s->uprev;
  } else if (s->has_solution) {
    ud = s->ufull;
  }
  if (p_prev) {
    CK(hipMemcpyAsync(s->tmpE, p_prev, sizeof(double) * s->np, hipMemcpyHostToDevice, st));
    pd = s->tmpE;
  } else if (s->has_solution) {
    pd = s->p;
  }
  CK(hipEventRecord(s->e0, st));
  CKS(setup_dev(s, ud, pd, yield_on));
  CK(hipEventRecord(s->e1, st));
  CK(hipStreamSynchronize(st));
  return GS_OK;
}

// Inner velocity solve K x = b (masked), Jacobi-preconditioned CG from a zero
// guess, stopping at |r| <= rtol |b|.
static int32_t inner_solve_jacobi(gs_solver* s, const double* b, double* x, double rtol, long long* its) {
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  const long long n = s->n;
  launch_cg_init(b, s->dinv, x, s->cg_r, s->cg_p, n, s->S, s->rd, st);
  CKS(read_scalars(s));
  const double bb = s->hS[S_BB];
  if (bb == 0.0) return GS_OK;
  if (!std::isfinite(bb)) return GS_E_NAN;
  const double target = rtol * rtol * bb;
  const int maxit = s->up.inner_max_iterations, every = s->up.check_every;
  int k = 0;
  while (k < maxit) {
    for (int j = 0; j < every && k < maxit; j++, k++) {
      apply_K_masked(s, s->cg_p, s->cg_Ap);
      launch_dot(s->cg_p, s->cg_Ap, n, s->S, S_PAP, s->rd, st);
      launch_cg_update(x, s->cg_r, s->cg_p, s->cg_Ap, s->dinv, n, s->S, s->rd, st);
      launch_cg_pupdate(s->cg_p, s->cg_r, s->dinv, n, s->S, st);
    }
    CKS(read_scalars(s));
    const double rr = s->hS[S_RR];
    if (!std::isfinite(rr)) return GS_E_NAN;
    if (rr <= target) break;
  }
  *its += k;
  if (k >= maxit) {
    CKS(read_scalars(s));
    if (s->hS[S_RR] > target) return GS_E_INNER_MAXIT;
  }
  return GS_OK;
}

// Inner velocity solve with the multigrid V-cycle as a (possibly FP32, hence
// slightly variable) preconditioner of FP64 flexible CG.  The operator is the
// FP64 moment form of K; vectors, dot products and the stopping test
// |r| <= rtol |b| stay FP64.
static int32_t inner_solve_mg(gs_solver* s, const double* b, double* x, double rtol, long long* its) {
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  const long long n = s->n;
  MGLevel& L0 = s->mg.lv[0];
  launch_fill(x, n, 0.0, st);
  CK(hipMemcpyAsync(s->cg_r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  // b.b is read back with the first convergence check, not before the first
  // iteration (one host round trip fewer per inner solve).  With b = 0 every
  // iterate stays exactly zero (the V-cycle of 0 is 0 and alpha, beta are
  // guarded), so that case still returns x = 0 after no counted iteration.
  launch_dot(b, b, n, s->S, S_BB, s->rd, st);
  mg_precond(s, s->cg_r, s->cg_z);
  CK(hipMemcpyAsync(s->cg_p, s->cg_z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  launch_dot(s->cg_r, s->cg_z, n, s->S, S_RZ, s->rd, st);
  const int maxit = s->up.inner_max_iterations;
  // Convergence is checked after every iteration: a V-cycle iteration costs
  // far more than the host round trip, and stopping one iteration earlier
  // pays for many of them (over ten 96^3 fields the learned solve took 5 %
  // less time than with a check every second iteration, faster on all ten;
  // the plain solve the same time).  check_every paces the Jacobi CG.
  const int every = 1;
  int k = 0;
  bool conv = false, first = true;
  double target = 0.0;
  while (k < maxit) {
    launch_mom_apply<double>(L0.g, s->cg_p, s->cg_Ap, L0.M64, pb->bcflag, L0.h64, st);
    s->k_applies++;
    launch_dot(s->cg_p, s->cg_Ap, n, s->S, S_PAP, s->rd, st);
    launch_fcg_update(x, s->cg_r, s->cg_p, s->cg_Ap, n, s->S, s->rd, st);
    k++;
    if (k % every == 0 || k >= maxit) {
      CKS(read_scalars(s));
      if (first) {
        first = false;
        const double bb = s->hS[S_BB];
        if (bb == 0.0) return GS_OK;
        if (!std::isfinite(bb)) return GS_E_NAN;
        target = rtol * rtol * bb;
      }
      const double rr = s->hS[S_RR];
      if (!std::isfinite(rr)) return GS_E_NAN;
      if (rr <= target) {
        conv = true;
        break;
      }
    }
    mg_precond(s, s->cg_r, s->cg_z);
    launch_fcg_dots(s->cg_z, s->cg_r, s->cg_Ap, n, s->S, s->rd, st);
    launch_fcg_p(s->cg_p, s->cg_z, n, s->S, st);
  }
  *its += k;
  if (!conv) return GS_E_INNER_MAXIT;
  return GS_OK;
}

static int32_t inner_solve(gs_solver* s, const double* b, double* x, double rtol, long long* its) {
  if (s->up.inner_preconditioner == GS_INNER_MULTIGRID) return inner_solve_mg(s, b, x, rtol, its);
  return inner_solve_jacobi(s, b, x, rtol, its);
}

static double dot_host(gs_solver* s, const double* a, const double* b, long long n, int32_t* err) {
  launch_dot(a, b, n, s->S, S_TMP1, s->rd, s->pb->st);
  *err = read_scalars(s);
  return s->hS[S_TMP1];
}

// Uzawa with the learned Schur preconditioner: truncated flexible CG.  Each
// preconditioned residual z = M(r) (network correction of Q^-1 r) is
// S-orthogonalised against the last m directions; alpha = s.r / s.q.  A
// network output with z.r <= 0 is replaced by Q^-1 r for that iteration; on
// stagnation (residual down by less than 5% over the window) the iteration
// restarts, and after learned_restarts restarts the rest of the solve uses Q.
// With inner_relax_max > 0 the inner solves are relaxed as the outer
// residual falls (inner rtol = min(inner_relax_max, gain irt |r_0| / |r_k|),
// the inexact-Krylov relaxation rule); the recursive residual then drifts from
// the true one, so convergence is re-checked with an accurate solve of
// u = K^-1 (f - G p), and the loop resumes from the true residual if needed.
// Stopping test, residual definition and firewalls are the reference's.
static int32_t uzawa_learned_loop(gs_solver* s, gs_report* rep, double uzawaRhsScale, double irt, bool recording,
                                  int proj, int* it_io, long long* inner_io, double* relRes_io, double* absRes_io) {
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  const long long n = s->n, np = s->np;
  const double tol = s->up.tolerance;
  const double qRecip = 1.0 / std::sqrt((double)np);
  int m = s->up.flexible_window > 0 ? s->up.flexible_window : 16;
  if (m > kMaxWindow) m = kMaxWindow;
  const int win = s->up.stagnation_window > 0 ? s->up.stagnation_window : 25;
  const double relax = s->up.inner_relax_max > 0 ? s->up.inner_relax_max : 0.0;
  const double gain = s->up.inner_relax_gain > 0 ? s->up.inner_relax_gain : 3.0;
  if (s->fl_m != m) {
    dfree(s->fl_s);
    dfree(s->fl_q);
    s->fl_m = 0;
    CK(dalloc(&s->fl_s, (long long)m * np));
    CK(dalloc(&s->fl_q, (long long)m * np));
    s->fl_m = m;
  }
  if (!s->fl_aux) CK(dalloc(&s->fl_aux, (long long)kMaxReduceBlocks * kMaxWindow + 2 * kMaxWindow));
  double* partials = s->fl_aux;
  double* beta = s->fl_aux + (long long)kMaxReduceBlocks * kMaxWindow;
  double* sqv = beta + kMaxWindow;
  int it = *it_io;
  long long inner = *inner_io;
  int32_t err = GS_OK;
  bool lp_on = true;
  int restarts_left = s->up.learned_restarts > 0 ? s->up.learned_restarts : 0;
  const int q_every = s->up.learned_q_every > 0 ? s->up.learned_q_every : 0;
  constexpr int kMaxBad = 5;
  int bad_left = kMaxBad;
  long long nbad_seen = 0;
  launch_fill(s->S + S_NBAD, 1, 0.0, st);
  int32_t e2 = GS_OK;
  double relRes = std::sqrt(dot_host(s, s->r, s->r, np, &e2)) * qRecip / uzawaRhsScale;
  CKS(e2);
  double absRes = relRes * uzawaRhsScale;
  constexpr int kMaxRounds = 4;
  for (int round = 0; round < kMaxRounds; round++) {
    const double relax_ref = relRes;  // residual the relaxation is measured from
    int base = it;  // first iteration of the current orthogonalisation window
    std::vector<double> hist;
    bool converged = (s->up.use_absolute_tolerance ? absRes : relRes) <= tol && it >= s->up.min_iterations;
    while (!converged) {
      const bool use_net = lp_on && !(q_every > 0 && (it + 1) % q_every == 0);
      if (use_net) {
        // a network output that is not a descent direction (z.r <= 0 or
        // not finite) is replaced by Q^-1 r on the device, without waiting
        // for z.r; the replacements are counted in S[S_NBAD] and read with
        // this iteration's update, and after kMaxBad of them the network is
        // dropped for the solve
        CKS(lpc_precond(s->lpc, pb->g, s->r, s->Qinv, s->logeta, nullptr, s->z, s->S, s->rd, st));
        launch_uz_precond_if_bad(s->r, s->Qinv, s->z, np, s->S, s->rd, st);
      } else {
        launch_uz_precond(s->r, s->Qinv, s->z, np, s->S, s->rd, st);
      }
      const int k = it - base;
      const int c = k < m ? k : m;
      const int slot = k % m;
      double* sk = s->fl_s + (long long)slot * np;
      double* qk = s->fl_q + (long long)slot * np;
      launch_orthogonalise(s->z, s->fl_s, s->fl_q, sqv, c, slot, np, partials, beta, st);
      launch_apply_G(pb->g, sk, s->fT, pb->bcflag, st);
      double irt_k = irt;
      if (relax > 0 && relRes > 0) irt_k = std::min(relax, std::max(irt, gain * irt * relax_ref / relRes));
      CKS(inner_solve(s, s->fT, s->vstar, irt_k, &inner));
      launch_apply_Gt(pb->g, s->vstar, qk, st);
      if (recording && s->rec_every > 0 && it % s->rec_every == 0)
        CKS(lpc_record(s->rec, pb->g, qk, sk, s->Qinv, s->logeta, proj, s->S, s->rd, st));
      launch_flex_alpha(sk, s->r, qk, np, s->S, sqv + slot, s->rd, st);
      launch_uz_update(s->p, sk, s->r, qk, np, s->S, s->rd, st);
      launch_uz_uupdate(s->u, s->vstar, n, s->S, st);
      CKS(read_scalars(s));
      if (use_net) {
        const long long nbad = (long long)s->hS[S_NBAD];
        if (nbad > nbad_seen) {  // this iteration took Q^-1 r
          nbad_seen = nbad;
          if (--bad_left <= 0) {
            lp_on = false;
            rep->learned_fallback = it + 1;
          }
        }
      }
      const double sq = s->hS[S_SQ], alpha = s->hS[S_TMP0];
      if (!std::isfinite(sq) || !std::isfinite(alpha)) {
        err = GS_E_NAN;
        break;
      }
      absRes = std::sqrt(s->hS[S_ORR]) * qRecip;
      relRes = absRes / uzawaRhsScale;
      if (!std::isfinite(absRes)) {
        err = GS_E_NAN;
        break;
      }
      if (s->up.monitor > 0 && (it + 1) % s->up.monitor == 0)
        std::fprintf(stderr, "  %04d uzawa residual norm %12.13e%s\n", it + 1, relRes, lp_on ? " (learned)" : "");
      if (lp_on) {
        hist.push_back(relRes);
        const size_t h = hist.size();
        if (h > (size_t)(2 * win) && relRes > 0.95 * hist[h - 1 - win]) {
          if (restarts_left > 0) {
            restarts_left--;
            hist.clear();
          } else {
            lp_on = false;
            rep->learned_fallback = it + 1;
          }
          base = it + 1;
        }
      }
      if (!(it < s->up.max_iterations)) {
        err = GS_E_MAXIT;
        break;
      }
      it++;
      converged = ((s->up.use_absolute_tolerance ? absRes : relRes) <= tol) && (it >= s->up.min_iterations);
    }
    if (err != GS_OK || relax <= 0) break;
    // true residual with an accurate inner solve (STEP 3 of the reference)
    launch_apply_G(pb->g, s->p, s->tmpN, pb->bcflag, st);
    launch_axpby(s->tmpN, s->f, 1.0, -1.0, n, st);
    CKS(inner_solve(s, s->tmpN, s->u, irt, &inner));
    launch_apply_Gt(pb->g, s->u, s->r, st);
    launch_axpby(s->r, s->h, -1.0, 1.0, np, st);
    absRes = std::sqrt(dot_host(s, s->r, s->r, np, &e2)) * qRecip;
    CKS(e2);
    relRes = absRes / uzawaRhsScale;
    rep->verify_rounds = round + 1;
    if (((s->up.use_absolute_tolerance ? absRes : relRes) <= tol) || !std::isfinite(relRes)) break;
    if (round == kMaxRounds - 1) err = GS_E_MAXIT;
  }
  *it_io = it;
  *inner_io = inner;
  *relRes_io = relRes;
  *absRes_io = absRes;
  return err;
}

// _Stokes_SLE_UzawaSolver_Solve, Stokes_SLE_UzawaSolver.cxx:393-938.
static int32_t uzawa(gs_solver* s, gs_report* rep) {
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  const long long n = s->n, np = s->np;
  const double tol = s->up.tolerance;
  const double qRecip = 1.0 / std::sqrt((double)np);
  long long inner = 0;
  int32_t err = GS_OK;
  int it = 0;
  double relRes = 0.0, absRes = 0.0;
  const bool learned = s->up.preconditioner == GS_PC_LEARNED;
  const double* Qinv = (s->up.preconditioner == GS_PC_DIAGONAL || learned) ? s->Qinv : nullptr;
  if (learned && (!s->lpc || !lpc_shape_ok(pb->g) || !s->logeta)) return GS_E_INVALID;
  const bool recording = s->rec && s->logeta && Qinv;
  // :476-493 zero-force firewall
  const double ff = dot_host(s, s->f, s->f, n, &err);
  CKS(err);
  if (std::sqrt(ff) / std::sqrt((double)(pb->n_free > 0 ? pb->n_free : 1)) <= 1e-99) return GS_E_ZERO_FORCE;
  // STEP 1 (:553-565): vStar = K^-1 f ; qTemp = G^T vStar - h
  const double rt1 = s->up.inner_rtol_override > 0 ? s->up.inner_rtol_override : tol;
  CKS(inner_solve(s, s->f, s->vstar, rt1, &inner));
  launch_apply_Gt(pb->g, s->vstar, s->z, st);
  launch_axpby(s->z, s->h, -1.0, 1.0, np, st);
  // STEP 2 (:587-605)
  const double irt = s->up.inner_rtol_override > 0 ? s->up.inner_rtol_override : s->up.inner_rtol_factor * tol;
  const double uzawaRhsScale = std::sqrt(dot_host(s, s->z, s->z, np, &err)) * qRecip;
  CKS(err);
  if (!std::isfinite(uzawaRhsScale)) return GS_E_NAN;
  // STEP 3 (:616-645): u = K^-1 (f - G q) ; r = G^T u - h
  launch_apply_G(pb->g, s->p, s->tmpN, pb->bcflag, st);
  launch_axpby(s->tmpN, s->f, 1.0, -1.0, n, st);  // tmpN = f - G p
  CKS(inner_solve(s, s->tmpN, s->u, irt, &inner));
  launch_apply_Gt(pb->g, s->u, s->r, st);
  {
    const double rn = dot_host(s, s->r, s->r, np, &err);
    CKS(err);
    const double un = dot_host(s, s->u, s->u, n, &err);
    CKS(err);
    if (!std::isfinite(std::sqrt(rn) / std::sqrt(un))) return GS_E_NAN;
  }
  launch_axpby(s->r, s->h, -1.0, 1.0, np, st);
  if (recording) {
    CK(hipMemcpyAsync(s->rec_r0, s->r, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    CK(hipMemcpyAsync(s->rec_p0, s->p, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
  }
  // Closed domains (every boundary normal fixed) leave the pressure defined up
  // to a constant: training pairs then compare mean-free fields.
  const int proj = pb->closed ? 1 : 0;
  // STEP 4 (:664-837)
  launch_fill(s->S + S_ZR_PREV, 1, 1.0, st);
  if (learned) {
    err = uzawa_learned_loop(s, rep, uzawaRhsScale, irt, recording, proj, &it, &inner, &relRes, &absRes);
  } else {
    do {
      launch_uz_precond(s->r, Qinv, s->z, np, s->S, s->rd, st);
      launch_uz_sdir(s->s, s->z, np, s->S, it == 0, st);
      launch_apply_G(pb->g, s->s, s->fT, pb->bcflag, st);
      CKS(inner_solve(s, s->fT, s->vstar, irt, &inner));
      launch_apply_Gt(pb->g, s->vstar, s->q, st);
      if (recording && s->rec_every > 0 && it % s->rec_every == 0)
        CKS(lpc_record(s->rec, pb->g, s->q, s->s, s->Qinv, s->logeta, proj, s->S, s->rd, st));
      launch_dot(s->s, s->q, np, s->S, S_SQ, s->rd, st);
      launch_uz_update(s->p, s->s, s->r, s->q, np, s->S, s->rd, st);
      launch_uz_uupdate(s->u, s->vstar, n, s->S, st);
      CKS(read_scalars(s));
      const double zr = s->hS[S_ZR], sq = s->hS[S_SQ], alpha = s->hS[S_TMP0];
      if (!std::isfinite(zr) || !std::isfinite(sq) || !std::isfinite(alpha)) {
        err = GS_E_NAN;
        break;
      }
      absRes = std::sqrt(s->hS[S_ORR]) * qRecip;
      relRes = absRes / uzawaRhsScale;
      if (!std::isfinite(absRes)) {
        err = GS_E_NAN;
        break;
      }
      if (s->up.monitor > 0 && (it + 1) % s->up.monitor == 0)
        std::fprintf(stderr, "  %04d uzawa residual norm %12.13e\n", it + 1, relRes);
      if (!(it < s->up.max_iterations)) {
        err = GS_E_MAXIT;
        break;
      }
      it++;
    } while (((s->up.use_absolute_tolerance ? absRes : relRes) > tol) || (it < s->up.min_iterations));
  }
  if (recording && err == GS_OK) {
    // pair (r0, p* - p0)
    CK(hipMemcpyAsync(s->tmpE, s->p, sizeof(double) * np, hipMemcpyDeviceToDevice, st));
    launch_axpby(s->tmpE, s->rec_p0, -1.0, 1.0, np, st);
    CKS(lpc_record(s->rec, pb->g, s->rec_r0, s->tmpE, s->Qinv, s->logeta, proj, s->S, s->rd, st));
  }
  rep->uzawa_iterations += it;
  rep->inner_iterations += inner;
  rep->rel_residual = relRes;
  rep->abs_residual = absRes;
  launch_add(s->ufull, s->u, pb->ubc, n, st);
  s->has_solution = true;
  if (err != GS_OK) return err;
  // STEP 5 summary (:844-870)
  launch_apply_Gt(pb->g, s->u, s->tmpE, st);
  const double dn = dot_host(s, s->tmpE, s->tmpE, np, &err);
  CKS(err);
  const double un = dot_host(s, s->u, s->u, n, &err);
  CKS(err);
  rep->div_u = std::sqrt(dn) / std::sqrt(un);
  apply_K_masked(s, s->u, s->fT);
  launch_apply_G(pb->g, s->p, s->tmpN, pb->bcflag, st);
  launch_add(s->fT, s->fT, s->tmpN, n, st);
  launch_axpby(s->fT, s->f, 1.0, -1.0, n, st);
  const double rn = dot_host(s, s->fT, s->fT, n, &err);
  CKS(err);
  rep->momentum_residual = std::sqrt(rn) / std::sqrt(ff);
  return GS_OK;
}

extern "C" int32_t gs_stokes_solve(gs_solver* s, double* u, double* p, gs_report* rep) {
  if (!s) return GS_E_INVALID;
  gs_report local{};
  gs_report* R = rep ? rep : &local;
  std::memset(R, 0, sizeof(*R));
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  const long long k0 = s->k_applies;
  if (needs_setup(s)) {
    CK(hipEventRecord(s->e0, st));
    CKS(setup_dev(s, s->prev_solution ? s->ufull : nullptr, s->prev_solution ? s->p : nullptr, s->prev_solution));
    CK(hipEventRecord(s->e1, st));
    CK(hipEventSynchronize(s->e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, s->e0, s->e1));
    R->setup_seconds = ms * 1e-3;
  }
  CK(hipEventRecord(s->e0, st));
  int32_t stt = uzawa(s, R);
  CK(hipEventRecord(s->e1, st));
  CK(hipEventSynchronize(s->e1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, s->e0, s->e1));
  R->solve_seconds = ms * 1e-3;
  R->k_applies = s->k_applies - k0;
  R->status = stt;
  if (stt == GS_OK || s->has_solution) {
    if (u) CK(hipMemcpy(u, s->ufull, sizeof(double) * s->n, hipMemcpyDeviceToHost));
    if (p) CK(hipMemcpy(p, s->p, sizeof(double) * s->np, hipMemcpyDeviceToHost));
  }
  return stt;
}

extern "C" int32_t gs_solver_set_solution(gs_solver* s, const double* u_full, const double* p) {
  if (!s || !u_full || !p) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  CK(hipMemcpyAsync(s->ufull, u_full, sizeof(double) * s->n, hipMemcpyHostToDevice, st));
  CK(hipMemcpyAsync(s->p, p, sizeof(double) * s->np, hipMemcpyHostToDevice, st));
  // the solver's own velocity holds the free dofs (Dirichlet dofs zero)
  launch_masked_copy(s->ufull, s->u, pb->bcdof, s->n, st);
  CK(hipStreamSynchronize(st));
  s->has_solution = true;
  s->prev_solution = 1;
  return GS_OK;
}

extern "C" int32_t gs_solver_set_pressure(gs_solver* s, const double* p) {
  if (!s) return GS_E_INVALID;
  if (p)
    CK(hipMemcpy(s->p, p, sizeof(double) * s->np, hipMemcpyHostToDevice));
  else {
    launch_fill(s->p, s->np, 0.0, s->pb->st);
    CK(hipStreamSynchronize(s->pb->st));
  }
  return GS_OK;
}

extern "C" int32_t gs_picard_solve(gs_solver* s, const gs_picard_params* pp, double* u, double* p, gs_report* rep) {
  if (!s || !pp) return GS_E_INVALID;
  gs_report local{};
  gs_report* R = rep ? rep : &local;
  std::memset(R, 0, sizeof(*R));
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  const long long n = s->n;
  const long long k0 = s->k_applies;
  float ms = 0.f;
  int32_t stt = GS_OK;
  // first solve: linear unless a previous solution exists (a previous time
  // step's or a restart's), which it then yields from (YieldRheology.cxx:203)
  CK(hipEventRecord(s->e0, st));
  CKS(setup_dev(s, s->prev_solution ? s->ufull : nullptr, s->prev_solution ? s->p : nullptr, s->prev_solution));
  stt = uzawa(s, R);
  CK(hipEventRecord(s->e1, st));
  CK(hipEventSynchronize(s->e1));
  CK(hipEventElapsedTime(&ms, s->e0, s->e1));
  R->solve_seconds += ms * 1e-3;
  s->prev_solution = 1;
  int it = 1;
  bool converged = false;
  double res = 0.0;
  for (; stt == GS_OK && it < pp->max_iterations; it++) {
    CK(hipMemcpyAsync(s->uprev, s->u, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    CK(hipEventRecord(s->e0, st));
    CKS(setup_dev(s, s->ufull, s->p, 1));
    stt = uzawa(s, R);
    CK(hipEventRecord(s->e1, st));
    CK(hipEventSynchronize(s->e1));
    CK(hipEventElapsedTime(&ms, s->e0, s->e1));
    R->solve_seconds += ms * 1e-3;
    if (stt != GS_OK) break;
    launch_axpby(s->uprev, s->u, 1.0, -1.0, n, st);  // uprev = u - uprev
    int32_t err = GS_OK;
    const double dn = dot_host(s, s->uprev, s->uprev, n, &err);
    CKS(err);
    const double cn = dot_host(s, s->u, s->u, n, &err);
    CKS(err);
    res = std::sqrt(dn) / std::sqrt(cn);
    converged = res < pp->tolerance;
    if (converged && it >= pp->min_iterations) break;
  }
  R->picard_iterations = (stt == GS_OK && converged && it < pp->max_iterations) ? it + 1 : it;
  R->picard_residual = res;
  if (stt == GS_OK && !converged && pp->kill_non_convergent) stt = GS_E_PICARD;
  R->status = stt;
  R->k_applies = s->k_applies - k0;
  if (u) CK(hipMemcpy(u, s->ufull, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (p) CK(hipMemcpy(p, s->p, sizeof(double) * s->np, hipMemcpyDeviceToHost));
  return stt;
}

extern "C" int32_t gs_get_rhs(gs_solver* s, double* fhat) {
  if (!s) return GS_E_INVALID;
  if (needs_setup(s)) CKS(setup_dev(s, nullptr, nullptr, 0));
  gs_problem* pb = s->pb;
  long long its = 0;
  const double rt = s->up.inner_rtol_override > 0 ? s->up.inner_rtol_override : s->up.tolerance;
  CKS(inner_solve(s, s->f, s->vstar, rt, &its));
  launch_apply_Gt(pb->g, s->vstar, s->tmpE, pb->st);
  launch_axpby(s->tmpE, s->h, -1.0, 1.0, s->np, pb->st);
  if (fhat) CK(hipMemcpy(fhat, s->tmpE, sizeof(double) * s->np, hipMemcpyDeviceToHost));
  return GS_OK;
}

extern "C" int32_t gs_form_residual(gs_solver* s, const double* p, double* r) {
  if (!s || !p) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  CKS(gs_get_rhs(s, nullptr));  // tmpE = fhat
  CK(hipMemcpy(s->z, p, sizeof(double) * s->np, hipMemcpyHostToDevice));
  launch_apply_G(pb->g, s->z, s->fT, pb->bcflag, st);
  long long its = 0;
  const double rt = s->up.inner_rtol_override > 0 ? s->up.inner_rtol_override : s->up.tolerance;
  CKS(inner_solve(s, s->fT, s->vstar, rt, &its));
  launch_apply_Gt(pb->g, s->vstar, s->q, st);
  launch_axpby(s->tmpE, s->q, -1.0, 1.0, s->np, st);
  if (r) CK(hipMemcpy(r, s->tmpE, sizeof(double) * s->np, hipMemcpyDeviceToHost));
  return GS_OK;
}

extern "C" int32_t gs_compute_dt(gs_solver* s, double dt_factor, double* dt) {
  if (!s || !dt) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  launch_vmax(s->ufull, pb->g.N, pb->dim, s->S, S_VMAX, s->rd, pb->st);
  CKS(read_scalars(s));
  const double vmax = s->hS[S_VMAX];
  double minsep = INFINITY;
  for (int d = 0; d < pb->dim; d++)
    for (int i = 0; i < pb->g.res[d]; i++) {
      auto vc = [&](int idx) {
        if (idx <= 0) return pb->g.lo[d];
        if (idx >= pb->g.res[d]) return pb->g.hi[d];
        return pb->g.lo[d] + ((double)idx / (double)pb->g.res[d]) * (pb->g.hi[d] - pb->g.lo[d]);
      };
      const double hh = vc(i + 1) - vc(i);
      if (hh < minsep) minsep = hh;
    }
  *dt = vmax == 0.0 ? 0.0 : dt_factor * 0.5 * minsep / vmax;
  return GS_OK;
}

extern "C" int32_t gs_advect_rk2(gs_solver* s, double dt) {
  if (!s) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  launch_fill(pb->S + S_ESC, 1, 0.0, pb->st);
  launch_advect_rk2(pb->g, pb->coord, pb->P, s->ufull, dt, pb->S, pb->st);
  CK(hipMemcpyAsync(pb->hS, pb->S, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, pb->st));
  CK(hipStreamSynchronize(pb->st));
  unsigned esc;
  std::memcpy(&esc, pb->hS + S_ESC, sizeof(unsigned));
  s->setup_done = false;
  // the other time integrand of the material swarm: StrainWeakening
  if (pb->pstr) launch_weakening_step(pb->pstr, pb->pinc, pb->coord, pb->P, dt, pb->slim, pb->st);
  const int32_t rs = relocate(pb);
  if (esc) return GS_E_ESCAPED;
  return rs;
}

extern "C" int32_t gs_apply_K(gs_solver* s, const double* x, double* y) {
  if (!s || !x || !y) return GS_E_INVALID;
  if (needs_setup(s)) CKS(setup_dev(s, nullptr, nullptr, 0));
  launch_apply_K(s->pb->g, x, y, s->pb->pk, s->pb->off, nullptr, s->pb->halo64, s->pb->st);
  s->k_applies++;
  return GS_OK;
}

extern "C" int32_t gs_apply_K_noalloc(gs_solver* s, const double* x, double* y) {
  if (!s || !x || !y) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  const size_t need = sizeof(double) * (size_t)tile_halo_size(pb->g);
  if (need > kNoAllocScratchBytes) {
    std::fprintf(stderr,
                 "gale_stokes: gs_apply_K_noalloc: this mesh needs %zu bytes of tile-face scratch, more than the "
                 "%zu-byte static scratch; use gs_apply_K\n", need, kNoAllocScratchBytes);
    return GS_E_INVALID;
  }
  if (needs_setup(s)) {
    std::fprintf(stderr, "gale_stokes: gs_apply_K_noalloc: the solver is not set up for the problem's current "
                         "particles (gs_solver_setup allocates; call it first)\n");
    return GS_E_INVALID;
  }
  if (!pb->scratch) return GS_E_HIP;
  launch_apply_K(pb->g, x, y, pb->pk, pb->off, nullptr, static_cast<double*>(pb->scratch), pb->st);
  s->k_applies++;
  return GS_OK;
}

static int32_t ensure_pk32(gs_solver* s) {
  gs_problem* pb = s->pb;
  if (needs_setup(s)) CKS(setup_dev(s, nullptr, nullptr, 0));
  if (!pb->pk32) CK(dalloc(&pb->pk32, pb->P));
  if (!pb->pk32_valid) {
    launch_pack32(pb->pk, pb->pk32, pb->P, pb->st);
    pb->pk32_valid = true;
  }
  return GS_OK;
}

extern "C" int32_t gs_apply_K_fp32(gs_solver* s, const float* x, float* y) {
  if (!s || !x || !y) return GS_E_INVALID;
  CKS(ensure_pk32(s));
  launch_apply_K32(s->pb->g, x, y, s->pb->pk32, s->pb->off, nullptr, s->pb->halo32, s->pb->st);
  s->k_applies++;
  return GS_OK;
}

extern "C" int32_t gs_apply_G(gs_solver* s, const double* p, double* y) {
  if (!s || !p || !y) return GS_E_INVALID;
  launch_apply_G(s->pb->g, p, y, nullptr, s->pb->st);
  return GS_OK;
}

extern "C" int32_t gs_apply_Gt(gs_solver* s, const double* x, double* q) {
  if (!s || !x || !q) return GS_E_INVALID;
  launch_apply_Gt(s->pb->g, x, q, s->pb->st);
  return GS_OK;
}

extern "C" int32_t gs_solver_sync(gs_solver* s) {
  if (!s) return GS_E_INVALID;
  CK(hipStreamSynchronize(s->pb->st));
  return GS_OK;
}

extern "C" int32_t gs_apply_K_moments(gs_solver* s, const double* x, double* y) {
  if (!s || !x || !y) return GS_E_INVALID;
  if (needs_setup(s) || !s->mg.structured) {
    CKS(setup_dev(s, nullptr, nullptr, 0));
    if (!s->mg.structured) CKS(mg_numeric(s));
  }
  launch_mom_apply<double>(s->mg.lv[0].g, x, y, s->mg.lv[0].M64, nullptr, s->mg.lv[0].h64, s->pb->st);
  s->k_applies++;
  return GS_OK;
}

// Average device time of n_rep applies of one operator form.  With cold != 0
// the applies cycle through several copies of their operands (input vector,
// output vector, moments or particle stream) whose total exceeds three times
// the chip's 256 MB last-level cache, so every apply streams its bytes from
// HBM as it does inside a solve, where other kernels run between applies.
extern "C" int32_t gs_time_apply(gs_solver* s, int32_t form, int32_t n_rep, int32_t cold, double* ms_per_apply) {
  if (!s || n_rep <= 0 || !ms_per_apply || form < 0 || form > GS_APPLY_GT) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  if (needs_setup(s)) CKS(setup_dev(s, nullptr, nullptr, 0));
  if ((form == 2 || form == 3) && !s->mg.structured) CKS(mg_numeric(s));
  if (form == 1) CKS(ensure_pk32(s));
  if (form == 1 || form == 3) {
    if (!s->x32) CK(dalloc(&s->x32, s->n));
    if (!s->y32) CK(dalloc(&s->y32, s->n));
  }
  const long long n = s->n, E = pb->g.E, P = pb->P;
  // operand set 0 is the solver's own buffers; x_bytes / y_bytes / in_bytes
  // are the sizes of the input vector, output vector and streamed operand
  void* x0 = nullptr;
  void* y0 = nullptr;
  const void* in0 = nullptr;
  size_t xb = 0, yb = 0, ib = 0;
  switch (form) {
    case 0: x0 = s->cg_p; y0 = s->cg_Ap; in0 = pb->pk; xb = yb = 8 * n; ib = sizeof(double4) * P; break;
    case 1: x0 = s->x32; y0 = s->y32; in0 = pb->pk32; xb = yb = 4 * n; ib = sizeof(float4) * P; break;
    case 2: x0 = s->cg_p; y0 = s->cg_Ap; in0 = s->mg.lv[0].M64; xb = yb = 8 * n; ib = 8ull * mom_count<double>(pb->g); break;
    case 3: x0 = s->x32; y0 = s->y32; in0 = s->mg.lv[0].M32; xb = yb = 4 * n; ib = 4ull * mom_count<float>(pb->g); break;
    case GS_APPLY_G: x0 = s->tmpE; y0 = s->cg_Ap; xb = 8 * E; yb = 8 * n; break;
    default: x0 = s->cg_p; y0 = s->tmpE; xb = 8 * n; yb = 8 * E; break;
  }
  int R = 1;
  if (cold) {
    const double set_bytes = (double)(xb + yb + ib);
    R = (int)std::ceil(768.0e6 / (set_bytes > 1.0 ? set_bytes : 1.0));
    if (R < 2) R = 2;
    if (R > 64) R = 64;
  }
  std::vector<void*> xs(R, nullptr), ys(R, nullptr), ins(R, nullptr);
  xs[0] = x0;
  ys[0] = y0;
  ins[0] = const_cast<void*>(in0);
  int32_t err = GS_OK;
  for (int i = 1; i < R && err == GS_OK; i++) {
    if (hipMalloc(&xs[i], xb) != hipSuccess || hipMalloc(&ys[i], yb) != hipSuccess ||
        (ib && hipMalloc(&ins[i], ib) != hipSuccess)) {
      err = GS_E_HIP;
      break;
    }
    if (ib && hipMemcpyAsync(ins[i], in0, ib, hipMemcpyDeviceToDevice, st) != hipSuccess) err = GS_E_HIP;
  }
  auto one = [&](int i) {
    void* x = xs[i];
    void* y = ys[i];
    void* in = ins[i];
    switch (form) {
      case 0: launch_apply_K(pb->g, (const double*)x, (double*)y, (const double4*)in, pb->off, nullptr, pb->halo64, st); break;
      case 1: launch_apply_K32(pb->g, (const float*)x, (float*)y, (const float4*)in, pb->off, nullptr, pb->halo32, st); break;
      case 2: launch_mom_apply<double>(s->mg.lv[0].g, (const double*)x, (double*)y, (const double*)in, nullptr, s->mg.lv[0].h64, st); break;
      case 3: launch_mom_apply<float>(s->mg.lv[0].g, (const float*)x, (float*)y, (const float*)in, nullptr, s->mg.lv[0].h32, st); break;
      case GS_APPLY_G: launch_apply_G(pb->g, (const double*)x, (double*)y, pb->bcflag, st); break;
      default: launch_apply_Gt(pb->g, (const double*)x, (double*)y, st); break;
    }
  };
  if (err == GS_OK) {
    for (int i = 0; i < R; i++) {
      if (form == 1 || form == 3)
        launch_fill_f((float*)xs[i], (long long)(xb / 4), 1.0f, st);
      else
        launch_fill((double*)xs[i], (long long)(xb / 8), 1.0, st);
      one(i);
    }
    float ms = 0.f;
    if (hipEventRecord(s->e0, st) != hipSuccess) err = GS_E_HIP;
    for (int i = 0; i < n_rep && err == GS_OK; i++) one(i % R);
    if (err == GS_OK && (hipEventRecord(s->e1, st) != hipSuccess || hipEventSynchronize(s->e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, s->e0, s->e1) != hipSuccess))
      err = GS_E_HIP;
    *ms_per_apply = (double)ms / n_rep;
    s->k_applies += n_rep + R;
  }
  (void)hipStreamSynchronize(st);
  for (int i = 1; i < R; i++) {
    if (xs[i]) (void)hipFree(xs[i]);
    if (ys[i]) (void)hipFree(ys[i]);
    if (ins[i]) (void)hipFree(ins[i]);
  }
  return err;
}

extern "C" int32_t gs_time_apply_K(gs_solver* s, int32_t form, int32_t n_rep, double* ms_per_apply) {
  return gs_time_apply(s, form, n_rep, 0, ms_per_apply);
}

extern "C" int32_t gs_solver_setup_arrays(gs_solver* s, double* diagK, double* Q, double* f, double* h, double* eta) {
  if (!s) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  if (needs_setup(s)) CKS(setup_dev(s, nullptr, nullptr, 0));
  CK(hipStreamSynchronize(pb->st));
  if (diagK) CK(hipMemcpy(diagK, s->diag, sizeof(double) * s->n, hipMemcpyDeviceToHost));
  if (Q) CK(hipMemcpy(Q, s->Q, sizeof(double) * s->np, hipMemcpyDeviceToHost));
  if (f) CK(hipMemcpy(f, s->f, sizeof(double) * s->n, hipMemcpyDeviceToHost));
  if (h) CK(hipMemcpy(h, s->h, sizeof(double) * s->np, hipMemcpyDeviceToHost));
  if (eta) {
    const long long P = pb->P;
    std::vector<double> e(P > 0 ? P : 1);
    std::vector<int> pid(P > 0 ? P : 1);
    CK(hipMemcpy(e.data(), s->eta, sizeof(double) * P, hipMemcpyDeviceToHost));
    CK(hipMemcpy(pid.data(), pb->pid, sizeof(int) * P, hipMemcpyDeviceToHost));
    int offE = 0;
    CK(hipMemcpy(&offE, pb->off + pb->g.E, sizeof(int), hipMemcpyDeviceToHost));
    for (long long i = 0; i < P; i++) eta[pid[i]] = i < offE ? e[i] : 0.0;
  }
  return GS_OK;
}

// Output fields of the material points in the caller's order (StoreViscosity
// / StoreStress / the strain-rate fields).  u: full nodal velocity (N*dim,
// host), or NULL for the solver's last solution.
extern "C" int32_t gs_solver_particle_fields(gs_solver* s, const double* u, double* strain_rate,
                                             double* strain_rate_inv, double* viscosity, double* stress) {
  if (!s) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  if (needs_setup(s)) CKS(setup_dev(s, nullptr, nullptr, 0));
  const long long P = pb->P;
  const int nst = pb->dim == 2 ? 3 : 6;
  hipStream_t st = pb->st;
  if (P <= 0) return GS_OK;
  double *d_sr = nullptr, *d_inv = nullptr, *d_st = nullptr;
  const double* ud = s->ufull;
  if (u) {
    CK(hipMemcpyAsync(s->tmpN, u, sizeof(double) * s->n, hipMemcpyHostToDevice, st));
    ud = s->tmpN;
  }
  int32_t rc = GS_OK;
  if (dalloc(&d_sr, P * nst) != hipSuccess || dalloc(&d_inv, P) != hipSuccess || dalloc(&d_st, P * nst) != hipSuccess) {
    rc = GS_E_HIP;
  } else {
    launch_particle_fields(pb->g, pb->pk, pb->keys, P, ud, s->eta, d_sr, d_inv, d_st, st);
    std::vector<double> sr(P * nst), inv(P), sv(P * nst);
    std::vector<int> pid(P);
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(sr.data(), d_sr, sizeof(double) * P * nst, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(inv.data(), d_inv, sizeof(double) * P, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(sv.data(), d_st, sizeof(double) * P * nst, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(pid.data(), pb->pid, sizeof(int) * P, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = GS_E_HIP;
    } else {
      std::vector<double> eta;
      if (viscosity) {
        eta.resize(P);
        if (hipMemcpy(eta.data(), s->eta, sizeof(double) * P, hipMemcpyDeviceToHost) != hipSuccess) rc = GS_E_HIP;
      }
      int offE = 0;
      if (hipMemcpy(&offE, pb->off + pb->g.E, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = GS_E_HIP;
      if (rc == GS_OK)
        for (long long i = 0; i < P; i++) {
          const long long o = pid[i];
          for (int k = 0; k < nst; k++) {
            if (strain_rate) strain_rate[o * nst + k] = sr[i * nst + k];
            if (stress) stress[o * nst + k] = sv[i * nst + k];
          }
          if (strain_rate_inv) strain_rate_inv[o] = inv[i];
          if (viscosity) viscosity[o] = i < offE ? eta[i] : 0.0;
        }
    }
  }
  dfree(d_sr);
  dfree(d_inv);
  dfree(d_st);
  return rc;
}

// ---------------------------------------------------------------------------
// learned preconditioner attachment (include/gale_learned.h)
// ---------------------------------------------------------------------------
extern "C" int32_t gs_solver_set_learned(gs_solver* s, gs_lpc* L) {
  if (!s) return GS_E_INVALID;
  if (L && !lpc_shape_ok(s->pb->g)) return GS_E_INVALID;
  s->lpc = L;
  if (L) s->setup_done = false;  // the element viscosity field is made at setup
  return GS_OK;
}

extern "C" int32_t gs_solver_record_pairs(gs_solver* s, gs_lpc_dataset* D, int32_t krylov_every) {
  if (!s || krylov_every < 0) return GS_E_INVALID;
  if (D && !lpc_shape_ok(s->pb->g)) return GS_E_INVALID;
  if (D && (!s->rec_r0 || !s->rec_p0)) {
    CK(dalloc(&s->rec_r0, s->np));
    CK(dalloc(&s->rec_p0, s->np));
  }
  if (D && !s->logeta) s->setup_done = false;
  s->rec = D;
  s->rec_every = krylov_every;
  return GS_OK;
}
