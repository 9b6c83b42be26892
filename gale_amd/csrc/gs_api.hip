// gs_api.hip -- host driver and C ABI (include/gale_stokes.h).
//
// The host side owns device buffers and drives the Uzawa / CG / Picard loops
// with all scalars kept on the device; it waits for the device only at
// convergence checks.  Reference behaviour each function follows is cited in
// include/gale_stokes.h.
#include <hip/hip_runtime.h>
#include <string.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "../../include/gale_stokes.h"
#include "gs_internal.h"

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
  double gravity = 0, ghat[3] = {0, 0, 0}, clamp_min = 0, clamp_max = 0;
  uint8_t* bcdof = nullptr;   // N*dim
  uint8_t* bcflag = nullptr;  // N (bit d = dof d is Dirichlet)
  double* ubc = nullptr;      // N*dim, Dirichlet values, zeros elsewhere
  bool has_bc = false;
  long long n_free = 0;
  long long P = 0, escaped = 0;
  double4* pk = nullptr;
  float4* pk32 = nullptr;
  bool pk32_valid = false;
  double* coord = nullptr;
  int* mat = nullptr;
  int* pid = nullptr;
  double* eta0p = nullptr;
  int* off = nullptr;
  unsigned* keys = nullptr;  // sorted owner of each sorted particle (E = outside)
  unsigned* keys2 = nullptr;
  int* vals = nullptr;
  int* vals2 = nullptr;
  double* coord2 = nullptr;
  double* eta0p2 = nullptr;
  int* mat2 = nullptr;
  int* pid2 = nullptr;
  void* sort_tmp = nullptr;
  size_t sort_bytes = 0;
  double* S = nullptr;
  double* hS = nullptr;  // pinned
  Reducer rd{};
  bool min_one_particle = true;
};

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
  bool setup_done = false;
  bool has_solution = false;
  int has_executed = 0;
  long long k_applies = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

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
  CK(dalloc(&rd->counter, 1));
  CK(hipMemset(rd->counter, 0, sizeof(unsigned)));
  return GS_OK;
}

// Carry the particle arrays along the sort permutation.
__global__ void k_gather_particles(const int* __restrict__ perm, long long P, const double* __restrict__ c,
                                   double* __restrict__ c2, const int* __restrict__ m, int* __restrict__ m2,
                                   const int* __restrict__ id, int* __restrict__ id2, const double* __restrict__ e0,
                                   double* __restrict__ e02) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= P) return;
  const long long j = perm[i];
  c2[3 * i] = c[3 * j];
  c2[3 * i + 1] = c[3 * j + 1];
  c2[3 * i + 2] = c[3 * j + 2];
  m2[i] = m[j];
  id2[i] = id[j];
  if (e0) e02[i] = e0[j];
}

// Locate every particle (bit-exact owner rule), sort by owner (stable radix
// sort: within an element particles keep their previous relative order),
// carry all particle arrays along, rebuild the element offsets and remap xi.
static int32_t relocate(gs_problem* pb) {
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
                       pb->coord, pb->coord2, pb->mat, pb->mat2, pb->pid, pb->pid2, pb->eta0p, pb->eta0p2);
    std::swap(pb->coord, pb->coord2);
    std::swap(pb->mat, pb->mat2);
    std::swap(pb->pid, pb->pid2);
    std::swap(pb->eta0p, pb->eta0p2);
  }
  launch_offsets(reinterpret_cast<const int*>(pb->keys), P, pb->g.E, pb->off, st);
  launch_remap(pb->g, pb->coord, pb->off, pb->pk, pb->g.E, st);
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
  return esc ? GS_E_ESCAPED : GS_OK;
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
  if (hipStreamCreate(&pb->st) != hipSuccess) return fail(GS_E_HIP);
  const long long ND = g.N * g.dim;
  std::vector<MatDev> md(d->n_materials);
  for (int i = 0; i < d->n_materials; i++) {
    const gs_material& m = d->materials[i];
    md[i] = MatDev{m.eta0, m.density, m.yield_type, 0, m.cohesion, m.friction, m.min_yield_stress, m.min_viscosity,
                   m.max_strain_rate};
  }
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
  const long long P = d->n_particles;
  pb->P = P;
  if (dalloc(&pb->mats_d, d->n_materials) || dalloc(&pb->bcdof, ND) || dalloc(&pb->bcflag, g.N) ||
      dalloc(&pb->ubc, ND) || dalloc(&pb->pk, P) || dalloc(&pb->coord, 3 * P) || dalloc(&pb->coord2, 3 * P) ||
      dalloc(&pb->mat, P) || dalloc(&pb->mat2, P) || dalloc(&pb->pid, P) || dalloc(&pb->pid2, P) ||
      dalloc(&pb->off, g.E + 1) || dalloc(&pb->keys, P) || dalloc(&pb->keys2, P) || dalloc(&pb->vals, P) ||
      dalloc(&pb->vals2, P) || dalloc(&pb->S, S_COUNT))
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
  dfree(pb->vals2); dfree(pb->S); dfree(pb->rd.partials); dfree(pb->rd.counter);
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
    for (long long e = 0; e < E; e++)
      for (int i = off[e]; i < off[e + 1]; i++) {
        const long long o = pid[i];
        if (xi) {
          xi[o * dim] = pk[i].x;
          xi[o * dim + 1] = pk[i].y;
          if (dim == 3) xi[o * dim + 2] = pk[i].z;
        }
        if (weight) weight[o] = (dim == 2 ? 4.0 : 8.0) / (double)(off[e + 1] - off[e]);
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
  if (pb->eta0p) {
    std::vector<double> e0(P), e0o(P);
    CK(hipMemcpy(e0.data(), pb->eta0p, sizeof(double) * P, hipMemcpyDeviceToHost));
    for (long long i = 0; i < P; i++) e0o[pid[i]] = e0[i];
    CK(hipMemcpy(pb->eta0p, e0o.data(), sizeof(double) * P, hipMemcpyHostToDevice));
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
  return p;
}

extern "C" int32_t gs_solver_set_params(gs_solver* s, const gs_uzawa_params* up) {
  if (!s) return GS_E_INVALID;
  s->up = up ? *up : default_params();
  if (s->up.check_every <= 0) s->up.check_every = 4;
  if (s->up.inner_max_iterations <= 0) s->up.inner_max_iterations = 50000;
  if (s->up.inner_rtol_factor <= 0) s->up.inner_rtol_factor = 0.1;
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
      dalloc(&s->cg_r, n) || dalloc(&s->cg_p, n) || dalloc(&s->cg_Ap, n) || dalloc(&s->tmpN, n) ||
      dalloc(&s->tmpE, np) || dalloc(&s->eta, P) || dalloc(&s->uprev, n) || dalloc(&s->yielded, P) ||
      dalloc(&s->S, S_COUNT))
    return fail(GS_E_HIP);
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
  dfree(s->cg_r); dfree(s->cg_p); dfree(s->cg_Ap); dfree(s->tmpN); dfree(s->tmpE); dfree(s->eta); dfree(s->uprev);
  dfree(s->yielded); dfree(s->x32); dfree(s->y32); dfree(s->S); dfree(s->rd.partials); dfree(s->rd.counter);
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
  launch_apply_K(pb->g, x, y, pb->pk, pb->off, pb->bcflag, pb->st);
  s->k_applies++;
}

// The fused constitutive pass + diagonal operators + right-hand sides.
static int32_t setup_dev(gs_solver* s, const double* u_prev, const double* p_prev, int yield_on) {
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  if (!pb->min_one_particle) return GS_E_INVALID;
  SetupArgs a{};
  a.u_prev = u_prev;
  a.p_prev = p_prev;
  a.yield_on = yield_on;
  a.mats = pb->mats_d;
  a.mat = pb->mat;
  a.eta0p = pb->eta0p;
  a.clamp_min = pb->clamp_min;
  a.clamp_max = pb->clamp_max;
  for (int d = 0; d < 3; d++) a.g[d] = pb->ghat[d] * pb->gravity;
  launch_setup(pb->g, a, pb->pk, pb->off, s->diag, s->f, s->Q, s->eta, s->yielded, st);
  pb->pk32_valid = false;
  launch_make_dinv(s->diag, pb->bcdof, s->dinv, s->n, st);
  launch_recip(s->Q, s->Qinv, s->np, st);
  if (pb->has_bc) {
    launch_apply_K(pb->g, pb->ubc, s->tmpN, pb->pk, pb->off, nullptr, st);
    s->k_applies++;
    launch_lift(s->f, s->tmpN, pb->bcdof, s->n, st);
    launch_apply_Gt(pb->g, pb->ubc, s->h, st);
    launch_scale(s->h, s->h, -1.0, s->np, st);
  } else {
    launch_fill(s->h, s->np, 0.0, st);
  }
  s->setup_done = true;
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
static int32_t inner_solve(gs_solver* s, const double* b, double* x, double rtol, long long* its) {
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

static double dot_host(gs_solver* s, const double* a, const double* b, long long n, int32_t* err) {
  launch_dot(a, b, n, s->S, S_TMP1, s->rd, s->pb->st);
  *err = read_scalars(s);
  return s->hS[S_TMP1];
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
  const double* Qinv = s->up.preconditioner == GS_PC_DIAGONAL ? s->Qinv : nullptr;
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
  // STEP 4 (:664-837)
  launch_fill(s->S + S_ZR_PREV, 1, 1.0, st);
  do {
    launch_uz_precond(s->r, Qinv, s->z, np, s->S, s->rd, st);
    launch_uz_sdir(s->s, s->z, np, s->S, it == 0, st);
    launch_apply_G(pb->g, s->s, s->fT, pb->bcflag, st);
    CKS(inner_solve(s, s->fT, s->vstar, irt, &inner));
    launch_apply_Gt(pb->g, s->vstar, s->q, st);
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
    if (s->up.monitor) std::fprintf(stderr, "  %04d uzawa residual norm %12.13e\n", it + 1, relRes);
    if (!(it < s->up.max_iterations)) {
      err = GS_E_MAXIT;
      break;
    }
    it++;
  } while (((s->up.use_absolute_tolerance ? absRes : relRes) > tol) || (it < s->up.min_iterations));
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
  if (!s->setup_done) {
    CK(hipEventRecord(s->e0, st));
    CKS(setup_dev(s, nullptr, nullptr, 0));
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
  s->has_executed = 1;
  return stt;
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
  // linear first solve (no yielding before a previous solution exists)
  CK(hipEventRecord(s->e0, st));
  CKS(setup_dev(s, nullptr, nullptr, 0));
  stt = uzawa(s, R);
  CK(hipEventRecord(s->e1, st));
  CK(hipEventSynchronize(s->e1));
  CK(hipEventElapsedTime(&ms, s->e0, s->e1));
  R->solve_seconds += ms * 1e-3;
  s->has_executed = 1;
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
  if (!s->setup_done) CKS(setup_dev(s, nullptr, nullptr, 0));
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
  const int32_t rs = relocate(pb);
  if (esc) return GS_E_ESCAPED;
  return rs;
}

extern "C" int32_t gs_apply_K(gs_solver* s, const double* x, double* y) {
  if (!s || !x || !y) return GS_E_INVALID;
  if (!s->setup_done) CKS(setup_dev(s, nullptr, nullptr, 0));
  launch_apply_K(s->pb->g, x, y, s->pb->pk, s->pb->off, nullptr, s->pb->st);
  s->k_applies++;
  return GS_OK;
}

static int32_t ensure_pk32(gs_solver* s) {
  gs_problem* pb = s->pb;
  if (!s->setup_done) CKS(setup_dev(s, nullptr, nullptr, 0));
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
  launch_apply_K32(s->pb->g, x, y, s->pb->pk32, s->pb->off, nullptr, s->pb->st);
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

extern "C" int32_t gs_time_apply_K(gs_solver* s, int32_t precision, int32_t n_rep, double* ms_per_apply) {
  if (!s || n_rep <= 0 || !ms_per_apply) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  hipStream_t st = pb->st;
  if (!s->setup_done) CKS(setup_dev(s, nullptr, nullptr, 0));
  float ms = 0.f;
  if (precision == GS_PREC_FP64) {
    launch_fill(s->cg_p, s->n, 1.0, st);
    launch_apply_K(pb->g, s->cg_p, s->cg_Ap, pb->pk, pb->off, nullptr, st);
    CK(hipEventRecord(s->e0, st));
    for (int i = 0; i < n_rep; i++) launch_apply_K(pb->g, s->cg_p, s->cg_Ap, pb->pk, pb->off, nullptr, st);
    CK(hipEventRecord(s->e1, st));
  } else {
    CKS(ensure_pk32(s));
    if (!s->x32) CK(dalloc(&s->x32, s->n));
    if (!s->y32) CK(dalloc(&s->y32, s->n));
    launch_fill_f(s->x32, s->n, 1.0f, st);
    launch_apply_K32(pb->g, s->x32, s->y32, pb->pk32, pb->off, nullptr, st);
    CK(hipEventRecord(s->e0, st));
    for (int i = 0; i < n_rep; i++) launch_apply_K32(pb->g, s->x32, s->y32, pb->pk32, pb->off, nullptr, st);
    CK(hipEventRecord(s->e1, st));
  }
  CK(hipEventSynchronize(s->e1));
  CK(hipEventElapsedTime(&ms, s->e0, s->e1));
  *ms_per_apply = (double)ms / n_rep;
  s->k_applies += n_rep + 1;
  return GS_OK;
}

extern "C" int32_t gs_solver_setup_arrays(gs_solver* s, double* diagK, double* Q, double* f, double* h, double* eta) {
  if (!s) return GS_E_INVALID;
  gs_problem* pb = s->pb;
  if (!s->setup_done) CKS(setup_dev(s, nullptr, nullptr, 0));
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
