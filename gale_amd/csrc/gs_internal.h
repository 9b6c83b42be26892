// gs_internal.h -- device data layout and kernel launchers shared by the
// host driver (gs_api.hip) and the kernels (gs_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gs {

// Regular Cartesian Q1 mesh.  Nodes and elements are numbered
// lexicographically, x fastest; element-local node a = ax + 2 ay + 4 az.
struct Geo {
  int dim;
  int res[3];
  long long nn[3];   // nodes per axis
  double lo[3], hi[3];
  long long N, E;    // nodes, elements
};

// Material parameters on the device (one RheologyMaterial each).
struct MatDev {
  double eta0, density;
  int yield_type, sw_on;  // sw_on: the yield rheology carries StrainWeakening
  double cohesion, friction, min_yield_stress, min_viscosity, max_strain_rate;
  // StrainWeakening (healingRate, initial/finalSofteningStrain) and the
  // after-softening cohesion / friction coefficient of the yield rheology
  double sw_healing, sw_eps0, sw_eps1, cohesion_after, friction_after;
};

// Scalar slots kept on the device so that the CG / Uzawa loops never wait
// for the host except at convergence checks.
enum Slot {
  S_RZ = 0,      // r . (D^-1 r) of the inner CG
  S_PAP = 1,     // p . K p
  S_RR = 2,      // r . r (inner)
  S_BB = 3,      // b . b (inner)
  S_BETA = 4,    // inner beta
  S_ALPHA = 5,   // inner alpha (flexible CG)
  S_ZR = 8,      // Uzawa z . r
  S_ZR_PREV = 9,
  S_UBETA = 10,  // Uzawa beta
  S_SQ = 11,     // s . q
  S_ORR = 12,    // Uzawa r . r
  S_NBAD = 13,   // learned Uzawa: network outputs replaced by Q^-1 r (stored as double)
  S_TMP0 = 16,
  S_TMP1 = 17,
  S_TMP2 = 18,
  S_VMAX = 20,
  S_ESC = 24,    // escaped particle count (stored as double)
  S_LP_TT = 26,  // learned preconditioner: sum (Q^-1 r)^2
  S_LP_ZQ = 27,  // learned preconditioner: z . q_prev
  S_DS_TT = 28,  // training-pair recording: sum (Q^-1 r)^2, sum e, sum e^2
  S_DS_E = 29,
  S_DS_EE = 30,
  S_COUNT = 32
};

struct Reducer {
  double* partials;   // >= kMaxReduceBlocks * 4 doubles (block partial sums)
};

constexpr int kBlock = 256;
constexpr int kMaxReduceBlocks = 1024;

// Fused constitutive pass inputs.
struct SetupArgs {
  const double* u_prev;   // N*dim or null
  const double* p_prev;   // E or null
  int yield_on;
  const MatDev* mats;
  const int* mat;         // P (sorted)
  const double* eta0p;    // P (sorted) or null
  const double* wp;       // P (sorted) integration weights, or null: 2^dim / count
  const double* pstr;     // P (sorted) accumulated plastic strain (StrainWeakening), or null
  double* pinc;           // P (sorted) its rate from this pass, written when pstr is set
  double clamp_min, clamp_max;
  double g[3];            // gravity vector
};

// ---- launchers (gs_kernels.hip) ----
void launch_fill(double* y, long long n, double v, hipStream_t st);
void launch_fill_f(float* y, long long n, float v, hipStream_t st);
// K apply on full nodal vectors; bcflag (per node bitmask) null = unconstrained.
// halo: tile_halo_size(g) values of scratch (gs_tile.h).
void launch_apply_K(const Geo& g, const double* x, double* y, const double4* pk, const int* off,
                    const uint8_t* bcflag, double* halo, hipStream_t st);
void launch_apply_K32(const Geo& g, const float* x, float* y, const float4* pk, const int* off,
                      const uint8_t* bcflag, float* halo, hipStream_t st);
void launch_apply_G(const Geo& g, const double* p, double* y, const uint8_t* bcflag, hipStream_t st);
void launch_apply_Gt(const Geo& g, const double* x, double* q, hipStream_t st);
void launch_setup(const Geo& g, const SetupArgs& a, double4* pk, const int* off, double* diag, double* f,
                  double* Q, double* eta_out, uint8_t* yielded, hipStream_t st);
void launch_pack32(const double4* pk, float4* pk32, long long P, hipStream_t st);
// StrainWeakening's strain-limited region: particles inside get `limit`
// after each step (type: -1 none, 0 everywhere, 1 sphere p[0..2] centre p[3]
// radius, 2 box p[0..2] start p[3..5] end).
struct StrainLimit {
  int type, dim;
  double p[6];
  double limit;
};
// StrainWeakening time step: w += dt * inc, the strain-limited region set,
// then negative values set to 0.  coord: P*3 (same order as w).
void launch_weakening_step(double* w, const double* inc, const double* coord, long long P, double dt,
                           const StrainLimit& lim, hipStream_t st);
// Strain rate (P x nst), its second invariant (P) and stress (P x nst) of the
// material points (sorted order) for velocity u and viscosity eta.
void launch_particle_fields(const Geo& g, const double4* pk, const unsigned* keys, long long P, const double* u,
                            const double* eta, double* sr, double* inv, double* st, hipStream_t st_);

// Reductions / vector kernels.  Results land in S[slot].
void launch_dot(const double* a, const double* b, long long n, double* S, int slot, Reducer r, hipStream_t st);
void launch_masked_copy(const double* src, double* dst, const uint8_t* bcdof, long long n, hipStream_t st);
void launch_lift(double* f, const double* Kubc, const uint8_t* bcdof, long long n, hipStream_t st);
void launch_make_dinv(const double* diag, const uint8_t* bcdof, double* dinv, long long n, hipStream_t st);
void launch_recip(const double* a, double* b, long long n, hipStream_t st);
void launch_scale(double* y, const double* x, double a, long long n, hipStream_t st);  // y = a*x
void launch_axpby(double* y, const double* x, double a, double b, long long n, hipStream_t st); // y = a x + b y
void launch_add(double* y, const double* a, const double* b, long long n, hipStream_t st);  // y = a + b
void launch_mul(const double* a, const double* b, double* c, long long n, hipStream_t st);  // c = a * b

// Inner CG (Jacobi).
void launch_cg_init(const double* b, const double* dinv, double* x, double* r, double* p, long long n,
                    double* S, Reducer rd, hipStream_t st);
void launch_cg_update(double* x, double* r, const double* p, const double* Ap, const double* dinv, long long n,
                      double* S, Reducer rd, hipStream_t st);
void launch_cg_pupdate(double* p, const double* r, const double* dinv, long long n, const double* S,
                       hipStream_t st);

// Flexible (Polak-Ribiere) CG with a variable preconditioner.
void launch_fcg_update(double* x, double* r, const double* p, const double* Ap, long long n, double* S, Reducer rd,
                       hipStream_t st);
void launch_fcg_dots(const double* z, const double* r, const double* Ap, long long n, double* S, Reducer rd,
                     hipStream_t st);
void launch_fcg_p(double* p, const double* z, long long n, const double* S, hipStream_t st);

// Uzawa outer loop.
void launch_uz_precond(const double* r, const double* Qinv, double* z, long long n, double* S, Reducer rd,
                       hipStream_t st);
// launch_uz_precond only where S[S_ZR] (z.r of a network output) is not
// positive and finite; counts those cases in S[S_NBAD].
void launch_uz_precond_if_bad(const double* r, const double* Qinv, double* z, long long n, double* S, Reducer rd,
                              hipStream_t st);
void launch_uz_sdir(double* s, const double* z, long long n, const double* S, int first, hipStream_t st);
void launch_uz_update(double* p, const double* s, double* r, const double* q, long long n, double* S,
                      Reducer rd, hipStream_t st);
void launch_uz_uupdate(double* u, const double* v, long long n, const double* S, hipStream_t st);

// Particles.
void launch_locate(const Geo& g, const double* coord, long long P, int* keys, int* vals, double* S,
                   hipStream_t st);
void launch_offsets(const int* keys_sorted, long long P, long long E, int* off, hipStream_t st);
void launch_remap(const Geo& g, const double* coord, const unsigned* owner, double4* pk, long long P,
                  hipStream_t st);
void launch_advect_rk2(const Geo& g, double* coord, long long P, const double* u, double dt, double* S,
                       hipStream_t st);
void launch_vmax(const double* u, long long nnode, int dim, double* S, int slot, Reducer rd, hipStream_t st);

// Element iteration helpers shared by kernels and host.
inline long long colour_count(const Geo& g, int c) {
  long long m = 1;
  for (int d = 0; d < g.dim; d++) m *= (g.res[d] - ((c >> d) & 1) + 1) / 2;
  return m;
}

}  // namespace gs
