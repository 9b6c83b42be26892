/*
 * gale_oracle.h -- CPU restatement of Gale's Q1-P0 Stokes hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation under gale_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the geodynamics/gale source tree).  Parity pins for each
 * function are listed in the header comment of gale_oracle.c.
 */
#ifndef GALE_ORACLE_H
#define GALE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_OK = 0,
  ORC_E_NAN = 1,         /* Journal_Firewall on NaN/Inf (Stokes_SLE_UzawaSolver.cxx:604,635,784,811) */
  ORC_E_MAXIT = 2,       /* Uzawa maxIterations firewall (Stokes_SLE_UzawaSolver.cxx:815) */
  ORC_E_ZERO_FORCE = 3,  /* zero momentum force vector (Stokes_SLE_UzawaSolver.cxx:476-493) */
  ORC_E_INVALID = 6,
  ORC_E_ESCAPED = 7,     /* particle left the box */
  ORC_E_INNER_MAXIT = 8, /* inner velocity solve did not reach its tolerance */
  ORC_E_PICARD = 9       /* Picard did not converge and killNonConvergent (SystemLinearEquations.cxx:899-903) */
};

enum { ORC_YIELD_NONE = 0, ORC_YIELD_DRUCKER_PRAGER = 1, ORC_YIELD_VON_MISES = 2 };

typedef struct orc_material {
  double eta0;              /* MaterialViscosity eta0 (MaterialViscosity.cxx:129-147) */
  double density;           /* Material density used by BuoyancyForceTerm */
  int32_t yield_type;       /* ORC_YIELD_* */
  int32_t pad_;
  double cohesion;          /* VonMises "cohesion" */
  double friction;          /* DruckerPrager "frictionCoefficient" = tan(phi) */
  double min_yield_stress;  /* DruckerPrager "minimumYieldStress" (0 -> cohesion) */
  double min_viscosity;     /* DruckerPrager "minimumViscosity" / YieldRheology minVisc */
  double max_strain_rate;   /* DruckerPrager "maxStrainRate" (0 -> off) */
} orc_material;

typedef struct orc_problem_desc {
  int32_t dim;
  int32_t res[3];
  double lo[3], hi[3];
  const uint8_t* bc_mask;   /* N*dim, nonzero = Dirichlet dof; NULL = no BCs */
  const double* bc_value;   /* N*dim; NULL = zeros */
  double gravity;
  double ghat[3];
  double eta_clamp_min;     /* 0 = off */
  double eta_clamp_max;     /* 0 = off */
  int32_t n_materials;
  int32_t pad_;
  const orc_material* materials;
  int64_t n_particles;
  const double* coord;      /* P*3 (Coord is double[3] even in 2D) */
  const int32_t* material;  /* P */
  const double* eta0_particle; /* optional per-particle eta0 override (P) */
} orc_problem_desc;

typedef struct orc_uzawa_params {
  double tolerance;          /* "tolerance" (default 1e-5, Stokes_SLE_UzawaSolver.cxx:251) */
  int32_t max_iterations;    /* "maxIterations" (1000) */
  int32_t min_iterations;    /* "minIterations" (1) */
  int32_t use_absolute_tolerance;
  int32_t use_preconditioner; /* 1 = diagonal Uzawa preconditioner Q */
  int32_t inner;             /* 0 = CG + Jacobi, 1 = GMRES(30) + ILU(0) */
  int32_t inner_max_iterations;
  double inner_rtol_factor;  /* inner rtol = factor * tolerance after the F-hat solve (0.1) */
  double inner_rtol_override;/* > 0: use for every inner solve */
  int32_t nthreads;
  int32_t pad_;
} orc_uzawa_params;

typedef struct orc_picard_params {
  double tolerance;          /* "nonLinearTolerance" (0.01) */
  int32_t max_iterations;    /* "nonLinearMaxIterations" (500) */
  int32_t min_iterations;    /* "nonLinearMinIterations" (1) */
  int32_t kill_non_convergent; /* "killNonConvergent" */
  int32_t pad_;
} orc_picard_params;

typedef struct orc_report {
  int32_t status;
  int32_t uzawa_iterations;
  int64_t inner_iterations;
  double rel_residual;
  double abs_residual;
  double div_u;              /* |G^T u| / |u| */
  double momentum_residual;  /* |f - K u - G p| / |f| */
  double solve_seconds;
  double assemble_seconds;
  int32_t picard_iterations;
  int32_t pad_;
  double picard_residual;
} orc_report;

typedef struct orc_problem orc_problem;

/* --- Park-Miller minimal standard generator (Random.cxx:45-74) --- */
typedef struct orc_rng { long state; } orc_rng;
void orc_rng_seed(orc_rng* g, long seed);
double orc_rng_next(orc_rng* g);                     /* returns the integer state as a double */
double orc_rng_range(orc_rng* g, double mn, double mx);

/* --- constitutive conventions --- */
void orc_isotropic_D(int32_t dim, double eta, double* D);
void orc_isotropic_stress(int32_t dim, double eta, const double* sr, double* st);

/* --- mesh --- */
double orc_vertex_coord(int32_t n, double lo, double hi, int32_t idx);
int32_t orc_search_element(int32_t dim, const int32_t* res, const double* lo, const double* hi,
                           const double* x, int32_t* elInd);

/* RandomParticleLayout + PerCellParticleLayout: writes P = ppc*E particles (P*3 coords). */
int64_t orc_layout_random(int32_t dim, const int32_t* res, const double* lo, const double* hi,
                          int32_t ppc, long seed, double* coord);

/* --- StrainWeakening (Underworld/Rheology/src/StrainWeakening.cxx) --- */
typedef struct orc_strain_weakening {
  double healing_rate, initial_softening_strain, final_softening_strain, cohesion_after, friction_after;
} orc_strain_weakening;
/* Attach (sw != NULL) or detach strain weakening on a material with a yield
 * rheology; the first attach allocates the plastic strain w and its rate (0). */
int32_t orc_set_strain_weakening(orc_problem* pb, int32_t material, const orc_strain_weakening* sw);
int32_t orc_set_plastic_strain(orc_problem* pb, const double* w);         /* P, caller order */
int32_t orc_plastic_strain(const orc_problem* pb, double* w, double* inc); /* P each, caller order */
/* The TimeIntegrand step of StrainWeakening and MakeValuesPositive. */
int32_t orc_weakening_step(orc_problem* pb, double dt);
/* The same with the strain-limited region: particles flagged in inside get
 * `limit` before the clamp at 0. */
int32_t orc_weakening_step_limited(orc_problem* pb, double dt, const uint8_t* inside, double limit);
/* _StrainWeakening_Initialise's initial damage with the C library's
 * srand/rand: prm = {fraction, factor, finalSofteningStrain, wavenumber,
 * sinI, cosI, sinK, cosK}; inside (may be NULL) flags the particles inside
 * initialStrainShape. */
int32_t orc_initial_damage(int64_t P, const double* coord, const double* prm, long seed, const uint8_t* inside,
                           double* w);

/* --- Sobol sequence (StgDomain/Utils/src/SobolGenerator.cxx) --- */
/* One generator: primitive polynomial of degree `degree` whose inner
 * coefficients are the bits of `coefficient`, initial direction numbers m[0..degree). */
typedef struct orc_sobol {
  unsigned degree, coefficient;
  unsigned v[32];
  int have[32];
  unsigned count;   /* numbers drawn so far */
  unsigned x;       /* last number, as a 30-bit integer */
} orc_sobol;
void orc_sobol_init(orc_sobol* s, unsigned degree, unsigned coefficient, const unsigned* m);
unsigned orc_sobol_direction(orc_sobol* s, unsigned i);      /* GetDirectionalNumber */
double orc_sobol_next(orc_sobol* s);                         /* GetNextNumber */
double orc_sobol_by_index(orc_sobol* s, unsigned index);     /* GetNumberByIndex */
/* count numbers of one generator: next[k] by GetNextNumber, by_index[k] by
 * GetNumberByIndex(k) (either may be NULL). */
int32_t orc_sobol_sequence(unsigned degree, unsigned coefficient, const unsigned* m, int64_t count, double* next,
                           double* by_index);

/* SpaceFillerParticleLayout: total particles (or E * avg_per_cell when total
 * is 0) at the Sobol points of generators (deg[d], coef[d], m + 9 d) scaled to
 * the box [lo, hi]; P*3 coordinates, the unused third one 0 in 2D.  Returns P. */
int64_t orc_layout_spacefiller(int32_t dim, const int32_t* res, const double* lo, const double* hi, int64_t total,
                               double avg_per_cell, const unsigned* deg, const unsigned* coef, const unsigned* m,
                               double* coord);

/* --- output fields of the material points (StoreStress / StoreViscosity) --- */
int32_t orc_particle_fields(orc_problem* pb, const double* u_full, double* sr, double* inv, double* eta,
                            double* stress);

/* --- discrete Voronoi cell weights (DVCWeights.cxx:202-786) --- */
/* One FE cell: sub-grid res (x, y[, z]) over [-1,1]^dim; xi: nump x dim local
 * coordinates in cell-table order; cent (nump x dim) and w (nump) out. */
int32_t orc_dvc_cell(int32_t dim, const int32_t* res, int32_t nump, const double* xi, double* cent, double* w);
/* the same, with the owning particle of every sub-cell (sub-cell k = i + r0 (j + r1 l)) */
int32_t orc_dvc_cell_owners(int32_t dim, const int32_t* res, int32_t nump, const double* xi, double* cent, double* w,
                            int32_t* owner);

/* --- problem --- */
orc_problem* orc_problem_new(const orc_problem_desc* d);
/* Weights of the material points: 0 ConstantWeights (ConstantWeights.cxx
 * :178-186), 1 DVCWeights with sub-grid res (integration points move to the
 * Voronoi centroids, DVCWeights.cxx:875-951).  Re-maps the particles. */
int32_t orc_set_weights(orc_problem* pb, int32_t kind, const int32_t* res);
/* PCDVC population control (Inflow off) followed by DVC weights. */
typedef struct orc_pcdvc_params {
  double upper_t, lower_t;            /* "upperT", "lowerT" (percent of the cell volume) */
  int32_t max_splits, max_deletions;  /* "maxSplits", "maxDeletions" */
  int32_t split_in_interface_cells, delete_in_interface_cells;
  int32_t res[3];
  int32_t inflow;                    /* "Inflow" */
  int32_t particles_per_cell;        /* the root "particlesPerCell" */
  int32_t pad_;
  double threshold, cent_pos_ratio;  /* "Threshold", "CentPosRatio" */
  int64_t rand_seed, rand_skip;      /* srand(rand_seed), then rand_skip draws discarded */
} orc_pcdvc_params;
/* counts (may be NULL) = {added, deleted, rand() calls} */
int32_t orc_population_control(orc_problem* pb, const orc_pcdvc_params* p, int64_t* counts);
void orc_problem_free(orc_problem* pb);
int64_t orc_num_nodes(const orc_problem* pb);
int64_t orc_num_elements(const orc_problem* pb);
int64_t orc_num_particles(const orc_problem* pb);
void orc_get_particles(const orc_problem* pb, int32_t* owner, double* xi, double* weight, double* coord);
/* cell tables: cell_off[E+1], cell_list[P] (particle indices in each cell's table order) */
void orc_get_cell_tables(const orc_problem* pb, int64_t* cell_off, int64_t* cell_list);
int32_t orc_set_coords(orc_problem* pb, const double* coord);  /* relocate + remap after moving particles */

/* per-particle effective viscosity (ConstitutiveMatrix_Assemble chain) */
int32_t orc_particle_viscosity(orc_problem* pb, const double* u_prev_full, const double* p_prev,
                               int32_t yield_on, double* eta, uint8_t* yielded);
/* element matrices */
int32_t orc_element_K(orc_problem* pb, int64_t e, double* Ke);   /* uses the last viscosity pass */
int32_t orc_element_G(orc_problem* pb, int64_t e, double* Ge);
int32_t orc_element_f(orc_problem* pb, int64_t e, double* fe);

/* global system with Dirichlet elimination */
int32_t orc_assemble(orc_problem* pb, const double* u_prev_full, const double* p_prev, int32_t yield_on,
                     int32_t nthreads);
void orc_system_sizes(const orc_problem* pb, int64_t* out4);     /* n_u, n_p, nnz(K), nnz(G) */
void orc_system_export(const orc_problem* pb, int64_t* Kptr, int32_t* Kcol, double* Kval,
                       int64_t* Gptr, int32_t* Gcol, double* Gval, double* f, double* h,
                       double* Q, int32_t* eq);

/* unconstrained operators on full nodal vectors (current viscosities) */
void orc_apply_K_full(orc_problem* pb, const double* x, double* y);
void orc_apply_G_full(orc_problem* pb, const double* p, double* y);
void orc_apply_Gt_full(orc_problem* pb, const double* x, double* q);
void orc_diag_K_full(orc_problem* pb, double* d);
void orc_uzawa_Q(orc_problem* pb, double* Q);

/* solvers */
int32_t orc_uzawa(orc_problem* pb, const orc_uzawa_params* up, const double* p0,
                  double* u_full, double* p, orc_report* rep);
int32_t orc_picard(orc_problem* pb, const orc_uzawa_params* up, const orc_picard_params* pp,
                   double* u_full, double* p, orc_report* rep);
/* The same with a previous solution (u_prev_full N*dim, p_prev E; NULL =
 * none): the first solve then yields from it and starts from p_prev. */
int32_t orc_picard_from(orc_problem* pb, const orc_uzawa_params* up, const orc_picard_params* pp,
                        const double* u_prev_full, const double* p_prev, double* u_full, double* p, orc_report* rep);

/* time stepping */
double orc_compute_dt(orc_problem* pb, const double* u_full, double dt_factor);
int32_t orc_interpolate_velocity(orc_problem* pb, const double* u_full, const double* x, double* v);
int32_t orc_advect_rk2(orc_problem* pb, const double* u_full, double dt, double* coord_out);

#ifdef __cplusplus
}
#endif
#endif
