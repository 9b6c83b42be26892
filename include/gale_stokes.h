/*
 * gale_stokes.h -- C ABI of the MI355X Q1-P0 PIC Stokes hot path.
 *
 * Plain C types only: host pointers unless an entry point says "device
 * pointer".  Every call returns a gs_status (0 = GS_OK); gs_strerror() gives
 * its text.  A solver owns one HIP stream and is not re-entrant; distinct
 * solvers may be driven from different host threads.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the geodynamics/gale source tree).
 */
#ifndef GALE_STOKES_H
#define GALE_STOKES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference aborts through Journal_Firewall instead of
 * returning; each code names the firewall it mirrors. */
typedef enum gs_status {
  GS_OK = 0,
  GS_E_NAN = 1,          /* NaN/Inf in alpha, zeta, residual (Stokes_SLE_UzawaSolver.cxx:604,635,784,811) */
  GS_E_MAXIT = 2,        /* maxIterations reached (Stokes_SLE_UzawaSolver.cxx:815) */
  GS_E_ZERO_FORCE = 3,   /* zero momentum force vector (Stokes_SLE_UzawaSolver.cxx:476-493) */
  GS_E_STAGNATION = 4,   /* learned/mixed path stagnated; caller falls back to plain Uzawa */
  GS_E_HIP = 5,          /* HIP runtime error */
  GS_E_INVALID = 6,      /* bad argument */
  GS_E_ESCAPED = 7,      /* a particle left the box (MaterialPointsSwarm.cxx:468-482) */
  GS_E_INNER_MAXIT = 8,  /* inner velocity solve hit its iteration cap */
  GS_E_PICARD = 9,       /* Picard not converged with killNonConvergent (SystemLinearEquations.cxx:899-903) */
  GS_E_NO_DEVICE = 10    /* no HIP device / kernels not loadable */
} gs_status;

enum { GS_YIELD_NONE = 0, GS_YIELD_DRUCKER_PRAGER = 1, GS_YIELD_VON_MISES = 2 };
enum { GS_PC_NONE = 0, GS_PC_DIAGONAL = 1 };
enum { GS_PREC_FP64 = 0, GS_PREC_MIXED = 1 };
/* Inner velocity preconditioner: Jacobi, or a geometric multigrid V-cycle whose
 * levels are the Galerkin coarse operators of the particle-integrated K. */
enum { GS_INNER_JACOBI = 0, GS_INNER_MULTIGRID = 1 };
/* Operator forms timed by gs_time_apply_K (G p and G^T u included, so every
 * operator of the Uzawa loop can be measured the same way). */
enum { GS_APPLY_PARTICLES_FP64 = 0, GS_APPLY_PARTICLES_FP32 = 1, GS_APPLY_MOMENTS_FP64 = 2, GS_APPLY_MOMENTS_FP32 = 3,
       GS_APPLY_G = 4, GS_APPLY_GT = 5 };

/* One RheologyMaterial: MaterialViscosity (+ optional yield rheology) and
 * the density used by BuoyancyForceTerm.  XML keys in comments. */
typedef struct gs_material {
  double eta0;              /* MaterialViscosity "eta0" */
  double density;           /* Material "density" */
  int32_t yield_type;       /* GS_YIELD_* */
  int32_t pad_;
  double cohesion;          /* VonMises "cohesion" */
  double friction;          /* DruckerPrager "frictionCoefficient" (tan phi) */
  double min_yield_stress;  /* DruckerPrager "minimumYieldStress" (0 -> cohesion) */
  double min_viscosity;     /* DruckerPrager "minimumViscosity" */
  double max_strain_rate;   /* DruckerPrager "maxStrainRate" (0 -> off) */
} gs_material;

/* StrainWeakening attached to a material's yield rheology
 * (Underworld/Rheology/src/StrainWeakening.cxx; YieldRheology "StrainWeakening"). */
typedef struct gs_strain_weakening {
  double healing_rate;              /* StrainWeakening "healingRate" */
  double initial_softening_strain;  /* "initialSofteningStrain" */
  double final_softening_strain;    /* "finalSofteningStrain" (older name "softeningStrain") */
  double cohesion_after;            /* VonMises / DruckerPrager "cohesionAfterSoftening" */
  double friction_after;            /* DruckerPrager "frictionCoefficientAfterSoftening" */
} gs_strain_weakening;

/* Mesh (CartesianGenerator regular box), boundary conditions (WallVC as a
 * per-dof mask), gravity (BuoyancyForceTerm) and the material swarm. */
typedef struct gs_problem_desc {
  int32_t dim;              /* 2 or 3 */
  int32_t res[3];           /* elementResI/J/K */
  double lo[3], hi[3];      /* minX.. / maxX.. */
  const uint8_t* bc_mask;   /* N*dim, nonzero = Dirichlet dof (node-major, x fastest); NULL = none */
  const double* bc_value;   /* N*dim values of Dirichlet dofs; NULL = zeros */
  double gravity;           /* "gravity" */
  double ghat[3];           /* gravity direction, reference default (0,-1,0) */
  double eta_clamp_min;     /* effective-viscosity clamp (0 = off); no reference key */
  double eta_clamp_max;
  int32_t n_materials;
  int32_t pad_;
  const gs_material* materials;
  int64_t n_particles;
  const double* coord;      /* P*3 global coordinates (Coord = double[3], also in 2D) */
  const int32_t* material;  /* P material indices */
  const double* eta0_particle; /* optional per-particle eta0 override (P), NULL = material eta0 */
} gs_problem_desc;

/* Stokes_SLE_UzawaSolver XML parameters (Stokes_SLE_UzawaSolver.cxx:241-264)
 * plus the inner-solver controls this build adds. */
typedef struct gs_uzawa_params {
  double tolerance;             /* "tolerance" (1e-5) */
  int32_t max_iterations;       /* "maxIterations" (1000) */
  int32_t min_iterations;       /* "minIterations" (1) */
  int32_t use_absolute_tolerance; /* "useAbsoluteTolerance" (0) */
  int32_t preconditioner;       /* "Preconditioner": GS_PC_NONE, GS_PC_DIAGONAL (or GS_PC_LEARNED, gale_learned.h) */
  int32_t precision;            /* GS_PREC_FP64 or GS_PREC_MIXED (FP32 K-apply preconditioning FP64 CG) */
  int32_t inner_max_iterations; /* cap per inner velocity solve (0 = 50000) */
  double inner_rtol_factor;     /* inner rtol = factor * tolerance (reference: 0.1, :591,596) */
  double inner_rtol_override;   /* > 0: this rtol for every inner solve */
  int32_t monitor;              /* "monitor": print the outer residual to stderr every monitor iterations (0 = off) */
  int32_t check_every;          /* Jacobi inner CG convergence check interval (0 = 4); the multigrid
                                   inner CG checks after every iteration */
  int32_t inner_preconditioner; /* GS_INNER_JACOBI or GS_INNER_MULTIGRID */
  int32_t mg_smooth;            /* Chebyshev-Jacobi steps before and after each coarse correction (0 = 2) */
  int32_t flexible_window;      /* learned preconditioner: previous directions each new one is
                                   S-orthogonalised against (0 = 16, max 32) */
  int32_t stagnation_window;    /* learned preconditioner: stagnation = the residual drops by less than 5%
                                   over this many iterations (0 = 25) */
  int32_t learned_restarts;     /* learned preconditioner: on stagnation, restart the flexible iteration
                                   (keeping the network) up to this many times before falling back to Q */
  int32_t learned_q_every;      /* learned preconditioner: every k-th iteration preconditions with Q alone
                                   (0 = never) */
  double inner_relax_max;       /* learned preconditioner: > 0 relaxes the inner tolerance as the outer
                                   residual falls, up to this rtol, and re-checks the true residual */
  double inner_relax_gain;      /* the relaxed inner rtol is min(inner_relax_max, gain * rtol0 * |r_0| / |r_k|),
                                   rtol0 the unrelaxed inner rtol (0 = 3) */
  int32_t mg_smooth_coarse;     /* Chebyshev-Jacobi steps on the levels below the finest (0 = 1 when the
                                   inner tolerance is relaxed, inner_relax_max > 0, else mg_smooth) */
} gs_uzawa_params;

/* SystemLinearEquations nonlinear keys (SystemLinearEquations.cxx:346-356). */
typedef struct gs_picard_params {
  double tolerance;             /* "nonLinearTolerance" (0.01) */
  int32_t max_iterations;       /* "nonLinearMaxIterations" (500) */
  int32_t min_iterations;       /* "nonLinearMinIterations" (1) */
  int32_t kill_non_convergent;  /* "killNonConvergent" */
  int32_t pad_;
} gs_picard_params;

typedef struct gs_report {
  int32_t status;
  int32_t uzawa_iterations;
  int64_t inner_iterations;
  double rel_residual;          /* (|r|/sqrt(n_p)) / uzawaRhsScale, Stokes_SLE_UzawaSolver.cxx:800-802 */
  double abs_residual;
  double div_u;                 /* |G^T u| / |u| */
  double momentum_residual;     /* |f - K u - G p| / |f| */
  double solve_seconds;         /* device time of the Uzawa solve(s) */
  double setup_seconds;         /* device time of viscosity pass + diagonal operators + RHS */
  int32_t picard_iterations;
  int32_t learned_fallback;     /* learned preconditioner: Uzawa iteration of the fallback to Q (0 = none) */
  double picard_residual;
  int64_t k_applies;            /* K applies issued (FP64 + FP32) */
  int32_t verify_rounds;        /* learned preconditioner with relaxed inner solves: true-residual checks */
  int32_t pad2_;
} gs_report;

typedef struct gs_problem gs_problem;
typedef struct gs_solver gs_solver;

const char* gs_strerror(int32_t status);
int32_t gs_version(void);
/* Number of HIP devices visible (0 if none). */
int32_t gs_device_count(void);

/* Mesh + swarm + materials: replaces the FeMesh / MaterialPointsSwarm /
 * IntegrationPointsSwarm construction (CartesianGenerator.cxx:894-955,
 * IntegrationPointsSwarm.cxx:391-449, CoincidentMapper.cxx:158-260,
 * ConstantWeights.cxx:178-186).  Particles are located (K9) and sorted by
 * owning element on the GPU. */
int32_t gs_problem_create(const gs_problem_desc* d, gs_problem** out);
void gs_problem_destroy(gs_problem* pb);
/* out5 = {N nodes, E elements, P particles, dim, escaped particles} */
int32_t gs_problem_info(const gs_problem* pb, int64_t* out5);
/* Particle state in the caller's original particle order: coord (P*3),
 * owner (P, -1 = outside), xi (P*dim), weight (P).  Any pointer may be NULL. */
int32_t gs_problem_particles(gs_problem* pb, double* coord, int32_t* owner, double* xi, double* weight);

/* Stokes_SLE + Stokes_SLE_UzawaSolver: _DefaultNew/_AssignFromXML/_Build
 * (Stokes_SLE_UzawaSolver.cxx:199-264). */
int32_t gs_solver_create(gs_problem* pb, const gs_uzawa_params* up, gs_solver** out);
int32_t gs_solver_set_params(gs_solver* s, const gs_uzawa_params* up);
void gs_solver_destroy(gs_solver* s);

/* MatrixSetup + VectorSetup of one linear execute (SystemLinearEquations.cxx
 * :573-603): the fused constitutive pass (ConstitutiveMatrix_Assemble chain,
 * ConstitutiveMatrix.cxx:280-330, with YieldRheology.cxx:190-231 when
 * yield_on), diag(K), the Uzawa preconditioner Q (UzawaPreconditionerTerm.cxx
 * :156-239), buoyancy f (BuoyancyForceTerm.cxx:414-533) and Dirichlet lifting
 * (StiffnessMatrix.cxx:2092-2144).  u_prev (N*dim) / p_prev (E) are host
 * arrays of the previous solution (NULL = use the solver's own). */
int32_t gs_solver_setup(gs_solver* s, const double* u_prev, const double* p_prev, int32_t yield_on);

/* _Stokes_SLE_UzawaSolver_Solve (Stokes_SLE_UzawaSolver.cxx:393-938).  The
 * initial pressure guess is the solver's pressure vector (zero after create,
 * the last solution afterwards, or what gs_solver_set_pressure set), as the
 * reference keeps pSolnVec.  u (N*dim, with Dirichlet values) and p (E) are
 * written if non-NULL. */
int32_t gs_stokes_solve(gs_solver* s, double* u, double* p, gs_report* rep);
int32_t gs_solver_set_pressure(gs_solver* s, const double* p);

/* SystemLinearEquations_NonLinearExecute (SystemLinearEquations.cxx:735-920):
 * setup + solve, then Picard.  The first solve yields only when a previous
 * solution exists -- after this solver's first nonlinear execute (the
 * reference's sle->hasExecuted, :794, read as previousSolutionExists by
 * ConstitutiveMatrixCartesian.cxx:270-273 and YieldRheology.cxx:203-204) or
 * after gs_solver_set_solution -- and then yields from the stored solution.
 * gs_stokes_solve follows the same rule when it runs the setup itself; it
 * does not mark a previous solution (a linear execute does not set
 * hasExecuted). */
int32_t gs_picard_solve(gs_solver* s, const gs_picard_params* pp, double* u, double* p, gs_report* rep);
/* Restart: load a previous solution (u: full nodal velocity N*dim with the
 * Dirichlet values; p: E) and mark it as existing, as a checkpoint restart
 * does (ConstitutiveMatrix.cxx:127-130, context->loadFromCheckPoint): the
 * next setup yields from it, and p is the next Uzawa initial guess. */
int32_t gs_solver_set_solution(gs_solver* s, const double* u, const double* p);

/* _Stokes_SLE_UzawaSolver_GetRhs: fhat = G^T K^-1 f - h  (E). */
int32_t gs_get_rhs(gs_solver* s, double* fhat);
/* _Stokes_SLE_UzawaSolver_FormResidual: r = fhat - G^T K^-1 G p  (E). */
int32_t gs_form_residual(gs_solver* s, const double* p, double* r);

/* Stokes_SLE_UpdateDt (UpdateDt.cxx:53-84) on the solver's velocity. */
int32_t gs_compute_dt(gs_solver* s, double dt_factor, double* dt);
/* TimeIntegrand_SecondOrder with SwarmAdvector (TimeIntegrand.cxx:265-335,
 * SwarmAdvector.cxx:261-280), then Swarm_UpdateAllParticleOwners
 * (SwarmClass.cxx:878-974) and the CoincidentMapper remap with
 * ConstantWeights, using the solver's velocity.  With strain weakening on,
 * the plastic strain of every particle is integrated over the same dt first
 * (StrainWeakening.cxx:432-468: w += dt * increment, negatives set to 0). */
int32_t gs_advect_rk2(gs_solver* s, double dt);
/* Owner update only (Mesh_RegularAlgorithms.cxx:191-224 rule, bit-exact);
 * owner_out (P, caller order) may be NULL. */
int32_t gs_relocate(gs_problem* pb, int32_t* owner_out);
/* Replace particle coordinates (caller order, P*3) and relocate. */
int32_t gs_problem_set_coords(gs_problem* pb, const double* coord);

/* Attach StrainWeakening to material `material`'s yield rheology (NULL
 * detaches it).  The yield criterion then uses cohesion and friction
 * interpolated towards their after-softening values by the ratio
 * clip((w - initial) / (final - initial), 0, 1) of each particle's
 * accumulated plastic strain w (VonMises.cxx:185-205,
 * DruckerPrager.cxx:448-489, StrainWeakening_CalcRatio
 * StrainWeakening.cxx:497-519), and every constitutive pass with a previous
 * solution stores the rate (tau_y / eta) (beta / (1 - beta) - healing w),
 * beta = max(0, 1 - tau_y / tau_II) (_StrainWeakening_CalcIncrementIsotropic,
 * :470-495; YieldRheology.cxx:190-228).  The first call allocates w and its
 * rate, both 0 (no initial damage).  GS_E_INVALID for a material without a
 * yield rheology or final <= initial softening strain. */
int32_t gs_problem_set_strain_weakening(gs_problem* pb, int32_t material, const gs_strain_weakening* sw);
/* StrainWeakening "strainLimitedShape" / "strainLimit"
 * (_StrainWeakening_MakeValuesPositive, StrainWeakening.cxx:443-468): after
 * each time step the particles inside the shape get the plastic strain
 * `limit` (then negatives become 0).  shape_type: -1 none, 0 everywhere,
 * 1 sphere (params: centre x, y, z, radius), 2 box (start x, y, z, end x, y,
 * z) — GS_SHAPE_* of gale_generate.h, containment as in Sphere.cxx:189-208
 * and Box.cxx:209-221. */
int32_t gs_problem_set_strain_limit(gs_problem* pb, int32_t shape_type, const double* shape_params, double limit);
/* Set the accumulated plastic strain (P, caller order); the reference's
 * initial damage (rand()-seeded, StrainWeakening.cxx:333-430) is set this way. */
int32_t gs_problem_set_plastic_strain(gs_problem* pb, const double* w);
/* Read the accumulated plastic strain and its rate from the last
 * constitutive pass (P each, caller order; either may be NULL). */
int32_t gs_problem_plastic_strain(gs_problem* pb, double* w, double* inc);

/* Operator applies on full nodal vectors, unconstrained (no Dirichlet
 * masking), for operator-level parity and measurement.  DEVICE pointers on
 * the solver's device; the call is asynchronous on the solver's stream until
 * gs_solver_sync.  K uses the viscosities of the last gs_solver_setup.
 *   K: x, y (N*dim)   G: p (E) -> y (N*dim)   G^T: x (N*dim) -> q (E) */
int32_t gs_apply_K(gs_solver* s, const double* x_dev, double* y_dev);
/* gs_apply_K that allocates nothing: no hipMalloc, hipFree or allocating
 * library call.  No counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).  The
 * apply's scratch (the tile-face partial sums, 8 * 405 * 3 bytes per 8x8x4
 * element tile in 3D) lives in a static 64 MiB device array, one per device,
 * resolved when the problem is created; larger meshes (more than 6904 such
 * tiles, about 121^3 elements) are refused with GS_E_INVALID and a message.  The
 * solver must already be set up for the problem's current particles
 * (gs_solver_setup, which does allocate), otherwise GS_E_INVALID.  Every call
 * on one device shares that device's scratch: calls that overlap on one
 * device (different streams or threads) race.  Same result as gs_apply_K,
 * bit for bit. */
int32_t gs_apply_K_noalloc(gs_solver* s, const double* x_dev, double* y_dev);
int32_t gs_apply_K_fp32(gs_solver* s, const float* x_dev, float* y_dev);
int32_t gs_apply_G(gs_solver* s, const double* p_dev, double* y_dev);
int32_t gs_apply_Gt(gs_solver* s, const double* x_dev, double* q_dev);
int32_t gs_solver_sync(gs_solver* s);
/* Average device time (ms) of n_rep applies of the given form (GS_APPLY_*:
 * K in four forms, G p, G^T u), timed with HIP events on the solver's stream. */
int32_t gs_time_apply_K(gs_solver* s, int32_t form, int32_t n_rep, double* ms_per_apply);
/* The same with cold != 0: the applies cycle through copies of their operands
 * (input and output vectors, moments or particle stream) totalling more than
 * three times the 256 MB last-level cache, so each apply streams from HBM as
 * it does inside a solve.  cold = 0 is gs_time_apply_K.  No reference
 * counterpart (measurement only). */
int32_t gs_time_apply(gs_solver* s, int32_t form, int32_t n_rep, int32_t cold, double* ms_per_apply);
/* K apply from the per-element particle moments (the form the solver uses;
 * equal to gs_apply_K up to rounding).  Device pointers, full nodal vectors. */
int32_t gs_apply_K_moments(gs_solver* s, const double* x_dev, double* y_dev);

/* Diagnostics of the last setup (host arrays): diag(K) on all dofs (N*dim),
 * Q (E), f after lifting (N*dim, zero on Dirichlet dofs), h (E),
 * per-particle effective viscosity (P, caller order).  NULL skips. */
int32_t gs_solver_setup_arrays(gs_solver* s, double* diagK, double* Q, double* f, double* h, double* eta);
/* Output fields of the material points, caller order, any pointer may be NULL:
 * strain rate of the velocity u at each integration point (nst = 3 in 2D:
 * xx, yy, xy; 6 in 3D: xx, yy, zz, xy, xz, yz; the symmetric part of the
 * velocity gradient, the reference's StrainRateField), its second invariant
 * (SymmetricTensor_2ndInvariant, TensorMath.cxx:431-459), the viscosity of
 * the last constitutive pass (StoreVisc, StoreViscosity.cxx:156-180) and the
 * stress of the isotropic constitutive matrix with that viscosity
 * (StoreStress, StoreStress.cxx:190-217 via ConstitutiveMatrix_CalculateStress).
 * u: full nodal velocity (N*dim), or NULL for the solver's last solution. */
int32_t gs_solver_particle_fields(gs_solver* s, const double* u, double* strain_rate, double* strain_rate_inv,
                                  double* viscosity, double* stress);

#ifdef __cplusplus
}
#endif
#endif
