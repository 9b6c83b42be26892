/* gale_batch.h -- batched solve of many independent small 2D Stokes problems
 * (libgale_stokes.so).
 *
 * No counterpart in the reference, which runs one model per process
 * (StgFEM/SLE/ProvidedSystems/StokesFlow/src/Stokes_SLE_UzawaSolver.cxx:393-938
 * per model).  Here every instance keeps the reference Uzawa loop (diagonal
 * Q of UzawaPreconditionerTerm.cxx:156-239, the same residual and stopping
 * rule) and the whole batch is one persistent launch: each workgroup takes
 * instances from a queue and runs their complete solve, so no grid-wide
 * synchronisation is needed.  The inner velocity solve is FP64 flexible CG
 * preconditioned by one FP32 geometric-multigrid V-cycle (the instance's
 * Galerkin levels built by the single-problem setup; FP32 K applies, FP64
 * vectors, dots and outer loop).  The element-tile shape of the K apply
 * (workgroup size x tile width) is chosen by gs_batch_autotune; the solve can
 * be captured once as a HIP graph and replayed.
 *
 * Instances: 2D, homogeneous Dirichlet conditions (e.g. free slip), no
 * yielding (first, linear solve).  Setup of each instance (particle
 * location, constitutive pass, element moments, right-hand side) reuses the
 * single-problem path of gale_stokes.h.
 */
#ifndef GALE_BATCH_H
#define GALE_BATCH_H

#include <stdint.h>

#include "gale_stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gs_batch gs_batch;

typedef struct gs_batch_params {
  double tolerance;          /* Uzawa "tolerance" (relative residual, as the reference) */
  int32_t max_iterations;    /* Uzawa "maxIterations" per instance */
  int32_t max_inner;         /* cap per inner solve */
  double inner_rtol_factor;  /* inner rtol = factor * tolerance (reference: 0.1) */
  int32_t mg_smooth;         /* Chebyshev-Jacobi steps before and after each coarse correction (0 = 2) */
  int32_t use_graph;         /* 1: capture the launch once as a HIP graph and replay it */
  double inner_relax_max;    /* > 0: inexact Uzawa -- the inner rtol grows as the outer residual falls,
                                min(inner_relax_max, max(irt, irt |r_0| / |r_k|)), the outer loop is flexible
                                CG on the Schur complement S-orthogonalised against the last flexible_window
                                directions, and convergence is re-checked on the true residual (accurate
                                inner solve); 0 = the reference loop */
  int32_t flexible_window;   /* directions of the flexible outer loop (0 = 8, at most 16) */
  int32_t pad_;
} gs_batch_params;

typedef struct gs_batch_result {
  int32_t status;            /* gs_status of the instance */
  int32_t uzawa_iterations;
  int64_t inner_iterations;
  double rel_residual;
  double seconds;            /* device time the instance's solve took inside the launch */
} gs_batch_result;

/* count problem descriptions (2D); each instance is set up on the device. */
int32_t gs_batch_create(int32_t count, const gs_problem_desc* descs, gs_batch** out);
void gs_batch_destroy(gs_batch* b);
/* Tile shape of the batched K apply: block in {256, 512, 1024}, tile width tx
 * dividing block (tx in {16, 32, 64}). */
int32_t gs_batch_set_tile(gs_batch* b, int32_t block, int32_t tx);
/* Times every tile shape on this batch (a few Uzawa iterations each, HIP
 * events) and keeps the fastest; ms receives one time per candidate in the
 * order (256,16) (256,32) (512,32) (512,64) (1024,32) (1024,64). */
int32_t gs_batch_autotune(gs_batch* b, const gs_batch_params* p, int32_t* block, int32_t* tx, float* ms);
/* Solves every instance (results: count entries, may be NULL); seconds = device
 * time of the launch. */
int32_t gs_batch_solve(gs_batch* b, const gs_batch_params* p, gs_batch_result* results, double* seconds);
/* Velocity (2 N_i) and pressure (E_i) of instance i. */
int32_t gs_batch_solution(gs_batch* b, int32_t i, double* u, double* p);

#ifdef __cplusplus
}
#endif

#endif
