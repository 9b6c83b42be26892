/* gale_learned.h -- learned Schur-complement preconditioner for the Uzawa
 * pressure loop (libgale_stokes.so).
 *
 * This part has no counterpart in the reference: Gale's Uzawa loop
 * (StgFEM/SLE/ProvidedSystems/StokesFlow/src/Stokes_SLE_UzawaSolver.cxx:664-837)
 * preconditions the Schur residual with the diagonal Q of
 * UzawaPreconditionerTerm.cxx:156-239 only.  Here a 3D convolutional
 * encoder-decoder (BF16 matrix-core convolutions, FP32 accumulation) maps the
 * element log10 viscosity and the diagonally preconditioned residual Q^-1 r to
 * a correction (5-level U-Net, channels 16/32/48/64/96, 1.03 M parameters): z = Q^-1 r + rms(Q^-1 r) * net(Q^-1 r / rms, log10 eta / 3).
 * The Uzawa loop then runs as truncated flexible CG (each direction
 * S-orthogonalised against the previous flexible_window ones); the first
 * preconditioned step is the network's initial pressure guess (scaled by the
 * line search).  When the residual stagnates, or z.r <= 0, the solve
 * falls back to the plain diagonal Q and continues to the same tolerance.
 *
 * The network is trained on the device (forward and backward kernels of the
 * library, Adam) from pairs that plain Uzawa solves record:
 *   (r0, p* - p0): initial Schur residual and converged pressure change, and
 *   (q_k, s_k):    Krylov pairs q_k = G^T K^-1 G s_k of the Uzawa iterations.
 * Both satisfy S e = r exactly (up to the inner solve tolerance).
 *
 * Requirements: 3D problems with element counts divisible by 16 per axis.
 * Device pointers are marked d_; everything else is host memory.
 */
#ifndef GALE_LEARNED_H
#define GALE_LEARNED_H

#include <stdint.h>

#include "gale_stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { GS_PC_LEARNED = 2 }; /* gs_uzawa_params.preconditioner */

typedef struct gs_lpc gs_lpc;                 /* network + optimiser state */
typedef struct gs_lpc_dataset gs_lpc_dataset; /* training pairs on the device */

typedef struct gs_lpc_train_params {
  int32_t steps;
  int32_t batch;
  float lr;        /* peak learning rate (cosine decay to lr_final) */
  float lr_final;
  float beta1, beta2, eps;
  float clip;      /* global gradient-norm clip (0 = none) */
  int32_t warmup;  /* linear warm-up steps */
  int32_t first;   /* train on samples [first, first + count) of the dataset */
  int32_t count;   /* 0 = all from first */
  uint32_t seed;   /* batch sampling */
} gs_lpc_train_params;

/* Network with He-normal initial weights from seed (FP32 master copy). */
int32_t gs_lpc_create(uint64_t seed, gs_lpc** out);
void gs_lpc_destroy(gs_lpc* L);
int64_t gs_lpc_num_params(void);
int32_t gs_lpc_get_params(gs_lpc* L, float* params);
int32_t gs_lpc_set_params(gs_lpc* L, const float* params);

/* Raw network map on a batch of grids: d_in is FP32 [batch][nz][ny][nx][2]
 * (channel 0 the scaled residual, 1 the scaled log viscosity), d_out FP32
 * [batch][nz][ny][nx]. */
int32_t gs_lpc_forward(gs_lpc* L, int32_t nx, int32_t ny, int32_t nz, int32_t batch, const float* d_in,
                       float* d_out);
/* Parameter gradient of sum(d_gout * net(d_in)) (forward included); grads
 * receives gs_lpc_num_params() floats. */
int32_t gs_lpc_backward(gs_lpc* L, int32_t nx, int32_t ny, int32_t nz, int32_t batch, const float* d_in,
                        const float* d_gout, float* grads);
/* Training loss of the batch (d_target FP32 [batch][V], weights and
 * mean-projection flags per sample, host) and its gradient d_gout (device). */
int32_t gs_lpc_loss(gs_lpc* L, int32_t nx, int32_t ny, int32_t nz, int32_t batch, const float* d_out,
                    const float* d_target, const float* weights, const int32_t* proj, float* d_gout,
                    double* loss);

int32_t gs_lpc_dataset_create(int32_t nx, int32_t ny, int32_t nz, int32_t capacity, gs_lpc_dataset** out);
void gs_lpc_dataset_destroy(gs_lpc_dataset* D);
int32_t gs_lpc_dataset_size(const gs_lpc_dataset* D);
/* Copy sample idx to host: in (V*2), target (V), weight, proj (any may be NULL). */
int32_t gs_lpc_dataset_get(gs_lpc_dataset* D, int32_t idx, float* in, float* target, float* weight,
                           int32_t* proj);
/* Record training pairs during the following gs_stokes_solve calls of s:
 * the (r0, p* - p0) pair of each converged solve and every
 * krylov_every-th Krylov pair (0 = none).  D = NULL stops recording. */
int32_t gs_solver_record_pairs(gs_solver* s, gs_lpc_dataset* D, int32_t krylov_every);

/* Adam on the device; loss_history receives params->steps values (or NULL). */
int32_t gs_lpc_train(gs_lpc* L, gs_lpc_dataset* D, const gs_lpc_train_params* params, float* loss_history);
/* Mean loss over samples [first, first + count). */
int32_t gs_lpc_eval(gs_lpc* L, gs_lpc_dataset* D, int32_t first, int32_t count, double* loss);

/* Attach the network to a solver; used when params.preconditioner ==
 * GS_PC_LEARNED.  The solver keeps a reference; L must outlive it. */
int32_t gs_solver_set_learned(gs_solver* s, gs_lpc* L);

/* Diagnostic: one 16x16x16 BF16 matrix-core product against exact values. */
int32_t gs_lpc_selftest(float* max_err);

#ifdef __cplusplus
}
#endif

#endif
