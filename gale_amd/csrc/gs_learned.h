// gs_learned.h -- internal interface between the Uzawa driver (gs_api.hip)
// and the learned preconditioner (gs_learned.hip, gs_net.hip).
#pragma once

#include "../../include/gale_learned.h"
#include "gs_internal.h"

namespace gs {

// Element log10 of the mean particle viscosity, divided by 3:
// eta_e = sum_p w_p detJ eta_p / volume (the moment M000 over the volume).
void launch_elem_logeta(const Geo& g, const double4* pk, const int* off, float* out, hipStream_t st);

// z = Q^-1 r + rms * net(Q^-1 r / rms, logeta); S[S_ZR] = z.r and, when
// qprev is given, S[S_UBETA] = -alpha_prev (z.qprev) / S[S_ZR_PREV]
// (Polak-Ribiere beta of the flexible iteration).
int lpc_precond(gs_lpc* L, const Geo& g, const double* r, const double* Qinv, const float* logeta,
                const double* qprev, double* z, double* S, Reducer rd, hipStream_t st);

// Append one training pair (S e = rlike) to D; a full dataset ignores it.
int lpc_record(gs_lpc_dataset* D, const Geo& g, const double* rlike, const double* elike, const double* Qinv,
               const float* logeta, int proj, double* S, Reducer rd, hipStream_t st);

// Truncated flexible CG on the Schur complement: with the c previous
// directions s_j (ring sring[j*n]) and q_j = S s_j (qring), sq[j] = s_j . q_j,
//   s = z - sum_j (z . q_j / sq[j]) s_j        (written into sring slot `out`)
constexpr int kMaxWindow = 32;
void launch_orthogonalise(const double* z, double* sring, const double* qring, const double* sq, int c, int out,
                          long long n, double* partials, double* beta, hipStream_t st);
// S[S_ZR] = s.r, S[S_SQ] = s.q, sq_out[0] = s.q
void launch_flex_alpha(const double* s, const double* r, const double* q, long long n, double* S, double* sq_out,
                       Reducer rd, hipStream_t st);

// Shape check: 3D and every element count divisible by 16 (the network's pooling depth).
bool lpc_shape_ok(const Geo& g);

}  // namespace gs
