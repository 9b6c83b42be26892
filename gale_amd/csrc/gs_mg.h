// gs_mg.h -- launchers of the multigrid kernels (gs_mg.hip).
#pragma once

#include "gs_internal.h"

namespace gs {

void launch_moments_fine(const Geo& g, const double4* pk, const int* off, double* M, hipStream_t st);
void launch_moments_coarsen(const Geo& gc, const Geo& gf, const double* Mf, double* Mc, hipStream_t st);
void launch_d2f(const double* a, float* b, long long n, hipStream_t st);
// FP32 copy of a level's moments, layout-aware (gs_tile.h mom_index).
void launch_mom_d2f(const Geo& g, const double* a, float* b, hipStream_t st);
void launch_f2d(const float* a, double* b, long long n, hipStream_t st);
// Moment-form K apply (tiled, gs_tile.hip); Mom in the layout of gs_tile.h
// (mom_index: tile-major in 3D, SoA in 2D); halo
// needs tile_halo_size(g) values.
template <typename T>
void launch_mom_apply(const Geo& g, const T* x, T* y, const T* Mom, const uint8_t* bcflag, T* halo, hipStream_t st);
void launch_mom_diag(const Geo& g, const double* Mom, double* diag, hipStream_t st);
template <typename T>
void launch_restrict(const Geo& gc, const Geo& gf, const T* rf, T* bc, const uint8_t* bcflag_c, hipStream_t st);
template <typename T>
void launch_prolong_add(const Geo& gc, const Geo& gf, const T* xc, T* xf, const uint8_t* bcflag_f, hipStream_t st);
template <typename T>
void launch_cheb(const T* b, const T* Ax, const T* dinv, T* d, T* x, long long n, T alpha, T beta, int first,
                 hipStream_t st);
template <typename T>
void launch_resid(const T* b, const T* Ax, T* r, long long n, hipStream_t st);
// Dense coarsest solve x = Kinv b on the free dofs.  ld == 0: Kinv packed
// nf x nf (one wavefront per row, the right-hand side gathered per row);
// ld > 0: rows padded to ld (a multiple of 4) for 16-byte row loads, the
// gathered right-hand side staged once per workgroup in LDS (nf at most
// kDenseWideMax).
constexpr int kDenseWideMax = 8192;
template <typename T>
void launch_dense_solve(const T* Kinv, const int* freeidx, int nf, int ld, const T* b, T* x, long long n,
                        hipStream_t st);
// Fills the upper triangle of a column-major n x n matrix (leading dimension
// ld) from its lower triangle.
void launch_symmetrize_lower(double* A, int n, int ld, hipStream_t st);

inline int n_moments(int dim) { return dim == 2 ? 9 : 27; }

}  // namespace gs
