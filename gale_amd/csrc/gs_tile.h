// gs_tile.h -- tiled K apply (gs_tile.hip).
#pragma once

#include "gs_internal.h"

namespace gs {

constexpr int kTileThreads = 256;
constexpr int kTile3X = 8, kTile3Y = 8, kTile3Z = 4;  // 256 elements, 9x9x5 node box
constexpr int kTile2X = 16, kTile2Y = 16;             // 256 elements, 17x17 node box

struct TileGrid {
  int n[3];
  int nb;            // nodes in a tile's node box
  long long ntiles;
};

// Node update fused into a K apply where the apply completes a node, so a
// multigrid smoothing step or residual needs no separate vector pass.  With
// ax = (K x)_i (zero on Dirichlet dofs):
//   kind 0: y_i = ax
//   kind 1: Chebyshev step  d_i = alpha d_i + beta dinv_i (b_i - ax),  xout_i = x_i + d_i
//   kind 2: residual  r_i = b_i - ax
// The update runs in the node-parallel launch that completes tile-face
// nodes; y holds the apply's partial sums.  xout must not alias x.
template <typename T>
struct TileEpi {
  int kind;
  const T* b;
  const T* dinv;
  T* d;
  T* xout;
  T* r;
  T alpha, beta;
};

// The bottom of a 3D multigrid V-cycle (the smallest levels and the coarsest
// solve) in one workgroup: each level's K applies accumulate into one LDS
// node array, so the whole bottom is a single launch instead of a dozen
// small ones per level.  Same smoothing, transfers and coarsest solve as the
// multi-launch cycle; sums are taken in a different (fixed) order.
template <typename T>
struct BotLevel {
  Geo g;
  const T* M;            // moments, SoA
  const T* dinv;         // masked inverse diagonal
  const uint8_t* bc;     // Dirichlet node flags (bit c: dof c)
  T *b, *x, *x2, *r, *d; // work vectors (x, x2: ping-pong)
  double lmax;           // Chebyshev bound of D^-1 K
};
constexpr int kBotMaxLevels = 6;
constexpr int kBotMaxDense = 1024;
template <typename T>
struct BotArgs {
  BotLevel<T> lv[kBotMaxLevels];
  int nlev;              // lv[nlev - 1] is the coarsest level
  int nu;                // Chebyshev steps per smoothing
  const T* Kinv;         // dense coarsest inverse (nfc x nfc), or nullptr
  const int* freeidx;
  int nfc;               // 0: eight Chebyshev steps on the coarsest level
};
// A level fits the bottom kernel when its node accumulator fits in 64 KiB of LDS.
template <typename T>
inline bool bot_fits(const Geo& g) {
  return g.dim == 3 && g.N * 3 * (long long)sizeof(T) <= 65536;
}
// Solves lv[0] (b given) by the V-cycle below it; the result is left in lv[0].x.
void mg_bottom3(const BotArgs<float>& a, hipStream_t st);
void mg_bottom3(const BotArgs<double>& a, hipStream_t st);

TileGrid tile_grid(const Geo& g);

// Element moments (27 per 3D element, 9 per 2D element).
// 3D: tile-major.  The 256 elements of one 8x8x4 apply tile are contiguous
// for every moment, and consecutive moments are grouped by G = 16 / sizeof(T)
// (pairs in FP64, quads in FP32) so one 16-byte load brings G moments of one
// element:
//   base(e) = tile(e) * 27 * 256
//   k < G * (27 / G):  base(e) + G * (k / G) * 256 + G * local(e) + k % G
//   otherwise:         base(e) + k * 256 + local(e)
// with tile(e) the apply tile of element e and local(e) = lx + 8 ly + 64 lz
// its place in it; tiles cut by the mesh edge keep all 256 slots.  A
// workgroup applying one tile streams 27 * 256 contiguous values.
// 2D: SoA, M[k * E + e].
template <typename T>
__host__ __device__ inline long long mom_count(const Geo& g) {
  if (g.dim != 3) return 9 * g.E;
  const long long tx = (g.res[0] + kTile3X - 1) / kTile3X, ty = (g.res[1] + kTile3Y - 1) / kTile3Y,
                  tz = (g.res[2] + kTile3Z - 1) / kTile3Z;
  return tx * ty * tz * 27 * kTileThreads;
}
template <typename T>
__host__ __device__ inline long long mom_index(const Geo& g, int k, long long e) {
  if (g.dim != 3) return (long long)k * g.E + e;
  constexpr int G = 16 / (int)sizeof(T);
  const long long ex = e % g.res[0], ey = (e / g.res[0]) % g.res[1], ez = e / ((long long)g.res[0] * g.res[1]);
  const long long tnx = (g.res[0] + kTile3X - 1) / kTile3X, tny = (g.res[1] + kTile3Y - 1) / kTile3Y;
  const long long tile = ex / kTile3X + tnx * (ey / kTile3Y + tny * (ez / kTile3Z));
  const int local = (int)(ex % kTile3X + kTile3X * (ey % kTile3Y + kTile3Y * (ez % kTile3Z)));
  const long long base = tile * 27 * kTileThreads;
  if (k < G * (27 / G)) return base + (long long)G * (k / G) * kTileThreads + G * local + k % G;
  return base + (long long)k * kTileThreads + local;
}
// Halo buffer length (values) for one apply on this mesh.
long long tile_halo_size(const Geo& g);

void tile_apply_particles(const Geo& g, const double* x, double* y, double* halo, const double4* pk, const int* off,
                          const uint8_t* bcflag, hipStream_t st);
void tile_apply_particles32(const Geo& g, const float* x, float* y, float* halo, const float4* pk, const int* off,
                            const uint8_t* bcflag, hipStream_t st);
void tile_apply_moments(const Geo& g, const double* x, double* y, double* halo, const double* Mom,
                        const uint8_t* bcflag, hipStream_t st);
void tile_apply_moments32(const Geo& g, const float* x, float* y, float* halo, const float* Mom,
                          const uint8_t* bcflag, hipStream_t st);
// Moment-form apply with a fused node update (kind 0 is tile_apply_moments).
void tile_apply_moments_epi(const Geo& g, const double* x, double* y, double* halo, const double* Mom,
                            const uint8_t* bcflag, const TileEpi<double>& ep, hipStream_t st);
void tile_apply_moments_epi(const Geo& g, const float* x, float* y, float* halo, const float* Mom,
                            const uint8_t* bcflag, const TileEpi<float>& ep, hipStream_t st);

}  // namespace gs
