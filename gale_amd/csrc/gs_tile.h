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

TileGrid tile_grid(const Geo& g);
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

}  // namespace gs
