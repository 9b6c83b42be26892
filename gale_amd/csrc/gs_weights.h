// gs_weights.h -- discrete Voronoi cell weights on the device (gs_weights.hip).
#pragma once

#include "gs_internal.h"

namespace gs {

constexpr int kDvcMaxRes = 32;              // sub-cells per axis
constexpr size_t kDvcMaxLds = 64 * 1024;    // LDS of one cell's workgroup

// Cell e's particles pk[off[e]..off[e+1]) (xi in pk.xyz, growth order by pid)
// get pk.xyz = centroid and w = weight.  maxp: largest particle count of a
// cell (cells with more are left untouched; callers check beforehand).
int dvc_weights(int dim, const int* res, long long ncell, const int* off, const int* pid, double4* pk, double* w,
                int maxp, hipStream_t st);
// Largest per-cell particle count whose LDS fits kDvcMaxLds.
long long dvc_max_particles(int dim, const int* res);

}  // namespace gs
