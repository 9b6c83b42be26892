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

// PCDVC decisions per cell (one thread per element, particles in pid order):
// del[i] = 1 for a deleted (sorted) particle, nsplit[e] new particles at
// scoord[(e * slots + k) * 3 ..] with parent (sorted index) sparent[e * slots + k].
struct PcdvcArgs {
  double max_w, min_w;
  int max_splits, max_deletions, split_if, delete_if, slots;
};
void pcdvc_decide(const Geo& g, const int* off, const int* pid, const int* mat, const double4* pk, const double* w,
                  const PcdvcArgs& a, unsigned char* del, int* nsplit, double* scoord, int* sparent, hipStream_t st);
// Rebuild the particle arrays in the new caller order: survivors to map[i]
// (-1 = deleted), new particles after n_surv in cell order.
// Optional per-particle double arrays that follow the particles (per-particle
// eta0, StrainWeakening plastic strain and its rate); null entries are skipped.
constexpr int kPartDoubles = 3;
struct PartDoubles {
  double* a[kPartDoubles];
};
void pcdvc_rebuild(long long P, long long E, const int* map, const double* coord, const int* mat, PartDoubles x,
                   const int* nsplit, const int* base, int slots, const double* scoord, const int* sparent,
                   long long n_surv, double* coordN, int* matN, PartDoubles xN, hipStream_t st);

}  // namespace gs
