/* gale_weights.h -- integration weights of the material points
 * (libgale_stokes.so).
 *
 * The reference picks the weights with the integration swarm's
 * "WeightsCalculator" component.  Two are provided:
 *   GS_WEIGHTS_CONSTANT  ConstantWeights (ConstantWeights.cxx:178-186):
 *                        w = 2^dim / particles-in-cell, xi = the particle's
 *                        local coordinates (the default).
 *   GS_WEIGHTS_DVC       DVCWeights (DVCWeights.cxx:693-786, 875-951): the
 *                        particles of a cell grow discrete Voronoi regions on
 *                        a fixed sub-grid of the local box [-1,1]^dim; each
 *                        integration point moves to the centroid of its region
 *                        and weighs the region's volume.  This is also the
 *                        weight PCDVC computes when it neither splits nor
 *                        deletes particles (PCDVC.cxx:1092-1512 with
 *                        maxSplits = maxDeletions = 0 and Inflow off).
 * Particles of a cell take part in the growth in the order of the caller's
 * particle indices (the order of the reference's cell particle table after
 * the initial layout).  Limits: sub-grid resolution 1..32 per axis; a cell
 * may hold at most gs_dvc_max_particles(dim, res) particles (GS_E_INVALID
 * otherwise).
 */
#ifndef GALE_WEIGHTS_H
#define GALE_WEIGHTS_H

#include <stdint.h>

#include "gale_stokes.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { GS_WEIGHTS_CONSTANT = 0, GS_WEIGHTS_DVC = 1 };

/* Selects the weights of pb's material points and re-maps them at once; they
 * are recomputed on every later relocation (gs_relocate, gs_advect_rk2,
 * gs_problem_set_coords).  res: "resolutionX/Y/Z" of DVCWeights (reference
 * default 10), ignored for GS_WEIGHTS_CONSTANT.  Replaces
 * IntegrationPointsSwarm's WeightsCalculator (IntegrationPointsSwarm.cxx:391-449). */
int32_t gs_problem_set_weights(gs_problem* pb, int32_t kind, const int32_t* res);

/* Discrete Voronoi centroids and weights of ncell independent FE cells on the
 * device (host buffers in and out): cell e holds particles off[e]..off[e+1]
 * of xi (dim values each, local coordinates in [-1,1]); centroid gets dim
 * values per particle, weight one.  _DVCWeights_CreateVoronoi(2D) +
 * _DVCWeights_GetCentroids(2D) (DVCWeights.cxx:202-258, 693-786) per cell. */
int32_t gs_dvc_cells(int32_t dim, const int32_t* res, int64_t ncell, const int32_t* off, const double* xi,
                     double* centroid, double* weight);

/* One PCDVC pass (PCDVC.cxx:656-1090 in 3D, 1092-1512 in 2D, Inflow
 * included) followed by the DVC weights it leaves.  Replaces the pass that
 * WeightsCalculator_CalculateAll runs over the cells in order
 * (WeightsCalculator.cxx:195-197) with PCDVC as the calculator.
 *
 * The reference's swarm is a global particle array plus one table per cell
 * (built in index order by the mapper); a split appends a copy of the parent
 * to both, a deletion moves the cell's last table entry into the freed slot
 * (SwarmClass.cxx:976-990) and the array's last particle into the freed
 * index (PCDVC.cxx:339-545).  Per cell, on the discrete Voronoi regions of
 * its particles in table order:
 *   Inflow (inflow != 0): when the last particle's squared distance to its
 *   centroid exceeds 2 cent_pos_ratio, or the count over particles_per_cell
 *   is below threshold, particles_per_cell particles are added at the centres
 *   of random sub-cells (copies of those sub-cells' owners) and, when the
 *   original count is above 5, that many are deleted at random table slots;
 *   the regions are redone.
 *   Then the first max_splits particles whose region exceeds upper_t percent
 *   of the cell volume split (at the region's centroid; with Inflow at the
 *   centre of a random sub-cell of the region) and, of those below lower_t
 *   percent, the ones with the highest indices are deleted up to
 *   max_deletions; a cell holding more than one material splits and deletes
 *   nothing unless split_in_interface_cells / delete_in_interface_cells;
 *   max_deletions is cut to nump / 2 (nump - 4 in 2D with Inflow) when it
 *   exceeds nump - 4.  The regions are redone after any change.
 * Afterwards the caller's particle order is the reference's array order and
 * the weights are those of the cells' final tables (until the next
 * relocation, whose tables list particles in index order again).
 * rand() is the C library's (glibc) stream after srand(rand_seed), entered
 * after rand_skip calls; a pass that continues the previous pass's stream is
 * resumed without replaying it.
 * counts (may be NULL) = {particles added, particles deleted, rand() calls,
 * re-populated cells, device launches over re-populated cells}.
 * GS_E_INVALID for a cell left with fewer than 3 particles after a change
 * (the reference's firewall), for a cell too large for the sub-grid's LDS
 * budget, and with Inflow for sub-grids whose region lists
 * ((int)(w / da' + 0.5) entries, da' = (int)(1 / da + 0.5)) could exceed the
 * owned count.  Reference defaults: upperT 25, lowerT 0.6, maxSplits 3,
 * maxDeletions 2, Inflow off, Threshold 0.8, CentPosRatio 0.01,
 * particlesPerCell 25, resolution 10. */
typedef struct gs_pcdvc_params {
  double upper_t;
  double lower_t;
  int32_t max_splits;
  int32_t max_deletions;
  int32_t split_in_interface_cells;
  int32_t delete_in_interface_cells;
  int32_t res[3];
  int32_t inflow;
  int32_t particles_per_cell;
  int32_t pad_;
  double threshold;
  double cent_pos_ratio;
  int64_t rand_seed;
  int64_t rand_skip;
} gs_pcdvc_params;
int32_t gs_problem_population_control(gs_problem* pb, const gs_pcdvc_params* p, int64_t* counts);

/* Largest particle count per cell that gs_dvc_cells and GS_WEIGHTS_DVC accept
 * for this sub-grid (0 for an invalid resolution). */
int64_t gs_dvc_max_particles(int32_t dim, const int32_t* res);

#ifdef __cplusplus
}
#endif

#endif
