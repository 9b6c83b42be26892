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

// Optional per-particle double arrays that follow the particles (per-particle
// eta0, StrainWeakening plastic strain and its rate); null entries are skipped.
constexpr int kPartDoubles = 3;
struct PartDoubles {
  double* a[kPartDoubles];
};

// The sub-grid of one cell's discrete Voronoi diagram.
struct DvcGrid {
  int dim;
  int r[3];
  int nc, nwords;
  double h[3];
  double vol;
};

// PCDVC population control on the device (gs_weights.hip).  One workgroup
// per cell runs the pass up to the deletions whose order depends on the
// swarm's global indices; the host decides that order and k_pc_final
// finishes the cells.  Staging area of cell e: soff[e] .. soff[e+1], one
// slot per table entry it can hold.
struct PcParams {
  DvcGrid G;
  double max_w, min_w, thresh, cpr;
  int max_splits, max_deletions, split_if, delete_if, inflow, ppc, maxcap, one_oda, probe;
};
struct PcData {
  const int* off;
  const int* pid;
  const int* mat;
  const double4* pk;        // local coordinates of the existing particles
  const long long* soff;    // staging offset of each cell
  int* tbl;                 // table of each cell (references)
  int* tcnt;                // its length
  int* rsrc;                // per record: source particle (sorted index)
  double* rxi;              // per record: local coordinate (3)
  int* nins;                // Inflow insertions
  int* nspl;                // splits
  int* ninfdel;             // Inflow deletions (references in infdel)
  int* ncand;               // deletion candidates (references in cand)
  int* cntdel;              // how many candidates go
  int* ndraw;               // rand() calls of the cell
  int* flags;               // kPc*
  int* infdel;
  int* cand;
  double* scent;            // final centroid per table slot (3)
  double* sw;               // final weight per table slot
  const int* rs;            // rand() values
  long long rs_n;
  int* ndraw_list;          // optional: rand() calls per listed cell (block order)
};
constexpr int kPcTrig = 1, kPcFinal = 2, kPcRedo = 4, kPcErr = 8;

bool pc_grid(int dim, const int* res, PcParams* A);
size_t pc_lds_bytes(const PcParams& A);
long long pc_max_capacity(const PcParams& A);  // largest maxcap whose LDS fits
void pc_cells(const PcParams& A, const PcData& D, long long ncells, const int* cells, const long long* roff,
              hipStream_t st);
void pc_final(const PcParams& A, const PcData& D, long long ncells, const int* cells, const int* dorder,
              const long long* doff, hipStream_t st);
void pc_gather(const PcData& D, long long ncells, const int* cells, const long long* goff, int* out, hipStream_t st);
void pc_rebuild(const Geo& g, const PcData& D, const int* off2, const int* fidx, const int* nbase,
                const int* fidx_new, const double* coord, const int* mat, PartDoubles x, double* coordN, int* matN,
                int* pidN, PartDoubles xN, double4* pkN, double* wN, unsigned* keysN, hipStream_t st);

}  // namespace gs
