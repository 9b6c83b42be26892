/*
 * gale_oracle.c -- CPU restatement of Gale's Q1-P0 Stokes hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see gale_oracle.h).  This is the checker that the
 * HIP implementation in gale_amd/ is compared against.  It follows the
 * reference algorithm step by step, assembling K, G and Q element by element
 * into CSR exactly as the reference's PETSc path does, and solving with the
 * reference's Uzawa loop.
 *
 * What pins each part (the tests under tests/ that check this file):
 *   - Park-Miller stream: the textbook known answer (seed 1 -> 10000th value
 *     1043618065, Park & Miller 1988) through the reference's seed mask.
 *   - K element integration + global assembly: the reference's own PETSc
 *     binary fixtures testIsoviscousSiffness.matrix (3x3 Q1) and
 *     testMaterialViscosity2D.matrix (10x10 Q1), evaluated at 2x2 Gauss
 *     points; plus rigid-body nullity and symmetry in 3D.
 *   - D-matrix / stress conventions: testConstitutiveMatrix_Cartesian2D.expected.
 *   - G: the divergence-theorem closed form (G^T u)_e = -(flux of u out of e).
 *   - Uzawa solve: a dense direct solve of the saddle-point system (numpy),
 *     and the hydrostatic closed form (u = 0, p = rho g (H - y) + c).
 *   - Drucker-Prager / von Mises viscosity: closed-form pure-shear cases.
 *   - Ownership: brute force over face/edge/corner lattices and +-1e-12 /
 *     +-1 ulp offsets.
 *   - RK2: rigid rotation closed form and the linear-field exactness of the
 *     midpoint rule.
 *   - DVC weights: the reference's DVCWeightsSuite fixtures (centroids and
 *     weights of 10 hard-coded particles, 2D and 3D) and its PCDVC element
 *     integration statistics (testPCDVC_*.expected).
 *   - Sobol sequence: the reference's testSobolGeneratorOutput.*.expected
 *     fixtures for all 100 tabulated generators, its own GetNextNumber /
 *     GetNumberByIndex consistency check, and the (0,1)-sequence property
 *     (every dyadic interval of length 2^-k holds one point of each run of
 *     2^k); SpaceFiller layout: testSpaceFillerParticleLayoutOutput.expected.
 *   - StrainWeakening: with no healing the increment of a yielding point is
 *     its plastic strain rate 2 e_II - tau_y / eta (linear velocity fields,
 *     exact strain rates); healing alone gives -(tau_y / eta) h w; a problem
 *     at or past the final softening strain solves bit-identically to the
 *     after-softening material (tests/test_weakening_oracle.py).  The
 *     initial damage calls the C library's srand/rand itself (the generator
 *     the reference uses), closed form at fraction 1.
 * Parity unpinned: the iterate counts of the GMRES(30)/ILU(0) inner solve
 * (PETSc's internals are not in the reference tree); the solution it returns
 * is pinned by the direct solve (tests/test_oracle_pins.py).
 */
#include "gale_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXNEN 8
#define MAXDOF 24

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Park-Miller minimal standard with Schrage's trick.                       */
/* StgDomain/Swarm/src/Random.cxx:45-74                                      */
/* ------------------------------------------------------------------------ */
static const long PM_M = 2147483647L, PM_A = 16807L, PM_Q = 127773L, PM_R = 2836L;
static const long PM_MASK = 123459876L;

void orc_rng_seed(orc_rng* g, long seed) { g->state = seed ^ PM_MASK; }

double orc_rng_next(orc_rng* g) {
  const long tmp = g->state / PM_Q;
  g->state = PM_A * (g->state - tmp * PM_Q) - PM_R * tmp;
  if (g->state < 0) g->state += PM_M;
  return (double)g->state;
}

/* Random.cxx:76-78 Swarm_Random_Random_WithMinMax */
double orc_rng_range(orc_rng* g, double mn, double mx) {
  return (mx - mn) * orc_rng_next(g) / (double)PM_M + mn;
}

/* ------------------------------------------------------------------------ */
/* Mesh: CartesianGenerator vertex coordinates (CartesianGenerator.cxx:     */
/* 2174-2215, contactDepth = 0) and regular element search                   */
/* (Mesh_RegularAlgorithms.cxx:155-224, Num_Approx epsilon 1e-12 from        */
/* StGermain/Base/Foundation/src/Numerics.cxx:36-41).                        */
/* ------------------------------------------------------------------------ */
double orc_vertex_coord(int32_t n, double lo, double hi, int32_t idx) {
  if (idx <= 0) return lo;
  if (idx >= n) return hi;
  return lo + ((double)idx / (double)n) * (hi - lo);
}

static int num_approx(double v, double val) { return v >= val - 1e-12 && v <= val + 1e-12; }

int32_t orc_search_element(int32_t dim, const int32_t* res, const double* lo, const double* hi,
                           const double* x, int32_t* elInd) {
  int32_t inds[3] = {0, 0, 0};
  for (int d = 0; d < dim; d++) {
    const double sep = (hi[d] - lo[d]) / res[d];
    if (num_approx(x[d] - hi[d], 0.0)) {
      inds[d] = res[d] - 1;
    } else if (x[d] < lo[d] || x[d] > hi[d]) {
      return 0;
    } else {
      double integer;
      const double out = (x[d] - lo[d]) / sep;
      const double frac = modf(out, &integer);
      inds[d] = (int32_t)(unsigned)integer;
      if (inds[d] > 0 && num_approx(frac, 0.0)) inds[d]--;
    }
  }
  /* Grid_Project: x fastest */
  *elInd = inds[0] + res[0] * (inds[1] + (dim == 3 ? res[1] * inds[2] : 0));
  return 1;
}

/* RandomParticleLayout.cxx:196-216 driven cell by cell as in
 * PerCellParticleLayout.cxx:147-168.  Three draws per particle, also in 2D
 * (the third coordinate is discarded here). */
int64_t orc_layout_random(int32_t dim, const int32_t* res, const double* lo, const double* hi,
                          int32_t ppc, long seed, double* coord) {
  orc_rng g;
  orc_rng_seed(&g, seed);
  const int64_t E = (int64_t)res[0] * res[1] * (dim == 3 ? res[2] : 1);
  int64_t p = 0;
  for (int64_t e = 0; e < E; e++) {
    int32_t ijk[3] = {(int32_t)(e % res[0]), (int32_t)((e / res[0]) % res[1]),
                      dim == 3 ? (int32_t)(e / ((int64_t)res[0] * res[1])) : 0};
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
      mn[d] = orc_vertex_coord(res[d], lo[d], hi[d], ijk[d]);
      mx[d] = orc_vertex_coord(res[d], lo[d], hi[d], ijk[d] + 1);
    }
    for (int c = 0; c < ppc; c++, p++) {
      for (int d = 0; d < 3; d++) {
        double v = orc_rng_range(&g, mn[d], mx[d]);
        coord[p * 3 + d] = d < dim ? v : 0.0;
      }
    }
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* Sobol sequence and the SpaceFiller layout                                 */
/* ------------------------------------------------------------------------ */
/* SobolGenerator.cxx:87-112: the first `degree` direction numbers are given;
 * the running number starts at 0 (the reference assigns 0.5 to an unsigned). */
void orc_sobol_init(orc_sobol* s, unsigned degree, unsigned coefficient, const unsigned* m) {
  memset(s, 0, sizeof(*s));
  s->degree = degree;
  s->coefficient = coefficient;
  for (unsigned i = 0; i < degree; i++) {
    s->v[i] = m[i];
    s->have[i] = 1;
  }
}

/* SobolGenerator.cxx:198-244, the direction-number recurrence
 * v_i = 2 a_1 v_{i-1} ^ 4 a_2 v_{i-2} ^ ... ^ 2^s v_{i-s} ^ v_{i-s},
 * a_k the bit degree-1-k of the coefficient, evaluated recursively. */
unsigned orc_sobol_direction(orc_sobol* s, unsigned i) {
  if (s->have[i]) return s->v[i];
  unsigned r = 0, last = 0;
  for (unsigned k = 0; k < s->degree; k++) {
    const int on = k == s->degree - 1 || ((s->coefficient >> (s->degree - 2 - k)) & 1u);
    if (on) {
      last = orc_sobol_direction(s, i - 1 - k);
      r ^= (2u << k) * last;
    }
  }
  r ^= last;
  s->have[i] = 1;
  return s->v[i] = r;
}

/* SobolGenerator.cxx:184-196 and :246-269: the Gray codes of n and n+1 differ
 * in the lowest zero bit c of n, so x_{n+1} = x_n ^ (v_c << (29 - c)). */
double orc_sobol_next(orc_sobol* s) {
  unsigned c = 0, n = s->count;
  while (n & 1u) {
    n >>= 1;
    c++;
  }
  s->count++;
  s->x ^= orc_sobol_direction(s, c) << (30 - c - 1);
  return (double)s->x / (double)(1L << 30);
}

/* SobolGenerator.cxx:271-294: x(index) = XOR over the set bits b of the Gray
 * code of index + 1 of v_b << (29 - b). */
double orc_sobol_by_index(orc_sobol* s, unsigned index) {
  const unsigned n = index + 1, g = n ^ (n >> 1);
  unsigned r = 0;
  for (unsigned b = 0; b < 30; b++)
    if (g & (1u << b)) r ^= orc_sobol_direction(s, b) << (30 - b - 1);
  return (double)r / (double)(1L << 30);
}

int32_t orc_sobol_sequence(unsigned degree, unsigned coefficient, const unsigned* m, int64_t count, double* next,
                           double* by_index) {
  if (degree < 1 || degree > 30 || count < 0 || count >= (1L << 30)) return -1;
  orc_sobol s;
  orc_sobol_init(&s, degree, coefficient, m);
  for (int64_t k = 0; k < count; k++) {
    const double a = orc_sobol_next(&s);
    if (next) next[k] = a;
    if (by_index) by_index[k] = orc_sobol_by_index(&s, (unsigned)k);
  }
  return 0;
}

/* SpaceFillerParticleLayout.cxx:211-240 through GlobalParticleLayout.cxx
 * (_SetInitialCounts: total = cells * average, truncated to an unsigned;
 * _InitialiseParticles: particles in generation order): coordinate d of
 * particle p is (max - min) * next(generator d) + min over the mesh's box. */
int64_t orc_layout_spacefiller(int32_t dim, const int32_t* res, const double* lo, const double* hi, int64_t total,
                               double avg_per_cell, const unsigned* deg, const unsigned* coef, const unsigned* m,
                               double* coord) {
  const int64_t E = (int64_t)res[0] * res[1] * (dim == 3 ? res[2] : 1);
  const int64_t P = total > 0 ? total : (int64_t)(unsigned)(E * avg_per_cell);
  orc_sobol s[3];
  for (int d = 0; d < dim; d++) orc_sobol_init(&s[d], deg[d], coef[d], m + 9 * d);
  for (int64_t p = 0; p < P; p++) {
    coord[3 * p + 2] = 0.0;
    for (int d = 0; d < dim; d++) {
      const double u = orc_sobol_next(&s[d]);
      coord[3 * p + d] = (hi[d] - lo[d]) * u + lo[d];
    }
  }
  return P;
}

/* ------------------------------------------------------------------------ */
/* Problem                                                                   */
/* ------------------------------------------------------------------------ */
struct orc_problem {
  int dim, nen, ndof;
  int32_t res[3];
  int64_t nn[3];
  double lo[3], hi[3];
  int64_t N, E, P;
  uint8_t* bc_mask;
  double* bc_value;
  double gravity, ghat[3];
  double clamp_min, clamp_max;
  int nmat;
  orc_material* mats;
  double* coord;
  int32_t* material;
  double* eta0p;
  double* pstr;                 /* StrainWeakening plastic strain (caller order), NULL when off */
  double* pinc;                 /* its rate from the last constitutive evaluation */
  orc_strain_weakening* sw;     /* per material */
  int* sw_on;
  int32_t* owner;
  double* xi;
  double* weight;
  int64_t* cell_off;
  int64_t* cell_list;
  double* eta;
  uint8_t* yielded;
  int has_visc;
  int wkind;          /* 0: ConstantWeights, 1: discrete Voronoi cell weights */
  int32_t wres[3];    /* DVC sub-grid resolution */
  /* reduced system */
  int32_t* eq;
  int64_t n_u;
  int64_t* Kptr;
  int32_t* Kcol;
  double* Kval;
  int64_t* Gptr;
  int32_t* Gcol;
  double* Gval;
  double* f;
  double* h;
  double* Q;
  double* Kdiag;
  double* ilu; /* ILU(0) factor values on the K pattern */
  int64_t* Kdiagpos;
  int assembled;
};

int64_t orc_num_nodes(const orc_problem* pb) { return pb->N; }
int64_t orc_num_elements(const orc_problem* pb) { return pb->E; }
int64_t orc_num_particles(const orc_problem* pb) { return pb->P; }

static void elem_ijk(const orc_problem* pb, int64_t e, int32_t* ijk) {
  ijk[0] = (int32_t)(e % pb->res[0]);
  ijk[1] = (int32_t)((e / pb->res[0]) % pb->res[1]);
  ijk[2] = pb->dim == 3 ? (int32_t)(e / ((int64_t)pb->res[0] * pb->res[1])) : 0;
}

/* CartesianGenerator.cxx:894-955: element vertices in lexicographic order,
 * x fastest. */
static void elem_nodes(const orc_problem* pb, int64_t e, int64_t* nodes) {
  int32_t ijk[3];
  elem_ijk(pb, e, ijk);
  for (int a = 0; a < pb->nen; a++) {
    const int64_t i = ijk[0] + (a & 1), j = ijk[1] + ((a >> 1) & 1), k = ijk[2] + ((a >> 2) & 1);
    nodes[a] = i + pb->nn[0] * (j + pb->nn[1] * k);
  }
}

static void node_coord(const orc_problem* pb, int64_t n, double* x) {
  const int64_t i = n % pb->nn[0], j = (n / pb->nn[0]) % pb->nn[1], k = n / (pb->nn[0] * pb->nn[1]);
  const int64_t ijk[3] = {i, j, k};
  for (int d = 0; d < pb->dim; d++) x[d] = orc_vertex_coord(pb->res[d], pb->lo[d], pb->hi[d], (int32_t)ijk[d]);
}

/* BilinearElementType.cxx:213-243, TrilinearElementType.cxx:250-300:
 * N_a = c * prod_d (1 + s_ad xi_d), local node a lexicographic. */
static void shape_funcs(int dim, const double* xi, double* N) {
  if (dim == 2) {
    N[0] = 0.25 * (1.0 - xi[0]) * (1.0 - xi[1]);
    N[1] = 0.25 * (1.0 + xi[0]) * (1.0 - xi[1]);
    N[2] = 0.25 * (1.0 - xi[0]) * (1.0 + xi[1]);
    N[3] = 0.25 * (1.0 + xi[0]) * (1.0 + xi[1]);
  } else {
    for (int a = 0; a < 8; a++) {
      const double sx = (a & 1) ? 1.0 : -1.0, sy = (a & 2) ? 1.0 : -1.0, sz = (a & 4) ? 1.0 : -1.0;
      N[a] = 0.125 * (1.0 + sx * xi[0]) * (1.0 + sy * xi[1]) * (1.0 + sz * xi[2]);
    }
  }
}

static void shape_local_derivs(int dim, const double* xi, double GNi[3][MAXNEN]) {
  const int nen = dim == 2 ? 4 : 8;
  const double c = dim == 2 ? 0.25 : 0.125;
  for (int a = 0; a < nen; a++) {
    double s[3], f[3];
    for (int d = 0; d < dim; d++) {
      s[d] = ((a >> d) & 1) ? 1.0 : -1.0;
      f[d] = 1.0 + s[d] * xi[d];
    }
    for (int d = 0; d < dim; d++) {
      double v = c * s[d];
      for (int d2 = 0; d2 < dim; d2++)
        if (d2 != d) v *= f[d2];
      GNi[d][a] = v;
    }
  }
}

/* ElementType.cxx:339-474 ElementType_ShapeFunctionsGlobalDerivs */
static double global_derivs(const orc_problem* pb, int64_t e, const double* xi, double GNx[3][MAXNEN]) {
  const int dim = pb->dim, nen = pb->nen;
  int64_t nodes[MAXNEN];
  double GNi[3][MAXNEN], jac[3][3] = {{0}}, inv[3][3], D;
  elem_nodes(pb, e, nodes);
  shape_local_derivs(dim, xi, GNi);
  for (int n = 0; n < nen; n++) {
    double x[3];
    node_coord(pb, nodes[n], x);
    for (int i = 0; i < dim; i++)
      for (int j = 0; j < dim; j++) jac[i][j] += GNi[i][n] * x[j];
  }
  if (dim == 2) {
    D = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
    inv[0][0] = jac[1][1] / D;
    inv[1][1] = jac[0][0] / D;
    inv[0][1] = -jac[0][1] / D;
    inv[1][0] = -jac[1][0] / D;
  } else {
    D = jac[0][0] * (jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1]) -
        jac[0][1] * (jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]) +
        jac[0][2] * (jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0]);
    double cof[3][3];
    cof[0][0] = jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1];
    cof[1][0] = -(jac[1][0] * jac[2][2] - jac[1][2] * jac[2][0]);
    cof[2][0] = jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0];
    cof[0][1] = -(jac[0][1] * jac[2][2] - jac[0][2] * jac[2][1]);
    cof[1][1] = jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0];
    cof[2][1] = -(jac[0][0] * jac[2][1] - jac[0][1] * jac[2][0]);
    cof[0][2] = jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1];
    cof[1][2] = -(jac[0][0] * jac[1][2] - jac[0][2] * jac[1][0]);
    cof[2][2] = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) inv[i][j] = cof[i][j] / D;
  }
  for (int dx = 0; dx < dim; dx++)
    for (int n = 0; n < nen; n++) {
      double v = 0.0;
      for (int dxi = 0; dxi < dim; dxi++) v += GNi[dxi][n] * inv[dx][dxi];
      GNx[dx][n] = v;
    }
  return D;
}

static void solve_small(int dim, double J[3][3], const double* b, double* x) {
  if (dim == 2) {
    const double D = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    x[0] = (b[0] * J[1][1] - J[0][1] * b[1]) / D;
    x[1] = (J[0][0] * b[1] - b[0] * J[1][0]) / D;
  } else {
    const double D = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    for (int c = 0; c < 3; c++) {
      double M[3][3];
      memcpy(M, J, sizeof(M));
      for (int r = 0; r < 3; r++) M[r][c] = b[r];
      x[c] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
              M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
              M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) / D;
    }
  }
}

/* ElementType.cxx:221-333 Newton-Raphson global -> local, initial guess 0,
 * stop when max |dxi| < 1e-4 or after 100 iterations. */
static void global_to_local(const orc_problem* pb, int64_t e, const double* x, double* xi) {
  const int dim = pb->dim, nen = pb->nen;
  int64_t nodes[MAXNEN];
  double xn[MAXNEN][3];
  elem_nodes(pb, e, nodes);
  for (int n = 0; n < nen; n++) node_coord(pb, nodes[n], xn[n]);
  for (int d = 0; d < dim; d++) xi[d] = 0.0;
  for (int it = 0; it < 100; it++) {
    double N[MAXNEN], GNi[3][MAXNEN], J[3][3] = {{0}}, rhs[3] = {0, 0, 0}, inc[3];
    shape_funcs(dim, xi, N);
    shape_local_derivs(dim, xi, GNi);
    for (int n = 0; n < nen; n++)
      for (int i = 0; i < dim; i++) {
        for (int j = 0; j < dim; j++) J[i][j] += GNi[j][n] * xn[n][i];
        rhs[i] -= N[n] * xn[n][i];
      }
    for (int i = 0; i < dim; i++) rhs[i] += x[i];
    solve_small(dim, J, rhs, inc);
    double mr = 0.0;
    for (int d = 0; d < dim; d++) {
      xi[d] += inc[d];
      if (fabs(inc[d]) > mr) mr = fabs(inc[d]);
    }
    if (mr < 1e-4) break;
  }
}

/* ------------------------------------------------------------------------ */
/* Discrete Voronoi cell weights                                             */
/*   PICellerator/Weights/src/DVCWeights.cxx:202-786, 875-951                */
/* ------------------------------------------------------------------------ */
/* One FE cell: a fixed sub-grid of res[0] x res[1] (x res[2]) sub-cells over
 * the local box [-1,1]^dim, sub-cell k = i + res0 (j + res1 l).  Sub-cell
 * centres accumulate one step at a time from (-1 - h/2) as
 * _DVCWeights_ConstructGrid(2D) does (:525-619).  Regions grow as
 * _DVCWeights_CreateVoronoi(2D) (:693-786): every particle, in order, seeds
 * the sub-cell it lies in (a later particle in the same sub-cell takes it
 * over) and its boundary becomes the face neighbours of that sub-cell it does
 * not own (UpdateBchain, :420-523).  Then, round after round until a round
 * claims nothing, every particle in order contests its boundary (ClaimCells,
 * :260-397): an unowned sub-cell is taken, an owned one is taken when the
 * contest (x1 + x2 - 2 x0).(x1 - x2) > 0, i.e. the claimant is strictly
 * closer to the centre x0 than the owner; the new boundary is the face
 * neighbours of the sub-cells it just took that it does not own.  Centroid
 * and weight as _DVCWeights_GetCentroids(2D) (:202-258): mean of the owned
 * centres in sub-cell order, weight = count x sub-cell volume (a particle
 * left with no sub-cell keeps centroid 0 and weight 0). */
static double dvc_contest(int dim, const double* c, const double* o, const double* q) {
  double s = 0.0;
  for (int d = 0; d < dim; d++) s += (o[d] + q[d] - c[d] - c[d]) * (o[d] - q[d]);
  return s;
}

/* owner_out (may be NULL) receives the owning particle of every sub-cell. */
static int32_t dvc_cell_impl(int32_t dim, const int32_t* res, int32_t nump, const double* xi, double* cent, double* w,
                             int32_t* owner_out) {
  if (dim != 2 && dim != 3) return ORC_E_INVALID;
  int r[3] = {res[0], res[1], dim == 3 ? res[2] : 1};
  for (int d = 0; d < dim; d++)
    if (r[d] < 1) return ORC_E_INVALID;
  const int nc = r[0] * r[1] * r[2];
  double h[3], *ctr[3];
  for (int d = 0; d < 3; d++) {
    h[d] = d < dim ? 2.0 / r[d] : 0.0;
    ctr[d] = (double*)malloc(sizeof(double) * r[d]);
    double x = -1.0 - h[d] / 2.0;
    for (int i = 0; i < r[d]; i++) {
      x = x + h[d];
      ctr[d][i] = x;
    }
  }
  const double vol = dim == 2 ? h[0] * h[1] : h[0] * h[1] * h[2];
  int* owner = (int*)malloc(sizeof(int) * nc);
  unsigned char* mark = (unsigned char*)calloc(nc, 1);
  int** bound = (int**)malloc(sizeof(int*) * (nump > 0 ? nump : 1));
  int* nbound = (int*)calloc(nump > 0 ? nump : 1, sizeof(int));
  int* took = (int*)malloc(sizeof(int) * nc);
  for (int c = 0; c < nc; c++) owner[c] = -1;
  for (int i = 0; i < nump; i++) bound[i] = (int*)malloc(sizeof(int) * nc);
#define DVC_CENTRE(cc, out)                                         \
  do {                                                              \
    const int ci_ = (cc) % r[0], cj_ = ((cc) / r[0]) % r[1];        \
    const int ck_ = (cc) / (r[0] * r[1]);                           \
    (out)[0] = ctr[0][ci_];                                         \
    (out)[1] = ctr[1][cj_];                                         \
    (out)[2] = dim == 3 ? ctr[2][ck_] : 0.0;                        \
  } while (0)
  /* boundary of particle i from the sub-cells it just took */
#define DVC_REBOUND(i, list, nlist)                                                        \
  do {                                                                                     \
    int nb_ = 0;                                                                           \
    for (int t_ = 0; t_ < (nlist); t_++) {                                                 \
      const int c_ = (list)[t_];                                                           \
      const int ci_ = c_ % r[0], cj_ = (c_ / r[0]) % r[1], ck_ = c_ / (r[0] * r[1]);       \
      int nbr_[6], nn_ = 0;                                                                \
      if (cj_ > 0) nbr_[nn_++] = c_ - r[0];                                                \
      if (cj_ < r[1] - 1) nbr_[nn_++] = c_ + r[0];                                         \
      if (ci_ < r[0] - 1) nbr_[nn_++] = c_ + 1;                                            \
      if (ci_ > 0) nbr_[nn_++] = c_ - 1;                                                   \
      if (dim == 3 && ck_ < r[2] - 1) nbr_[nn_++] = c_ + r[0] * r[1];                      \
      if (dim == 3 && ck_ > 0) nbr_[nn_++] = c_ - r[0] * r[1];                             \
      for (int k_ = 0; k_ < nn_; k_++) {                                                   \
        const int m_ = nbr_[k_];                                                           \
        if (owner[m_] != (i) && !mark[m_]) {                                               \
          mark[m_] = 1;                                                                    \
          bound[i][nb_++] = m_;                                                            \
        }                                                                                  \
      }                                                                                    \
    }                                                                                      \
    for (int t_ = 0; t_ < nb_; t_++) mark[bound[i][t_]] = 0;                               \
    nbound[i] = nb_;                                                                       \
  } while (0)
  for (int i = 0; i < nump; i++) {
    int idx[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
      int k = (int)((xi[i * dim + d] - (-1.0)) / h[d]);
      if (k == r[d]) k--;
      idx[d] = k;
    }
    const int c = idx[0] + r[0] * (idx[1] + r[1] * idx[2]);
    owner[c] = i;
    took[0] = c;
    DVC_REBOUND(i, took, 1);
  }
  int claimed = 1;
  while (claimed) {
    claimed = 0;
    for (int i = 0; i < nump; i++) {
      int nt = 0;
      for (int t = 0; t < nbound[i]; t++) {
        const int c = bound[i][t];
        const int o = owner[c];
        int take = 0;
        if (o == -1) {
          take = 1;
        } else if (o != i) {
          double x0[3], x1[3], x2[3];
          DVC_CENTRE(c, x0);
          for (int d = 0; d < 3; d++) {
            x1[d] = d < dim ? xi[o * dim + d] : 0.0;
            x2[d] = d < dim ? xi[i * dim + d] : 0.0;
          }
          take = dvc_contest(dim, x0, x1, x2) > 0.0;
        }
        if (take) {
          owner[c] = i;
          took[nt++] = c;
        }
      }
      claimed += nt;
      DVC_REBOUND(i, took, nt);
    }
  }
  for (int i = 0; i < nump; i++) {
    double sc[3] = {0.0, 0.0, 0.0};
    int cnt = 0;
    for (int c = 0; c < nc; c++)
      if (owner[c] == i) {
        double x0[3];
        DVC_CENTRE(c, x0);
        for (int d = 0; d < dim; d++) sc[d] += x0[d];
        cnt++;
      }
    w[i] = cnt * vol;
    for (int d = 0; d < dim; d++) cent[i * dim + d] = cnt ? sc[d] / cnt : 0.0;
  }
#undef DVC_CENTRE
#undef DVC_REBOUND
  if (owner_out)
    for (int c = 0; c < nc; c++) owner_out[c] = owner[c];
  for (int i = 0; i < nump; i++) free(bound[i]);
  free(bound); free(nbound); free(took); free(owner); free(mark);
  for (int d = 0; d < 3; d++) free(ctr[d]);
  return ORC_OK;
}

int32_t orc_dvc_cell(int32_t dim, const int32_t* res, int32_t nump, const double* xi, double* cent, double* w) {
  return dvc_cell_impl(dim, res, nump, xi, cent, w, NULL);
}

int32_t orc_dvc_cell_owners(int32_t dim, const int32_t* res, int32_t nump, const double* xi, double* cent, double* w,
                            int32_t* owner) {
  return dvc_cell_impl(dim, res, nump, xi, cent, w, owner);
}

/* Stable counting sort of particles by owning element: the cell tables that
 * CoincidentMapper builds by visiting material points in index order
 * (CoincidentMapper.cxx:207-240), weights from ConstantWeights.cxx:178-186
 * with cellLocalVolume = 2^dim (WeightsCalculator.cxx:84). */
static int relocate_all(orc_problem* pb) {
  const int dim = pb->dim;
  for (int64_t p = 0; p < pb->P; p++) {
    int32_t el;
    if (!orc_search_element(dim, pb->res, pb->lo, pb->hi, pb->coord + 3 * p, &el)) return ORC_E_ESCAPED;
    pb->owner[p] = el;
  }
  memset(pb->cell_off, 0, sizeof(int64_t) * (pb->E + 1));
  for (int64_t p = 0; p < pb->P; p++) pb->cell_off[pb->owner[p] + 1]++;
  for (int64_t e = 0; e < pb->E; e++) pb->cell_off[e + 1] += pb->cell_off[e];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * pb->E);
  memcpy(fill, pb->cell_off, sizeof(int64_t) * pb->E);
  for (int64_t p = 0; p < pb->P; p++) pb->cell_list[fill[pb->owner[p]]++] = p;
  free(fill);
  const double vol = dim == 2 ? 4.0 : 8.0;
  int64_t maxc = 0;
  for (int64_t e = 0; e < pb->E; e++) {
    const int64_t cnt = pb->cell_off[e + 1] - pb->cell_off[e];
    if (cnt > maxc) maxc = cnt;
    for (int64_t q = pb->cell_off[e]; q < pb->cell_off[e + 1]; q++) {
      const int64_t p = pb->cell_list[q];
      pb->weight[p] = vol / (double)cnt;
      global_to_local(pb, e, pb->coord + 3 * p, pb->xi + dim * p);
    }
  }
  if (pb->wkind == 1) {
    /* DVCWeights: integration point at the Voronoi centroid, weight = its
     * volume (DVCWeights.cxx:875-951), particles in cell-table order */
    double* xs = (double*)malloc(sizeof(double) * dim * (maxc > 0 ? maxc : 1));
    double* cs = (double*)malloc(sizeof(double) * dim * (maxc > 0 ? maxc : 1));
    double* ws = (double*)malloc(sizeof(double) * (maxc > 0 ? maxc : 1));
    for (int64_t e = 0; e < pb->E; e++) {
      const int64_t b = pb->cell_off[e], n = pb->cell_off[e + 1] - b;
      for (int64_t q = 0; q < n; q++)
        for (int d = 0; d < dim; d++) xs[q * dim + d] = pb->xi[dim * pb->cell_list[b + q] + d];
      orc_dvc_cell(dim, pb->wres, (int32_t)n, xs, cs, ws);
      for (int64_t q = 0; q < n; q++) {
        const int64_t p = pb->cell_list[b + q];
        for (int d = 0; d < dim; d++) pb->xi[dim * p + d] = cs[q * dim + d];
        pb->weight[p] = ws[q];
      }
    }
    free(xs); free(cs); free(ws);
  }
  pb->assembled = 0;
  return ORC_OK;
}

/* PCDVC population control followed by the DVC weights it leaves
 * (PCDVC.cxx:656-1090 in 3D, 1092-1512 in 2D), restated as the reference
 * runs it: one pass over the cells in order (WeightsCalculator.cxx:195-197)
 * on a swarm held as a global particle array (indexOnCPU) plus one table per
 * cell, both built as CoincidentMapper builds them (particles added to their
 * cell in index order, CoincidentMapper.cxx:207-240).
 *
 * Swarm operations (PCDVC.cxx:339-610, SwarmClass.cxx:976-990):
 *  - split: a copy of the parent appended to the global array and to the end
 *    of the cell's table, at local coordinate xi (global position
 *    0.5 ((1 - xi) a + (1 + xi) b) per axis);
 *  - delete: the cell's last table entry takes the freed table slot, and the
 *    global array's last particle takes the freed index (its table entry is
 *    rewritten to that index).
 *
 * Per cell, on the DVC of the particles' local coordinates in table order:
 *  Inflow (when on): the last particle's squared distance to its centroid is
 *  compared with CentPosRatio x 2 (the loop keeps only the last one, :783-787);
 *  when it is larger, or nump / ParticlesPerCell < Threshold, ParticlesPerCell
 *  particles are added, each a split of the owner of a random sub-cell j =
 *  (int)(nc rand() / (RAND_MAX + 1.0)) placed at that sub-cell's centre, then
 *  when the original count is above 5 that many particles are deleted at
 *  random table slots (int)(nump rand() / (RAND_MAX + 1.0)), and the DVC is
 *  redone.  Each particle's list of owned sub-cells is then built as the
 *  reference builds it: (int)(w / da' + 0.5) entries (da' = round(1 / da)),
 *  filled from the end in sub-cell order (:898-918).
 *  Then: maxW = upperT %, minW = lowerT % of the cell volume; a cell with more
 *  than one material splits (deletes) nothing unless splitInInterfaceCells
 *  (deleteInInterfaceCells); maxDeletions above nump - 4 becomes nump / 2
 *  (nump - 4 in 2D with Inflow, :1387); the first maxSplits particles above
 *  maxW split, at their region's centroid, or with Inflow at the centre of a
 *  random entry of their sub-cell list; the candidates below minW, sorted by
 *  index, highest first (compare_indexOnCPU, :611-620), are deleted up to
 *  maxDeletions; the DVC is redone when anything changed.  Every particle of
 *  the cell then moves to its region's centroid with its region's volume as
 *  weight.
 *
 * Afterwards the caller order is the global array's order and the cell
 * tables keep the order the pass left (the weights belong to that order until
 * the next relocation rebuilds the tables).  rand() is the C library's,
 * seeded with srand(rand_seed) and advanced rand_skip times first.
 * counts = {added, deleted, rand() calls}. */
typedef struct {
  int64_t n, cap;
  int64_t* v;
} pc_list;

static void pc_push(pc_list* l, int64_t x) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 8;
    l->v = (int64_t*)realloc(l->v, sizeof(int64_t) * l->cap);
  }
  l->v[l->n++] = x;
}

typedef struct {
  int dim;
  int64_t n, cap;
  double* coord;  /* 3 per particle */
  int32_t* mat;
  double* eta0;   /* optional per-particle values, NULL when absent */
  double* pstr;
  double* pinc;
  double* xi;     /* 3 per particle: the integration point's local coordinate */
  double* w;
  int64_t* cell;
} pc_swarm;

static void pc_reserve(pc_swarm* s, int64_t need) {
  if (need <= s->cap) return;
  int64_t c = s->cap ? s->cap : 16;
  while (c < need) c *= 2;
  s->coord = (double*)realloc(s->coord, sizeof(double) * 3 * c);
  s->mat = (int32_t*)realloc(s->mat, sizeof(int32_t) * c);
  if (s->eta0) s->eta0 = (double*)realloc(s->eta0, sizeof(double) * c);
  if (s->pstr) s->pstr = (double*)realloc(s->pstr, sizeof(double) * c);
  if (s->pinc) s->pinc = (double*)realloc(s->pinc, sizeof(double) * c);
  s->xi = (double*)realloc(s->xi, sizeof(double) * 3 * c);
  s->w = (double*)realloc(s->w, sizeof(double) * c);
  s->cell = (int64_t*)realloc(s->cell, sizeof(int64_t) * c);
  s->cap = c;
}

static void pc_copy(pc_swarm* s, int64_t to, int64_t from) {
  memcpy(s->coord + 3 * to, s->coord + 3 * from, sizeof(double) * 3);
  s->mat[to] = s->mat[from];
  if (s->eta0) s->eta0[to] = s->eta0[from];
  if (s->pstr) {
    s->pstr[to] = s->pstr[from];
    s->pinc[to] = s->pinc[from];
  }
  memcpy(s->xi + 3 * to, s->xi + 3 * from, sizeof(double) * 3);
  s->w[to] = s->w[from];
  s->cell[to] = s->cell[from];
}

/* splitIntParticleByIndexWithinCell (PCDVC.cxx:547-610) */
static void pc_split(const orc_problem* pb, pc_swarm* s, pc_list* T, int64_t e, int64_t slot, const double* xi) {
  pc_reserve(s, s->n + 1);
  const int64_t k = s->n++;
  pc_copy(s, k, T[e].v[slot]);
  int32_t ijk[3];
  elem_ijk(pb, e, ijk);
  for (int d = 0; d < 3; d++) {
    double x = 0.0;
    if (d < s->dim) {
      const double a = orc_vertex_coord(pb->res[d], pb->lo[d], pb->hi[d], ijk[d]);
      const double b = orc_vertex_coord(pb->res[d], pb->lo[d], pb->hi[d], ijk[d] + 1);
      x = 0.5 * ((1.0 - xi[d]) * a + (1.0 + xi[d]) * b);
    }
    s->coord[3 * k + d] = x;
    s->xi[3 * k + d] = d < s->dim ? xi[d] : 0.0;
  }
  pc_push(&T[e], k);
}

/* Removal of table slot `slot` of cell e and of its particle from the global
 * array (Swarm_RemoveParticleFromCell, then the last particle takes the freed
 * index: PCDVC.cxx:363-437). */
static void pc_delete_slot(pc_swarm* s, pc_list* T, int64_t e, int64_t slot) {
  const int64_t gi = T[e].v[slot];
  T[e].v[slot] = T[e].v[T[e].n - 1];
  T[e].n--;
  const int64_t last = s->n - 1;
  if (gi != last) {
    pc_list* tl = &T[s->cell[last]];
    for (int64_t q = 0; q < tl->n; q++)
      if (tl->v[q] == last) {
        tl->v[q] = gi;
        break;
      }
    pc_copy(s, gi, last);
  }
  s->n--;
}

static double pc_draw(int64_t* ndraw) {
  (*ndraw)++;
  return rand() / (RAND_MAX + 1.0);
}

/* the cell's DVC on its table order; owner of every sub-cell into own */
static void pc_dvc(const pc_swarm* s, const pc_list* t, const int32_t* res, double* xs, double* cs, double* ws,
                   int32_t* own) {
  const int dim = s->dim;
  for (int64_t q = 0; q < t->n; q++)
    for (int d = 0; d < dim; d++) xs[q * dim + d] = s->xi[3 * t->v[q] + d];
  dvc_cell_impl(dim, res, (int32_t)t->n, xs, cs, ws, own);
}

int32_t orc_population_control(orc_problem* pb, const orc_pcdvc_params* pp, int64_t* counts) {
  const int dim = pb->dim;
  const int32_t res[3] = {pp->res[0], pp->res[1], dim == 3 ? pp->res[2] : 1};
  for (int d = 0; d < 3; d++)
    if (res[d] < 1) return ORC_E_INVALID;
  if (pp->max_splits < 0 || pp->max_deletions < 0) return ORC_E_INVALID;
  const int inflow = pp->inflow != 0;
  const int64_t ppc = pp->particles_per_cell;
  if (inflow && ppc < 0) return ORC_E_INVALID;
  double upT = pp->upper_t, lowT = pp->lower_t;
  if (upT < lowT) { /* PCDVC.cxx:265-270 */
    lowT = 0.6;
    upT = 25.0;
  }
  const double V = dim == 2 ? 4.0 : 8.0;
  const double maxW = upT * V / 100.0, minW = lowT * V / 100.0;
  const int nc = res[0] * res[1] * res[2];
  /* sub-cell centres and volume, as _DVCWeights_ConstructGrid(2D) builds them */
  double h[3], *ctr[3];
  for (int d = 0; d < 3; d++) {
    h[d] = d < dim ? 2.0 / res[d] : 0.0;
    ctr[d] = (double*)malloc(sizeof(double) * res[d]);
    double x = -1.0 - h[d] / 2.0;
    for (int i = 0; i < res[d]; i++) {
      x = x + h[d];
      ctr[d][i] = x;
    }
  }
  const double da = dim == 2 ? h[0] * h[1] : h[0] * h[1] * h[2];
  const int one_oda = (int)(1.0 / da + 0.5);
  if (inflow) {
    /* a list length above the owned count would read entries the reference
     * never wrote: such sub-grids are refused */
    for (int k = 1; k <= nc; k++)
      if ((int)(k * da * one_oda + 0.5) > k) {
        for (int d = 0; d < 3; d++) free(ctr[d]);
        return ORC_E_INVALID;
      }
  }
  /* the swarm and the cell tables as the mapper leaves them */
  pc_swarm s;
  memset(&s, 0, sizeof s);
  s.dim = dim;
  if (pb->eta0p) s.eta0 = (double*)malloc(sizeof(double));
  if (pb->pstr) {
    s.pstr = (double*)malloc(sizeof(double));
    s.pinc = (double*)malloc(sizeof(double));
  }
  pc_reserve(&s, pb->P > 0 ? pb->P : 1);
  s.n = pb->P;
  pc_list* T = (pc_list*)calloc((size_t)(pb->E > 0 ? pb->E : 1), sizeof(pc_list));
  for (int64_t e = 0; e < pb->E; e++)
    for (int64_t q = pb->cell_off[e]; q < pb->cell_off[e + 1]; q++) {
      const int64_t p = pb->cell_list[q];
      memcpy(s.coord + 3 * p, pb->coord + 3 * p, sizeof(double) * 3);
      s.mat[p] = pb->material[p];
      if (s.eta0) s.eta0[p] = pb->eta0p[p];
      if (s.pstr) {
        s.pstr[p] = pb->pstr[p];
        s.pinc[p] = pb->pinc[p];
      }
      double x[3] = {0.0, 0.0, 0.0};
      global_to_local(pb, e, pb->coord + 3 * p, x);
      memcpy(s.xi + 3 * p, x, sizeof x);
      s.w[p] = 0.0;
      s.cell[p] = e;
      pc_push(&T[e], p);
    }
  for (int64_t e = 0; e < pb->E; e++) /* index order, as the mapper adds them */
    for (int64_t a = 1; a < T[e].n; a++)
      for (int64_t b = a; b > 0 && T[e].v[b - 1] > T[e].v[b]; b--) {
        const int64_t t = T[e].v[b];
        T[e].v[b] = T[e].v[b - 1];
        T[e].v[b - 1] = t;
      }
  srand((unsigned)pp->rand_seed);
  for (int64_t i = 0; i < pp->rand_skip; i++) (void)rand();
  int64_t ndraw = 0, added = 0, deleted = 0;
  int32_t st = ORC_OK;
  int64_t cap = 16;
  double* xs = (double*)malloc(sizeof(double) * dim * cap);
  double* cs = (double*)malloc(sizeof(double) * dim * cap);
  double* ws = (double*)malloc(sizeof(double) * cap);
  int32_t* own = (int32_t*)malloc(sizeof(int32_t) * nc);
  int32_t* lcount = (int32_t*)malloc(sizeof(int32_t) * cap);
  int32_t* lsize = (int32_t*)malloc(sizeof(int32_t) * cap);
  int32_t* lists = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap * nc);
  int64_t* dcand = (int64_t*)malloc(sizeof(int64_t) * cap);
#define PC_FIT(nn)                                                              \
  do {                                                                          \
    if ((nn) > cap) {                                                           \
      while (cap < (nn)) cap *= 2;                                              \
      xs = (double*)realloc(xs, sizeof(double) * dim * cap);                    \
      cs = (double*)realloc(cs, sizeof(double) * dim * cap);                    \
      ws = (double*)realloc(ws, sizeof(double) * cap);                          \
      lcount = (int32_t*)realloc(lcount, sizeof(int32_t) * cap);                \
      lsize = (int32_t*)realloc(lsize, sizeof(int32_t) * cap);                  \
      lists = (int32_t*)realloc(lists, sizeof(int32_t) * (size_t)cap * nc);     \
      dcand = (int64_t*)realloc(dcand, sizeof(int64_t) * cap);                  \
    }                                                                           \
  } while (0)
#define PC_CENTRE(j, out)                                     \
  do {                                                        \
    const int ci_ = (j) % res[0], cj_ = ((j) / res[0]) % res[1]; \
    const int ck_ = (j) / (res[0] * res[1]);                  \
    (out)[0] = ctr[0][ci_];                                   \
    (out)[1] = ctr[1][cj_];                                   \
    (out)[2] = dim == 3 ? ctr[2][ck_] : 0.0;                  \
  } while (0)
  for (int64_t e = 0; e < pb->E && st == ORC_OK; e++) {
    pc_list* t = &T[e];
    int64_t nump = t->n;
    if (nump == 0) continue;
    PC_FIT(nump + ppc + pp->max_splits + 1);
    pc_dvc(&s, t, res, xs, cs, ws, own);
    if (inflow) {
      const int64_t nump_orig = nump;
      double dist = 0.0;
      {
        const int64_t i = nump_orig - 1;
        for (int d = 0; d < dim; d++) dist += (cs[i * dim + d] - xs[i * dim + d]) * (cs[i * dim + d] - xs[i * dim + d]);
      }
      const int flag = dist > pp->cent_pos_ratio * 2.0;
      if ((1.0 * nump_orig) / ppc < pp->threshold || flag) {
        int64_t changes = 0;
        for (int64_t i = 0; i < ppc; i++) {
          const int j = (int)(nc * pc_draw(&ndraw));
          double x[3];
          PC_CENTRE(j, x);
          pc_split(pb, &s, T, e, own[j], x);
          nump++;
          changes++;
          added++;
        }
        const int64_t del_num = nump - ppc;
        if (del_num > 5)
          for (int64_t i = 0; i < del_num; i++) {
            const int64_t j = (int64_t)(nump * pc_draw(&ndraw));
            pc_delete_slot(&s, T, e, j);
            nump--;
            changes++;
            deleted++;
          }
        if (changes) {
          if (nump < 3) {
            st = ORC_E_INVALID; /* the reference's under-resolved-cell firewall */
            break;
          }
          pc_dvc(&s, t, res, xs, cs, ws, own);
        }
      }
      /* owned sub-cell lists, filled from the end in sub-cell order */
      for (int64_t i = 0; i < nump; i++) {
        lcount[i] = (int32_t)(ws[i] * one_oda + 0.5);
        lsize[i] = lcount[i];
      }
      for (int c = 0; c < nc; c++) {
        const int o = own[c];
        if (lcount[o] > 0) {
          lists[(size_t)o * nc + lcount[o] - 1] = c;
          lcount[o]--;
        }
      }
    }
    /* the normal pass */
    int64_t mS = pp->max_splits, mD = pp->max_deletions;
    const int32_t m0 = s.mat[t->v[0]];
    for (int64_t i = 0; i < nump; i++)
      if (s.mat[t->v[i]] != m0) {
        if (!pp->delete_in_interface_cells) mD = 0;
        if (!pp->split_in_interface_cells) mS = 0;
        break;
      }
    int64_t nsplit = 0, ndc = 0;
    int64_t* splits = (int64_t*)malloc(sizeof(int64_t) * (nump > 0 ? nump : 1));
    for (int64_t i = 0; i < nump; i++) {
      if (ws[i] > maxW) splits[nsplit++] = i;
      if (ws[i] < minW) dcand[ndc++] = t->v[i];
    }
    for (int64_t a = 1; a < ndc; a++) /* indexOnCPU, highest first */
      for (int64_t b = a; b > 0 && dcand[b - 1] < dcand[b]; b--) {
        const int64_t x = dcand[b];
        dcand[b] = dcand[b - 1];
        dcand[b - 1] = x;
      }
    if (mD > nump - 4) mD = (dim == 2 && inflow) ? nump - 4 : nump / 2;
    const int64_t cs_n = mS < nsplit ? mS : nsplit;
    for (int64_t i = 0; i < cs_n; i++) {
      const int64_t q = splits[i];
      double x[3] = {0.0, 0.0, 0.0};
      if (!inflow) {
        for (int d = 0; d < dim; d++) x[d] = cs[q * dim + d];
      } else {
        const int j = (int)(lsize[q] * pc_draw(&ndraw));
        PC_CENTRE(lists[(size_t)q * nc + j], x);
      }
      pc_split(pb, &s, T, e, q, x);
      nump++;
      added++;
    }
    free(splits);
    const int64_t cd_n = mD < ndc ? mD : ndc;
    for (int64_t i = 0; i < cd_n; i++) {
      int64_t slot = 0;
      while (slot < t->n && t->v[slot] != dcand[i]) slot++;
      if (slot == t->n) { /* cannot happen: candidates are deleted highest index first */
        st = ORC_E_INVALID;
        break;
      }
      pc_delete_slot(&s, T, e, slot);
      nump--;
      deleted++;
    }
    if (st != ORC_OK) break;
    if (cs_n > 0 || cd_n > 0) {
      if (nump < 3) {
        st = ORC_E_INVALID;
        break;
      }
      pc_dvc(&s, t, res, xs, cs, ws, own);
    }
    for (int64_t i = 0; i < nump; i++) {
      const int64_t p = t->v[i];
      for (int d = 0; d < dim; d++) s.xi[3 * p + d] = cs[i * dim + d];
      s.w[p] = ws[i];
    }
  }
#undef PC_FIT
#undef PC_CENTRE
  free(xs); free(cs); free(ws); free(own); free(lcount); free(lsize); free(lists); free(dcand);
  for (int d = 0; d < 3; d++) free(ctr[d]);
  if (st != ORC_OK) {
    for (int64_t e = 0; e < pb->E; e++) free(T[e].v);
    free(T);
    free(s.coord); free(s.mat); free(s.eta0); free(s.pstr); free(s.pinc); free(s.xi); free(s.w); free(s.cell);
    return st;
  }
  /* the problem takes the swarm: caller order = global order, cell lists in
   * table order, integration points and weights from the pass */
  const int64_t P2 = s.n;
  const int64_t PA = P2 > 0 ? P2 : 1;
  free(pb->coord); free(pb->material); free(pb->eta0p); free(pb->pstr); free(pb->pinc);
  pb->coord = (double*)malloc(sizeof(double) * 3 * PA);
  memcpy(pb->coord, s.coord, sizeof(double) * 3 * P2);
  pb->material = (int32_t*)malloc(sizeof(int32_t) * PA);
  memcpy(pb->material, s.mat, sizeof(int32_t) * P2);
  pb->eta0p = NULL;
  pb->pstr = NULL;
  pb->pinc = NULL;
  if (s.eta0) {
    pb->eta0p = (double*)malloc(sizeof(double) * PA);
    memcpy(pb->eta0p, s.eta0, sizeof(double) * P2);
  }
  if (s.pstr) {
    pb->pstr = (double*)malloc(sizeof(double) * PA);
    pb->pinc = (double*)malloc(sizeof(double) * PA);
    memcpy(pb->pstr, s.pstr, sizeof(double) * P2);
    memcpy(pb->pinc, s.pinc, sizeof(double) * P2);
  }
  pb->P = P2;
  pb->owner = (int32_t*)realloc(pb->owner, sizeof(int32_t) * PA);
  pb->xi = (double*)realloc(pb->xi, sizeof(double) * dim * PA);
  pb->weight = (double*)realloc(pb->weight, sizeof(double) * PA);
  pb->cell_list = (int64_t*)realloc(pb->cell_list, sizeof(int64_t) * PA);
  pb->cell_off[0] = 0;
  for (int64_t e = 0; e < pb->E; e++) {
    pb->cell_off[e + 1] = pb->cell_off[e] + T[e].n;
    for (int64_t q = 0; q < T[e].n; q++) pb->cell_list[pb->cell_off[e] + q] = T[e].v[q];
  }
  for (int64_t p = 0; p < P2; p++) {
    pb->owner[p] = (int32_t)s.cell[p];
    for (int d = 0; d < dim; d++) pb->xi[dim * p + d] = s.xi[3 * p + d];
    pb->weight[p] = s.w[p];
  }
  for (int64_t e = 0; e < pb->E; e++) free(T[e].v);
  free(T);
  free(s.coord); free(s.mat); free(s.eta0); free(s.pstr); free(s.pinc); free(s.xi); free(s.w); free(s.cell);
  free(pb->eta); free(pb->yielded);
  pb->eta = (double*)calloc(PA, sizeof(double));
  pb->yielded = (uint8_t*)calloc(PA, 1);
  pb->has_visc = 0;
  pb->assembled = 0;
  pb->wkind = 1;
  for (int d = 0; d < 3; d++) pb->wres[d] = res[d];
  if (counts) {
    counts[0] = added;
    counts[1] = deleted;
    counts[2] = ndraw;
  }
  return ORC_OK;
}

int32_t orc_set_strain_weakening(orc_problem* pb, int32_t material, const orc_strain_weakening* sw) {
  if (material < 0 || material >= pb->nmat) return ORC_E_INVALID;
  if (!pb->sw) {
    pb->sw = (orc_strain_weakening*)calloc((size_t)pb->nmat, sizeof(orc_strain_weakening));
    pb->sw_on = (int*)calloc((size_t)pb->nmat, sizeof(int));
  }
  if (!sw) {
    pb->sw_on[material] = 0;
    return ORC_OK;
  }
  if (pb->mats[material].yield_type == ORC_YIELD_NONE) return ORC_E_INVALID;
  if (!(sw->final_softening_strain > sw->initial_softening_strain) || sw->healing_rate < 0.0) return ORC_E_INVALID;
  pb->sw[material] = *sw;
  pb->sw_on[material] = 1;
  if (!pb->pstr) {
    pb->pstr = (double*)calloc((size_t)(pb->P > 0 ? pb->P : 1), sizeof(double));
    pb->pinc = (double*)calloc((size_t)(pb->P > 0 ? pb->P : 1), sizeof(double));
  }
  pb->has_visc = 0;
  return ORC_OK;
}

int32_t orc_set_plastic_strain(orc_problem* pb, const double* w) {
  if (!pb->pstr) return ORC_E_INVALID;
  memcpy(pb->pstr, w, sizeof(double) * pb->P);
  pb->has_visc = 0;
  return ORC_OK;
}

int32_t orc_plastic_strain(const orc_problem* pb, double* w, double* inc) {
  if (!pb->pstr) return ORC_E_INVALID;
  if (w) memcpy(w, pb->pstr, sizeof(double) * pb->P);
  if (inc) memcpy(inc, pb->pinc, sizeof(double) * pb->P);
  return ORC_OK;
}

/* StrainWeakening.cxx:432-468: the TimeIntegrand's derivative is the stored
 * increment, so each step adds dt * increment; then values below 0 become 0. */
int32_t orc_weakening_step_limited(orc_problem* pb, double dt, const uint8_t* inside, double limit) {
  if (!pb->pstr) return ORC_E_INVALID;
  for (int64_t p = 0; p < pb->P; p++) {
    pb->pstr[p] += dt * pb->pinc[p];
    if (inside && inside[p]) pb->pstr[p] = limit; /* StrainWeakening.cxx:455-461 */
    if (pb->pstr[p] < 0.0) pb->pstr[p] = 0.0;
  }
  return ORC_OK;
}

int32_t orc_weakening_step(orc_problem* pb, double dt) { return orc_weakening_step_limited(pb, dt, NULL, 0.0); }

/* StrainWeakening.cxx:333-430 (loadFromCheckPoint false): srand(randomSeed),
 * then per particle one draw against initialDamageFraction and, when it
 * passes and the particle is inside the shape, a second draw scaled by the
 * factor and the final softening strain, times the wavenumber profiles. */
int32_t orc_initial_damage(int64_t P, const double* coord, const double* prm, long seed, const uint8_t* inside,
                           double* w) {
  const double kPi = 3.14159265358979323846; /* M_PI of <math.h> */
  srand((unsigned)seed);
  for (int64_t p = 0; p < P; p++) {
    const double* x = coord + 3 * p;
    double v = 0.0;
    if (rand() < RAND_MAX * prm[0]) {
      if (inside && !inside[p]) {
        w[p] = v;
        continue;
      }
      v = prm[1] * rand() * prm[2] / RAND_MAX;
      if (prm[3] > 0.0 && prm[5] == -1.0) v *= pow(0.5 + 0.5 * cos(kPi * x[0] * prm[3]), 2.0);
      if (prm[5] > 0.0) v *= pow(0.5 + 0.5 * cos(kPi * x[0] * prm[5]), 2.0);
      if (prm[4] > 0.0) v *= pow(0.5 + 0.5 * sin(kPi * x[0] * prm[4]), 2.0);
      if (prm[7] > 0.0) v *= pow(0.5 + 0.5 * cos(kPi * x[2] * prm[7]), 2.0);
      if (prm[6] > 0.0) v *= pow(0.5 + 0.5 * sin(kPi * x[2] * prm[6]), 2.0);
    }
    w[p] = v;
  }
  return ORC_OK;
}

int32_t orc_set_weights(orc_problem* pb, int32_t kind, const int32_t* res) {
  if (kind != 0 && kind != 1) return ORC_E_INVALID;
  pb->wkind = kind;
  for (int d = 0; d < 3; d++) pb->wres[d] = res ? res[d] : 10;
  return relocate_all(pb);
}

orc_problem* orc_problem_new(const orc_problem_desc* d) {
  if (!d || (d->dim != 2 && d->dim != 3)) return NULL;
  orc_problem* pb = (orc_problem*)calloc(1, sizeof(orc_problem));
  pb->dim = d->dim;
  pb->nen = d->dim == 2 ? 4 : 8;
  pb->ndof = pb->nen * d->dim;
  pb->N = 1;
  pb->E = 1;
  for (int i = 0; i < 3; i++) {
    pb->res[i] = i < d->dim ? d->res[i] : 1;
    pb->nn[i] = i < d->dim ? d->res[i] + 1 : 1;
    pb->lo[i] = d->lo[i];
    pb->hi[i] = d->hi[i];
    pb->N *= pb->nn[i];
    pb->E *= pb->res[i];
  }
  const int64_t ND = pb->N * pb->dim;
  pb->bc_mask = (uint8_t*)calloc(ND, 1);
  pb->bc_value = (double*)calloc(ND, sizeof(double));
  if (d->bc_mask) memcpy(pb->bc_mask, d->bc_mask, ND);
  if (d->bc_value) memcpy(pb->bc_value, d->bc_value, ND * sizeof(double));
  pb->gravity = d->gravity;
  for (int i = 0; i < 3; i++) pb->ghat[i] = d->ghat[i];
  pb->clamp_min = d->eta_clamp_min;
  pb->clamp_max = d->eta_clamp_max;
  pb->nmat = d->n_materials;
  pb->mats = (orc_material*)malloc(sizeof(orc_material) * (pb->nmat > 0 ? pb->nmat : 1));
  memcpy(pb->mats, d->materials, sizeof(orc_material) * pb->nmat);
  pb->P = d->n_particles;
  pb->coord = (double*)malloc(sizeof(double) * 3 * pb->P);
  memcpy(pb->coord, d->coord, sizeof(double) * 3 * pb->P);
  pb->material = (int32_t*)malloc(sizeof(int32_t) * pb->P);
  memcpy(pb->material, d->material, sizeof(int32_t) * pb->P);
  if (d->eta0_particle) {
    pb->eta0p = (double*)malloc(sizeof(double) * pb->P);
    memcpy(pb->eta0p, d->eta0_particle, sizeof(double) * pb->P);
  }
  pb->owner = (int32_t*)malloc(sizeof(int32_t) * pb->P);
  pb->xi = (double*)malloc(sizeof(double) * pb->dim * pb->P);
  pb->weight = (double*)malloc(sizeof(double) * pb->P);
  pb->cell_off = (int64_t*)malloc(sizeof(int64_t) * (pb->E + 1));
  pb->cell_list = (int64_t*)malloc(sizeof(int64_t) * (pb->P > 0 ? pb->P : 1));
  pb->eta = (double*)calloc(pb->P > 0 ? pb->P : 1, sizeof(double));
  pb->yielded = (uint8_t*)calloc(pb->P > 0 ? pb->P : 1, 1);
  for (int64_t p = 0; p < pb->P; p++)
    if (pb->material[p] < 0 || pb->material[p] >= pb->nmat) {
      orc_problem_free(pb);
      return NULL;
    }
  if (relocate_all(pb) != ORC_OK) {
    orc_problem_free(pb);
    return NULL;
  }
  return pb;
}

static void free_system(orc_problem* pb) {
  free(pb->eq); free(pb->Kptr); free(pb->Kcol); free(pb->Kval);
  free(pb->Gptr); free(pb->Gcol); free(pb->Gval);
  free(pb->f); free(pb->h); free(pb->Q); free(pb->Kdiag); free(pb->ilu); free(pb->Kdiagpos);
  pb->eq = NULL; pb->Kptr = NULL; pb->Kcol = NULL; pb->Kval = NULL;
  pb->Gptr = NULL; pb->Gcol = NULL; pb->Gval = NULL;
  pb->f = NULL; pb->h = NULL; pb->Q = NULL; pb->Kdiag = NULL; pb->ilu = NULL; pb->Kdiagpos = NULL;
  pb->assembled = 0;
}

void orc_problem_free(orc_problem* pb) {
  if (!pb) return;
  free_system(pb);
  free(pb->bc_mask); free(pb->bc_value); free(pb->mats); free(pb->coord); free(pb->material);
  free(pb->eta0p); free(pb->pstr); free(pb->pinc); free(pb->sw); free(pb->sw_on); free(pb->owner); free(pb->xi); free(pb->weight); free(pb->cell_off);
  free(pb->cell_list); free(pb->eta); free(pb->yielded);
  free(pb);
}

void orc_get_cell_tables(const orc_problem* pb, int64_t* cell_off, int64_t* cell_list) {
  memcpy(cell_off, pb->cell_off, sizeof(int64_t) * (pb->E + 1));
  memcpy(cell_list, pb->cell_list, sizeof(int64_t) * pb->P);
}

void orc_get_particles(const orc_problem* pb, int32_t* owner, double* xi, double* weight, double* coord) {
  if (owner) memcpy(owner, pb->owner, sizeof(int32_t) * pb->P);
  if (xi) memcpy(xi, pb->xi, sizeof(double) * pb->dim * pb->P);
  if (weight) memcpy(weight, pb->weight, sizeof(double) * pb->P);
  if (coord) memcpy(coord, pb->coord, sizeof(double) * 3 * pb->P);
}

int32_t orc_set_coords(orc_problem* pb, const double* coord) {
  memcpy(pb->coord, coord, sizeof(double) * 3 * pb->P);
  return relocate_all(pb);
}

/* ------------------------------------------------------------------------ */
/* Constitutive evaluation per material point                               */
/*   MaterialViscosity.cxx:129-147, YieldRheology.cxx:190-231,               */
/*   VonMises.cxx:180-291, DruckerPrager.cxx:296-465,                        */
/*   TensorMath.cxx:431-459 (second invariant)                              */
/* ------------------------------------------------------------------------ */
/* ConstitutiveMatrix_SetIsotropicViscosity (ConstitutiveMatrix.cxx:358-364)
 * with the Cartesian isotropic correction (ConstitutiveMatrixCartesian.cxx
 * :536-556): 2D D = diag(2eta, 2eta, eta), 3D diag(2eta x3, eta x3), in the
 * engineering-shear (Voigt) layout.  D is nst x nst row-major. */
void orc_isotropic_D(int32_t dim, double eta, double* D) {
  const int nst = dim == 2 ? 3 : 6;
  for (int i = 0; i < nst * nst; i++) D[i] = 0.0;
  for (int k = 0; k < nst; k++) D[k * nst + k] = k < dim ? 2.0 * eta : eta;
}

/* ConstitutiveMatrixCartesian{2D,3D}_CalculateStress, diagonal branch
 * (ConstitutiveMatrixCartesian.cxx:727-775): the strain-rate tensor's shear
 * components enter doubled.  Layouts: 2D [00,11,01], 3D [00,11,22,01,02,12]. */
void orc_isotropic_stress(int32_t dim, double eta, const double* sr, double* st) {
  const int nst = dim == 2 ? 3 : 6;
  double D[36];
  orc_isotropic_D(dim, eta, D);
  for (int k = 0; k < nst; k++) st[k] = k < dim ? D[k * nst + k] * sr[k] : D[k * nst + k] * 2.0 * sr[k];
}

static double sym_invariant(int dim, const double* t) {
  /* t layout: 2D [00,11,01]; 3D [00,11,22,01,02,12] (TensorMath.h:94-98) */
  double inv;
  if (dim == 2)
    inv = 0.5 * (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2];
  else
    inv = 0.5 * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) + t[3] * t[3] + t[4] * t[4] + t[5] * t[5];
  return sqrt(inv);
}

static double particle_viscosity(orc_problem* pb, int64_t p, int64_t e, const double* u_prev,
                                 const double* p_prev, int yield_on, uint8_t* yielded) {
  const orc_material* m = &pb->mats[pb->material[p]];
  double eta = pb->eta0p ? pb->eta0p[p] : m->eta0;
  *yielded = 0;
  const int swon = pb->pstr && pb->sw_on[pb->material[p]];
  if (pb->pinc) pb->pinc[p] = 0.0;
  if (yield_on && m->yield_type != ORC_YIELD_NONE && u_prev) {
    /* StrainWeakening_CalcRatio (StrainWeakening.cxx:497-519) clamps the
     * stored value at 0 and clips the ratio to [0, 1] */
    double ratio = 0.0, w = 0.0;
    const orc_strain_weakening* sw = swon ? &pb->sw[pb->material[p]] : NULL;
    if (sw) {
      if (pb->pstr[p] < 0.0) pb->pstr[p] = 0.0;
      w = pb->pstr[p];
      ratio = (w - sw->initial_softening_strain) / (sw->final_softening_strain - sw->initial_softening_strain);
      if (ratio > 1.0) ratio = 1.0;
      if (ratio < 0.0) ratio = 0.0;
    }
    const double eta_before = eta;
    const int dim = pb->dim, nen = pb->nen;
    const double* xi = pb->xi + dim * p;
    double GNx[3][MAXNEN];
    int64_t nodes[MAXNEN];
    global_derivs(pb, e, xi, GNx);
    elem_nodes(pb, e, nodes);
    double L[3][3] = {{0}};
    for (int a = 0; a < nen; a++)
      for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++) L[i][j] += GNx[j][a] * u_prev[nodes[a] * dim + i];
    double sr[6], st[6];
    if (dim == 2) {
      sr[0] = L[0][0];
      sr[1] = L[1][1];
      sr[2] = 0.5 * (L[0][1] + L[1][0]);
    } else {
      sr[0] = L[0][0];
      sr[1] = L[1][1];
      sr[2] = L[2][2];
      sr[3] = 0.5 * (L[0][1] + L[1][0]);
      sr[4] = 0.5 * (L[0][2] + L[2][0]);
      sr[5] = 0.5 * (L[1][2] + L[2][1]);
    }
    /* ConstitutiveMatrixCartesian CalculateStress, isotropic D:
     * normal stress = 2 eta e_ii, shear = eta * 2 * e_ij */
    orc_isotropic_stress(dim, eta, sr, st);
    double trs = 0.0, trr = 0.0;
    for (int k = 0; k < dim; k++) {
      trs += st[k];
      trr += sr[k];
    }
    for (int k = 0; k < dim; k++) {
      sr[k] -= trr / dim;
      st[k] -= trs / dim;
    }
    const double srII = sym_invariant(dim, sr);
    const double tauII = sym_invariant(dim, st);
    double crit;
    if (m->yield_type == ORC_YIELD_DRUCKER_PRAGER) {
      const double pres = p_prev ? p_prev[e] : 0.0;
      /* DruckerPrager.cxx:448-489: effective cohesion and friction */
      const double mu = sw ? m->friction * (1.0 - ratio) + sw->friction_after * ratio : m->friction;
      const double C = sw ? m->cohesion * (1.0 - ratio) + sw->cohesion_after * ratio : m->cohesion;
      double factor, fs;
      if (dim == 2) {
        factor = 1 / sqrt(1 + mu * mu);
        fs = mu * pres * factor + C * factor;
      } else {
        const double cphi = 1 / sqrt(1 + mu * mu);
        const double sphi = mu * cphi;
        factor = 2 * cphi / (sqrt(3.0) * (3 - sphi));
        fs = mu * factor * pres + C * 3 * factor;
      }
      double minys = m->min_yield_stress;
      if (minys == 0.0) minys = C;
      if (fs < minys * factor) fs = minys * factor;
      crit = fs;
      if (crit < tauII) {
        double v;
        if (m->max_strain_rate > 0 && srII > m->max_strain_rate)
          v = crit / (2 * m->max_strain_rate);
        else
          v = crit / (2 * srII);
        if (v < m->min_viscosity) v = m->min_viscosity;
        eta = v;
        *yielded = 1;
      }
    } else {
      /* VonMises.cxx:196-199 */
      crit = sw ? m->cohesion * (1.0 - ratio) + sw->cohesion_after * ratio : m->cohesion;
      if (crit < tauII) {
        const double beta = 1.0 - crit / tauII;
        double corr = -eta * beta;
        if (eta + corr < m->min_viscosity) corr = m->min_viscosity - eta;
        eta = eta + corr;
        *yielded = 1;
      }
    }
    if (sw) {
      /* YieldRheology.cxx:190-228 assigns the increment before yielding
       * changes the viscosity; StrainWeakening.cxx:470-495 */
      double beta = 1.0 - crit / tauII;
      if (beta < 0.0) beta = 0.0;
      pb->pinc[p] = (crit / eta_before) * (beta / (1.0 - beta) - sw->healing_rate * w);
    }
  }
  if (pb->clamp_min > 0 && eta < pb->clamp_min) eta = pb->clamp_min;
  if (pb->clamp_max > 0 && eta > pb->clamp_max) eta = pb->clamp_max;
  return eta;
}

/* Output fields of the material points (StoreStress.cxx:190-217,
 * StoreViscosity.cxx:156-180, the StrainRateField): strain rate of u_full at
 * each integration point, its second invariant (TensorMath.cxx:431-459), the
 * viscosity of the last constitutive pass and the isotropic stress
 * (ConstitutiveMatrixCartesian CalculateStress).  Any output may be NULL. */
int32_t orc_particle_fields(orc_problem* pb, const double* u_full, double* sr_out, double* inv_out,
                            double* eta_out, double* st_out) {
  if (!pb->has_visc) return ORC_E_INVALID;
  const int dim = pb->dim, nen = pb->nen, nst = dim == 2 ? 3 : 6;
  for (int64_t e = 0; e < pb->E; e++)
    for (int64_t q = pb->cell_off[e]; q < pb->cell_off[e + 1]; q++) {
      const int64_t p = pb->cell_list[q];
      double GNx[3][MAXNEN];
      int64_t nodes[MAXNEN];
      global_derivs(pb, e, pb->xi + dim * p, GNx);
      elem_nodes(pb, e, nodes);
      double L[3][3] = {{0}};
      for (int a = 0; a < nen; a++)
        for (int i = 0; i < dim; i++)
          for (int j = 0; j < dim; j++) L[i][j] += GNx[j][a] * u_full[nodes[a] * dim + i];
      double sr[6], st[6];
      if (dim == 2) {
        sr[0] = L[0][0];
        sr[1] = L[1][1];
        sr[2] = 0.5 * (L[0][1] + L[1][0]);
      } else {
        sr[0] = L[0][0];
        sr[1] = L[1][1];
        sr[2] = L[2][2];
        sr[3] = 0.5 * (L[0][1] + L[1][0]);
        sr[4] = 0.5 * (L[0][2] + L[2][0]);
        sr[5] = 0.5 * (L[1][2] + L[2][1]);
      }
      orc_isotropic_stress(dim, pb->eta[p], sr, st);
      for (int k = 0; k < nst; k++) {
        if (sr_out) sr_out[p * nst + k] = sr[k];
        if (st_out) st_out[p * nst + k] = st[k];
      }
      if (inv_out) inv_out[p] = sym_invariant(dim, sr);
      if (eta_out) eta_out[p] = pb->eta[p];
    }
  return ORC_OK;
}

int32_t orc_particle_viscosity(orc_problem* pb, const double* u_prev_full, const double* p_prev,
                               int32_t yield_on, double* eta, uint8_t* yielded) {
  for (int64_t e = 0; e < pb->E; e++)
    for (int64_t q = pb->cell_off[e]; q < pb->cell_off[e + 1]; q++) {
      const int64_t p = pb->cell_list[q];
      pb->eta[p] = particle_viscosity(pb, p, e, u_prev_full, p_prev, yield_on, &pb->yielded[p]);
      if (!isfinite(pb->eta[p])) return ORC_E_NAN;
    }
  pb->has_visc = 1;
  pb->assembled = 0;
  if (eta) memcpy(eta, pb->eta, sizeof(double) * pb->P);
  if (yielded) memcpy(yielded, pb->yielded, pb->P);
  return ORC_OK;
}

static void ensure_visc(orc_problem* pb) {
  if (!pb->has_visc) orc_particle_viscosity(pb, NULL, NULL, 0, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* Element matrices                                                          */
/* ------------------------------------------------------------------------ */
/* ConstitutiveMatrixCartesian.cxx:363-485 with the isotropic D of
 * ConstitutiveMatrixCartesian.cxx:536-556 and D.B of :631-724.  Row-major
 * Ke[(col node*dim + i)*ndof + row node*dim + j]. */
int32_t orc_element_K(orc_problem* pb, int64_t e, double* Ke) {
  ensure_visc(pb);
  const int dim = pb->dim, nen = pb->nen, nd = pb->ndof, nst = dim == 2 ? 3 : 6;
  memset(Ke, 0, sizeof(double) * nd * nd);
  for (int64_t q = pb->cell_off[e]; q < pb->cell_off[e + 1]; q++) {
    const int64_t p = pb->cell_list[q];
    double GNx[3][MAXNEN];
    const double detJ = global_derivs(pb, e, pb->xi + dim * p, GNx);
    const double eta = pb->eta[p];
    double Dm[36], Dd[6];
    orc_isotropic_D(dim, eta, Dm);
    for (int k = 0; k < nst; k++) Dd[k] = Dm[k * nst + k] * (detJ * pb->weight[p]);
    for (int r = 0; r < nen; r++) {
      /* B_r (nst x dim) */
      double Br[6][3] = {{0}};
      if (dim == 2) {
        Br[0][0] = GNx[0][r]; Br[1][1] = GNx[1][r]; Br[2][0] = GNx[1][r]; Br[2][1] = GNx[0][r];
      } else {
        Br[0][0] = GNx[0][r]; Br[1][1] = GNx[1][r]; Br[2][2] = GNx[2][r];
        Br[3][0] = GNx[1][r]; Br[3][1] = GNx[0][r];
        Br[4][0] = GNx[2][r]; Br[4][2] = GNx[0][r];
        Br[5][1] = GNx[2][r]; Br[5][2] = GNx[1][r];
      }
      for (int c = 0; c < nen; c++) {
        double Bc[6][3] = {{0}};
        if (dim == 2) {
          Bc[0][0] = GNx[0][c]; Bc[1][1] = GNx[1][c]; Bc[2][0] = GNx[1][c]; Bc[2][1] = GNx[0][c];
        } else {
          Bc[0][0] = GNx[0][c]; Bc[1][1] = GNx[1][c]; Bc[2][2] = GNx[2][c];
          Bc[3][0] = GNx[1][c]; Bc[3][1] = GNx[0][c];
          Bc[4][0] = GNx[2][c]; Bc[4][2] = GNx[0][c];
          Bc[5][1] = GNx[2][c]; Bc[5][2] = GNx[1][c];
        }
        for (int i = 0; i < dim; i++)
          for (int j = 0; j < dim; j++) {
            double s = 0.0;
            for (int k = 0; k < nst; k++) s += Bc[k][i] * (Dd[k] * Br[k][j]);
            Ke[(c * dim + i) * nd + r * dim + j] += s;
          }
      }
    }
  }
  return ORC_OK;
}

/* GradientStiffnessMatrixTerm.cxx:161-249 on the 2x2(x2) Gauss swarm
 * (GaussParticleLayout, weights 1):  G_e[a*dim+d] += w * (-detJ) * dN_a/dx_d */
int32_t orc_element_G(orc_problem* pb, int64_t e, double* Ge) {
  const int dim = pb->dim, nen = pb->nen, ng = dim == 2 ? 4 : 8;
  const double g = 1.0 / sqrt(3.0);
  memset(Ge, 0, sizeof(double) * pb->ndof);
  for (int q = 0; q < ng; q++) {
    double xi[3];
    for (int d = 0; d < dim; d++) xi[d] = ((q >> d) & 1) ? g : -g;
    double GNx[3][MAXNEN];
    const double detJ = global_derivs(pb, e, xi, GNx);
    for (int a = 0; a < nen; a++)
      for (int d = 0; d < dim; d++) Ge[a * dim + d] += 1.0 * (-detJ) * (GNx[d][a] * 1.0);
  }
  return ORC_OK;
}

/* BuoyancyForceTerm.cxx:474-533 (no temperature, no hydrostatic term,
 * adjust off), gravity vector gHat*gravity (:535-541). */
int32_t orc_element_f(orc_problem* pb, int64_t e, double* fe) {
  const int dim = pb->dim, nen = pb->nen;
  memset(fe, 0, sizeof(double) * pb->ndof);
  for (int64_t q = pb->cell_off[e]; q < pb->cell_off[e + 1]; q++) {
    const int64_t p = pb->cell_list[q];
    double GNx[3][MAXNEN], N[MAXNEN];
    const double detJ = global_derivs(pb, e, pb->xi + dim * p, GNx);
    shape_funcs(dim, pb->xi + dim * p, N);
    const double rho = pb->mats[pb->material[p]].density;
    const double factor = detJ * pb->weight[p] * 1.0 * rho;
    for (int a = 0; a < nen; a++)
      for (int d = 0; d < dim; d++) fe[a * dim + d] += (pb->ghat[d] * pb->gravity) * factor * N[a];
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Global assembly with Dirichlet elimination                                */
/*   StiffnessMatrix.cxx:2016-2200 (element loop, lifting f -= K_e u_bc,      */
/*   h -= G_e^T u_bc, MatSetValues ADD), ForceVector.cxx:303-520,            */
/*   UzawaPreconditionerTerm.cxx:156-239 (Q_ee = sum_v G_v^2 / K_vv).        */
/* Free dofs are numbered node by node, dof by dof (FeEquationNumber).       */
/* ------------------------------------------------------------------------ */
static int64_t find_col(const int32_t* cols, int64_t b, int64_t e, int32_t c) {
  while (b < e) {
    const int64_t m = (b + e) / 2;
    if (cols[m] < c)
      b = m + 1;
    else
      e = m;
  }
  return b;
}

int32_t orc_assemble(orc_problem* pb, const double* u_prev_full, const double* p_prev, int32_t yield_on,
                     int32_t nthreads) {
  const int dim = pb->dim, nen = pb->nen, nd = pb->ndof;
  int32_t st = orc_particle_viscosity(pb, u_prev_full, p_prev, yield_on, NULL, NULL);
  if (st) return st;
  free_system(pb);
  const int64_t ND = pb->N * dim;
  pb->eq = (int32_t*)malloc(sizeof(int32_t) * ND);
  int64_t nu = 0;
  for (int64_t k = 0; k < ND; k++) pb->eq[k] = pb->bc_mask[k] ? -1 : (int32_t)nu++;
  pb->n_u = nu;
  /* K pattern: neighbour nodes (3^dim) x dofs, free only, ascending. */
  pb->Kptr = (int64_t*)calloc(nu + 1, sizeof(int64_t));
  const int64_t nx = pb->nn[0], ny = pb->nn[1], nz = pb->nn[2];
  for (int pass = 0; pass < 2; pass++) {
    int64_t nnz = 0;
    for (int64_t n = 0; n < pb->N; n++) {
      const int64_t i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
      for (int c = 0; c < dim; c++) {
        const int32_t row = pb->eq[n * dim + c];
        if (row < 0) continue;
        if (pass == 1) pb->Kptr[row] = nnz;
        for (int64_t kk = k - 1; kk <= k + 1; kk++) {
          if (kk < 0 || kk >= nz) continue;
          for (int64_t jj = j - 1; jj <= j + 1; jj++) {
            if (jj < 0 || jj >= ny) continue;
            for (int64_t ii = i - 1; ii <= i + 1; ii++) {
              if (ii < 0 || ii >= nx) continue;
              const int64_t m = ii + nx * (jj + ny * kk);
              for (int c2 = 0; c2 < dim; c2++) {
                const int32_t col = pb->eq[m * dim + c2];
                if (col < 0) continue;
                if (pass == 1) pb->Kcol[nnz] = col;
                nnz++;
              }
            }
          }
        }
      }
    }
    if (pass == 0) {
      pb->Kcol = (int32_t*)malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
      pb->Kval = (double*)calloc(nnz > 0 ? nnz : 1, sizeof(double));
    } else {
      pb->Kptr[nu] = nnz;
    }
  }
  /* G pattern: each free velocity row touches its (up to 2^dim) elements. */
  pb->Gptr = (int64_t*)calloc(nu + 1, sizeof(int64_t));
  for (int pass = 0; pass < 2; pass++) {
    int64_t nnz = 0;
    for (int64_t n = 0; n < pb->N; n++) {
      const int64_t i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
      for (int c = 0; c < dim; c++) {
        const int32_t row = pb->eq[n * dim + c];
        if (row < 0) continue;
        if (pass == 1) pb->Gptr[row] = nnz;
        const int64_t k0 = dim == 3 ? k - 1 : 0, k1 = dim == 3 ? k : 0;
        for (int64_t kk = k0; kk <= k1; kk++) {
          if (kk < 0 || kk >= pb->res[2]) continue;
          for (int64_t jj = j - 1; jj <= j; jj++) {
            if (jj < 0 || jj >= pb->res[1]) continue;
            for (int64_t ii = i - 1; ii <= i; ii++) {
              if (ii < 0 || ii >= pb->res[0]) continue;
              if (pass == 1) pb->Gcol[nnz] = (int32_t)(ii + pb->res[0] * (jj + pb->res[1] * kk));
              nnz++;
            }
          }
        }
      }
    }
    if (pass == 0) {
      pb->Gcol = (int32_t*)malloc(sizeof(int32_t) * (nnz > 0 ? nnz : 1));
      pb->Gval = (double*)calloc(nnz > 0 ? nnz : 1, sizeof(double));
    } else {
      pb->Gptr[nu] = nnz;
    }
  }
  pb->f = (double*)calloc(nu > 0 ? nu : 1, sizeof(double));
  pb->h = (double*)calloc(pb->E, sizeof(double));
  pb->Q = (double*)calloc(pb->E, sizeof(double));
  /* element loop coloured by index parity so that scatters never collide */
  const int ncol = dim == 2 ? 4 : 8;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (int colour = 0; colour < ncol; colour++) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t e = 0; e < pb->E; e++) {
      int32_t ijk[3];
      elem_ijk(pb, e, ijk);
      const int ec = (ijk[0] & 1) | ((ijk[1] & 1) << 1) | ((ijk[2] & 1) << 2);
      if (ec != colour) continue;
      double Ke[MAXDOF * MAXDOF], Ge[MAXDOF], fe[MAXDOF];
      int64_t nodes[MAXNEN];
      int32_t leq[MAXDOF];
      double ubc[MAXDOF];
      orc_element_K(pb, e, Ke);
      orc_element_G(pb, e, Ge);
      orc_element_f(pb, e, fe);
      elem_nodes(pb, e, nodes);
      for (int a = 0; a < nen; a++)
        for (int d = 0; d < dim; d++) {
          leq[a * dim + d] = pb->eq[nodes[a] * dim + d];
          ubc[a * dim + d] = pb->bc_mask[nodes[a] * dim + d] ? pb->bc_value[nodes[a] * dim + d] : 0.0;
        }
      double q = 0.0, hl = 0.0;
      for (int r = 0; r < nd; r++) {
        q += Ge[r] * (Ge[r] / Ke[r * nd + r]);
        const int32_t row = leq[r];
        if (row < 0) {
          hl -= ubc[r] * Ge[r];
          continue;
        }
        double lift = 0.0;
        for (int c = 0; c < nd; c++) {
          const int32_t col = leq[c];
          if (col < 0) {
            lift -= ubc[c] * Ke[r * nd + c];
            continue;
          }
          const int64_t pos = find_col(pb->Kcol, pb->Kptr[row], pb->Kptr[row + 1], col);
          pb->Kval[pos] += Ke[r * nd + c];
        }
        pb->f[row] += fe[r] + lift;
        const int64_t gpos = find_col(pb->Gcol, pb->Gptr[row], pb->Gptr[row + 1], (int32_t)e);
        pb->Gval[gpos] += Ge[r];
      }
      pb->h[e] += hl;
      pb->Q[e] += q;
    }
  }
  pb->Kdiag = (double*)malloc(sizeof(double) * (nu > 0 ? nu : 1));
  pb->Kdiagpos = (int64_t*)malloc(sizeof(int64_t) * (nu > 0 ? nu : 1));
  for (int64_t r = 0; r < nu; r++) {
    const int64_t pos = find_col(pb->Kcol, pb->Kptr[r], pb->Kptr[r + 1], (int32_t)r);
    pb->Kdiagpos[r] = pos;
    pb->Kdiag[r] = pb->Kval[pos];
  }
  pb->assembled = 1;
  return ORC_OK;
}

void orc_system_sizes(const orc_problem* pb, int64_t* out4) {
  out4[0] = pb->n_u;
  out4[1] = pb->E;
  out4[2] = pb->Kptr ? pb->Kptr[pb->n_u] : 0;
  out4[3] = pb->Gptr ? pb->Gptr[pb->n_u] : 0;
}

void orc_system_export(const orc_problem* pb, int64_t* Kptr, int32_t* Kcol, double* Kval, int64_t* Gptr,
                       int32_t* Gcol, double* Gval, double* f, double* h, double* Q, int32_t* eq) {
  const int64_t nu = pb->n_u, nk = pb->Kptr[nu], ng = pb->Gptr[nu];
  if (Kptr) memcpy(Kptr, pb->Kptr, sizeof(int64_t) * (nu + 1));
  if (Kcol) memcpy(Kcol, pb->Kcol, sizeof(int32_t) * nk);
  if (Kval) memcpy(Kval, pb->Kval, sizeof(double) * nk);
  if (Gptr) memcpy(Gptr, pb->Gptr, sizeof(int64_t) * (nu + 1));
  if (Gcol) memcpy(Gcol, pb->Gcol, sizeof(int32_t) * ng);
  if (Gval) memcpy(Gval, pb->Gval, sizeof(double) * ng);
  if (f) memcpy(f, pb->f, sizeof(double) * nu);
  if (h) memcpy(h, pb->h, sizeof(double) * pb->E);
  if (Q) memcpy(Q, pb->Q, sizeof(double) * pb->E);
  if (eq) memcpy(eq, pb->eq, sizeof(int32_t) * pb->N * pb->dim);
}

/* Unconstrained element-by-element operators on full nodal vectors. */
void orc_apply_K_full(orc_problem* pb, const double* x, double* y) {
  const int dim = pb->dim, nen = pb->nen, nd = pb->ndof;
  memset(y, 0, sizeof(double) * pb->N * dim);
  for (int64_t e = 0; e < pb->E; e++) {
    double Ke[MAXDOF * MAXDOF];
    int64_t nodes[MAXNEN];
    orc_element_K(pb, e, Ke);
    elem_nodes(pb, e, nodes);
    for (int r = 0; r < nd; r++) {
      double s = 0.0;
      for (int c = 0; c < nd; c++) s += Ke[r * nd + c] * x[nodes[c / dim] * dim + c % dim];
      y[nodes[r / dim] * dim + r % dim] += s;
    }
  }
  (void)nen;
}

void orc_diag_K_full(orc_problem* pb, double* dg) {
  const int dim = pb->dim, nd = pb->ndof;
  memset(dg, 0, sizeof(double) * pb->N * dim);
  for (int64_t e = 0; e < pb->E; e++) {
    double Ke[MAXDOF * MAXDOF];
    int64_t nodes[MAXNEN];
    orc_element_K(pb, e, Ke);
    elem_nodes(pb, e, nodes);
    for (int r = 0; r < nd; r++) dg[nodes[r / dim] * dim + r % dim] += Ke[r * nd + r];
  }
}

void orc_apply_G_full(orc_problem* pb, const double* p, double* y) {
  const int dim = pb->dim, nd = pb->ndof;
  memset(y, 0, sizeof(double) * pb->N * dim);
  for (int64_t e = 0; e < pb->E; e++) {
    double Ge[MAXDOF];
    int64_t nodes[MAXNEN];
    orc_element_G(pb, e, Ge);
    elem_nodes(pb, e, nodes);
    for (int r = 0; r < nd; r++) y[nodes[r / dim] * dim + r % dim] += Ge[r] * p[e];
  }
}

void orc_apply_Gt_full(orc_problem* pb, const double* x, double* q) {
  const int dim = pb->dim, nd = pb->ndof;
  for (int64_t e = 0; e < pb->E; e++) {
    double Ge[MAXDOF];
    int64_t nodes[MAXNEN];
    orc_element_G(pb, e, Ge);
    elem_nodes(pb, e, nodes);
    double s = 0.0;
    for (int r = 0; r < nd; r++) s += Ge[r] * x[nodes[r / dim] * dim + r % dim];
    q[e] = s;
  }
}

void orc_uzawa_Q(orc_problem* pb, double* Q) {
  const int nd = pb->ndof;
  for (int64_t e = 0; e < pb->E; e++) {
    double Ke[MAXDOF * MAXDOF], Ge[MAXDOF];
    orc_element_K(pb, e, Ke);
    orc_element_G(pb, e, Ge);
    double q = 0.0;
    for (int r = 0; r < nd; r++) q += Ge[r] * (Ge[r] / Ke[r * nd + r]);
    Q[e] = q;
  }
}

/* ------------------------------------------------------------------------ */
/* Sparse kernels and inner solvers                                          */
/* ------------------------------------------------------------------------ */
static void spmv(int64_t n, const int64_t* ptr, const int32_t* col, const double* val, const double* x,
                 double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    double s = 0.0;
    for (int64_t k = ptr[r]; k < ptr[r + 1]; k++) s += val[k] * x[col[k]];
    y[r] = s;
  }
}

/* y = G^T x  (G is n_u x E in CSR by rows) */
static void spmv_t(int64_t nrows, int64_t ncols, const int64_t* ptr, const int32_t* col, const double* val,
                   const double* x, double* y) {
  memset(y, 0, sizeof(double) * ncols);
  for (int64_t r = 0; r < nrows; r++)
    for (int64_t k = ptr[r]; k < ptr[r + 1]; k++) y[col[k]] += val[k] * x[r];
}

static double dot(int64_t n, const double* a, const double* b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static double nrm2(int64_t n, const double* a) { return sqrt(dot(n, a, a)); }

/* CG with Jacobi preconditioner, zero initial guess, stop on
 * |r|_2 <= rtol |b|_2. */
static int cg_jacobi(orc_problem* pb, const double* b, double* x, double rtol, int maxit, int64_t* its) {
  const int64_t n = pb->n_u;
  double *r = (double*)malloc(sizeof(double) * n * 4 + 8), *z = r + n, *p = z + n, *Ap = p + n;
  memset(x, 0, sizeof(double) * n);
  memcpy(r, b, sizeof(double) * n);
  const double bn = nrm2(n, b);
  int k = 0;
  int st = ORC_OK;
  if (bn == 0.0) goto done;
  for (int64_t i = 0; i < n; i++) z[i] = r[i] / pb->Kdiag[i];
  memcpy(p, z, sizeof(double) * n);
  double rz = dot(n, r, z);
  for (k = 0; k < maxit; k++) {
    if (nrm2(n, r) <= rtol * bn) break;
    spmv(n, pb->Kptr, pb->Kcol, pb->Kval, p, Ap);
    const double alpha = rz / dot(n, p, Ap);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      z[i] = r[i] / pb->Kdiag[i];
    }
    const double rz2 = dot(n, r, z);
    const double beta = rz2 / rz;
    rz = rz2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  if (k >= maxit && nrm2(n, r) > rtol * bn) st = ORC_E_INNER_MAXIT;
done:
  *its += k;
  free(r);
  return st;
}

/* ILU(0) on the pattern of K (PETSc's serial default preconditioner). */
static void ilu0_factor(orc_problem* pb) {
  const int64_t n = pb->n_u;
  if (pb->ilu) return;
  pb->ilu = (double*)malloc(sizeof(double) * (pb->Kptr[n] > 0 ? pb->Kptr[n] : 1));
  memcpy(pb->ilu, pb->Kval, sizeof(double) * pb->Kptr[n]);
  double* a = pb->ilu;
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) pos[i] = -1;
  for (int64_t i = 0; i < n; i++) {
    for (int64_t k = pb->Kptr[i]; k < pb->Kptr[i + 1]; k++) pos[pb->Kcol[k]] = k;
    for (int64_t kk = pb->Kptr[i]; kk < pb->Kptr[i + 1]; kk++) {
      const int32_t k = pb->Kcol[kk];
      if (k >= i) break;
      a[kk] /= a[pb->Kdiagpos[k]];
      for (int64_t jj = pb->Kdiagpos[k] + 1; jj < pb->Kptr[k + 1]; jj++) {
        const int64_t p = pos[pb->Kcol[jj]];
        if (p >= 0) a[p] -= a[kk] * a[jj];
      }
    }
    for (int64_t k = pb->Kptr[i]; k < pb->Kptr[i + 1]; k++) pos[pb->Kcol[k]] = -1;
  }
  free(pos);
}

static void ilu0_solve(const orc_problem* pb, const double* b, double* x) {
  const int64_t n = pb->n_u;
  const double* a = pb->ilu;
  for (int64_t i = 0; i < n; i++) {
    double s = b[i];
    for (int64_t k = pb->Kptr[i]; k < pb->Kdiagpos[i]; k++) s -= a[k] * x[pb->Kcol[k]];
    x[i] = s;
  }
  for (int64_t i = n - 1; i >= 0; i--) {
    double s = x[i];
    for (int64_t k = pb->Kdiagpos[i] + 1; k < pb->Kptr[i + 1]; k++) s -= a[k] * x[pb->Kcol[k]];
    x[i] = s / a[pb->Kdiagpos[i]];
  }
}

/* Restarted GMRES(30), left preconditioned by ILU(0), modified Gram-Schmidt,
 * zero initial guess, stop when |M^-1 r| <= rtol |M^-1 b| (PETSc 3.2 KSP
 * defaults used by the reference's velocity solve). */
static int gmres_ilu(orc_problem* pb, const double* b, double* x, double rtol, int maxit, int64_t* its) {
  const int64_t n = pb->n_u;
  const int m = 30;
  ilu0_factor(pb);
  double* V = (double*)malloc(sizeof(double) * n * (m + 1));
  double* w = (double*)malloc(sizeof(double) * n * 2);
  double* t = w + n;
  double H[31][30], cs[30], sn[30], g[31], y[30];
  memset(x, 0, sizeof(double) * n);
  ilu0_solve(pb, b, t);
  const double r0 = nrm2(n, t);
  int total = 0, st = ORC_OK;
  if (r0 == 0.0) goto done;
  for (;;) {
    /* r = M^-1 (b - K x) */
    spmv(n, pb->Kptr, pb->Kcol, pb->Kval, x, w);
    for (int64_t i = 0; i < n; i++) w[i] = b[i] - w[i];
    ilu0_solve(pb, w, t);
    double beta = nrm2(n, t);
    if (beta <= rtol * r0) break;
    if (total >= maxit) {
      st = ORC_E_INNER_MAXIT;
      break;
    }
    for (int64_t i = 0; i < n; i++) V[i] = t[i] / beta;
    memset(g, 0, sizeof(g));
    g[0] = beta;
    int j;
    int conv = 0;
    for (j = 0; j < m && total < maxit; j++) {
      spmv(n, pb->Kptr, pb->Kcol, pb->Kval, V + (int64_t)j * n, w);
      ilu0_solve(pb, w, t);
      for (int i = 0; i <= j; i++) {
        H[i][j] = dot(n, t, V + (int64_t)i * n);
        const double hij = H[i][j];
        const double* vi = V + (int64_t)i * n;
        for (int64_t k = 0; k < n; k++) t[k] -= hij * vi[k];
      }
      H[j + 1][j] = nrm2(n, t);
      if (H[j + 1][j] != 0.0) {
        double* vj = V + (int64_t)(j + 1) * n;
        for (int64_t k = 0; k < n; k++) vj[k] = t[k] / H[j + 1][j];
      }
      for (int i = 0; i < j; i++) {
        const double tmp = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
        H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
        H[i][j] = tmp;
      }
      const double den = sqrt(H[j][j] * H[j][j] + H[j + 1][j] * H[j + 1][j]);
      cs[j] = den == 0.0 ? 1.0 : H[j][j] / den;
      sn[j] = den == 0.0 ? 0.0 : H[j + 1][j] / den;
      H[j][j] = cs[j] * H[j][j] + sn[j] * H[j + 1][j];
      H[j + 1][j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      total++;
      if (fabs(g[j + 1]) <= rtol * r0) {
        j++;
        conv = 1;
        break;
      }
    }
    for (int i = j - 1; i >= 0; i--) {
      double s = g[i];
      for (int k = i + 1; k < j; k++) s -= H[i][k] * y[k];
      y[i] = s / H[i][i];
    }
    for (int i = 0; i < j; i++) {
      const double* vi = V + (int64_t)i * n;
      for (int64_t k = 0; k < n; k++) x[k] += y[i] * vi[k];
    }
    if (conv) break;
  }
done:
  *its += total;
  free(V);
  free(w);
  return st;
}

static int inner_solve(orc_problem* pb, const orc_uzawa_params* up, const double* b, double* x, double rtol,
                       int64_t* its) {
  const int maxit = up->inner_max_iterations > 0 ? up->inner_max_iterations : 100000;
  if (up->inner == 1) return gmres_ilu(pb, b, x, rtol, maxit, its);
  return cg_jacobi(pb, b, x, rtol, maxit, its);
}

/* ------------------------------------------------------------------------ */
/* Uzawa: Stokes_SLE_UzawaSolver.cxx:393-938                                 */
/* ------------------------------------------------------------------------ */
static int good(double v) { return isfinite(v); }

static void expand_u(const orc_problem* pb, const double* uf, double* ufull) {
  const int64_t ND = pb->N * pb->dim;
  for (int64_t k = 0; k < ND; k++) ufull[k] = pb->eq[k] < 0 ? pb->bc_value[k] : uf[pb->eq[k]];
}

static int uzawa_reduced(orc_problem* pb, const orc_uzawa_params* up, double* u, double* q, orc_report* rep) {
  const int64_t nu = pb->n_u, np = pb->E;
  const double tol = up->tolerance;
  double* buf = (double*)calloc((size_t)(3 * np + 2 * nu + 8), sizeof(double));
  double *qTemp = buf, *r = qTemp + np, *s = r + np, *fTemp = s + np, *vStar = fTemp + nu;
  int64_t inner = 0;
  int st = ORC_OK;
  double zr_prev = 1.0, relRes = 0.0, absRes = 0.0;
  int it = 0;
  const double qScale = sqrt((double)np), qRecip = 1.0 / qScale;
  const double t0 = now_s();
  /* :476-493 zero force firewall */
  if (nrm2(nu, pb->f) / sqrt((double)nu) <= 1e-99) {
    st = ORC_E_ZERO_FORCE;
    goto out;
  }
  /* STEP 1 :553-565 */
  {
    const double rt = up->inner_rtol_override > 0 ? up->inner_rtol_override : tol;
    int s1 = inner_solve(pb, up, pb->f, vStar, rt, &inner);
    if (s1) { st = s1; goto out; }
  }
  spmv_t(nu, np, pb->Gptr, pb->Gcol, pb->Gval, vStar, qTemp);
  for (int64_t i = 0; i < np; i++) qTemp[i] -= pb->h[i];
  /* STEP 2 :587-605 */
  const double irt = up->inner_rtol_override > 0 ? up->inner_rtol_override
                                                 : (up->inner_rtol_factor > 0 ? up->inner_rtol_factor : 0.1) * tol;
  const double uzawaRhsScale = nrm2(np, qTemp) * qRecip;
  if (!good(uzawaRhsScale)) { st = ORC_E_NAN; goto out; }
  /* STEP 3 :616-645  fTemp = f - G q ; u = K^-1 fTemp ; r = G^T u - h */
  spmv(nu, pb->Gptr, pb->Gcol, pb->Gval, q, fTemp);
  for (int64_t i = 0; i < nu; i++) fTemp[i] = pb->f[i] - fTemp[i];
  {
    int s3 = inner_solve(pb, up, fTemp, u, irt, &inner);
    if (s3) { st = s3; goto out; }
  }
  spmv_t(nu, np, pb->Gptr, pb->Gcol, pb->Gval, u, r);
  {
    const double divU = nrm2(np, r) / nrm2(nu, u);
    if (!good(divU)) { st = ORC_E_NAN; goto out; }
  }
  for (int64_t i = 0; i < np; i++) r[i] -= pb->h[i];
  /* STEP 4 :664-837 */
  do {
    if (up->use_preconditioner)
      for (int64_t i = 0; i < np; i++) qTemp[i] = r[i] / pb->Q[i];
    else
      memcpy(qTemp, r, sizeof(double) * np);
    const double zr = dot(np, qTemp, r);
    if (it == 0)
      memcpy(s, qTemp, sizeof(double) * np);
    else {
      const double beta = zr / zr_prev;
      for (int64_t i = 0; i < np; i++) s[i] = qTemp[i] + beta * s[i];
    }
    spmv(nu, pb->Gptr, pb->Gcol, pb->Gval, s, fTemp);
    {
      int s4 = inner_solve(pb, up, fTemp, vStar, irt, &inner);
      if (s4) { st = s4; goto out; }
    }
    spmv_t(nu, np, pb->Gptr, pb->Gcol, pb->Gval, vStar, qTemp);
    const double sq = dot(np, s, qTemp);
    const double alpha = zr / sq;
    if (!good(zr) || !good(sq) || !good(alpha)) { st = ORC_E_NAN; goto out; }
    for (int64_t i = 0; i < np; i++) q[i] += alpha * s[i];
    for (int64_t i = 0; i < nu; i++) u[i] -= alpha * vStar[i];
    for (int64_t i = 0; i < np; i++) r[i] -= alpha * qTemp[i];
    zr_prev = zr;
    absRes = nrm2(np, r) * qRecip;
    relRes = absRes / uzawaRhsScale;
    if (!good(absRes)) { st = ORC_E_NAN; goto out; }
    if (!(it < up->max_iterations)) { st = ORC_E_MAXIT; goto out; }
    it++;
  } while (((up->use_absolute_tolerance ? absRes : relRes) > tol) || (it < up->min_iterations));
out:
  rep->status = st;
  rep->uzawa_iterations = it;
  rep->inner_iterations += inner;
  rep->rel_residual = relRes;
  rep->abs_residual = absRes;
  rep->solve_seconds += now_s() - t0;
  if (st == ORC_OK) {
    /* STEP 5 summary :844-870 */
    spmv_t(nu, np, pb->Gptr, pb->Gcol, pb->Gval, u, r);
    rep->div_u = nrm2(np, r) / nrm2(nu, u);
    spmv(nu, pb->Gptr, pb->Gcol, pb->Gval, q, vStar);
    spmv(nu, pb->Kptr, pb->Kcol, pb->Kval, u, fTemp);
    for (int64_t i = 0; i < nu; i++) fTemp[i] = pb->f[i] - (fTemp[i] + vStar[i]);
    rep->momentum_residual = nrm2(nu, fTemp) / nrm2(nu, pb->f);
  }
  free(buf);
  return st;
}

int32_t orc_uzawa(orc_problem* pb, const orc_uzawa_params* up, const double* p0, double* u_full, double* p,
                  orc_report* rep) {
  memset(rep, 0, sizeof(*rep));
  if (!pb->assembled) {
    const double ta = now_s();
    int32_t st = orc_assemble(pb, NULL, NULL, 0, up->nthreads);
    rep->assemble_seconds = now_s() - ta;
    if (st) return rep->status = st;
  }
#ifdef _OPENMP
  if (up->nthreads > 0) omp_set_num_threads(up->nthreads);
#endif
  double* uf = (double*)calloc(pb->n_u > 0 ? pb->n_u : 1, sizeof(double));
  double* q = (double*)calloc(pb->E, sizeof(double));
  if (p0) memcpy(q, p0, sizeof(double) * pb->E);
  int32_t st = uzawa_reduced(pb, up, uf, q, rep);
  if (u_full) expand_u(pb, uf, u_full);
  if (p) memcpy(p, q, sizeof(double) * pb->E);
  free(uf);
  free(q);
  return st;
}

/* Picard loop: SystemLinearEquations.cxx:735-920.  The first solve yields
 * only when a previous solution exists (YieldRheology.cxx:203-204 with
 * previousSolutionExists = sle->hasExecuted, ConstitutiveMatrixCartesian.cxx
 * :270-273; hasExecuted is set after the first solve of the first nonlinear
 * execute, SystemLinearEquations.cxx:794, and stays set for later time steps,
 * or is set up front on a restart, ConstitutiveMatrix.cxx:127-130).
 * u_prev_full / p_prev: that previous solution (NULL = none: the first solve
 * is linear and starts from p = 0); p_prev is also the first Uzawa initial
 * guess, as the reference keeps pSolnVec.  The pressure vector carries over
 * between linear solves as the Uzawa initial guess; the residual is
 * |u_prev - u| / |u| over the free velocity dofs. */
int32_t orc_picard_from(orc_problem* pb, const orc_uzawa_params* up, const orc_picard_params* pp,
                        const double* u_prev_full, const double* p_prev, double* u_full, double* p, orc_report* rep) {
  memset(rep, 0, sizeof(*rep));
  const int64_t ND = pb->N * pb->dim;
  double* uf_full = (double*)calloc(ND, sizeof(double));
  double* q = (double*)calloc(pb->E, sizeof(double));
  double *uprev = NULL, *ucur = NULL;
  orc_report r1;
  int32_t st;
  if (u_prev_full && p_prev) memcpy(q, p_prev, sizeof(double) * pb->E);
  double ta = now_s();
  if (u_prev_full && p_prev)
    st = orc_assemble(pb, u_prev_full, p_prev, 1, up->nthreads);
  else
    st = orc_assemble(pb, NULL, NULL, 0, up->nthreads);
  rep->assemble_seconds += now_s() - ta;
  if (st) goto out;
  memset(&r1, 0, sizeof(r1));
  st = orc_uzawa(pb, up, q, uf_full, q, &r1);
  rep->inner_iterations += r1.inner_iterations;
  rep->uzawa_iterations += r1.uzawa_iterations;
  rep->solve_seconds += r1.solve_seconds;
  rep->rel_residual = r1.rel_residual;
  if (st) goto out;
  {
    const int64_t nu = pb->n_u;
    uprev = (double*)malloc(sizeof(double) * (nu > 0 ? nu : 1));
    ucur = (double*)malloc(sizeof(double) * (nu > 0 ? nu : 1));
    int converged = 0;
    int it;
    double res = 0.0;
    for (it = 1; it < pp->max_iterations; it++) {
      for (int64_t k = 0; k < ND; k++)
        if (pb->eq[k] >= 0) uprev[pb->eq[k]] = uf_full[k];
      ta = now_s();
      st = orc_assemble(pb, uf_full, q, 1, up->nthreads);
      rep->assemble_seconds += now_s() - ta;
      if (st) goto out;
      st = orc_uzawa(pb, up, q, uf_full, q, &r1);
      rep->inner_iterations += r1.inner_iterations;
      rep->uzawa_iterations += r1.uzawa_iterations;
      rep->solve_seconds += r1.solve_seconds;
      rep->rel_residual = r1.rel_residual;
      if (st) goto out;
      for (int64_t k = 0; k < ND; k++)
        if (pb->eq[k] >= 0) ucur[pb->eq[k]] = uf_full[k];
      double dn = 0.0, cn = 0.0;
      for (int64_t i = 0; i < nu; i++) {
        dn += (uprev[i] - ucur[i]) * (uprev[i] - ucur[i]);
        cn += ucur[i] * ucur[i];
      }
      res = sqrt(dn) / sqrt(cn);
      converged = res < pp->tolerance;
      if (converged && it >= pp->min_iterations) break;
    }
    rep->picard_iterations = it < pp->max_iterations ? it + 1 : it;
    rep->picard_residual = res;
    if (!converged && pp->kill_non_convergent) st = ORC_E_PICARD;
  }
out:
  rep->status = st;
  if (u_full) memcpy(u_full, uf_full, sizeof(double) * ND);
  if (p) memcpy(p, q, sizeof(double) * pb->E);
  free(uf_full);
  free(q);
  free(uprev);
  free(ucur);
  return st;
}

int32_t orc_picard(orc_problem* pb, const orc_uzawa_params* up, const orc_picard_params* pp, double* u_full,
                   double* p, orc_report* rep) {
  return orc_picard_from(pb, up, pp, NULL, NULL, u_full, p, rep);
}

/* ------------------------------------------------------------------------ */
/* dt and advection                                                           */
/* ------------------------------------------------------------------------ */
/* UpdateDt.cxx:53-84: dt = dtFactor * 0.5 * minSeparation / max|v| */
double orc_compute_dt(orc_problem* pb, const double* u, double dt_factor) {
  double vmax = 0.0;
  for (int64_t n = 0; n < pb->N; n++) {
    double s = 0.0;
    for (int d = 0; d < pb->dim; d++) s += u[n * pb->dim + d] * u[n * pb->dim + d];
    s = sqrt(s);
    if (s > vmax) vmax = s;
  }
  double minsep = INFINITY;
  for (int d = 0; d < pb->dim; d++)
    for (int32_t i = 0; i < pb->res[d]; i++) {
      const double h = orc_vertex_coord(pb->res[d], pb->lo[d], pb->hi[d], i + 1) -
                       orc_vertex_coord(pb->res[d], pb->lo[d], pb->hi[d], i);
      if (h < minsep) minsep = h;
    }
  if (vmax == 0.0) return 0.0;
  return dt_factor * 0.5 * minsep / vmax;
}

/* FeVariable.cxx:904-950 + InterpolateWithinElement: 0 if outside. */
int32_t orc_interpolate_velocity(orc_problem* pb, const double* u, const double* x, double* v) {
  int32_t e;
  if (!orc_search_element(pb->dim, pb->res, pb->lo, pb->hi, x, &e)) return 0;
  double xi[3], N[MAXNEN];
  int64_t nodes[MAXNEN];
  global_to_local(pb, e, x, xi);
  shape_funcs(pb->dim, xi, N);
  elem_nodes(pb, e, nodes);
  for (int d = 0; d < pb->dim; d++) {
    double s = 0.0;
    for (int a = 0; a < pb->nen; a++) s += N[a] * u[nodes[a] * pb->dim + d];
    v[d] = s;
    if (isinf(s)) return 0;
  }
  return 1;
}

/* TimeIntegrand.cxx:265-335 midpoint RK2 with first-order fallback
 * (:317-331, _TimeIntegrand_RewindToStartAndApplyFirstOrderUpdate :459-483)
 * and SwarmAdvector.cxx:261-280 as the time derivative. */
int32_t orc_advect_rk2(orc_problem* pb, const double* u, double dt, double* coord_out) {
  const int dim = pb->dim;
  for (int64_t p = 0; p < pb->P; p++) {
    double x0[3], x[3], v[3];
    memcpy(x0, pb->coord + 3 * p, sizeof(x0));
    memcpy(x, x0, sizeof(x));
    if (!orc_interpolate_velocity(pb, u, x, v)) return ORC_E_ESCAPED;
    for (int d = 0; d < dim; d++) x[d] = x0[d] + 0.5 * dt * v[d];
    if (orc_interpolate_velocity(pb, u, x, v)) {
      for (int d = 0; d < dim; d++) x[d] = x0[d] + dt * v[d];
    } else {
      orc_interpolate_velocity(pb, u, x0, v);
      for (int d = 0; d < dim; d++) x[d] = x0[d] + dt * v[d];
    }
    memcpy(coord_out + 3 * p, x, sizeof(x));
  }
  return ORC_OK;
}
