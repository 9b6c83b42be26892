/*
 * A plain C caller of the C ABI (include/gale_stokes.h, gale_generate.h): the
 * 2D sinker of the reference's examples -- unit square, 32 x 32 Q1-P0, free
 * slip, background rho 1 eta 1, a disc of radius 0.1 at (0.5, 0.6) with
 * rho 1.1 eta 1e3, 20 random particles per cell (Park-Miller seed 13), one
 * Uzawa solve to 1e-6, then one RK2 step with dt = 0.5 x CFL.  This is what a
 * StGermain-side component does through the library, in C only.
 *
 * Build:  gcc -std=c11 -I include tests/c_abi/sinker_c1.c -L gale_amd -lgale_stokes -Wl,-rpath,<dir> -lm
 * Exit status: 0 on success, 3 when no GPU is visible (after checking the
 * no-device error path), 1 on any other failure.  Prints one summary line.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gale_generate.h"
#include "gale_stokes.h"

#define RES 32
#define PPC 20

static int fail(const char* what, int32_t st) {
  fprintf(stderr, "sinker_c1: %s: %s (%d)\n", what, gs_strerror(st), (int)st);
  return 1;
}

int main(void) {
  const int32_t dim = 2, res[3] = {RES, RES, 1};
  const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {1.0, 1.0, 1.0};
  const int64_t N = (int64_t)(RES + 1) * (RES + 1), E = (int64_t)RES * RES;
  const int64_t P = gs_layout_random(dim, res, lo, hi, PPC, 13, NULL);
  if (P != E * PPC) return fail("gs_layout_random (count)", GS_E_INVALID);
  double* coord = malloc(sizeof(double) * 3 * P);
  int32_t* mat = malloc(sizeof(int32_t) * P);
  uint8_t* mask = calloc((size_t)N * dim, 1);
  double* vals = calloc((size_t)N * dim, sizeof(double));
  double* u = malloc(sizeof(double) * N * dim);
  double* p = malloc(sizeof(double) * E);
  if (!coord || !mat || !mask || !vals || !u || !p) return 1;
  if (gs_layout_random(dim, res, lo, hi, PPC, 13, coord) != P) return fail("gs_layout_random", GS_E_INVALID);
  gs_shape shapes[2];
  memset(shapes, 0, sizeof shapes);
  shapes[0].type = GS_SHAPE_EVERYWHERE;
  shapes[0].material = 0;
  shapes[1].type = GS_SHAPE_SPHERE;
  shapes[1].material = 1;
  shapes[1].p[0] = 0.5;
  shapes[1].p[1] = 0.6;
  shapes[1].p[3] = 0.1;
  int32_t st = gs_shape_layout(dim, P, coord, 2, shapes, mat);
  if (st != GS_OK) return fail("gs_shape_layout", st);
  /* free slip: the normal component fixed to 0 on every wall */
  for (int32_t w = 0; w < 4; w++)
    if ((st = gs_wall_bc(dim, res, w, w / 2, 0.0, mask, vals)) != GS_OK) return fail("gs_wall_bc", st);

  gs_material mats[2];
  memset(mats, 0, sizeof mats);
  mats[0].eta0 = 1.0;
  mats[0].density = 1.0;
  mats[1].eta0 = 1e3;
  mats[1].density = 1.1;
  gs_problem_desc d;
  memset(&d, 0, sizeof d);
  d.dim = dim;
  for (int i = 0; i < 3; i++) {
    d.res[i] = res[i];
    d.lo[i] = lo[i];
    d.hi[i] = hi[i];
  }
  d.bc_mask = mask;
  d.bc_value = vals;
  d.gravity = 1.0;
  d.ghat[1] = -1.0;
  d.n_materials = 2;
  d.materials = mats;
  d.n_particles = P;
  d.coord = coord;
  d.material = mat;

  if (gs_device_count() <= 0) {
    /* no GPU: the library must refuse, not fall back */
    gs_problem* none = NULL;
    st = gs_problem_create(&d, &none);
    if (st != GS_E_NO_DEVICE || none != NULL) return fail("gs_problem_create without a device", st);
    printf("sinker_c1: no HIP device; gs_problem_create refused with \"%s\"\n", gs_strerror(st));
    return 3;
  }
  gs_problem* pb = NULL;
  if ((st = gs_problem_create(&d, &pb)) != GS_OK) return fail("gs_problem_create", st);
  gs_uzawa_params up;
  memset(&up, 0, sizeof up); /* zero = library defaults */
  up.tolerance = 1e-6;
  up.max_iterations = 1000;
  up.min_iterations = 1;
  up.preconditioner = GS_PC_DIAGONAL;
  up.precision = GS_PREC_FP64;
  up.inner_rtol_factor = 0.1;
  up.inner_preconditioner = GS_INNER_MULTIGRID;
  gs_solver* s = NULL;
  if ((st = gs_solver_create(pb, &up, &s)) != GS_OK) return fail("gs_solver_create", st);
  gs_report rep;
  if ((st = gs_stokes_solve(s, u, p, &rep)) != GS_OK) return fail("gs_stokes_solve", st);
  double dt = 0.0;
  if ((st = gs_compute_dt(s, 0.5, &dt)) != GS_OK) return fail("gs_compute_dt", st);
  if ((st = gs_advect_rk2(s, dt)) != GS_OK) return fail("gs_advect_rk2", st);
  int64_t info[5];
  if ((st = gs_problem_info(pb, info)) != GS_OK) return fail("gs_problem_info", st);
  double vmax = 0.0;
  for (int64_t i = 0; i < N; i++) {
    const double v = sqrt(u[2 * i] * u[2 * i] + u[2 * i + 1] * u[2 * i + 1]);
    if (v > vmax) vmax = v;
  }
  printf("sinker_c1: uzawa %d its, rel residual %.3e, div %.2e, |v|max %.6e, dt %.6e, particles %lld, escaped %lld\n",
         (int)rep.uzawa_iterations, rep.rel_residual, rep.div_u, vmax, dt, (long long)info[2], (long long)info[4]);
  gs_solver_destroy(s);
  gs_problem_destroy(pb);
  free(coord); free(mat); free(mask); free(vals); free(u); free(p);
  return (rep.rel_residual <= 1e-6 && info[4] == 0) ? 0 : 1;
}
