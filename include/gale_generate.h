/*
 * gale_generate.h -- host-side problem generation for the Stokes hot path:
 * particle layouts, wall boundary conditions and material shapes, the parts
 * of the reference's XML component graph that run once before the first
 * solve.  Exported by libgale_stokes.so next to gale_stokes.h.
 */
#ifndef GALE_GENERATE_H
#define GALE_GENERATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Park-Miller stream with the reference's seed mask; replaces
 * Swarm_Random_Seed / Swarm_Random_Random (StgDomain/Swarm/src/Random.cxx:45-74).
 * Writes count raw draws to out (may be NULL) and returns the last one. */
double gs_rng_stream(int64_t seed, int64_t count, double* out);

/* RandomParticleLayout (StgDomain/Swarm/src/RandomParticleLayout.cxx:196-216)
 * driven cell by cell as PerCellParticleLayout (PerCellParticleLayout.cxx:147-168):
 * ppc particles per element, coordinates P*3.  Returns P (or -1). */
int64_t gs_layout_random(int32_t dim, const int32_t* res, const double* lo, const double* hi, int32_t ppc,
                         int64_t seed, double* coord);

/* GaussParticleLayout with 2 points per axis (GaussParticleLayout.cxx:175-180)
 * placed as global coordinates; with ConstantWeights they carry weight 1. */
int64_t gs_layout_gauss(int32_t dim, const int32_t* res, const double* lo, const double* hi, double* coord);

/* Sobol sequence of tabulated generator `generator` (0..99, the rows of
 * StgDomain/Utils/src/SobolGenerator.cxx:307-420): out[k] is the number the
 * reference's SobolGenerator_GetNumberByIndex(first + k) returns, i.e. the
 * (first+k+1)-th SobolGenerator_GetNextNumber (SobolGenerator.cxx:246-294).
 * Needs first + count < 2^30.  Returns 0, or -1 on bad arguments. */
int32_t gs_sobol_sequence(int32_t generator, int64_t first, int64_t count, double* out);

/* SpaceFillerParticleLayout (StgDomain/Swarm/src/SpaceFillerParticleLayout.cxx:88-113,
 * :211-240) through GlobalParticleLayout (GlobalParticleLayout.cxx
 * _SetInitialCounts / _InitialiseParticles): exactly one of total > 0 or
 * avg_per_cell > 0 (then P = (unsigned)(cells * avg_per_cell)); particle p's
 * coordinate d is (hi[d] - lo[d]) * u + lo[d], u the (p+1)-th number of
 * generator first_generator + d (the reference creates its x, y[, z]
 * generators from the table in construction order: 0, 1, 2 for the first
 * layout of a run).  Coordinates P*3 in generation order, the third 0 in 2D;
 * coord NULL returns P only.  Returns P, or -1 on bad arguments. */
int64_t gs_layout_spacefiller(int32_t dim, const int32_t* res, const double* lo, const double* hi, int64_t total,
                              double avg_per_cell, int32_t first_generator, double* coord);

/* WallVC (StgDomain/Utils/src/WallVC.cxx): wall 0 left, 1 right, 2 bottom,
 * 3 top, 4 back, 5 front; marks component comp of every node on the wall in
 * mask (N*dim) and writes value into values (N*dim, may be NULL). */
int32_t gs_wall_bc(int32_t dim, const int32_t* res, int32_t wall, int32_t comp, double value, uint8_t* mask,
                   double* values);

enum {
  GS_SHAPE_EVERYWHERE = 0,   /* Everywhere */
  GS_SHAPE_SPHERE = 1,       /* Sphere (Sphere.cxx:189-208): p[0..2] centre, p[3] radius */
  GS_SHAPE_BOX = 2,          /* Box (Box.cxx:209-221): p[0..2] start, p[3..5] end */
  GS_SHAPE_BELOW_SINES = 3   /* below x_axis = p[0] + sum_k modes[4k] sin(2pi(modes[4k+1] a + modes[4k+2] b) + modes[4k+3]) */
};

typedef struct gs_shape {
  int32_t type;
  int32_t material;
  int32_t axis;       /* vertical axis for GS_SHAPE_BELOW_SINES */
  int32_t n_modes;
  double p[6];
  double modes[32];
} gs_shape;

/* Material_Layout (PICellerator/MaterialPoints/src/Material.cxx:210-228):
 * shapes are applied in order and the last one containing a particle sets
 * its material. */
int32_t gs_shape_layout(int32_t dim, int64_t P, const double* coord, int32_t n_shapes, const gs_shape* shapes,
                        int32_t* material);

/* The C library's rand() sequence after srand(seed) (glibc's generator, the
 * one the reference's StrainWeakening initial damage draws from): count
 * values into out.  Returns 0 or 6 (bad arguments). */
int32_t gs_crand_stream(int64_t seed, int64_t count, int32_t* out);

/* StrainWeakening initial damage (StrainWeakening.cxx:232-290 keys,
 * _StrainWeakening_Initialise :333-430).  A wavenumber <= 0 is off (-1 is
 * the reference's default); `wavenumber` applies only when wavenumber_cos_i
 * is -1, as in the reference. */
typedef struct gs_initial_damage_params {
  double fraction;           /* "initialDamageFraction" */
  double factor;             /* "initialDamageFactor" (1) */
  double final_softening_strain; /* "finalSofteningStrain" */
  double wavenumber;         /* "initialDamageWavenumber" */
  double wavenumber_sin_i;   /* "initialDamageWavenumberSinI" */
  double wavenumber_cos_i;   /* "initialDamageWavenumberCosI" */
  double wavenumber_sin_k;   /* "initialDamageWavenumberSinK" */
  double wavenumber_cos_k;   /* "initialDamageWavenumberCosK" */
  int64_t seed;              /* "randomSeed" (srand argument) */
} gs_initial_damage_params;

/* Initial plastic strain of P particles (coordinates P*3, in the order the
 * layout created them) into w; shape (may be NULL) is "initialStrainShape".
 * Feed w to gs_problem_set_plastic_strain.  Returns 0 or 6. */
int32_t gs_initial_damage(int32_t dim, int64_t P, const double* coord, const gs_initial_damage_params* d,
                          const gs_shape* shape, double* w);

/* Seeded Gaussian random field on a node grid (white noise filtered by a
 * separable Gaussian with covariance length corr), min-max scaled to
 * [vmin, vmax].  No reference counterpart: used to build viscosity fields. */
int32_t gs_gaussian_field(int32_t dim, const int32_t* nodes, const double* spacing, double corr, int64_t seed,
                          double vmin, double vmax, double* out);

#ifdef __cplusplus
}
#endif
#endif
