// gs_generate.cpp -- host-side problem generation: the particle layouts,
// material shapes and wall boundary conditions that the reference builds
// from its XML components before the first solve.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/gale_generate.h"
#include "gs_crand.h"

using gs::CRand;

namespace {

// Park-Miller "minimal standard" with Schrage's factorisation and the
// reference's seed mask (StgDomain/Swarm/src/Random.cxx:45-74).
struct ParkMiller {
  long state;
  explicit ParkMiller(long seed) : state(seed ^ 123459876L) {}
  double next() {
    const long tmp = state / 127773L;
    state = 16807L * (state - tmp * 127773L) - 2836L * tmp;
    if (state < 0) state += 2147483647L;
    return (double)state;
  }
  // Random.cxx:76-78
  double range(double mn, double mx) { return (mx - mn) * next() / (double)2147483647L + mn; }
};

// CartesianGenerator.cxx:2174-2215 (contactDepth 0).
double vertex_coord(int n, double lo, double hi, long long idx) {
  if (idx <= 0) return lo;
  if (idx >= n) return hi;
  return lo + ((double)idx / (double)n) * (hi - lo);
}

}  // namespace

extern "C" double gs_rng_stream(int64_t seed, int64_t count, double* out) {
  ParkMiller g((long)seed);
  double v = 0.0;
  for (int64_t i = 0; i < count; i++) {
    v = g.next();
    if (out) out[i] = v;
  }
  return v;
}

extern "C" int64_t gs_layout_random(int32_t dim, const int32_t* res, const double* lo, const double* hi, int32_t ppc,
                                    int64_t seed, double* coord) {
  if ((dim != 2 && dim != 3) || ppc <= 0 || !res || !lo || !hi) return -1;
  ParkMiller g((long)seed);
  const long long E = (long long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
  long long p = 0;
  for (long long e = 0; e < E; e++) {
    const long long ijk[3] = {e % res[0], (e / res[0]) % res[1], dim == 3 ? e / ((long long)res[0] * res[1]) : 0};
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
      mn[d] = vertex_coord(res[d], lo[d], hi[d], ijk[d]);
      mx[d] = vertex_coord(res[d], lo[d], hi[d], ijk[d] + 1);
    }
    // RandomParticleLayout.cxx:196-216: three draws per particle, also in 2D
    for (int c = 0; c < ppc; c++, p++)
      for (int d = 0; d < 3; d++) {
        const double v = g.range(mn[d], mx[d]);
        if (coord) coord[3 * p + d] = d < dim ? v : 0.0;
      }
  }
  return p;
}

extern "C" int64_t gs_layout_gauss(int32_t dim, const int32_t* res, const double* lo, const double* hi,
                                   double* coord) {
  if ((dim != 2 && dim != 3) || !res || !lo || !hi) return -1;
  const double gp = 1.0 / std::sqrt(3.0);
  const long long E = (long long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
  const int ng = dim == 2 ? 4 : 8;
  long long p = 0;
  for (long long e = 0; e < E; e++) {
    const long long ijk[3] = {e % res[0], (e / res[0]) % res[1], dim == 3 ? e / ((long long)res[0] * res[1]) : 0};
    for (int q = 0; q < ng; q++, p++)
      for (int d = 0; d < 3; d++) {
        double v = 0.0;
        if (d < dim) {
          const double a = vertex_coord(res[d], lo[d], hi[d], ijk[d]);
          const double b = vertex_coord(res[d], lo[d], hi[d], ijk[d] + 1);
          const double xi = ((q >> d) & 1) ? gp : -gp;
          v = a + (1.0 + xi) * (b - a) * 0.5;
        }
        if (coord) coord[3 * p + d] = v;
      }
  }
  return p;
}

// WallVC (StgDomain/Utils/src/WallVC.cxx): every node on the named wall,
// corners and edges included.
namespace {

// Sobol generators: the primitive polynomials and initial direction numbers
// published by Joe & Kuo (ACM TOMS 29(1), 2003, remark on Bratley & Fox's
// Algorithm 659), the first 100 rows in the order of the reference's table
// (StgDomain/Utils/src/SobolGenerator.cxx:307-420).  The layouts draw from
// them in this order, so the rows must match the reference's one for one.
// `coefficient` holds the polynomial's inner coefficients a_1..a_{s-1} as
// bits s-2..0.
struct SobolRow {
  uint8_t degree, coefficient;
  uint16_t m[9];
};
constexpr int kSobolGenerators = 100;
constexpr int kSobolBits = 30;   // numbers are 30-bit binary fractions
const SobolRow kSobolTable[kSobolGenerators] = {
    // 0
    {1, 0, {1}},
    {2, 1, {1, 1}},
    {3, 1, {1, 3, 7}},
    {3, 2, {1, 1, 5}},
    {4, 1, {1, 3, 1, 1}},
    {4, 4, {1, 1, 3, 7}},
    {5, 2, {1, 3, 3, 9, 9}},
    {5, 13, {1, 3, 7, 13, 3}},
    {5, 7, {1, 1, 5, 11, 27}},
    {5, 14, {1, 3, 5, 1, 15}},
    // 10
    {5, 11, {1, 1, 7, 3, 29}},
    {5, 4, {1, 3, 7, 7, 21}},
    {6, 1, {1, 1, 1, 9, 23, 37}},
    {6, 16, {1, 3, 3, 5, 19, 33}},
    {6, 13, {1, 1, 3, 13, 11, 7}},
    {6, 22, {1, 1, 7, 13, 25, 5}},
    {6, 19, {1, 3, 5, 11, 7, 11}},
    {6, 25, {1, 1, 1, 3, 13, 39}},
    {7, 1, {1, 3, 1, 15, 17, 63, 13}},
    {7, 32, {1, 1, 5, 5, 1, 59, 33}},
    // 20
    {7, 4, {1, 3, 3, 3, 25, 17, 115}},
    {7, 8, {1, 1, 7, 15, 29, 15, 41}},
    {7, 7, {1, 3, 1, 7, 3, 23, 79}},
    {7, 56, {1, 3, 7, 9, 31, 29, 17}},
    {7, 14, {1, 1, 5, 13, 11, 3, 29}},
    {7, 28, {1, 1, 1, 9, 5, 21, 119}},
    {7, 19, {1, 1, 3, 1, 23, 13, 75}},
    {7, 50, {1, 3, 7, 11, 27, 31, 73}},
    {7, 21, {1, 1, 7, 7, 19, 25, 105}},
    {7, 42, {1, 3, 1, 5, 21, 9, 7}},
    // 30
    {7, 31, {1, 1, 1, 15, 5, 49, 59}},
    {7, 62, {1, 3, 1, 1, 1, 33, 65}},
    {7, 37, {1, 3, 5, 15, 17, 19, 21}},
    {7, 41, {1, 1, 7, 11, 13, 29, 3}},
    {7, 55, {1, 3, 7, 5, 7, 11, 113}},
    {7, 59, {1, 1, 5, 11, 15, 19, 61}},
    {8, 14, {1, 1, 1, 1, 9, 27, 89, 7}},
    {8, 56, {1, 1, 3, 7, 31, 15, 45, 23}},
    {8, 21, {1, 3, 3, 9, 25, 25, 107, 39}},
    {8, 22, {1, 1, 7, 7, 3, 63, 21, 217}},
    // 40
    {8, 38, {1, 3, 5, 7, 5, 55, 71, 141}},
    {8, 47, {1, 1, 5, 1, 23, 17, 79, 27}},
    {8, 49, {1, 1, 5, 15, 7, 63, 19, 53}},
    {8, 50, {1, 1, 3, 15, 3, 49, 71, 181}},
    {8, 52, {1, 3, 3, 15, 17, 19, 61, 169}},
    {8, 67, {1, 3, 3, 13, 23, 41, 41, 35}},
    {8, 70, {1, 1, 1, 3, 3, 59, 57, 15}},
    {8, 84, {1, 3, 1, 3, 3, 3, 121, 207}},
    {8, 97, {1, 3, 5, 15, 21, 57, 87, 45}},
    {8, 103, {1, 1, 1, 5, 25, 33, 119, 247}},
    // 50
    {8, 115, {1, 1, 1, 9, 25, 49, 55, 185}},
    {8, 122, {1, 3, 5, 7, 23, 53, 85, 117}},
    {9, 8, {1, 3, 3, 13, 11, 57, 121, 41, 235}},
    {9, 13, {1, 1, 3, 3, 19, 57, 119, 81, 307}},
    {9, 16, {1, 3, 3, 7, 3, 39, 11, 223, 495}},
    {9, 22, {1, 3, 3, 5, 11, 21, 23, 151, 417}},
    {9, 25, {1, 3, 1, 11, 31, 7, 61, 81, 57}},
    {9, 44, {1, 1, 3, 9, 7, 53, 11, 189, 151}},
    {9, 47, {1, 3, 7, 1, 9, 9, 35, 61, 19}},
    {9, 52, {1, 1, 5, 9, 5, 55, 33, 95, 119}},
    // 60
    {9, 55, {1, 3, 7, 1, 17, 15, 43, 185, 375}},
    {9, 59, {1, 1, 3, 5, 23, 59, 107, 23, 451}},
    {9, 62, {1, 1, 7, 7, 17, 19, 113, 73, 55}},
    {9, 67, {1, 3, 1, 13, 17, 49, 101, 113, 449}},
    {9, 74, {1, 3, 3, 9, 25, 31, 29, 239, 501}},
    {9, 81, {1, 1, 3, 9, 13, 3, 87, 85, 53}},
    {9, 82, {1, 1, 5, 1, 11, 39, 119, 9, 185}},
    {9, 87, {1, 1, 1, 7, 31, 5, 97, 201, 317}},
    {9, 91, {1, 1, 3, 3, 27, 5, 29, 83, 17}},
    {9, 94, {1, 3, 5, 5, 19, 41, 17, 53, 21}},
    // 70
    {9, 103, {1, 1, 5, 1, 17, 9, 89, 183, 487}},
    {9, 104, {1, 1, 7, 11, 23, 19, 5, 203, 13}},
    {9, 109, {1, 3, 7, 11, 7, 9, 127, 91, 347}},
    {9, 122, {1, 1, 7, 13, 5, 57, 89, 149, 393}},
    {9, 124, {1, 1, 1, 7, 11, 25, 119, 101, 15}},
    {9, 137, {1, 1, 1, 7, 19, 1, 117, 13, 391}},
    {9, 138, {1, 3, 3, 9, 19, 15, 103, 111, 307}},
    {9, 143, {1, 3, 3, 9, 7, 51, 105, 239, 189}},
    {9, 145, {1, 1, 1, 1, 13, 11, 41, 3, 381}},
    {9, 152, {1, 3, 1, 1, 21, 19, 83, 205, 71}},
    // 80
    {9, 157, {1, 3, 5, 3, 21, 61, 25, 253, 163}},
    {9, 167, {1, 1, 1, 9, 7, 53, 41, 247, 99}},
    {9, 173, {1, 3, 5, 15, 9, 29, 55, 121, 467}},
    {9, 176, {1, 3, 7, 1, 11, 19, 69, 189, 167}},
    {9, 181, {1, 3, 5, 5, 1, 11, 117, 169, 433}},
    {9, 182, {1, 1, 1, 13, 5, 9, 49, 179, 337}},
    {9, 185, {1, 3, 7, 1, 21, 21, 127, 197, 257}},
    {9, 191, {1, 3, 5, 9, 11, 19, 29, 175, 179}},
    {9, 194, {1, 3, 3, 9, 13, 43, 1, 217, 47}},
    {9, 199, {1, 1, 3, 9, 25, 13, 99, 249, 385}},
    // 90
    {9, 218, {1, 3, 1, 9, 9, 13, 53, 195, 23}},
    {9, 220, {1, 3, 5, 9, 7, 41, 83, 95, 117}},
    {9, 227, {1, 1, 7, 13, 7, 25, 15, 63, 369}},
    {9, 229, {1, 3, 1, 11, 27, 31, 31, 19, 425}},
    {9, 230, {1, 3, 7, 3, 15, 9, 73, 7, 207}},
    {9, 234, {1, 3, 5, 5, 25, 11, 115, 5, 433}},
    {9, 236, {1, 1, 1, 11, 15, 19, 35, 75, 301}},
    {9, 241, {1, 3, 7, 11, 21, 5, 21, 217, 147}},
    {9, 244, {1, 1, 3, 13, 17, 53, 89, 245, 333}},
    {9, 253, {1, 3, 1, 5, 19, 37, 5, 111, 85}},
};

// Direction numbers of one generator as 30-bit fractions, V[b] = v_b << (29 - b),
// from v_i = 2 a_1 v_{i-1} ^ 4 a_2 v_{i-2} ^ ... ^ 2^s v_{i-s} ^ v_{i-s}
// (SobolGenerator.cxx:198-244 evaluates the same recurrence recursively).
void sobol_directions(int gen, uint32_t V[kSobolBits]) {
  const SobolRow& r = kSobolTable[gen];
  const int s = r.degree;
  uint32_t v[kSobolBits];
  for (int i = 0; i < kSobolBits; i++) {
    if (i < s) {
      v[i] = r.m[i];
    } else {
      uint32_t x = v[i - s] ^ (v[i - s] << s);
      for (int j = 1; j < s; j++)
        if ((r.coefficient >> (s - 1 - j)) & 1u) x ^= v[i - j] << j;
      v[i] = x;
    }
    V[i] = v[i] << (kSobolBits - 1 - i);
  }
}

// The n-th number of the sequence (n >= 1; SobolGenerator.cxx:271-294 with
// n = index + 1): XOR of V[b] over the set bits of the Gray code of n.
uint32_t sobol_at(const uint32_t V[kSobolBits], uint64_t n) {
  const uint64_t g = n ^ (n >> 1);
  uint32_t x = 0;
  for (int b = 0; b < kSobolBits; b++)
    if ((g >> b) & 1u) x ^= V[b];
  return x;
}

// Numbers n0+1 .. n0+count of one generator written with stride `stride` as
// (hi - lo) * u + lo.  Successive Gray codes differ in the lowest zero bit of
// n, so x_{n+1} = x_n ^ V[ctz(~n)] (SobolGenerator.cxx:246-269).
void sobol_fill(const uint32_t V[kSobolBits], uint64_t n0, int64_t count, double lo, double hi, double* out,
                int64_t stride) {
#pragma clang fp contract(off)
  uint32_t x = sobol_at(V, n0);
  const double span = hi - lo;
  for (int64_t k = 0; k < count; k++) {
    const uint64_t n = n0 + (uint64_t)k;
    x ^= V[__builtin_ctzll(~n)];
    const double u = (double)x / (double)(1u << kSobolBits);
    const double t = span * u;
    out[k * stride] = t + lo;
  }
}

}  // namespace

extern "C" int32_t gs_sobol_sequence(int32_t generator, int64_t first, int64_t count, double* out) {
  if (generator < 0 || generator >= kSobolGenerators || first < 0 || count < 0 || !out ||
      first + count >= (1LL << kSobolBits))
    return -1;
  uint32_t V[kSobolBits];
  sobol_directions(generator, V);
  sobol_fill(V, (uint64_t)first, count, 0.0, 1.0, out, 1);
  return 0;
}

extern "C" int64_t gs_layout_spacefiller(int32_t dim, const int32_t* res, const double* lo, const double* hi,
                                         int64_t total, double avg_per_cell, int32_t first_generator,
                                         double* coord) {
  if ((dim != 2 && dim != 3) || !res || !lo || !hi) return -1;
  if ((total > 0) == (avg_per_cell > 0.0) || total < 0 || avg_per_cell < 0.0) return -1;
  if (first_generator < 0 || first_generator + dim > kSobolGenerators) return -1;
  const long long E = (long long)res[0] * res[1] * (dim == 3 ? res[2] : 1);
  // GlobalParticleLayout.cxx:_SetInitialCounts stores cells * average in an unsigned int
  const double want = total > 0 ? (double)total : (double)E * avg_per_cell;
  if (want >= (double)((1LL << kSobolBits) - 1)) return -1;
  const int64_t P = total > 0 ? total : (int64_t)(uint32_t)want;
  if (!coord) return P;
  uint32_t V[3][kSobolBits];
  for (int d = 0; d < dim; d++) sobol_directions(first_generator + d, V[d]);
  const int64_t kChunk = 1 << 20;
  const int64_t chunks = (P + kChunk - 1) / kChunk;
  const int nthreads = (int)std::min<int64_t>(chunks, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  auto work = [&](int t) {
    for (int64_t c = t; c < chunks; c += nthreads) {
      const int64_t p0 = c * kChunk, n = std::min(kChunk, P - p0);
      for (int d = 0; d < dim; d++) sobol_fill(V[d], (uint64_t)p0, n, lo[d], hi[d], coord + 3 * p0 + d, 3);
      if (dim == 2)
        for (int64_t k = 0; k < n; k++) coord[3 * (p0 + k) + 2] = 0.0;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; t++) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  return P;
}

extern "C" int32_t gs_wall_bc(int32_t dim, const int32_t* res, int32_t wall, int32_t comp, double value,
                              uint8_t* mask, double* values) {
  if ((dim != 2 && dim != 3) || comp < 0 || comp >= dim || wall < 0 || wall > 5 || !mask) return 6;
  const int axis = wall / 2;
  if (axis >= dim) return 6;
  const long long nn[3] = {res[0] + 1LL, res[1] + 1LL, dim == 3 ? res[2] + 1LL : 1LL};
  const long long target = (wall % 2 == 0) ? 0 : nn[axis] - 1;
  const long long N = nn[0] * nn[1] * nn[2];
  for (long long n = 0; n < N; n++) {
    const long long ijk[3] = {n % nn[0], (n / nn[0]) % nn[1], n / (nn[0] * nn[1])};
    if (ijk[axis] != target) continue;
    mask[n * dim + comp] = 1;
    if (values) values[n * dim + comp] = value;
  }
  return 0;
}

namespace {

// Stg_Shape_IsCoordInside for the shapes of gs_shape; -1 for an unknown type.
int shape_contains(int dim, const gs_shape& sh, const double* x) {
  switch (sh.type) {
    case GS_SHAPE_EVERYWHERE:
      return 1;
    case GS_SHAPE_SPHERE: {  // Sphere.cxx:189-208
      double v = 0.0;
      for (int d = 0; d < dim; d++) {
        const double t = x[d] - sh.p[d];
        v += t * t;
      }
      return v <= sh.p[3] * sh.p[3];
    }
    case GS_SHAPE_BOX: {  // Box.cxx:209-221: |x - centre| <= width/2 on every axis
      for (int d = 0; d < dim; d++) {
        const double c = 0.5 * (sh.p[d] + sh.p[3 + d]);
        const double w = sh.p[3 + d] - sh.p[d];
        if (std::fabs(x[d] - c) > 0.5 * w) return 0;
      }
      return 1;
    }
    case GS_SHAPE_BELOW_SINES: {
      // below the surface x_v = p[0] + sum_k amp_k sin(2 pi (kx x + kz z) + phase_k)
      const int v = sh.axis;
      double level = sh.p[0];
      for (int k = 0; k < sh.n_modes && k < 8; k++) {
        const double* m = sh.modes + 4 * k;
        const int a0 = v == 0 ? 1 : 0;
        const int a1 = dim == 3 ? (v == 2 ? 1 : 2) : -1;
        double arg = m[1] * x[a0];
        if (a1 >= 0) arg += m[2] * x[a1];
        level += m[0] * std::sin(2.0 * M_PI * arg + m[3]);
      }
      return x[v] < level;
    }
    default:
      return -1;
  }
}


}  // namespace

// Material layout: for each shape in order, particles inside take its
// material (Material.cxx:210-228, the last registered material wins).
extern "C" int32_t gs_shape_layout(int32_t dim, int64_t P, const double* coord, int32_t n_shapes,
                                   const gs_shape* shapes, int32_t* material) {
  if (!coord || !material || (n_shapes > 0 && !shapes)) return 6;
  for (int s = 0; s < n_shapes; s++) {
    const gs_shape& sh = shapes[s];
    for (int64_t p = 0; p < P; p++) {
      const int in = shape_contains(dim, sh, coord + 3 * p);
      if (in < 0) return 6;
      if (in) material[p] = sh.material;
    }
  }
  return 0;
}

extern "C" int32_t gs_crand_stream(int64_t seed, int64_t count, int32_t* out) {
  if (seed < 0 || seed > 0xffffffffLL || count < 0 || (count > 0 && !out)) return 6;
  CRand g((unsigned)seed);
  for (int64_t i = 0; i < count; i++) out[i] = g.next();
  return 0;
}

// _StrainWeakening_Initialise (StrainWeakening.cxx:333-430), the initial
// damage loop: particles in order; with probability "initialDamageFraction"
// (one draw) a particle inside the optional shape gets
// initialDamageFactor * draw * finalSofteningStrain / RAND_MAX, modulated by
// the squared half-cosine / half-sine profiles of the wavenumbers that are set.
extern "C" int32_t gs_initial_damage(int32_t dim, int64_t P, const double* coord, const gs_initial_damage_params* d,
                                     const gs_shape* shape, double* w) {
#pragma clang fp contract(off)
  if ((dim != 2 && dim != 3) || P < 0 || !d || (P > 0 && (!coord || !w))) return 6;
  if (d->seed < 0 || d->seed > 0xffffffffLL) return 6;
  if (shape && shape_contains(dim, *shape, coord) < 0 && P > 0) return 6;
  const double kRandMax = 2147483647.0;
  CRand g((unsigned)d->seed);
  for (int64_t p = 0; p < P; p++) {
    const double* x = coord + 3 * p;
    double v = 0.0;
    if ((double)g.next() < kRandMax * d->fraction) {
      if (shape && !shape_contains(dim, *shape, x)) {
        w[p] = 0.0;
        continue;
      }
      const double a = d->factor * (double)g.next();
      const double b = a * d->final_softening_strain;
      v = b / kRandMax;
      if (d->wavenumber > 0.0 && d->wavenumber_cos_i == -1.0)
        v *= std::pow(0.5 + 0.5 * std::cos(M_PI * x[0] * d->wavenumber), 2.0);
      if (d->wavenumber_cos_i > 0.0) v *= std::pow(0.5 + 0.5 * std::cos(M_PI * x[0] * d->wavenumber_cos_i), 2.0);
      if (d->wavenumber_sin_i > 0.0) v *= std::pow(0.5 + 0.5 * std::sin(M_PI * x[0] * d->wavenumber_sin_i), 2.0);
      if (d->wavenumber_cos_k > 0.0) v *= std::pow(0.5 + 0.5 * std::cos(M_PI * x[2] * d->wavenumber_cos_k), 2.0);
      if (d->wavenumber_sin_k > 0.0) v *= std::pow(0.5 + 0.5 * std::sin(M_PI * x[2] * d->wavenumber_sin_k), 2.0);
    }
    w[p] = v;
  }
  return 0;
}

// Gaussian random field on the nodes of a regular grid: Park-Miller white
// noise (Box-Muller), filtered by a separable Gaussian of standard deviation
// corr/sqrt(2) (covariance exp(-d^2 / (2 corr^2))), then min-max scaled to
// [vmin, vmax].
extern "C" int32_t gs_gaussian_field(int32_t dim, const int32_t* nodes, const double* spacing, double corr,
                                     int64_t seed, double vmin, double vmax, double* out) {
  if ((dim != 2 && dim != 3) || !nodes || !spacing || !out || corr <= 0) return 6;
  const long long n[3] = {nodes[0], nodes[1], dim == 3 ? nodes[2] : 1};
  const long long N = n[0] * n[1] * n[2];
  std::vector<double> a(N), b(N);
  ParkMiller g((long)seed);
  for (long long i = 0; i < N; i += 2) {
    const double u1 = (g.next() + 1.0) / 2147483648.0, u2 = g.next() / 2147483647.0;
    const double rr = std::sqrt(-2.0 * std::log(u1));
    a[i] = rr * std::cos(2.0 * M_PI * u2);
    if (i + 1 < N) a[i + 1] = rr * std::sin(2.0 * M_PI * u2);
  }
  const double sig = corr / std::sqrt(2.0);
  for (int d = 0; d < dim; d++) {
    const int hw = (int)std::ceil(3.0 * sig / spacing[d]);
    std::vector<double> w(2 * hw + 1);
    for (int k = -hw; k <= hw; k++) {
      const double t = k * spacing[d] / sig;
      w[k + hw] = std::exp(-0.5 * t * t);
    }
    const long long stride = d == 0 ? 1 : (d == 1 ? n[0] : n[0] * n[1]);
    for (long long i = 0; i < N; i++) {
      const long long c = (i / stride) % n[d];
      double s = 0.0, ws = 0.0;
      for (int k = -hw; k <= hw; k++) {
        const long long cc = c + k;
        if (cc < 0 || cc >= n[d]) continue;
        s += w[k + hw] * a[i + k * stride];
        ws += w[k + hw];
      }
      b[i] = s / ws;
    }
    a.swap(b);
  }
  double lo = a[0], hi = a[0];
  for (long long i = 0; i < N; i++) {
    lo = std::fmin(lo, a[i]);
    hi = std::fmax(hi, a[i]);
  }
  for (long long i = 0; i < N; i++) out[i] = hi > lo ? vmin + (a[i] - lo) * (vmax - vmin) / (hi - lo) : vmin;
  return 0;
}
