// gs_generate.cpp -- host-side problem generation: the particle layouts,
// material shapes and wall boundary conditions that the reference builds
// from its XML components before the first solve.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/gale_generate.h"

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

// Material layout: for each shape in order, particles inside take its
// material (Material.cxx:210-228, the last registered material wins).
extern "C" int32_t gs_shape_layout(int32_t dim, int64_t P, const double* coord, int32_t n_shapes,
                                   const gs_shape* shapes, int32_t* material) {
  if (!coord || !material || (n_shapes > 0 && !shapes)) return 6;
  for (int s = 0; s < n_shapes; s++) {
    const gs_shape& sh = shapes[s];
    for (int64_t p = 0; p < P; p++) {
      const double* x = coord + 3 * p;
      bool in = false;
      switch (sh.type) {
        case GS_SHAPE_EVERYWHERE:
          in = true;
          break;
        case GS_SHAPE_SPHERE: {  // Sphere.cxx:189-208
          double v = 0.0;
          for (int d = 0; d < dim; d++) {
            const double t = x[d] - sh.p[d];
            v += t * t;
          }
          in = v <= sh.p[3] * sh.p[3];
          break;
        }
        case GS_SHAPE_BOX: {  // Box.cxx:209-221: |x - centre| <= width/2 on every axis
          in = true;
          for (int d = 0; d < dim; d++) {
            const double c = 0.5 * (sh.p[d] + sh.p[3 + d]);
            const double w = sh.p[3 + d] - sh.p[d];
            if (std::fabs(x[d] - c) > 0.5 * w) in = false;
          }
          break;
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
          in = x[v] < level;
          break;
        }
        default:
          return 6;
      }
      if (in) material[p] = sh.material;
    }
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
