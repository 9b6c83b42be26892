// gs_net.hip -- BF16 matrix-core kernels of the learned Schur-complement
// preconditioner network (forward, data gradient, weight gradient), its
// pooling / upsampling / loss kernels and the Adam optimiser.
//
// Convolutions are implicit GEMMs on the BF16 matrix cores: a wavefront
// computes 16 consecutive voxels (flat index over batch x grid) times all
// output channels; the K dimension is the flattened (tap, input channel)
// index kk = tap*Cin + ci, consumed KW at a time -- 32 with gfx950's
// v_mfma_f32_16x16x32_bf16 wherever Cin is a multiple of 8 (the forward
// passes), else 16 with v_mfma_f32_16x16x16_bf16.  Lane l holds the KW/4 K
// values KW/4 (l/16) .. of a step for row l%16 of A and column l%16 of B, and
// receives D[4(l/16)+i][l%16], i = 0..3 (a product over K does not depend on
// the order in which A and B share K).  Weights are pre-packed into exactly
// that per-lane B order, so a lane's B operand is one 8- or 16-byte load;
// with Cin a multiple of KW/4 a lane's A values are consecutive channels of
// one neighbour voxel (one load of the same width).
#include "gs_net.h"

#include <math.h>
#include <string.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/gale_stokes.h"
#include "gs_reduce.h"

namespace gs {
namespace net {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// One K step of the implicit GEMM (16 or 32 K values) on the matrix cores.
__device__ __forceinline__ f4 mfma_bf16(const s4& a, const s4& b, const f4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma_bf16(const s8& a, const s8& b, const f4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a), __builtin_bit_cast(bf8v, b), c, 0, 0,
                                                 0);
}
template <int KW>
struct KOp {
  typedef s4 type;
};
template <>
struct KOp<32> {
  typedef s8 type;
};

static LayerSpec make_layer(int l) {
  // encoder / bottom
  if (l < 2 * kLevels) {
    const int lev = l / 2;
    const bool a = (l % 2) == 0;
    if (a) {
      const int cin = lev == 0 ? 4 : kChan[lev - 1];
      return LayerSpec{cin, kChan[lev], lev, 1, lev == 0 ? 2 : cin};
    }
    const int cout = lev == kLevels - 1 ? kChan[kLevels - 2] : kChan[lev];
    return LayerSpec{kChan[lev], cout, lev, 1, kChan[lev]};
  }
  // decoder
  const int k = l - 2 * kLevels;
  const int lev = kLevels - 2 - k / 2;
  if ((k % 2) == 0) return LayerSpec{kChan[lev], kChan[lev], lev, 1, kChan[lev]};
  return lev == 0 ? LayerSpec{kChan[0], 1, 0, 0, kChan[0]} : LayerSpec{kChan[lev], kChan[lev - 1], lev, 1, kChan[lev]};
}

static const LayerSpec* build_spec() {
  static LayerSpec t[kLayers];
  for (int l = 0; l < kLayers; l++) t[l] = make_layer(l);
  return t;
}
const LayerSpec* const kSpec = build_spec();

long long weight_offset(int l) {
  long long o = 0;
  for (int i = 0; i < l; i++) o += 27LL * kSpec[i].cin * kSpec[i].cout + kSpec[i].cout;
  return o;
}
long long bias_offset(int l) { return weight_offset(l) + 27LL * kSpec[l].cin * kSpec[l].cout; }
long long num_params() { return weight_offset(kLayers); }

static inline int kchunks(int cin) { return (27 * cin + 15) / 16; }
static inline int kchunks32(int cin) { return (27 * cin + 31) / 32; }
static inline int ntiles(int cout) { return (cout + 15) / 16; }

#define NCK(x)                                                                                        \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess) {                                                                           \
      std::fprintf(stderr, "gale_stokes(net): HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, \
                   __LINE__);                                                                         \
      return GS_E_HIP;                                                                                \
    }                                                                                                 \
  } while (0)

// ---------------------------------------------------------------------------
// BF16 helpers (round to nearest even; NaN kept quiet)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void vox_xyz(long long vox, long long V, int nx, int ny, long long* b, int* x, int* y,
                                        int* z) {
  const long long bb = vox / V;
  const long long v = vox - bb * V;
  *b = bb;
  *x = (int)(v % nx);
  const long long t = v / nx;
  *y = (int)(t % ny);
  *z = (int)(t / ny);
}

// ---------------------------------------------------------------------------
// Convolution (forward and data gradient share this kernel)
// ---------------------------------------------------------------------------
// out[vox][co] = act(bias[co] + sum_kk in[nbr(vox, tap(kk))][ci(kk)] * W[kk][co])
// Wp: packed [KC][NT][64 lanes] of KW/4 BF16.  Grid: ceil(BV/64) blocks x 256.
template <int CIN, int NT, bool OUT_F32, int KW>
__global__ __launch_bounds__(256) void k_conv(const uint16_t* __restrict__ in,
                                              const typename KOp<KW>::type* __restrict__ Wp,
                                              const float* __restrict__ bias, void* __restrict__ out,
                                              int cout_store, int nx, int ny, int nz, long long BV, int relu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long v0 = ((long long)blockIdx.x * 4 + wave) * 16;
  if (v0 >= BV) return;
  const int row = lane & 15, kg = lane >> 4;
  const long long V = (long long)nx * ny * nz;
  const long long vox = v0 + row;
  const bool valid = vox < BV;
  long long b = 0;
  int x = 0, y = 0, z = 0;
  if (valid) vox_xyz(vox, V, nx, ny, &b, &x, &y, &z);
  const long long bbase = b * V;
  f4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  typedef typename KOp<KW>::type AT;
  constexpr int KT = 27 * CIN;
  constexpr int KC = (KT + KW - 1) / KW;
  // K steps in groups of U: all operand loads of a group are issued before
  // its MFMAs, so a wave keeps U x (1 + NT) loads in flight instead of
  // waiting on one load per step.
  constexpr int U = KC < 9 ? KC : 9;
  for (int kc0 = 0; kc0 < KC; kc0 += U) {
    AT a[U];
    AT w[U][NT];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int kc = kc0 + u;
      const int kk = kc * KW + kg * (KW / 4);
      a[u] = AT{};
      if (valid && kc < KC && kk < KT) {
        const int tap = kk / CIN;
        const int ci = kk - tap * CIN;
        const int xx = x + tap % 3 - 1, yy = y + (tap / 3) % 3 - 1, zz = z + tap / 9 - 1;
        if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz) {
          const long long nb = bbase + xx + (long long)nx * (yy + (long long)ny * zz);
          a[u] = *reinterpret_cast<const AT*>(in + nb * CIN + ci);
        }
      }
#pragma unroll
      for (int t = 0; t < NT; t++) w[u][t] = kc < KC ? Wp[(kc * NT + t) * 64 + lane] : AT{};
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[t] = mfma_bf16(a[u], w[u][t], acc[t]);
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int co = t * 16 + row;
    if (co >= cout_store) continue;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const long long vo = v0 + kg * 4 + i;
      if (vo >= BV) continue;
      float val = acc[t][i] + bv;
      if (relu) val = fmaxf(val, 0.f);
      if (OUT_F32)
        reinterpret_cast<float*>(out)[vo * cout_store + co] = val;
      else
        reinterpret_cast<uint16_t*>(out)[vo * cout_store + co] = f2bf(val);
    }
  }
}

// Weight gradient: P[chunk][kk][co] = sum_{vox in chunk} X[nbr(vox,tap)][ci] * G[vox][co]
// A = X (rows kk, K = voxels), B = G (K = voxels, columns co).  One wavefront
// per 16-row kk tile and voxel chunk; grid (ceil(KC/4), nchunks) x 256.
// Split-K form of k_conv for the small levels, where a wave per 16 voxels
// leaves most of the chip idle and each wave would walk all 27 x CIN / 16 K
// steps in sequence: S waves share one 16-voxel tile, wave w taking K steps
// in groups of four interleaved with the others, and their partial sums are
// added in wave order through LDS.
template <int CIN, int NT, bool OUT_F32, int S, int KW>
__global__ __launch_bounds__(64 * S) void k_conv_splitk(const uint16_t* __restrict__ in,
                                                        const typename KOp<KW>::type* __restrict__ Wp,
                                                        const float* __restrict__ bias, void* __restrict__ out,
                                                        int cout_store, int nx, int ny, int nz, long long BV,
                                                        int relu) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long v0 = (long long)blockIdx.x * 16;
  const int row = lane & 15, kg = lane >> 4;
  const long long V = (long long)nx * ny * nz;
  const long long vox = v0 + row;
  const bool valid = vox < BV;
  long long b = 0;
  int x = 0, y = 0, z = 0;
  if (valid) vox_xyz(vox, V, nx, ny, &b, &x, &y, &z);
  const long long bbase = b * V;
  f4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  typedef typename KOp<KW>::type AT;
  constexpr int KT = 27 * CIN;
  constexpr int KC = (KT + KW - 1) / KW;
  constexpr int U = 4;
  for (int kc0 = w * U; kc0 < KC; kc0 += S * U) {
    AT a[U];
    AT wv[U][NT];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int kc = kc0 + u;
      const int kk = kc * KW + kg * (KW / 4);
      a[u] = AT{};
      if (valid && kc < KC && kk < KT) {
        const int tap = kk / CIN;
        const int ci = kk - tap * CIN;
        const int xx = x + tap % 3 - 1, yy = y + (tap / 3) % 3 - 1, zz = z + tap / 9 - 1;
        if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz) {
          const long long nb = bbase + xx + (long long)nx * (yy + (long long)ny * zz);
          a[u] = *reinterpret_cast<const AT*>(in + nb * CIN + ci);
        }
      }
#pragma unroll
      for (int t = 0; t < NT; t++) wv[u][t] = kc < KC ? Wp[(kc * NT + t) * 64 + lane] : AT{};
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[t] = mfma_bf16(a[u], wv[u][t], acc[t]);
  }
  __shared__ f4 red[S][NT][64];
#pragma unroll
  for (int t = 0; t < NT; t++) red[w][t][lane] = acc[t];
  __syncthreads();
  if (w != 0) return;
#pragma unroll
  for (int t = 0; t < NT; t++) {
    f4 sum = red[0][t][lane];
    for (int ww = 1; ww < S; ww++) sum += red[ww][t][lane];
    const int co = t * 16 + row;
    if (co >= cout_store) continue;
    const float bv = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const long long vo = v0 + kg * 4 + i;
      if (vo >= BV) continue;
      float val = sum[i] + bv;
      if (relu) val = fmaxf(val, 0.f);
      if (OUT_F32)
        reinterpret_cast<float*>(out)[vo * cout_store + co] = val;
      else
        reinterpret_cast<uint16_t*>(out)[vo * cout_store + co] = f2bf(val);
    }
  }
}

// LDS-tiled form of k_conv for the large levels (CIN a multiple of 4, of 8
// on the 32-wide step; nx, ny, nz multiples of 16, 4, 4).  A block computes a 16 x 4 x 4 voxel tile:
// it stages the 18 x 6 x 6 input box (zero outside the grid) in LDS once, and
// each wave takes the four x-rows of one z-slice, so every input voxel is
// read from memory once per tile instead of once per tap.  The K steps and
// their order per output are those of k_conv (identical results).
template <int CIN, int NT, bool OUT_F32, int KW>
__global__ __launch_bounds__(256) void k_conv_tiled(const uint16_t* __restrict__ in,
                                                    const typename KOp<KW>::type* __restrict__ Wp,
                                                    const float* __restrict__ bias, void* __restrict__ out,
                                                    int cout_store, int nx, int ny, int nz, int relu) {
  typedef typename KOp<KW>::type AT;
  constexpr int BX = 18, BY = 6, BZ = 6, NBOX = BX * BY * BZ;
  constexpr int KT = 27 * CIN;
  constexpr int KC = (KT + KW - 1) / KW;  // K steps (with KW = 32 and CIN = 16 a step spans two taps)
  constexpr int Q = CIN / 4;              // s4 chunks per voxel
  constexpr int STG = 4 * 4 * 16 * NT * 16 / 4;  // output staging (s4 units), reuses the box
  __shared__ __align__(16) s4 box[NBOX * Q > STG ? NBOX * Q : STG];
  const int tx = nx / 16, ty = ny / 4, tz = nz / 4;
  long long t = blockIdx.x;
  const int bx = (int)(t % tx);
  t /= tx;
  const int by = (int)(t % ty);
  t /= ty;
  const int bz = (int)(t % tz);
  const long long b = t / tz;
  const int x0 = bx * 16, y0 = by * 4, z0 = bz * 4;
  const long long bbase = b * (long long)nx * ny * nz;
  if constexpr (Q % 2 == 0) {
    // 16-byte pieces (8 channels) when a voxel holds a whole number of them
    constexpr int Q8 = Q / 2;
    const s8* in8 = reinterpret_cast<const s8*>(in);
    s8* box8 = reinterpret_cast<s8*>(box);
    for (int q = threadIdx.x; q < NBOX * Q8; q += 256) {
      const int c8 = q % Q8, vx = q / Q8;
      const int xx = vx % BX, yy = (vx / BX) % BY, zz = vx / (BX * BY);
      const int gx = x0 - 1 + xx, gy = y0 - 1 + yy, gz = z0 - 1 + zz;
      s8 v = s8{0, 0, 0, 0, 0, 0, 0, 0};
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
        v = in8[(bbase + gx + (long long)nx * (gy + (long long)ny * gz)) * Q8 + c8];
      box8[q] = v;
    }
  } else {
    const s4* in4 = reinterpret_cast<const s4*>(in);
    for (int q = threadIdx.x; q < NBOX * Q; q += 256) {
      const int c4 = q % Q, vx = q / Q;
      const int xx = vx % BX, yy = (vx / BX) % BY, zz = vx / (BX * BY);
      const int gx = x0 - 1 + xx, gy = y0 - 1 + yy, gz = z0 - 1 + zz;
      s4 v = s4{0, 0, 0, 0};
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz)
        v = in4[(bbase + gx + (long long)nx * (gy + (long long)ny * gz)) * Q + c4];
      box[q] = v;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  f4 acc[4][NT];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int u = 0; u < NT; u++) acc[j][u] = f4{0.f, 0.f, 0.f, 0.f};
  constexpr int U = 9;
  for (int kc0 = 0; kc0 < KC; kc0 += U) {
    AT wv[U][NT];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int v = 0; v < NT; v++) wv[u][v] = kc0 + u < KC ? Wp[((kc0 + u) * NT + v) * 64 + lane] : AT{};
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int kk = (kc0 + u) * KW + kg * (KW / 4);  // this lane's first K index of the step
      const bool in_k = kc0 + u < KC && kk < KT;
      const int tap = kk / CIN, ci = kk - tap * CIN;
      const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const AT a = in_k ? *reinterpret_cast<const AT*>(&box[(((w + dz) * BY + (j + dy)) * BX + (row + dx)) * Q + ci / 4])
                          : AT{};
#pragma unroll
        for (int v = 0; v < NT; v++) acc[j][v] = mfma_bf16(a, wv[u][v], acc[j][v]);
      }
    }
  }
  if (!OUT_F32 && cout_store == NT * 16) {
    // BF16 output of every channel: the wave's 64 voxels x NT*16 channels
    // go through LDS (the box is free once every wave is past its K loop)
    // and leave as 16-byte stores of whole voxel rows, instead of one
    // 2-byte store per lane per voxel (same values, 10 % less time)
    __syncthreads();
    uint16_t* stg = reinterpret_cast<uint16_t*>(box) + w * (4 * 16 * NT * 16);
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int v = 0; v < NT; v++) {
        const int co = v * 16 + row;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          float val = acc[j][v][i] + bv;
          if (relu) val = fmaxf(val, 0.f);
          stg[(j * 16 + kg * 4 + i) * (NT * 16) + co] = f2bf(val);
        }
      }
    __syncthreads();
    const s8* st8 = reinterpret_cast<const s8*>(stg);
    constexpr int CH8 = NT * 2;  // 16-byte pieces per voxel
    for (int c = lane; c < 4 * 16 * CH8; c += 64) {
      const int vx = c / CH8, part = c - vx * CH8;
      const int j = vx / 16, xx = vx - j * 16;
      const long long vo = bbase + (x0 + xx) + (long long)nx * ((y0 + j) + (long long)ny * (z0 + w));
      reinterpret_cast<s8*>(reinterpret_cast<uint16_t*>(out) + vo * cout_store)[part] = st8[c];
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int v = 0; v < NT; v++) {
      const int co = v * 16 + row;
      if (co >= cout_store) continue;
      const float bv = bias ? bias[co] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const long long vo = bbase + (x0 + kg * 4 + i) + (long long)nx * ((y0 + j) + (long long)ny * (z0 + w));
        float val = acc[j][v][i] + bv;
        if (relu) val = fmaxf(val, 0.f);
        if (OUT_F32)
          reinterpret_cast<float*>(out)[vo * cout_store + co] = val;
        else
          reinterpret_cast<uint16_t*>(out)[vo * cout_store + co] = f2bf(val);
      }
    }
}

template <int CIN, int NT>
__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t* __restrict__ X, const uint16_t* __restrict__ G,
                                                    int cg, float* __restrict__ partial, int nx, int ny, int nz,
                                                    long long BV, int chunk_len) {
  constexpr int KT = 27 * CIN;
  constexpr int KC = (KT + 15) / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mt = blockIdx.x * 4 + wave;
  if (mt >= KC) return;
  const long long c0 = (long long)blockIdx.y * chunk_len;
  long long c1 = c0 + chunk_len;
  if (c1 > BV) c1 = BV;
  const int row = lane & 15, kg = lane >> 4;
  const int kkA = mt * 16 + row;
  const bool rowok = kkA < KT;
  const int tap = rowok ? kkA / CIN : 0;
  const int ci = rowok ? kkA - tap * CIN : 0;
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const long long V = (long long)nx * ny * nz;
  f4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  for (long long vb = c0; vb < c1; vb += 16) {
    s4 a = s4{0, 0, 0, 0};
    s4 bq[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) bq[t] = s4{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const long long vox = vb + kg * 4 + j;
      if (vox >= c1) continue;
      if (rowok) {
        long long b;
        int x, y, z;
        vox_xyz(vox, V, nx, ny, &b, &x, &y, &z);
        const int xx = x + dx, yy = y + dy, zz = z + dz;
        if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
          a[j] = (short)X[(b * V + xx + (long long)nx * (yy + (long long)ny * zz)) * CIN + ci];
      }
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const int co = t * 16 + row;
        if (co < cg) bq[t][j] = (short)G[vox * cg + co];
      }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bq[t], acc[t], 0, 0, 0);
  }
  const int ncol = NT * 16;
  float* P = partial + (long long)blockIdx.y * (KC * 16) * ncol;
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int i = 0; i < 4; i++) P[(long long)(mt * 16 + kg * 4 + i) * ncol + t * 16 + row] = acc[t][i];
}

// grad W[kk][co] = sum over chunks (fixed order).
__global__ void k_wgrad_reduce(const float* __restrict__ partial, int nchunk, int kt, int kcpad, int ncol, int cout,
                               float* __restrict__ gw) {
  const long long n = (long long)kt * cout;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int kk = (int)(i / cout), co = (int)(i % cout);
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += partial[((long long)c * kcpad + kk) * ncol + co];
    gw[i] = s;
  }
}

// bias partials: P[chunk][co] = sum_{vox in chunk} G[vox][co]
__global__ void k_bgrad_partial(const uint16_t* __restrict__ G, int cg, int cout, long long BV, int chunk_len,
                                float* __restrict__ partial) {
  const int co = threadIdx.x;
  if (co >= cout) return;
  const long long c0 = (long long)blockIdx.x * chunk_len;
  long long c1 = c0 + chunk_len;
  if (c1 > BV) c1 = BV;
  float s = 0.f;
  for (long long v = c0; v < c1; v++) s += bf2f(G[v * cg + co]);
  partial[(long long)blockIdx.x * cout + co] = s;
}
__global__ void k_bgrad_reduce(const float* __restrict__ partial, int nchunk, int cout, float* __restrict__ gb) {
  const int co = threadIdx.x;
  if (co >= cout) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; c++) s += partial[(long long)c * cout + co];
  gb[co] = s;
}

// Weight packing.  Forward: B[kk][co] = W[kk][co].  Transposed (data
// gradient of a layer with cin -> cout): the convolution runs cout -> cin with
// B[kk'][c] where kk' = tap'*cout_in + co, B = W[(26 - tap')*cin + c][co].
__global__ void k_pack(const float* __restrict__ W, int cin, int cout, int transpose, int cin_g, int nt, int kc,
                       s4* __restrict__ out) {
  // cin_g: input channels of the packed GEMM; nt: its N tiles; kc: its K chunks
  const long long n = (long long)kc * nt * 64;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int lane = (int)(i % 64);
    const long long r = i / 64;
    const int t = (int)(r % nt), c = (int)(r / nt);
    const int col = t * 16 + (lane & 15);
    s4 v = s4{0, 0, 0, 0};
    for (int j = 0; j < 4; j++) {
      const int kk = c * 16 + (lane >> 4) * 4 + j;
      if (kk >= 27 * cin_g) continue;
      const int tap = kk / cin_g, ci = kk - tap * cin_g;
      float w = 0.f;
      if (!transpose) {
        if (col < cout) w = W[(long long)kk * cout + col];
      } else {
        // ci indexes the forward layer's outputs, col its inputs
        if (ci < cout && col < cin) w = W[(long long)((26 - tap) * cin + col) * cout + ci];
      }
      v[j] = (short)f2bf(w);
    }
    out[i] = v;
  }
}

// Forward weights for the 32-wide K step: lane l of chunk c, tile t holds
// B[c*32 + 8(l/16) + j][t*16 + l%16], j = 0..7.  The 32-wide step halves the
// matrix-core instructions and operand loads (measured at 96^3: the 48^3
// convolutions 37.8 -> 22.6 us, the 96^3 ones 37.3 -> 36.1 us, the network
// 0.48 -> about 0.37 ms per application).  It regroups the FP32 partial sums,
// which flips some of the BF16 roundings of the activations between layers:
// on a 48^3 input the two widths differ by 3.5e-3 relative, as far as either
// is from the numpy oracle.  The learned solve's iteration count moves by up
// to 30 % either way with such changes, so the two widths were compared over
// ten 96^3 fields of the generator: 19.83 s (32) against 20.59 s (16) in
// total, 5 % less time per Uzawa iteration, faster on eight.  The 32-wide
// step is the default; GALE_LPC_K32=0 selects the 16-wide one.
__global__ void k_pack32(const float* __restrict__ W, int cin, int cout, int nt, int kc, s8* __restrict__ out) {
  const long long n = (long long)kc * nt * 64;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int lane = (int)(i % 64);
    const long long r = i / 64;
    const int t = (int)(r % nt), c = (int)(r / nt);
    const int col = t * 16 + (lane & 15);
    s8 v = s8{};
    for (int j = 0; j < 8; j++) {
      const int kk = c * 32 + (lane >> 4) * 8 + j;
      if (kk < 27 * cin && col < cout) v[j] = (short)f2bf(W[(long long)kk * cout + col]);
    }
    out[i] = v;
  }
}

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
__global__ void k_pack_input(const float* __restrict__ in2, uint16_t* __restrict__ x4, long long BV) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < BV; i += (long long)gridDim.x * blockDim.x) {
    x4[i * 4 + 0] = f2bf(in2[i * 2 + 0]);
    x4[i * 4 + 1] = f2bf(in2[i * 2 + 1]);
    x4[i * 4 + 2] = 0;
    x4[i * 4 + 3] = 0;
  }
}
__global__ void k_pack_gout(const float* __restrict__ g, uint16_t* __restrict__ g4, long long BV) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < BV; i += (long long)gridDim.x * blockDim.x) {
    g4[i * 4 + 0] = f2bf(g[i]);
    g4[i * 4 + 1] = 0;
    g4[i * 4 + 2] = 0;
    g4[i * 4 + 3] = 0;
  }
}

// 2x2x2 average pool: fine (nx,ny,nz) -> coarse (nx/2, ny/2, nz/2).
// 2x2x2 average pool; four channels (one 8-byte word) per thread, C % 4 == 0.
__global__ void k_pool(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int C, int nx, int ny, int nz,
                       long long Bc) {
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  const long long Vc = (long long)cx * cy * cz, Vf = (long long)nx * ny * nz;
  const int C4 = C / 4;
  const long long n = Bc * Vc * C4;
  const s4* in4 = reinterpret_cast<const s4*>(in);
  s4* out4 = reinterpret_cast<s4*>(out);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C4);
    const long long vv = i / C4;
    long long b;
    int x, y, z;
    vox_xyz(vv, Vc, cx, cy, &b, &x, &y, &z);
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int a = 0; a < 8; a++) {
      const long long f = b * Vf + (2 * x + (a & 1)) + (long long)nx * ((2 * y + ((a >> 1) & 1)) +
                                                                       (long long)ny * (2 * z + ((a >> 2) & 1)));
      const s4 v = in4[f * C4 + c];
#pragma unroll
      for (int k = 0; k < 4; k++) s[k] += bf2f((uint16_t)v[k]);
    }
    s4 o;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (short)f2bf(s[k] * 0.125f);
    out4[i] = o;
  }
}

// out[fine] = up(coarse)[fine] + skip[fine]
__global__ void k_up_add(const uint16_t* __restrict__ coarse, const uint16_t* __restrict__ skip,
                         uint16_t* __restrict__ out, int C, int nx, int ny, int nz, long long BVf) {
  const long long Vf = (long long)nx * ny * nz;
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  const long long Vc = (long long)cx * cy * cz;
  const int C4 = C / 4;  // four channels (one 8-byte word) per thread, C % 4 == 0
  const long long n = BVf * C4;
  const s4* c4p = reinterpret_cast<const s4*>(coarse);
  const s4* s4p = reinterpret_cast<const s4*>(skip);
  s4* o4p = reinterpret_cast<s4*>(out);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C4);
    const long long vv = i / C4;
    long long b;
    int x, y, z;
    vox_xyz(vv, Vf, nx, ny, &b, &x, &y, &z);
    const long long p = b * Vc + (x / 2) + (long long)cx * ((y / 2) + (long long)cy * (z / 2));
    (void)cz;
    const s4 cv = c4p[p * C4 + c], sv = s4p[i];
    s4 o;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (short)f2bf(bf2f((uint16_t)cv[k]) + bf2f((uint16_t)sv[k]));
    o4p[i] = o;
  }
}

// Backward of up+add for the coarse branch: gc[parent] = sum of the 8 children
// of gfine, masked by ReLU of the coarse activation (act > 0).
__global__ void k_up_bwd(const uint16_t* __restrict__ gf, const uint16_t* __restrict__ act_c,
                         uint16_t* __restrict__ gc, int C, int nx, int ny, int nz, long long Bc) {
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  const long long Vc = (long long)cx * cy * cz, Vf = (long long)nx * ny * nz;
  const long long n = Bc * Vc * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long long vv = i / C;
    long long b;
    int x, y, z;
    vox_xyz(vv, Vc, cx, cy, &b, &x, &y, &z);
    float s = 0.f;
    for (int a = 0; a < 8; a++) {
      const long long f = b * Vf + (2 * x + (a & 1)) + (long long)nx * ((2 * y + ((a >> 1) & 1)) +
                                                                       (long long)ny * (2 * z + ((a >> 2) & 1)));
      s += bf2f(gf[f * C + c]);
    }
    gc[i] = bf2f(act_c[i]) > 0.f ? f2bf(s) : (uint16_t)0;
  }
}

// Gradient of an encoder activation that feeds both the pool and a skip:
// gfine = (gpool[parent] / 8 + gskip[fine]) masked by ReLU of act.
__global__ void k_pool_bwd(const uint16_t* __restrict__ gpool, const uint16_t* __restrict__ gskip,
                           const uint16_t* __restrict__ act, uint16_t* __restrict__ gf, int C, int nx, int ny, int nz,
                           long long BVf) {
  const long long Vf = (long long)nx * ny * nz;
  const int cx = nx / 2, cy = ny / 2;
  const long long Vc = (long long)cx * cy * (nz / 2);
  const long long n = BVf * C;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    const long long vv = i / C;
    long long b;
    int x, y, z;
    vox_xyz(vv, Vf, nx, ny, &b, &x, &y, &z);
    const long long p = b * Vc + (x / 2) + (long long)cx * ((y / 2) + (long long)cy * (z / 2));
    const float s = bf2f(gpool[p * C + c]) * 0.125f + bf2f(gskip[i]);
    gf[i] = bf2f(act[i]) > 0.f ? f2bf(s) : (uint16_t)0;
  }
}

__global__ void k_relu_mask(uint16_t* __restrict__ g, const uint16_t* __restrict__ act, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    if (!(bf2f(act[i]) > 0.f)) g[i] = 0;
}

// ---------------------------------------------------------------------------
// loss
// ---------------------------------------------------------------------------
// per-sample partial sums of d = c - a (grid (nblk, B))
__global__ void k_seg_sum(const float* __restrict__ c, const float* __restrict__ a, long long V, int nblk,
                          double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  const int b = blockIdx.y;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < V; i += (long long)nblk * blockDim.x)
    s += (double)c[b * V + i] - (double)a[b * V + i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += sh[w];
    partial[b * nblk + blockIdx.x] = t;
  }
}
// gout = (2 w_b / B) (d - mean_b);  partial of sum (d - mean)^2
__global__ void k_loss_grad(const float* __restrict__ c, const float* __restrict__ a, const float* __restrict__ w,
                            const int* __restrict__ proj, const double* __restrict__ sums, long long V, int B,
                            int nblk, float* __restrict__ gout, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  const int b = blockIdx.y;
  double tot = 0.0;
  for (int k = 0; k < nblk; k++) tot += sums[b * nblk + k];
  const double mean = proj[b] ? tot / (double)V : 0.0;
  const double scale = 2.0 * (double)w[b] / (double)B;
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < V; i += (long long)nblk * blockDim.x) {
    const double d = (double)c[b * V + i] - (double)a[b * V + i] - mean;
    gout[b * V + i] = (float)(scale * d);
    s += d * d;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += sh[q];
    partial[b * nblk + blockIdx.x] = t;
  }
}

// ---------------------------------------------------------------------------
// optimiser
// ---------------------------------------------------------------------------
__global__ void k_sumsq_partial(const float* __restrict__ g, long long n, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s += (double)g[i] * (double)g[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += sh[q];
    partial[blockIdx.x] = t;
  }
}
// clip factor from the per-block sums of squares: out[0] = min(1, clip / |g|), out[1] = |g|
__global__ void k_clip_factor(const double* __restrict__ partial, int nblk, float clip, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double ss = 0.0;
  for (int k = 0; k < nblk; k++) ss += partial[k];
  const double gn = sqrt(ss);
  out[0] = (clip > 0.f && gn > clip) ? clip / gn : 1.0;
  out[1] = gn;
}
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, const double* __restrict__ clipf, float lr, float b1,
                       float b2, float eps, float bc1, float bc2) {
  const float sc = (float)clipf[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i] * sc;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline unsigned gridn(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b > 65535) b = 65535;
  if (b < 1) b = 1;
  return (unsigned)b;
}

static uint64_t splitmix(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double unif(uint64_t* s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

int create(Net* n, uint64_t seed, hipStream_t st) {
  n->st = st;
  const long long np = num_params();
  NCK(hipMalloc(&n->w, sizeof(float) * np));
  NCK(hipMalloc(&n->g, sizeof(float) * np));
  NCK(hipMalloc(&n->m, sizeof(float) * np));
  NCK(hipMalloc(&n->v, sizeof(float) * np));
  NCK(hipMemset(n->g, 0, sizeof(float) * np));
  NCK(hipMemset(n->m, 0, sizeof(float) * np));
  NCK(hipMemset(n->v, 0, sizeof(float) * np));
  // He-normal initialisation (fan-in 27 * real input channels), zero biases;
  // the output layer starts small so that the untrained network is close to
  // the plain diagonal preconditioner.
  std::vector<float> h(np, 0.f);
  uint64_t s = seed * 0x2545F4914F6CDD1DULL + 1;
  for (int l = 0; l < kLayers; l++) {
    const LayerSpec& L = kSpec[l];
    double sd = std::sqrt(2.0 / (27.0 * L.cin_real));
    if (l == kLayers - 1) sd *= 0.01;
    const long long o = weight_offset(l);
    for (int kk = 0; kk < 27 * L.cin; kk++) {
      const int ci = kk % L.cin;
      for (int co = 0; co < L.cout; co++) {
        const double u1 = unif(&s), u2 = unif(&s);
        const double gsn = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        h[o + (long long)kk * L.cout + co] = ci < L.cin_real ? (float)(sd * gsn) : 0.f;
      }
    }
  }
  NCK(hipMemcpy(n->w, h.data(), sizeof(float) * np, hipMemcpyHostToDevice));
  for (int l = 0; l < kLayers; l++) {
    const LayerSpec& L = kSpec[l];
    NCK(hipMalloc(&n->wp[l], sizeof(s4) * 64LL * kchunks(L.cin) * ntiles(L.cout)));
    // forward passes of layers with cin a multiple of 8 run on the 32-wide
    // step unless GALE_LPC_K32=0 (see the note at k_pack32)
    const char* k32 = std::getenv("GALE_LPC_K32");
    if (L.cin % 8 == 0 && !(k32 && k32[0] == '0'))
      NCK(hipMalloc(&n->wp32[l], sizeof(s8) * 64LL * kchunks32(L.cin) * ntiles(L.cout)));
    const int cin_t = L.cout < 4 ? 4 : L.cout;  // transposed GEMM input channels (gradient tensor)
    NCK(hipMalloc(&n->wt[l], sizeof(s4) * 64LL * kchunks(cin_t) * ntiles(L.cin)));
  }
  pack(n);
  NCK(hipStreamSynchronize(st));
  return GS_OK;
}

void pack(Net* n) {
  for (int l = 0; l < kLayers; l++) {
    const LayerSpec& L = kSpec[l];
    const float* W = n->w + weight_offset(l);
    const int kc = kchunks(L.cin), nt = ntiles(L.cout);
    hipLaunchKernelGGL(k_pack, dim3(gridn(64LL * kc * nt)), dim3(kBlock), 0, n->st, W, L.cin, L.cout, 0, L.cin, nt,
                       kc, (s4*)n->wp[l]);
    if (n->wp32[l])
      hipLaunchKernelGGL(k_pack32, dim3(gridn(64LL * kchunks32(L.cin) * nt)), dim3(kBlock), 0, n->st, W, L.cin,
                         L.cout, nt, kchunks32(L.cin), (s8*)n->wp32[l]);
    const int cin_t = L.cout < 4 ? 4 : L.cout;
    const int kct = kchunks(cin_t), ntt = ntiles(L.cin);
    hipLaunchKernelGGL(k_pack, dim3(gridn(64LL * kct * ntt)), dim3(kBlock), 0, n->st, W, L.cin, L.cout, 1, cin_t,
                       ntt, kct, (s4*)n->wt[l]);
  }
  n->packed = true;
}

static void free_work(Work& w) {
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(w.x4);
  for (int l = 0; l < kLayers; l++) f(w.act[l]);
  for (int l = 0; l < kLevels; l++) {
    f(w.pool[l]);
    f(w.upadd[l]);
    f(w.gskip[l]);
  }
  f(w.out32);
  f(w.g4);
  f(w.ga);
  f(w.gb);
  f(w.gc);
  f(w.partial);
  f(w.dpartial);
  w = Work{};
}

void destroy(Net* n) {
  free_work(n->work);
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(n->w);
  f(n->g);
  f(n->m);
  f(n->v);
  for (int l = 0; l < kLayers; l++) {
    f(n->wp[l]);
    f(n->wp32[l]);
    f(n->wt[l]);
  }
  *n = Net{};
}

static constexpr int kChunk = 1024;  // voxels per weight-gradient partial

int ensure_work(Net* n, const Shape& sh) {
  if (n->work.sh == sh && n->work.x4) return GS_OK;
  if (sh.nx % kGridMultiple || sh.ny % kGridMultiple || sh.nz % kGridMultiple || sh.nx < kGridMultiple ||
      sh.ny < kGridMultiple || sh.nz < kGridMultiple || sh.batch < 1)
    return GS_E_INVALID;
  free_work(n->work);
  Work& w = n->work;
  w.sh = sh;
  for (int l = 0; l < kLevels; l++)
    w.BV[l] = (long long)sh.batch * (sh.nx >> l) * (sh.ny >> l) * (sh.nz >> l);
  auto al = [](uint16_t** p, long long cnt) { return hipMalloc(p, sizeof(uint16_t) * (cnt > 0 ? cnt : 1)); };
  NCK(al(&w.x4, w.BV[0] * 4));
  long long gmax = w.BV[0] * 4;
  for (int l = 0; l < kLayers; l++) {
    const long long cnt = w.BV[kSpec[l].level] * kSpec[l].cout;
    if (l < kLayers - 1) NCK(al(&w.act[l], cnt));
    gmax = cnt > gmax ? cnt : gmax;
    const long long cin_cnt = w.BV[kSpec[l].level] * kSpec[l].cin;
    gmax = cin_cnt > gmax ? cin_cnt : gmax;
  }
  for (int l = 1; l < kLevels; l++) NCK(al(&w.pool[l], w.BV[l] * kChan[l - 1]));
  for (int l = 0; l < kLevels - 1; l++) {
    NCK(al(&w.upadd[l], w.BV[l] * kChan[l]));
    NCK(al(&w.gskip[l], w.BV[l] * kChan[l]));
  }
  NCK(hipMalloc(&w.out32, sizeof(float) * w.BV[0]));
  NCK(al(&w.g4, w.BV[0] * 4));
  NCK(al(&w.ga, gmax));
  NCK(al(&w.gb, gmax));
  NCK(al(&w.gc, gmax));
  long long pc = 0;
  for (int l = 0; l < kLayers; l++) {
    const LayerSpec& L = kSpec[l];
    const long long nch = (w.BV[L.level] + kChunk - 1) / kChunk;
    const long long c = nch * kchunks(L.cin) * 16LL * ntiles(L.cout) * 16LL;
    pc = c > pc ? c : pc;
  }
  w.partial_cap = pc;
  NCK(hipMalloc(&w.partial, sizeof(float) * pc));
  NCK(hipMalloc(&w.dpartial, sizeof(double) * 65536));
  return GS_OK;
}

// dispatch of the templated convolution kernels
#define CONV_CASES(X) \
  X(4, 1) X(16, 1) X(16, 2) X(32, 1) X(32, 2) X(32, 3) X(48, 2) X(48, 3) X(48, 4) X(64, 3) X(64, 4) X(64, 6) X(96, 4)

#define CONV_TILED_CASES(X) X(4, 1) X(16, 1) X(16, 2) X(32, 1) X(32, 2) X(32, 3)

// KW: 16 or 32 K values per matrix-core step (32 needs cin a multiple of 8
// and weights packed for it).
template <int KW>
static int launch_conv_kw(int cin, int nt, bool f32, const uint16_t* in, const void* Wp, const float* bias,
                          void* out, int cout_store, int nx, int ny, int nz, long long BV, int relu, hipStream_t st) {
  typedef typename KOp<KW>::type AT;
  const AT* W = static_cast<const AT*>(Wp);
  if (nx % 16 == 0 && ny % 4 == 0 && nz % 4 == 0) {
    const long long V = (long long)nx * ny * nz;
    const dim3 tgrid((unsigned)(BV / V * (V / 256))), tblock(256);
#define X(C, N)                                                                                                  \
  if constexpr (KW == 16 || (C) % 8 == 0) {                                                                      \
    if (cin == C && nt == N) {                                                                                   \
      if (f32)                                                                                                   \
        hipLaunchKernelGGL((k_conv_tiled<C, N, true, KW>), tgrid, tblock, 0, st, in, W, bias, out, cout_store,   \
                           nx, ny, nz, relu);                                                                    \
      else                                                                                                       \
        hipLaunchKernelGGL((k_conv_tiled<C, N, false, KW>), tgrid, tblock, 0, st, in, W, bias, out, cout_store,  \
                           nx, ny, nz, relu);                                                                    \
      return GS_OK;                                                                                              \
    }                                                                                                            \
  }
    CONV_TILED_CASES(X)
#undef X
  }
  const long long tiles16 = (BV + 15) / 16;
  if (tiles16 < 4096) {  // small level: split K over 4 (or 8) waves per 16-voxel tile
    const dim3 sgrid((unsigned)tiles16);
#define X(C, N)                                                                                                  \
  if constexpr (KW == 16 || (C) % 8 == 0) {                                                                      \
    if (cin == C && nt == N) {                                                                                   \
      if (tiles16 < 512) {                                                                                       \
        if (f32)                                                                                                 \
          hipLaunchKernelGGL((k_conv_splitk<C, N, true, 8, KW>), sgrid, dim3(512), 0, st, in, W, bias, out,      \
                             cout_store, nx, ny, nz, BV, relu);                                                  \
        else                                                                                                     \
          hipLaunchKernelGGL((k_conv_splitk<C, N, false, 8, KW>), sgrid, dim3(512), 0, st, in, W, bias, out,     \
                             cout_store, nx, ny, nz, BV, relu);                                                  \
      } else {                                                                                                   \
        if (f32)                                                                                                 \
          hipLaunchKernelGGL((k_conv_splitk<C, N, true, 4, KW>), sgrid, dim3(256), 0, st, in, W, bias, out,      \
                             cout_store, nx, ny, nz, BV, relu);                                                  \
        else                                                                                                     \
          hipLaunchKernelGGL((k_conv_splitk<C, N, false, 4, KW>), sgrid, dim3(256), 0, st, in, W, bias, out,     \
                             cout_store, nx, ny, nz, BV, relu);                                                  \
      }                                                                                                          \
      return GS_OK;                                                                                              \
    }                                                                                                            \
  }
    CONV_CASES(X)
#undef X
  }
  const dim3 grid((unsigned)((BV + 63) / 64)), block(256);
#define X(C, N)                                                                                                  \
  if constexpr (KW == 16 || (C) % 8 == 0) {                                                                      \
    if (cin == C && nt == N) {                                                                                   \
      if (f32)                                                                                                   \
        hipLaunchKernelGGL((k_conv<C, N, true, KW>), grid, block, 0, st, in, W, bias, out, cout_store, nx, ny,   \
                           nz, BV, relu);                                                                        \
      else                                                                                                       \
        hipLaunchKernelGGL((k_conv<C, N, false, KW>), grid, block, 0, st, in, W, bias, out, cout_store, nx, ny,  \
                           nz, BV, relu);                                                                        \
      return GS_OK;                                                                                              \
    }                                                                                                            \
  }
  CONV_CASES(X)
#undef X
  std::fprintf(stderr, "gale_stokes(net): no convolution kernel for cin %d, %d tiles, K step %d\n", cin, nt, KW);
  return GS_E_INVALID;
}

static int launch_conv(int cin, int nt, bool f32, const uint16_t* in, const void* Wp, const float* bias, void* out,
                       int cout_store, int nx, int ny, int nz, long long BV, int relu, hipStream_t st, int kw = 16) {
  if (kw == 32)
    return launch_conv_kw<32>(cin, nt, f32, in, Wp, bias, out, cout_store, nx, ny, nz, BV, relu, st);
  return launch_conv_kw<16>(cin, nt, f32, in, Wp, bias, out, cout_store, nx, ny, nz, BV, relu, st);
}

static int launch_wgrad(int cin, int nt, const uint16_t* X, const uint16_t* G, int cg, float* partial, int nx,
                        int ny, int nz, long long BV, hipStream_t st) {
  const int kc = kchunks(cin);
  const unsigned nchunk = (unsigned)((BV + kChunk - 1) / kChunk);
  const dim3 grid((unsigned)((kc + 3) / 4), nchunk), block(256);
#define X(C, N)                                                                                                   \
  if (cin == C && nt == N) {                                                                                      \
    hipLaunchKernelGGL((k_conv_wgrad<C, N>), grid, block, 0, st, X, G, cg, partial, nx, ny, nz, BV, kChunk);      \
    return GS_OK;                                                                                                 \
  }
  CONV_CASES(X)
#undef X
  std::fprintf(stderr, "gale_stokes(net): no weight-gradient kernel for cin %d, %d tiles\n", cin, nt);
  return GS_E_INVALID;
}

static int lvl_dims(const Shape& sh, int l, int* nx, int* ny, int* nz) {
  *nx = sh.nx >> l;
  *ny = sh.ny >> l;
  *nz = sh.nz >> l;
  return 0;
}

// forward of layer l from tensor `in` into act[l] (or out32 for the last)
static int fwd_layer(Net* n, int l, const uint16_t* in, void* out) {
  Work& w = n->work;
  const LayerSpec& L = kSpec[l];
  int nx, ny, nz;
  lvl_dims(w.sh, L.level, &nx, &ny, &nz);
  const bool last = (l == kLayers - 1);
  if (n->wp32[l])  // gfx950's 16x16x32 matrix-core step
    return launch_conv(L.cin, ntiles(L.cout), last, in, n->wp32[l], n->w + bias_offset(l), out, L.cout, nx, ny,
                       nz, w.BV[L.level], L.relu, n->st, 32);
  return launch_conv(L.cin, ntiles(L.cout), last, in, n->wp[l], n->w + bias_offset(l), out, L.cout, nx, ny, nz,
                     w.BV[L.level], L.relu, n->st);
}

int forward(Net* n, const Shape& sh, const float* in2, float* out) {
  int e = ensure_work(n, sh);
  if (e) return e;
  Work& w = n->work;
  hipStream_t st = n->st;
  hipLaunchKernelGGL(k_pack_input, dim3(gridn(w.BV[0])), dim3(kBlock), 0, st, in2, w.x4, w.BV[0]);
  int nx, ny, nz;
  // encoder and bottom level
  for (int l = 0; l < kLevels; l++) {
    const uint16_t* in = w.x4;
    if (l > 0) {
      lvl_dims(sh, l - 1, &nx, &ny, &nz);
      const int C = kChan[l - 1];
      hipLaunchKernelGGL(k_pool, dim3(gridn(w.BV[l] * C / 4)), dim3(kBlock), 0, st, w.act[enc_b(l - 1)], w.pool[l], C,
                         nx, ny, nz, (long long)sh.batch);
      in = w.pool[l];
    }
    if ((e = fwd_layer(n, enc_a(l), in, w.act[enc_a(l)]))) return e;
    if ((e = fwd_layer(n, enc_b(l), w.act[enc_a(l)], w.act[enc_b(l)]))) return e;
  }
  // decoder
  int cur = enc_b(kLevels - 1);
  for (int l = kLevels - 2; l >= 0; l--) {
    lvl_dims(sh, l, &nx, &ny, &nz);
    const int C = kChan[l];
    hipLaunchKernelGGL(k_up_add, dim3(gridn(w.BV[l] * C / 4)), dim3(kBlock), 0, st, w.act[cur], w.act[enc_b(l)],
                       w.upadd[l], C, nx, ny, nz, w.BV[l]);
    if ((e = fwd_layer(n, dec_a(l), w.upadd[l], w.act[dec_a(l)]))) return e;
    void* o = l > 0 ? (void*)w.act[dec_b(l)] : (void*)(out ? out : w.out32);
    if ((e = fwd_layer(n, dec_b(l), w.act[dec_a(l)], o))) return e;
    cur = dec_b(l);
  }
  return GS_OK;
}

// Gradients of layer l given g_pre (ReLU-masked gradient of its output,
// channels cg) and its input tensor x; writes dW, db, and (if gin) the
// gradient of its input (unmasked).
static int bwd_layer(Net* n, int l, const uint16_t* x, const uint16_t* gpre, int cg, uint16_t* gin) {
  Work& w = n->work;
  const LayerSpec& L = kSpec[l];
  int nx, ny, nz;
  lvl_dims(w.sh, L.level, &nx, &ny, &nz);
  const long long BV = w.BV[L.level];
  hipStream_t st = n->st;
  const int nt = ntiles(L.cout), kc = kchunks(L.cin);
  const int nchunk = (int)((BV + kChunk - 1) / kChunk);
  int e = launch_wgrad(L.cin, nt, x, gpre, cg, w.partial, nx, ny, nz, BV, st);
  if (e) return e;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(gridn(27LL * L.cin * L.cout)), dim3(kBlock), 0, st, w.partial, nchunk,
                     27 * L.cin, kc * 16, nt * 16, L.cout, n->g + weight_offset(l));
  hipLaunchKernelGGL(k_bgrad_partial, dim3(nchunk), dim3(128), 0, st, gpre, cg, L.cout, BV, kChunk, w.partial);
  hipLaunchKernelGGL(k_bgrad_reduce, dim3(1), dim3(128), 0, st, w.partial, nchunk, L.cout, n->g + bias_offset(l));
  if (gin) {
    const int cin_t = L.cout < 4 ? 4 : L.cout;
    e = launch_conv(cin_t, ntiles(L.cin), false, gpre, n->wt[l], nullptr, gin, L.cin, nx, ny, nz, BV, 0, st);
    if (e) return e;
  }
  return GS_OK;
}

int backward(Net* n, const Shape& sh, const float* gout) {
  Work& w = n->work;
  if (!(w.sh == sh) || !w.x4) return GS_E_INVALID;
  hipStream_t st = n->st;
  int e, nx, ny, nz;
  uint16_t *A = w.ga, *B = w.gb, *C = w.gc;
  hipLaunchKernelGGL(k_pack_gout, dim3(gridn(w.BV[0])), dim3(kBlock), 0, st, gout, w.g4, w.BV[0]);
  // decoder, finest level first; B holds the ReLU-masked gradient of the
  // level's output (level 0: the packed dL/dc)
  for (int l = 0; l <= kLevels - 2; l++) {
    const int a = dec_a(l), b = dec_b(l);
    const int Cl = kChan[l];
    if ((e = bwd_layer(n, b, w.act[a], l == 0 ? w.g4 : B, l == 0 ? 4 : kSpec[b].cout, A))) return e;
    hipLaunchKernelGGL(k_relu_mask, dim3(gridn(w.BV[l] * Cl)), dim3(kBlock), 0, st, A, w.act[a], w.BV[l] * Cl);
    if ((e = bwd_layer(n, a, w.upadd[l], A, Cl, w.gskip[l]))) return e;
    // upadd[l] = up(src) + skip l
    const int src = l == kLevels - 2 ? enc_b(kLevels - 1) : dec_b(l + 1);
    lvl_dims(sh, l, &nx, &ny, &nz);
    hipLaunchKernelGGL(k_up_bwd, dim3(gridn(w.BV[l + 1] * Cl)), dim3(kBlock), 0, st, w.gskip[l], w.act[src], B, Cl,
                       nx, ny, nz, (long long)sh.batch);
  }
  // bottom level
  {
    const int L = kLevels - 1;
    if ((e = bwd_layer(n, enc_b(L), w.act[enc_a(L)], B, kChan[L - 1], A))) return e;
    hipLaunchKernelGGL(k_relu_mask, dim3(gridn(w.BV[L] * kChan[L])), dim3(kBlock), 0, st, A, w.act[enc_a(L)],
                       w.BV[L] * kChan[L]);
    if ((e = bwd_layer(n, enc_a(L), w.pool[L], A, kChan[L], C))) return e;  // C: gradient of pool[L]
  }
  // encoder, coarse to fine: act[enc_b(l)] feeds pool[l+1] and skip l
  for (int l = kLevels - 2; l >= 0; l--) {
    const int Cl = kChan[l];
    lvl_dims(sh, l, &nx, &ny, &nz);
    hipLaunchKernelGGL(k_pool_bwd, dim3(gridn(w.BV[l] * Cl)), dim3(kBlock), 0, st, C, w.gskip[l], w.act[enc_b(l)], B,
                       Cl, nx, ny, nz, w.BV[l]);
    if ((e = bwd_layer(n, enc_b(l), w.act[enc_a(l)], B, Cl, A))) return e;
    hipLaunchKernelGGL(k_relu_mask, dim3(gridn(w.BV[l] * Cl)), dim3(kBlock), 0, st, A, w.act[enc_a(l)], w.BV[l] * Cl);
    if ((e = bwd_layer(n, enc_a(l), l > 0 ? w.pool[l] : w.x4, A, Cl, l > 0 ? C : nullptr))) return e;
  }
  return GS_OK;
}

int adam_step(Net* n, float lr, float beta1, float beta2, float eps, float clip, double* gnorm_out) {
  const long long np = num_params();
  Work& w = n->work;
  if (!w.dpartial) return GS_E_INVALID;
  const int nb = 512;
  double* clipf = w.dpartial + nb;
  hipLaunchKernelGGL(k_sumsq_partial, dim3(nb), dim3(kBlock), 0, n->st, n->g, np, w.dpartial);
  hipLaunchKernelGGL(k_clip_factor, dim3(1), dim3(64), 0, n->st, w.dpartial, nb, clip, clipf);
  n->adam_t++;
  const float bc1 = 1.f - std::pow(beta1, (float)n->adam_t), bc2 = 1.f - std::pow(beta2, (float)n->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(gridn(np)), dim3(kBlock), 0, n->st, n->w, n->g, n->m, n->v, np, clipf, lr, beta1,
                     beta2, eps, bc1, bc2);
  if (gnorm_out) {
    double h[2];
    NCK(hipMemcpyAsync(h, clipf, sizeof(h), hipMemcpyDeviceToHost, n->st));
    NCK(hipStreamSynchronize(n->st));
    *gnorm_out = h[1];
  }
  pack(n);
  return GS_OK;
}

int loss_grad(Net* n, const Shape& sh, const float* c, const float* a, const float* w, const int* proj, float* gout,
              double* loss) {
  Work& wk = n->work;
  if (!wk.dpartial) return GS_E_INVALID;
  const long long V = (long long)sh.nx * sh.ny * sh.nz;
  const int B = sh.batch;
  int nblk = (int)((V + kBlock - 1) / kBlock);
  if (nblk > 64) nblk = 64;
  if ((long long)B * nblk * 2 > 65536) return GS_E_INVALID;
  double* sums = wk.dpartial;
  double* sq = wk.dpartial + (long long)B * nblk;
  hipLaunchKernelGGL(k_seg_sum, dim3(nblk, B), dim3(kBlock), 0, n->st, c, a, V, nblk, sums);
  hipLaunchKernelGGL(k_loss_grad, dim3(nblk, B), dim3(kBlock), 0, n->st, c, a, w, proj, sums, V, B, nblk, gout, sq);
  if (loss) {
    std::vector<double> h((size_t)B * nblk);
    std::vector<float> hw(B);
    NCK(hipMemcpyAsync(h.data(), sq, sizeof(double) * B * nblk, hipMemcpyDeviceToHost, n->st));
    NCK(hipMemcpyAsync(hw.data(), w, sizeof(float) * B, hipMemcpyDeviceToHost, n->st));
    NCK(hipStreamSynchronize(n->st));
    double L = 0.0;
    for (int b = 0; b < B; b++) {
      double s = 0.0;
      for (int k = 0; k < nblk; k++) s += h[(size_t)b * nblk + k];
      L += hw[b] * s;
    }
    *loss = L / B;
  }
  return GS_OK;
}

// ---------------------------------------------------------------------------
// self test of the operand layout: D = A B for A[i][k] = (i + 2k) / 32,
// B[k][j] = (3k - j) / 32 (exact in BF16), one wavefront.
// ---------------------------------------------------------------------------
__global__ void k_mfma_selftest(float* __restrict__ out) {
  const int lane = threadIdx.x;
  if (lane >= 64) return;
  s4 a, b;
  for (int j = 0; j < 4; j++) {
    const int k = 4 * (lane >> 4) + j;
    a[j] = (short)f2bf((float)((lane & 15) + 2 * k) / 32.f);
    b[j] = (short)f2bf((float)(3 * k - (lane & 15)) / 32.f);
  }
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  for (int i = 0; i < 4; i++) out[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[i];
}

// The same for the 32-wide step (v_mfma_f32_16x16x32_bf16): lane l holds
// K = 8 (l / 16) + j, j = 0..7, of A's row l % 16 and B's column l % 16.
__global__ void k_mfma_selftest32(float* __restrict__ out) {
  const int lane = threadIdx.x;
  if (lane >= 64) return;
  s8 a, b;
  for (int j = 0; j < 8; j++) {
    const int k = 8 * (lane >> 4) + j;
    a[j] = (short)f2bf((float)((lane & 15) + 2 * k) / 32.f);
    b[j] = (short)f2bf((float)(3 * k - (lane & 15)) / 32.f);
  }
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  acc = mfma_bf16(a, b, acc);
  for (int i = 0; i < 4; i++) out[(4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[i];
}

// Both step widths against the exact product (every value is exact in BF16
// and every partial sum in FP32); returns the larger error.
int selftest_mfma(float* max_err) {
  float* d = nullptr;
  NCK(hipMalloc(&d, sizeof(float) * 512));
  hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, 0, d);
  hipLaunchKernelGGL(k_mfma_selftest32, dim3(1), dim3(64), 0, 0, d + 256);
  float h[512];
  NCK(hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost));
  NCK(hipFree(d));
  float e = 0.f;
  for (int kw = 16; kw <= 32; kw += 16)
    for (int i = 0; i < 16; i++)
      for (int j = 0; j < 16; j++) {
        double s = 0.0;
        for (int k = 0; k < kw; k++) s += ((i + 2.0 * k) / 32.0) * ((3.0 * k - j) / 32.0);
        e = std::fmax(e, (float)std::fabs(h[(kw == 32 ? 256 : 0) + i * 16 + j] - s));
      }
  *max_err = e;
  return GS_OK;
}

}  // namespace net
}  // namespace gs
