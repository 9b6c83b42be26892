// gs_net.h -- the learned Schur-complement preconditioner network (gs_net.hip).
//
// A 3D convolutional encoder-decoder on the element grid (pressure lives on
// elements).  Input channels per element: the diagonally preconditioned
// pressure residual Q^-1 r scaled to unit RMS, and log10 of the element mean
// viscosity / 3.  Output: one correction channel c; the preconditioned
// residual is  z = Q^-1 r + rms(Q^-1 r) * c.
//
// U-Net with kLevels grid levels and channels kChan[level] (3x3x3
// convolutions, zero padding, ReLU except on the output layer):
//   encoder level l: [avgpool,] conv C(l-1)->C(l), conv C(l)->C(l)   (skip l)
//   bottom level L-1: avgpool, conv C(L-2)->C(L-1), conv C(L-1)->C(L-2)
//   decoder level l: nearest up + skip l, conv C(l)->C(l), conv C(l)->C(l-1)
//   (level 0 ends with conv C(0)->1), the input has 4 stored channels (2 used).
// With C = {16, 32, 48, 64, 96} this is 18 convolutions, 1.03 M parameters;
// grids must be divisible by 2^(kLevels-1) = 16 along each axis.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "gs_internal.h"

namespace gs {
namespace net {

constexpr int kLevels = 5;
constexpr int kLayers = 4 * kLevels - 2;
constexpr int kChan[kLevels] = {16, 32, 48, 64, 96};
constexpr int kGridMultiple = 1 << (kLevels - 1);
inline int enc_a(int l) { return 2 * l; }
inline int enc_b(int l) { return 2 * l + 1; }
inline int dec_a(int l) { return 2 * kLevels + 2 * (kLevels - 2 - l); }
inline int dec_b(int l) { return dec_a(l) + 1; }

struct LayerSpec {
  int cin;    // input channels as stored (multiple of 4)
  int cout;   // output channels (real)
  int level;  // grid level (0 = finest)
  int relu;
  int cin_real;  // channels that carry data (fan-in for the initialisation)
};

extern const LayerSpec* const kSpec;  // kLayers entries

// Parameter layout: per layer W[27*cin][cout] (row kk = tap*cin + ci, tap =
// (dz+1)*9 + (dy+1)*3 + (dx+1)) followed by bias[cout].
long long num_params();
long long weight_offset(int l);
long long bias_offset(int l);

struct Shape {
  int nx, ny, nz, batch;
  bool operator==(const Shape& o) const {
    return nx == o.nx && ny == o.ny && nz == o.nz && batch == o.batch;
  }
};

// Per-shape buffers: activations of every layer (kept for the backward
// pass), gradients, weight-gradient partial sums.
struct Work {
  Shape sh{0, 0, 0, 0};
  long long BV[kLevels] = {};
  uint16_t* x4 = nullptr;             // input, 4 channels
  uint16_t* act[kLayers] = {};        // conv outputs (bf16); the last layer writes out32
  uint16_t* pool[kLevels] = {};       // pooled inputs of levels 1..L-1
  uint16_t* upadd[kLevels] = {};      // up + skip sums of levels 0..L-2
  float* out32 = nullptr;             // network output c [B*V0]
  uint16_t* g4 = nullptr;             // dL/dc packed to 4 channels
  uint16_t* ga = nullptr;             // gradient scratch (3 buffers of the largest tensor)
  uint16_t* gb = nullptr;
  uint16_t* gc = nullptr;
  uint16_t* gskip[kLevels] = {};      // gradient reaching each skip tensor through the decoder
  float* partial = nullptr;           // weight-gradient partial sums
  long long partial_cap = 0;
  double* dpartial = nullptr;         // reductions
};

struct Net {
  float* w = nullptr;     // FP32 master parameters
  float* g = nullptr;     // gradients
  float* m = nullptr;     // Adam first moment
  float* v = nullptr;     // Adam second moment
  void* wp[kLayers] = {};   // packed BF16 forward weights (16-wide K steps)
  void* wp32[kLayers] = {}; // the same for 32-wide K steps (layers with cin a multiple of 8; null otherwise)
  void* wt[kLayers] = {};   // packed BF16 transposed weights (data gradient)
  long long adam_t = 0;
  Work work;
  hipStream_t st = nullptr;
  bool packed = false;
};

int create(Net* n, uint64_t seed, hipStream_t st);
void destroy(Net* n);
int ensure_work(Net* n, const Shape& sh);
void pack(Net* n);
// in2: FP32 [B][V][2] (device).  out: FP32 [B][V] network output c.
int forward(Net* n, const Shape& sh, const float* in2, float* out);
// Gradient of the parameters for dL/dc = gout (FP32 [B][V]); uses the
// activations of the preceding forward() on the same shape.  Overwrites n->g.
int backward(Net* n, const Shape& sh, const float* gout);
// Global-norm clip and one Adam step on the master weights, then repack.
int adam_step(Net* n, float lr, float beta1, float beta2, float eps, float clip, double* gnorm_out);

// Loss of a batch against targets a (FP32 [B][V]) with per-sample weights w:
// L = (1/B) sum_b w_b |P_b (c_b - a_b)|^2, P_b removes the sample mean when
// proj[b] != 0.  Writes dL/dc into gout and the loss into *loss (host).
int loss_grad(Net* n, const Shape& sh, const float* c, const float* a, const float* w, const int* proj,
              float* gout, double* loss);

// Diagnostic: runs one MFMA tile on known data; max |error| vs FP32 FMA.
int selftest_mfma(float* max_err);

}  // namespace net
}  // namespace gs
