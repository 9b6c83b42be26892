// gs_reduce.h -- wavefront / block reductions with a deterministic last-block
// finish, shared by the kernels of the library.
#pragma once

#include <math.h>

#include "gs_internal.h"

namespace gs {

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Deterministic two-launch grid reduction.  Each block stores its NV block
// sums (or maxima) to partials[block * NV + k] with plain stores; a one-block
// finishing launch adds them in block order and hands the totals to Op on
// thread 0.  The launch boundary orders the two, so no cross-block fence or
// arrival counter is needed: on a chip whose XCDs each have their own L2, an
// agent-scope fence per block costs more than the reduction itself.
template <int NV, bool MAX>
__device__ void block_partials(double (&v)[NV], Reducer rd) {
  __shared__ double sh[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = MAX ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = sh[k][0];
      for (int w = 1; w < nw; w++) s = MAX ? fmax(s, sh[k][w]) : s + sh[k][w];
      rd.partials[blockIdx.x * NV + k] = s;
    }
}

template <int NV, bool MAX, typename Op>
__global__ void k_reduce_finish(Reducer rd, int nb, Op op) {
  __shared__ double sh[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = MAX ? -INFINITY : 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      const double pv = rd.partials[b * NV + k];
      s = MAX ? fmax(s, pv) : s + pv;
    }
    s = MAX ? wave_max(s) : wave_sum(s);
    if (lane == 0) sh[k][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = sh[k][0];
      for (int w = 1; w < nw; w++) s = MAX ? fmax(s, sh[k][w]) : s + sh[k][w];
      tot[k] = s;
    }
    op(tot);
  }
}

template <int NV, bool MAX, typename Op>
inline void reduce_finish(Reducer rd, int nb, const Op& op, hipStream_t st) {
  hipLaunchKernelGGL((k_reduce_finish<NV, MAX, Op>), dim3(1), dim3(kBlock), 0, st, rd, nb, op);
}

// S[slot] = total
struct OpStore {
  double* S;
  int slot;
  __device__ void operator()(const double* t) const { S[slot] = t[0]; }
};

inline int reduce_grid(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

inline unsigned grid_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace gs
