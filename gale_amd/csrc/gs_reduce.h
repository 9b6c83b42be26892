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

// Block reduction of NV values followed by a deterministic cross-block pass
// in the last block to finish (partials are summed in block order).  Returns
// true in the last block, with the totals in tot[] for every thread.
template <int NV, bool MAX>
__device__ bool reduce_last(double (&v)[NV], Reducer rd, double (&tot)[NV]) {
  __shared__ double sh[NV][kBlock / 64];
  __shared__ int last;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = MAX ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double s = sh[k][0];
      for (int w = 1; w < nw; w++) s = MAX ? fmax(s, sh[k][w]) : s + sh[k][w];
      __hip_atomic_store(&rd.partials[blockIdx.x * NV + k], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    const unsigned prev = atomicAdd(rd.counter, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = MAX ? -INFINITY : 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
      const double pv = __hip_atomic_load(&rd.partials[b * NV + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s = MAX ? fmax(s, pv) : s + pv;
    }
    v[k] = MAX ? wave_max(s) : wave_sum(s);
  }
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][wid] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = sh[k][0];
    for (int w = 1; w < nw; w++) s = MAX ? fmax(s, sh[k][w]) : s + sh[k][w];
    tot[k] = s;
  }
  if (threadIdx.x == 0) __hip_atomic_store(rd.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

inline int reduce_grid(long long n) {
  long long b = (n + kBlock - 1) / kBlock;
  if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

inline unsigned grid_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace gs
