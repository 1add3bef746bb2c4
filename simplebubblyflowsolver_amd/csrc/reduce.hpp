// Deterministic wavefront / workgroup reductions (fixed combination order,
// so a sum is the same on every run).
#pragma once

#include <hip/hip_runtime.h>

namespace sbfs {

constexpr int kRedThreads = 256;  // threads of every reduction workgroup

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// workgroup-wide reduction of kRedThreads values (4 wavefronts):
// OP 0 = min, 1 = max, 2 = sum
template <int OP>
__device__ double block_reduce(double v, double* sh) {
  v = OP == 0 ? wave_min(v) : OP == 1 ? wave_max(v) : wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = sh[0];
  for (int k = 1; k < kRedThreads / 64; k++)
    r = OP == 0 ? fmin(r, sh[k]) : OP == 1 ? fmax(r, sh[k]) : r + sh[k];
  return r;
}

}  // namespace sbfs
