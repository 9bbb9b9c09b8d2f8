// Wave and workgroup prefix sums shared by the feedback kernels and the
// tile scheduler of the plain renderer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rayt {

__device__ __forceinline__ uint32_t WaveInclusiveScan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread (BLOCK a multiple of 64).
template <int BLOCK>
__device__ __forceinline__ uint32_t BlockExclusiveScan(uint32_t v, uint32_t* lds_waves, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = WaveInclusiveScan(v);
  if (lane == 63) lds_waves[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < BLOCK / 64; ++w) {
      const uint32_t t = lds_waves[w];
      lds_waves[w] = run;
      run += t;
    }
    lds_waves[BLOCK / 64] = run;
  }
  __syncthreads();
  const uint32_t res = lds_waves[wave] + inc - v;
  total = lds_waves[BLOCK / 64];
  __syncthreads();
  return res;
}

}  // namespace rayt
