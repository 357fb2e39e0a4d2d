// Device helpers shared by the particle-side translation units (pmx_particles.hip, pmx_sky.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace {

__device__ __forceinline__ float3 ld3(const float* a, unsigned p) {
  return make_float3(a[3 * (size_t)p], a[3 * (size_t)p + 1], a[3 * (size_t)p + 2]);
}
__device__ __forceinline__ void st3(float* a, unsigned p, float3 v) {
  a[3 * (size_t)p] = v.x;
  a[3 * (size_t)p + 1] = v.y;
  a[3 * (size_t)p + 2] = v.z;
}

// Hardware dispatches workgroup b to XCD b % 8; the logical id remaps so each XCD walks one
// contiguous range of the work (one slab of the Lagrangian grid, one range of pixels) and the
// lines it touches stay in that XCD's L2.
__device__ __forceinline__ unsigned xcd_logical(unsigned b, unsigned G) {
  return (G % 8u == 0u) ? (b % 8u) * (G / 8u) + b / 8u : b;
}

// Workgroup sum of a double written to *dst by thread 0 (callers reduce the per-workgroup values
// in a fixed order afterwards: deterministic, and no same-address atomics from many workgroups).
__device__ __forceinline__ void block_sum_to(double v, double* dst) {
  __shared__ double part[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int wv = threadIdx.x >> 6;
  if (__lane_id() == 0) part[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += part[i];
    *dst = t;
  }
  __syncthreads();
}

}  // namespace

// Workgroup cap of the grid-stride launches (pmx_set_grid_cap; 2^20 by default).
extern unsigned pmx_grid_cap;

namespace {
inline int grid_for(size_t n, int block) {
  size_t g = (n + block - 1) / block;
  const size_t gmax = pmx_grid_cap;
  return (int)(g < 1 ? 1 : (g > gmax ? gmax : g));
}
}  // namespace
