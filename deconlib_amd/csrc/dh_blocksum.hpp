// Deterministic two-stage sums for elementwise kernels launched with
// kSumThreads threads: per-workgroup float64 partials in a fixed grid, then
// one workgroup sums them in a fixed order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dh {

constexpr int kSumThreads = 256;

// Per-workgroup sums of NV values into part[blockIdx.x * NV + j].
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ part) {
  __shared__ double sh[NV][kSumThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double x = v[j];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) sh[j][w] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < kSumThreads / 64; ++q) r += sh[threadIdx.x][q];
    part[(int64_t)blockIdx.x * NV + threadIdx.x] = r;
  }
}

// Sums the NV partial columns [block][NV] in a fixed order; the result is
// valid in every thread of the (single, kSumThreads-wide) finishing workgroup.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int nb, double (&out)[NV]) {
  __shared__ double sh[NV][kSumThreads / 64];
  __shared__ double res[NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double x = 0.0;
    for (int b = threadIdx.x; b < nb; b += kSumThreads) x += part[(int64_t)b * NV + j];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) sh[j][w] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < kSumThreads / 64; ++q) r += sh[threadIdx.x][q];
    res[threadIdx.x] = r;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) out[j] = res[j];
}

}  // namespace dh
