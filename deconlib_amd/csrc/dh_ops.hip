// Elementwise steps and deterministic reductions for the RL / PDHG loops.
//
// Reductions are two-stage: a fixed grid writes float64 partials, one
// workgroup sums them in a fixed order, so a result never depends on timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <math.h>
#include <stdint.h>

#include "dh_ops.hpp"

namespace dh {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

__device__ static double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ static double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

__device__ static double block_reduce(double v, int op, double* sh) {
  v = (op == RED_MAX) ? wave_max(v) : wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = (op == RED_MAX) ? -INFINITY : 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r = (op == RED_MAX) ? fmax(r, sh[i]) : r + sh[i];
  }
  return r;
}

__global__ void k_reduce_stage1(int op, const float* __restrict__ x, int64_t n, double* partials) {
  __shared__ double sh[16];
  double acc = (op == RED_MAX) ? -INFINITY : 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = x[i];
    if (op == RED_MAX) acc = fmax(acc, v);
    else if (op == RED_SUMSQ) acc += v * v;
    else acc += v;
  }
  double r = block_reduce(acc, op, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ void k_reduce_stage2(int op, const double* __restrict__ partials, int64_t n, double* out) {
  __shared__ double sh[16];
  double acc = (op == RED_MAX) ? -INFINITY : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
    acc = (op == RED_MAX) ? fmax(acc, partials[i]) : acc + partials[i];
  double r = block_reduce(acc, op == RED_SUMSQ ? RED_SUM : op, sh);
  if (threadIdx.x == 0) out[0] = r;
}

int64_t reduce_nblocks(int64_t n) {
  int64_t b = (n + kRedThreads - 1) / kRedThreads;
  if (b > kRedBlocks) b = kRedBlocks;
  if (b < 1) b = 1;
  return b;
}

hipError_t launch_reduce(int op, const float* x, int64_t n, double* partials, double* out, hipStream_t st) {
  const int64_t nb = reduce_nblocks(n);
  hipLaunchKernelGGL(k_reduce_stage1, dim3((unsigned)nb), dim3(kRedThreads), 0, st, op, x, n, partials);
  hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(1024), 0, st, op, (const double*)partials, nb, out);
  return hipGetLastError();
}

// complex max |z|^2
__global__ void k_cmax_stage1(const float2* __restrict__ z, int64_t n, double* partials) {
  __shared__ double sh[16];
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float2 v = z[i];
    acc = fmax(acc, (double)v.x * v.x + (double)v.y * v.y);
  }
  double r = block_reduce(acc, RED_MAX, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

hipError_t launch_cmax_abs2(const float2* z, int64_t n, double* partials, double* out, hipStream_t st) {
  const int64_t nb = reduce_nblocks(n);
  hipLaunchKernelGGL(k_cmax_stage1, dim3((unsigned)nb), dim3(kRedThreads), 0, st, z, n, partials);
  hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(1024), 0, st, (int)RED_MAX, (const double*)partials, nb, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// RL elementwise steps
// ---------------------------------------------------------------------------

static dim3 ew_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65535LL * 64) b = 65535LL * 64;
  return dim3((unsigned)b);
}

#define DH_EW_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

__global__ void k_rl_support(float* sens, float* x, int64_t n, const double* maxval) {
  const float mx = (float)maxval[0];
  const float floor_ = fmaxf(mx * 1e-6f, 1e-10f);
  DH_EW_LOOP(i, n) {
    const float s = sens[i];
    const bool act = s > floor_;
    sens[i] = act ? s : INFINITY;
    if (x) x[i] = act ? x[i] : 0.0f;
  }
}

// x /= total (a device float64 sum; unchanged when it is not positive)
__global__ void k_div_total(float* x, int64_t n, const double* total) {
  const double tv = *total;
  if (!(tv > 0.0)) return;
  const float inv = (float)(1.0 / tv);
  DH_EW_LOOP(i, n) x[i] = x[i] * inv;
}

__global__ void k_clamp_min(const float* in, float* out, int64_t n, float lo) {
  DH_EW_LOOP(i, n) out[i] = fmaxf(in[i], lo);
}

__global__ void k_sub_floor(const float* in, float* out, int64_t n, float a, float b) {
  DH_EW_LOOP(i, n) out[i] = fmaxf(in[i] - a, b);
}

__device__ __forceinline__ float idiv_term_ew(float o, float mdl) {
  const float os = fmaxf(o, 1e-6f), ms = fmaxf(mdl, 1e-6f);
  return o * logf(os / ms) - (o - ms);
}

__global__ void k_rl_ratio(const float* model, const float* obs, float* out, int64_t n, float bg, float eps,
                           double* partials) {
  __shared__ double sh[16];
  double acc = 0.0;
  DH_EW_LOOP(i, n) {
    const float m = model[i] + bg;
    const float o = obs[i];
    if (partials) acc += (double)idiv_term_ew(o, m);
    out[i] = o / fmaxf(m, eps);
  }
  if (partials) {
    double r = block_reduce(acc, RED_SUM, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
  }
}

__global__ void k_add_scalar(const float* in, const float* obs, float* out, int64_t n, float a,
                             double* partials) {
  __shared__ double sh[16];
  double acc = 0.0;
  DH_EW_LOOP(i, n) {
    const float m = in[i] + a;
    if (partials && obs) acc += (double)idiv_term_ew(obs[i], m);
    out[i] = m;
  }
  if (partials) {
    double r = block_reduce(acc, RED_SUM, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
  }
}

__global__ void k_scale_double(const double* in, double s, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = in[0] * s;
}

hipError_t launch_scale_double(const double* in, double s, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_double, dim3(1), dim3(64), 0, st, in, s, out);
  return hipGetLastError();
}

// sum of obs log(max(obs, eps) / max(model, eps)) - (obs - max(model, eps))  (rl_mlx.py:75-79)
__global__ void k_poisson_idiv(const float* obs, const float* model, int64_t n, float eps, double* partials) {
  __shared__ double sh[16];
  double acc = 0.0;
  DH_EW_LOOP(i, n) {
    const float o = obs[i];
    const float os = fmaxf(o, eps), ms = fmaxf(model[i], eps);
    acc += (double)(o * logf(os / ms) - (o - ms));
  }
  double r = block_reduce(acc, RED_SUM, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ void k_rl_update(float* x, const float* corr, const float* sens, int64_t n, float eps) {
  DH_EW_LOOP(i, n) x[i] = fmaxf(x[i] * (corr[i] / sens[i]), eps);
}

hipError_t launch_rl_support(float* sens, float* x, int64_t n, const double* maxval, hipStream_t st) {
  hipLaunchKernelGGL(k_rl_support, ew_grid(n), dim3(256), 0, st, sens, x, n, maxval);
  return hipGetLastError();
}
hipError_t launch_div_total(float* x, int64_t n, const double* total, hipStream_t st) {
  hipLaunchKernelGGL(k_div_total, ew_grid(n), dim3(256), 0, st, x, n, total);
  return hipGetLastError();
}

hipError_t launch_clamp_min(const float* in, float* out, int64_t n, float lo, hipStream_t st) {
  hipLaunchKernelGGL(k_clamp_min, ew_grid(n), dim3(256), 0, st, in, out, n, lo);
  return hipGetLastError();
}
hipError_t launch_sub_floor(const float* in, float* out, int64_t n, float a, float b, hipStream_t st) {
  hipLaunchKernelGGL(k_sub_floor, ew_grid(n), dim3(256), 0, st, in, out, n, a, b);
  return hipGetLastError();
}
hipError_t launch_poisson_idiv(const float* obs, const float* model, int64_t n, float eps, double* partials,
                               double* out, hipStream_t st) {
  const int64_t nb = reduce_nblocks(n);
  hipLaunchKernelGGL(k_poisson_idiv, dim3((unsigned)nb), dim3(kRedThreads), 0, st, obs, model, n, eps, partials);
  hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(1024), 0, st, (int)RED_SUM, (const double*)partials, nb, out);
  return hipGetLastError();
}
hipError_t launch_rl_ratio(const float* model, const float* obs, float* out, int64_t n, float bg, float eps,
                           double* partials, double* reduce_out, hipStream_t st) {
  const int64_t nb = reduce_nblocks(n);
  hipLaunchKernelGGL(k_rl_ratio, dim3((unsigned)nb), dim3(kRedThreads), 0, st, model, obs, out, n, bg, eps,
                     reduce_out ? partials : nullptr);
  if (reduce_out)
    hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(1024), 0, st, (int)RED_SUM, (const double*)partials, nb,
                       reduce_out);
  return hipGetLastError();
}
hipError_t launch_add_scalar(const float* in, const float* obs, float* out, int64_t n, float a,
                             double* partials, double* reduce_out, hipStream_t st) {
  const int64_t nb = reduce_nblocks(n);
  hipLaunchKernelGGL(k_add_scalar, dim3((unsigned)nb), dim3(kRedThreads), 0, st, in, obs, out, n, a,
                     reduce_out ? partials : nullptr);
  if (reduce_out)
    hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(1024), 0, st, (int)RED_SUM, (const double*)partials, nb,
                       reduce_out);
  return hipGetLastError();
}
hipError_t launch_rl_update(float* x, const float* corr, const float* sens, int64_t n, float eps,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_rl_update, ew_grid(n), dim3(256), 0, st, x, corr, sens, n, eps);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Layout operators
// ---------------------------------------------------------------------------

struct Box {
  int64_t ss[3], ds[3], s0[3], d0[3], ext[3];
};

__global__ void k_box_copy(const float* __restrict__ src, float* __restrict__ dst, Box b) {
  const int64_t n = b.ds[0] * b.ds[1] * b.ds[2];
  DH_EW_LOOP(i, n) {
    const int64_t x = i % b.ds[2];
    const int64_t y = (i / b.ds[2]) % b.ds[1];
    const int64_t z = i / (b.ds[2] * b.ds[1]);
    const int64_t rz = z - b.d0[0], ry = y - b.d0[1], rx = x - b.d0[2];
    float v = 0.0f;
    if (rz >= 0 && rz < b.ext[0] && ry >= 0 && ry < b.ext[1] && rx >= 0 && rx < b.ext[2])
      v = src[((rz + b.s0[0]) * b.ss[1] + (ry + b.s0[1])) * b.ss[2] + (rx + b.s0[2])];
    dst[i] = v;
  }
}

hipError_t launch_box_copy(const float* src, const int64_t* ss, float* dst, const int64_t* ds,
                           const int64_t* s0, const int64_t* d0, const int64_t* ext, hipStream_t st) {
  Box b;
  for (int a = 0; a < 3; ++a) {
    b.ss[a] = ss[a];
    b.ds[a] = ds[a];
    b.s0[a] = s0[a];
    b.d0[a] = d0[a];
    b.ext[a] = ext[a];
  }
  const int64_t n = ds[0] * ds[1] * ds[2];
  hipLaunchKernelGGL(k_box_copy, ew_grid(n), dim3(256), 0, st, src, dst, b);
  return hipGetLastError();
}

__global__ void k_banded(const float* __restrict__ src, float* __restrict__ dst, int64_t outer, int64_t n_in,
                         int64_t inner, int64_t n_out, const int32_t* __restrict__ idx,
                         const float* __restrict__ wt, int width) {
  const int64_t n = outer * n_out * inner;
  DH_EW_LOOP(i, n) {
    const int64_t q = i % inner;
    const int64_t o = (i / inner) % n_out;
    const int64_t b = i / (inner * n_out);
    const float* sb = src + b * n_in * inner + q;
    float acc = 0.0f;
    for (int k = 0; k < width; ++k) {
      const int32_t j = idx[o * width + k];
      acc += sb[(int64_t)j * inner] * wt[o * width + k];
    }
    dst[i] = acc;
  }
}

// Integer-factor block sum (FractionalAreaDownsample with scales that divide the
// extents, core_operators.py:205-219 applied along z, then y, then x): out[o0, o1, o2]
// = sum_c sum_b sum_a in[f0 o0 + a, f1 o1 + b, f2 o2 + c], each sum sequential from 0 in
// the order the three banded passes accumulate, so the value is the passes' bit for bit
// (their zero-weight taps add +0).  One thread per output sample, rows of the output
// along the grid's y dimension.
__global__ void k_block_sum(const float* __restrict__ in, float* __restrict__ out, int64_t rows, int n1, int n2,
                            int m1, int m2, int f0, int f1, int f2) {
  const int o2 = blockIdx.x * blockDim.x + threadIdx.x;
  if (o2 >= m2) return;
  for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
    const int64_t o0 = r / m1;
    const int o1 = (int)(r - o0 * m1);
    float r3 = 0.0f;
    for (int c = 0; c < f2; ++c) {
      float r2 = 0.0f;
      for (int b = 0; b < f1; ++b) {
        const float* p = in + ((o0 * f0) * n1 + (int64_t)o1 * f1 + b) * n2 + (int64_t)o2 * f2 + c;
        float r1 = 0.0f;
        for (int a = 0; a < f0; ++a) r1 += p[(int64_t)a * n1 * n2];
        r2 += r1;
      }
      r3 += r2;
    }
    __builtin_nontemporal_store(r3, out + r * m2 + o2);
  }
}

// k_block_sum for factors 2 x 2 x 2 and an output row length divisible by 2: a thread
// forms two outputs from four 16-byte loads (rows (2 o0 + a, 2 o1 + b), a, b in {0, 1}),
// adding in k_block_sum's order, so the values are the same bit for bit.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
__global__ void k_block_sum2(const float* __restrict__ in, float* __restrict__ out, int64_t rows, int n1, int n2,
                             int m1, int m2) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;  // output pair
  if (2 * h >= m2) return;
  for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
    const int64_t o0 = r / m1;
    const int o1 = (int)(r - o0 * m1);
    const float* p = in + ((2 * o0) * n1 + 2 * (int64_t)o1) * n2 + 4 * (int64_t)h;
    const int64_t pl = (int64_t)n1 * n2;
    const f4v v00 = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));            // a 0, b 0
    const f4v v10 = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p + pl));       // a 1, b 0
    const f4v v01 = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p + n2));       // a 0, b 1
    const f4v v11 = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p + pl + n2));  // a 1, b 1
    f2v o;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      float r3 = 0.0f;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int i = 2 * q + c;
        float r2 = 0.0f;
        float r1 = 0.0f;
        r1 += v00[i];
        r1 += v10[i];
        r2 += r1;
        r1 = 0.0f;
        r1 += v01[i];
        r1 += v11[i];
        r2 += r1;
        r3 += r2;
      }
      o[q] = r3;
    }
    __builtin_nontemporal_store(o, reinterpret_cast<f2v*>(out + r * m2 + 2 * h));
  }
}

// Its adjoint (the banded adjoint passes with one unit tap each): in[z, y, x] of the fine
// grid = coarse[z / f0, y / f1, x / f2].  A thread writes four consecutive samples (one
// 16-byte store) of a fine row; rows are dealt out grid-stride.
// vec: fine 16-byte and coarse 8-byte aligned.
__global__ void k_block_spread(const float* __restrict__ coarse, float* __restrict__ fine, int64_t rows, int n1,
                               int n2, int m1, int m2, int f0, int f1, int f2, int vec) {
  const int x = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (x >= n2) return;
  for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {  // fine rows (z, y)
    const int64_t z = r / n1;
    const int y = (int)(r - z * n1);
    const float* c = coarse + ((z / f0) * m1 + y / f1) * (int64_t)m2;
    float* d = fine + r * n2 + x;
    if (vec && f2 == 2 && (n2 & 3) == 0) {  // x + 3 < n2; one 8-byte load
      const f2v w = *reinterpret_cast<const f2v*>(c + (x >> 1));
      const f4v v = {w[0], w[0], w[1], w[1]};
      __builtin_nontemporal_store(v, reinterpret_cast<f4v*>(d));
    } else if (vec && x + 3 < n2 && (n2 & 3) == 0) {
      const f4v v = {c[x / f2], c[(x + 1) / f2], c[(x + 2) / f2], c[(x + 3) / f2]};
      __builtin_nontemporal_store(v, reinterpret_cast<f4v*>(d));
    } else {
      for (int j = 0; j < 4 && x + j < n2; ++j) __builtin_nontemporal_store(c[(x + j) / f2], d + j);
    }
  }
}

hipError_t launch_block_sum(const float* in, float* out, int n0, int n1, int n2, int f0, int f1, int f2,
                            hipStream_t st) {
  const int m1 = n1 / f1, m2 = n2 / f2;
  const int64_t rows = (int64_t)(n0 / f0) * m1;
  const unsigned gy = (unsigned)std::min<int64_t>(rows, 65535);
  if (f0 == 2 && f1 == 2 && f2 == 2 && (m2 & 1) == 0 && (reinterpret_cast<uintptr_t>(in) & 15) == 0 &&
      (reinterpret_cast<uintptr_t>(out) & 7) == 0) {
    hipLaunchKernelGGL(k_block_sum2, dim3((unsigned)((m2 / 2 + 127) / 128), gy), dim3(128), 0, st, in, out, rows,
                       n1, n2, m1, m2);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_block_sum, dim3((unsigned)((m2 + 127) / 128), gy), dim3(128), 0, st, in, out, rows, n1, n2,
                     m1, m2, f0, f1, f2);
  return hipGetLastError();
}

hipError_t launch_block_spread(const float* coarse, float* fine, int n0, int n1, int n2, int f0, int f1, int f2,
                               hipStream_t st) {
  const int m1 = n1 / f1, m2 = n2 / f2;
  const int64_t rows = (int64_t)n0 * n1;
  const int quads = (n2 + 3) / 4;
  const unsigned gx = (unsigned)((quads + 127) / 128);
  // about 8 rows per workgroup row slot: enough workgroups to fill the GPU, few enough to amortise them
  const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>((rows + 7) / 8, 65535));
  const int vec = (reinterpret_cast<uintptr_t>(fine) & 15) == 0 && (reinterpret_cast<uintptr_t>(coarse) & 7) == 0;
  hipLaunchKernelGGL(k_block_spread, dim3(gx, gy), dim3(128), 0, st, coarse, fine, rows, n1, n2, m1, m2, f0, f1,
                     f2, vec);
  return hipGetLastError();
}

hipError_t launch_banded(const float* src, float* dst, int64_t outer, int64_t n_in, int64_t inner, int64_t n_out,
                         const int32_t* idx, const float* wt, int width, hipStream_t st) {
  const int64_t n = outer * n_out * inner;
  hipLaunchKernelGGL(k_banded, ew_grid(n), dim3(256), 0, st, src, dst, outer, n_in, inner, n_out, idx, wt, width);
  return hipGetLastError();
}

}  // namespace dh
