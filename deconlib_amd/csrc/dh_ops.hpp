// Elementwise steps and deterministic reductions (host launchers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dh {

enum RedOp { RED_SUM = 0, RED_MAX = 1, RED_SUMSQ = 2 };

int64_t reduce_nblocks(int64_t n);
hipError_t launch_reduce(int op, const float* x, int64_t n, double* partials, double* out, hipStream_t st);
hipError_t launch_cmax_abs2(const float2* z, int64_t n, double* partials, double* out, hipStream_t st);
hipError_t launch_rl_support(float* sens, float* x, int64_t n, const double* maxval, hipStream_t st);
hipError_t launch_clamp_min(const float* in, float* out, int64_t n, float lo, hipStream_t st);
hipError_t launch_div_total(float* x, int64_t n, const double* total, hipStream_t st);
hipError_t launch_sub_floor(const float* in, float* out, int64_t n, float a, float b, hipStream_t st);
hipError_t launch_rl_ratio(const float* model, const float* obs, float* out, int64_t n, float bg, float eps,
                           double* partials, double* reduce_out, hipStream_t st);
hipError_t launch_add_scalar(const float* in, const float* obs, float* out, int64_t n, float a,
                             double* partials, double* reduce_out, hipStream_t st);
// out[0] = in[0] * s (a device scalar rescaled on the device)
hipError_t launch_scale_double(const double* in, double s, double* out, hipStream_t st);
hipError_t launch_poisson_idiv(const float* obs, const float* model, int64_t n, float eps, double* partials,
                               double* out, hipStream_t st);
hipError_t launch_rl_update(float* x, const float* corr, const float* sens, int64_t n, float eps,
                            hipStream_t st);

hipError_t launch_box_copy(const float* src, const int64_t* ss, float* dst, const int64_t* ds,
                           const int64_t* s0, const int64_t* d0, const int64_t* ext, hipStream_t st);
hipError_t launch_block_sum(const float* in, float* out, int n0, int n1, int n2, int f0, int f1, int f2,
                            hipStream_t st);
hipError_t launch_block_spread(const float* coarse, float* fine, int n0, int n1, int n2, int f0, int f1, int f2,
                               hipStream_t st);
hipError_t launch_banded(const float* src, float* dst, int64_t outer, int64_t n_in, int64_t inner, int64_t n_out,
                         const int32_t* idx, const float* wt, int width, hipStream_t st);

}  // namespace dh
