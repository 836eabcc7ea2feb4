// NLCG (nonlinear conjugate gradients, Poisson likelihood, f = s^2) kernels:
// host launchers.  Scalars live in a caller-owned float64 block (DH_NL_* slots).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/deconhip.h"

namespace dh {

int64_t nl_partials_needed();
hipError_t nl_init(const float* f0, float* s, float* f, int64_t n, float floor_, hipStream_t st);
hipError_t nl_model(const float* m, const float* g, float* w, int64_t n, float eps, double* sc, double* part,
                    hipStream_t st);
hipError_t nl_begin_step(const float* s, const float* d, float* sd, float* dd, int64_t n, double* sc,
                         int newton_iters, hipStream_t st);
hipError_t nl_reg_dots(const float* cf, const float* csd, const float* cdd, int64_t n, double* sc, double* part,
                       hipStream_t st);
hipError_t nl_newton(const float* m, const float* ksd, const float* kdd, const float* g, int64_t n, float bg,
                     double beta, double eps, double* sc, double* part, hipStream_t st);
hipError_t nl_trial(const float* m, const float* ksd, const float* kdd, const float* g, int64_t n, float bg,
                    double beta, double eps, double* sc, float* m_new, float* w, double* part, hipStream_t st);
hipError_t nl_update(float* s, const float* d, float* f, int64_t n, double eps, double* sc, double* part,
                     hipStream_t st);
hipError_t nl_gradient(const float* s, const float* ktw, const float* ctcf, float c4b, float* r, int64_t n, int mode,
                       double eps, double* sc, double* part, hipStream_t st);
hipError_t nl_direction(const float* r, float* d, int64_t n, const double* sc, hipStream_t st);
hipError_t nl_quad(const float* g, const float* m, const float* ksd, const float* v, const float* d, int64_t n,
                   double eps, double* out2, double* part, hipStream_t st);
hipError_t nl_idiv(const float* fp, const float* fc, int64_t n, double eps, double* out2, double* part,
                   hipStream_t st);

}  // namespace dh
