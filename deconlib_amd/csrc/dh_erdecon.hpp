// ER-Decon (Hessian-log regularised Gauss-Newton-CG) kernels: host launchers.
// Scalars live in a caller-owned float64 block (DH_ER_* slots).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/deconhip.h"

namespace dh {

int64_t er_partials_needed();
hipError_t er_hess_weights(const float* g, int nz, int ny, int nx, const float* w6, float lam, float eps_reg,
                           float w_floor, float iw, float* w_out, float* u_out, double* sc, double* part,
                           hipStream_t st);
hipError_t er_weights_stacked(const float* hg, int nc, const float* g, int64_t n, float lam, float eps_reg,
                              float w_floor, float iw, float* w_out, float* u_out, double* sc, double* part,
                              hipStream_t st);
hipError_t er_hess_apply_w(const float* v, const float* w, int nz, int ny, int nx, const float* w6, float* u_out,
                           hipStream_t st);
hipError_t er_scale_stacked(const float* x, const float* w, int nc, int64_t n, float* out, hipStream_t st);
hipError_t er_data(const float* m, const float* f, int64_t n, int poisson, float* score, float* hessw, double* sc,
                   double* part, hipStream_t st);
hipError_t er_gradient(const float* s, const float* dg, const float* rh, const float* w, const float* g, float iw,
                       int64_t n, double cg_tol, float* grad, float* x, float* r, float* p, float* sv, double* sc,
                       double* part, hipStream_t st);
hipError_t er_cg_ap(const float* s, const float* kt, const float* htu, const float* w, const float* sv, float iw,
                    const float* p, float* ap, int64_t n, double* sc, double* part, hipStream_t st);
hipError_t er_cg_update(float* x, float* r, float* p, const float* ap, const float* s, float* sv, int64_t n,
                        double* sc, double* part, hipStream_t st);
hipError_t er_dot(const float* a, const float* b, int64_t n, int slot, double* sc, double* part, hipStream_t st);
hipError_t er_ew(int op, const float* a, const float* b, float alpha, float beta, float* out, float* out2, int64_t n,
                 hipStream_t st);

}  // namespace dh
