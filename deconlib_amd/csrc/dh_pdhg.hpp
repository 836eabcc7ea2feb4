// PDHG-Hessian kernels: device scalar block layout and host launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/deconhip.h"

namespace dh {

// float64 scalar block slots (see DH_PD_* in deconhip.h)
enum PdSlot {
  PD_TAU = DH_PD_TAU,
  PD_SIGMA = DH_PD_SIGMA,
  PD_THETA = DH_PD_THETA,
  PD_STEP = DH_PD_STEP,
  PD_COUNT = DH_PD_COUNT,
  PD_DONE = DH_PD_DONE,
  PD_FINAL = DH_PD_FINAL,
  PD_DX2 = DH_PD_DX2,
  PD_XX = DH_PD_XX,
  PD_XNXN = DH_PD_XNXN,
  PD_DA2 = DH_PD_DA2,
  PD_DL2 = DH_PD_DL2,
  PD_TMP0 = DH_PD_TMP0,
  PD_TMP1 = DH_PD_TMP1,
};

typedef dh_pd_params PdParams;

hipError_t pd_y1(float* y1, const float* ax, const float* axp, const float* data, int64_t n, const double* sc,
                 float bg, int loss_type, hipStream_t st);
hipError_t pd_dual(float* y2, const float* x, const float* xp, int nz, int ny, int nx, const double* sc,
                   const float* w, float alpha, int l12, hipStream_t st);
hipError_t pd_primal(float* xn, const float* x, const float* g, const float* y2, int nz, int ny, int nx,
                     double* sc, const float* w, double* part, hipStream_t st);
hipError_t pd_dual_primal(float* y2_new, float* xn, const float* x, const float* xp, const float* y2, const float* g,
                          int nz, int ny, int nx, double* sc, const float* w, float alpha, int l12, double* part,
                          hipStream_t st);
hipError_t pd_hess_sums(const float* a, const float* b, int nz, int ny, int nx, const float* w, int want_abs,
                        double* sc, double* loss_out, double* part, hipStream_t st);
hipError_t pd_data_sums(const float* a, const float* b, const float* data, int64_t n, float bg, int want_loss,
                        int loss_type, double* sc, double* loss_out, double* part, hipStream_t st);
hipError_t pd_move(double* sc, int from0, int to0, int from1, double* out1, hipStream_t st);
hipError_t pd_init(double* sc, const double* blur_norm_sq, double reg_norm_sq, hipStream_t st);
hipError_t pd_scalars(double* sc, double* hist, int k, const PdParams& pp, hipStream_t st);
int64_t pd_partials_needed(int nz, int ny, int nx);
// gradient (DH_REG_GRADIENT, weights w[3]) and identity (DH_REG_IDENTITY) regularizers
hipError_t pr_dual(int kind, float* y2, const float* x, const float* xp, int nz, int ny, int nx, const double* sc,
                   const float* w, float alpha, int l12, hipStream_t st);
hipError_t pr_primal(int kind, float* xn, const float* x, const float* g, const float* y2, int nz, int ny, int nx,
                     double* sc, const float* w, double* part, hipStream_t st);
hipError_t pr_sums(int kind, const float* a, const float* b, int nz, int ny, int nx, const float* w, int want_abs,
                   double* sc, double* loss_out, double* part, hipStream_t st);
hipError_t grad_forward(const float* f, float* out, int nz, int ny, int nx, const float* w, hipStream_t st);
hipError_t grad_adjoint(const float* y, float* out, int nz, int ny, int nx, const float* w, hipStream_t st);
hipError_t hess_forward(const float* f, float* out, int nz, int ny, int nx, const float* w, hipStream_t st);
hipError_t hess_adjoint(const float* y2, float* out, int nz, int ny, int nx, const float* w, hipStream_t st);

}  // namespace dh
