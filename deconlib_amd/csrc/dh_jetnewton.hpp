// JetNewton (active-set projected Newton, exact Hessian) kernels: host launchers.
// Scalars live in a caller-owned float64 block (DH_JN_* slots).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/deconhip.h"

namespace dh {

int64_t jn_partials_needed();
hipError_t jn_penalty(const float* x, int nz, int ny, int nx, const float* w6, float beta, float eta, float* r_out,
                      float* w_out, float* c_out, float* wr_out, float* u_out, double* sc, double* part,
                      hipStream_t st);
hipError_t jn_penalty_stacked(const float* r, int nc, int64_t n, float beta, float eta, float* w_out, float* c_out,
                              float* wr_out, float* u_out, double* sc, double* part, hipStream_t st);
hipError_t jn_hvp_operand(const float* v, const float* r, const float* w, const float* c, int nz, int ny, int nx,
                          const float* w6, float* out, hipStream_t st);
hipError_t jn_hvp_operand_stacked(const float* q, const float* r, const float* w, const float* c, int nc, int64_t n,
                                  float* out, hipStream_t st);
hipError_t jn_kkt(const float* x, const float* g, int64_t n, double eps_bar, double* sc, double* part,
                  hipStream_t st);
hipError_t jn_active(const float* x, const float* g, int64_t n, float tau, float delta, uint8_t* frozen,
                     uint8_t* active, float* r, float* p, float* xcg, double eps, double* sc, double* part,
                     hipStream_t st);
hipError_t jn_cg_ap(const float* kt, float scale, const float* htu, const uint8_t* active, const float* p, float* ap,
                    int64_t n, double eps, double* sc, double* part, hipStream_t st);
hipError_t jn_cg_update(float* x, float* r, float* p, const float* ap, int64_t n, double* sc, double* part,
                        hipStream_t st);
hipError_t jn_direction(const float* xcg, const float* g, const uint8_t* active, float* d, int64_t n, double* sc,
                        double* part, hipStream_t st);
hipError_t jn_trial(const float* x, const float* d, float t, const float* g, float* xt, int64_t n, double* sc,
                    double* part, hipStream_t st);

}  // namespace dh
