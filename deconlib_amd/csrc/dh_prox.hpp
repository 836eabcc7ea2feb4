// PDHG data-term dual step on one sample (pdhg_mlx.py:115-189 with A xbar formed by
// linearity, :717-723).  Shared by k_pd_y1 and the correlation's PRO_PD_Y1 prologue so
// that both round alike: every multiply-add that could be contracted is written as fmaf.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/deconhip.h"

namespace dh {

struct PdY1Consts {
  float sf, th, one_pth, sb, s4, s2, inv1ps, bg;
};

// From the device scalar block (sigma, theta in float64) and the background.
__device__ __forceinline__ PdY1Consts pd_y1_consts(const double* sc, float bg) {
  const double sigma = sc[DH_PD_SIGMA], theta = sc[DH_PD_THETA];
  return PdY1Consts{(float)sigma,         (float)theta,        (float)(1.0 + theta), (float)(sigma * (double)bg),
                    (float)(4.0 * sigma), (float)(2.0 * sigma), (float)(1.0 + sigma), bg};
}

// y1' = prox_{sigma F*}(y1 + sigma A xbar), A xbar = (1 + theta) A x_k - theta A x_{k-1};
// loss 0: Poisson (the stable root, D ~ 0 handled as the reference), 1: Gaussian.
__device__ __forceinline__ float pd_y1_prox(float y, float ax, float axp, float d, const PdY1Consts& k, int loss) {
  const float axb = fmaf(k.one_pth, ax, -(k.th * axp));
  const float arg = fmaf(k.sf, axb, y);
  if (loss == 0) {
    const float c = (1.0f - arg) - k.sb;
    const float ds = fmaxf(d, 0.0f);
    const float disc = fmaf(c, c, k.s4 * ds);
    const float den = (sqrtf(disc) - c) + 1e-12f;
    float z = (k.s2 * ds) / den;
    if (ds < 1e-6f) z = fmaxf(c, 1e-6f);
    z = fmaxf(z, 1e-6f);
    return 1.0f - z;
  }
  return fmaf(-k.sf, d - k.bg, arg) / k.inv1ps;
}

}  // namespace dh
