// ER-Decon: edge-preserving Hessian-log deconvolution by Gauss-Newton-CG
// (erdecon_mlx.py:144-692).  The kernels around the convolutions: the
// per-voxel curvature q = sum_i (H_i g)^2 and IRLS weight w = lam / (eps + q)
// computed from the Hessian stencil on the fly (the six components are only
// written as w * H g, the operand of the adjoint), the data-term score and
// Gauss-Newton diagonal, the gradient, and the inner conjugate-gradient
// updates with their scalar logic on the device (float64 scalar block), so a
// CG step enqueues without a host round trip.
//
// Elementwise arithmetic is float32 in the reference's operation order; sums
// are two-stage and deterministic (dh_blocksum.hpp).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "dh_blocksum.hpp"
#include "dh_erdecon.hpp"
#include "dh_stencil.hpp"

namespace dh {

constexpr int kErThreads = kSumThreads;
constexpr int kErBlocks = 1024;
constexpr int kErMaxSums = 2;

#define ER_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * kErThreads + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kErThreads)

static int er_blocks(int64_t n) {
  int64_t b = (n + kErThreads - 1) / kErThreads;
  if (b > kErBlocks) b = kErBlocks;
  return (int)(b < 1 ? 1 : b);
}

__device__ __forceinline__ float er_idiv_term(float o, float mdl) {  // rl_mlx.py:58-80, eps 1e-6
  const float os = fmaxf(o, 1e-6f), ms = fmaxf(mdl, 1e-6f);
  return o * logf(os / ms) - (o - ms);
}

struct ErWeights {
  float lam, eps, w_floor, iw;
};

// q from the channel responses h[0..nc) and g; w = lam / (eps + q) [+ floor].
__device__ __forceinline__ float er_q(const float* h, int nc, float gv, float iw) {
  float q = h[0] * h[0];
  for (int c = 1; c < nc; ++c) q += h[c] * h[c];
  if (iw > 0.0f) q = q + iw * (gv * gv);
  return q;
}

// ---------------------------------------------------------------------------
// regulariser: weights and sums
// ---------------------------------------------------------------------------

// Hessian3D (weights w6; Hessian2D is the one-plane case) of g at every voxel:
// sums log(eps + q) and q; with w_out, w and u = w * H g (6, n).
__global__ void __launch_bounds__(kErThreads) k_er_hess_weights(const float* __restrict__ g, Dims d, W6 w6,
                                                                ErWeights p, float* __restrict__ w_out,
                                                                float* __restrict__ u_out,
                                                                double* __restrict__ part) {
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  auto gv = [&](int z, int y, int x) { return g[((int64_t)z * d.ny + y) * d.nx + x]; };
  double acc[2] = {0.0, 0.0};
  ER_LOOP(i, n) {
    const int x = (int)(i % d.nx);
    const int y = (int)((i / d.nx) % d.ny);
    const int z = (int)(i / ((int64_t)d.nx * d.ny));
    float h[6];
    hess6(gv, d, z, y, x, w6.w, h);
    const float q = er_q(h, 6, g[i], p.iw);
    acc[0] += (double)logf(p.eps + q);
    acc[1] += (double)q;
    if (w_out) {
      float wv = p.lam / (p.eps + q);
      if (p.w_floor != 0.0f) wv = wv + p.w_floor;
      w_out[i] = wv;
#pragma unroll
      for (int c = 0; c < 6; ++c) u_out[c * n + i] = wv * h[c];
    }
  }
  block_partials<2>(acc, part);
}

// The same for any operator's stacked responses hg (nc, n); u_out may alias hg.
__global__ void __launch_bounds__(kErThreads) k_er_weights_stacked(const float* hg, int nc,
                                                                   const float* __restrict__ g, int64_t n,
                                                                   ErWeights p, float* __restrict__ w_out, float* u_out,
                                                                   double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  ER_LOOP(i, n) {
    float q = 0.0f;
    for (int c = 0; c < nc; ++c) {
      const float h = hg[c * n + i];
      q = (c == 0) ? h * h : q + h * h;
    }
    if (p.iw > 0.0f) {
      const float gg = g[i];
      q = q + p.iw * (gg * gg);
    }
    acc[0] += (double)logf(p.eps + q);
    acc[1] += (double)q;
    if (w_out) {
      float wv = p.lam / (p.eps + q);
      if (p.w_floor != 0.0f) wv = wv + p.w_floor;
      w_out[i] = wv;
      for (int c = 0; c < nc; ++c) u_out[c * n + i] = wv * hg[c * n + i];
    }
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kErThreads) k_er_reg_finish(const double* __restrict__ part, int nb, double* sc) {
  double s[2];
  sum_partials<2>(part, nb, s);
  if (threadIdx.x == 0) {
    sc[DH_ER_LOGQ] = s[0];
    sc[DH_ER_Q] = s[1];
  }
}

// u = w * H(v) (six weighted components) through the sliding LDS plane window of v
// (each input voxel read from HBM about once, the six outputs streamed).
__global__ void __launch_bounds__(kTileThreads) k_er_hess_apply_w_tiled(const float* __restrict__ v,
                                                                        const float* __restrict__ w, Dims d, W6 w6,
                                                                        float* __restrict__ u) {
  __shared__ float pl[3 * kPlane];
  const Tile t = tile_of(d);
  const PlanePos pp(d, t);
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  float a0, a1;
  for (int zz = t.z0 - 1; zz <= t.z0; ++zz) {
    pp.load(v, pp.base(d, zz), a0, a1);
    pp.store(pl + slot_of(zz, t.z0) * kPlane, a0, a1);
  }
  pp.load(v, pp.base(d, t.z0 + 1), a0, a1);
  for (int z = t.z0; z < t.z1; ++z) {
    pp.store(pl + slot_of(z + 1, t.z0) * kPlane, a0, a1);
    __syncthreads();
    if (z + 1 < t.z1) pp.load(v, pp.base(d, z + 2), a0, a1);
    if (t.in) {
      auto f = [&](int zz, int yy, int xx) { return lds_at(pl, slot_of(zz, t.z0), t, yy, xx); };
      float h[6];
      hess6_halo(f, z, t.y, t.x, w6.w, h);
      const int64_t i = ((int64_t)z * d.ny + t.y) * d.nx + t.x;
      const float wv = __builtin_nontemporal_load(w + i);
#pragma unroll
      for (int c = 0; c < 6; ++c) __builtin_nontemporal_store(wv * h[c], u + c * n + i);
    }
    __syncthreads();
  }
}

__global__ void k_er_scale_stacked(const float* x, const float* __restrict__ w, int nc, int64_t n, float* out) {
  ER_LOOP(i, n) {
    const float wv = w[i];
    for (int c = 0; c < nc; ++c) out[c * n + i] = wv * x[c * n + i];
  }
}

// ---------------------------------------------------------------------------
// data term
// ---------------------------------------------------------------------------

// sums: the misfit (Gaussian sum (m - f)^2, Poisson sum(mm - f ln mm), mm = max(m, 1e-8))
// and the Poisson I-divergence of m against f; score / Gauss-Newton diagonal if asked.
__global__ void __launch_bounds__(kErThreads) k_er_data(const float* __restrict__ m, const float* __restrict__ f,
                                                        int64_t n, int poisson, float* __restrict__ score,
                                                        float* __restrict__ hessw, double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  ER_LOOP(i, n) {
    const float mv = m[i], fv = f[i];
    if (poisson) {
      const float mm = fmaxf(mv, 1e-8f);
      acc[0] += (double)mm - (double)(fv * logf(mm));
      if (score) {
        score[i] = 1.0f - fv / mm;
        hessw[i] = fv / (mm * mm);
      }
    } else {
      const float r = mv - fv;
      acc[0] += (double)(r * r);
      if (score) score[i] = 2.0f * (mv - fv);
    }
    acc[1] += (double)er_idiv_term(fv, mv);
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kErThreads) k_er_data_finish(const double* __restrict__ part, int nb, double* sc) {
  double s[2];
  sum_partials<2>(part, nb, s);
  if (threadIdx.x == 0) {
    sc[DH_ER_DATA] = s[0];
    sc[DH_ER_IDIV] = s[1];
  }
}

// ---------------------------------------------------------------------------
// gradient and the inner CG
// ---------------------------------------------------------------------------

// grad = 2 s K^T(score) + 2 s H^T(w H g) [+ 2 s (iw w g)]; CG starts at x = 0,
// r = p = -grad, sv = s p; sum grad^2.
__global__ void __launch_bounds__(kErThreads) k_er_gradient(const float* __restrict__ s,
                                                            const float* __restrict__ dg,
                                                            const float* __restrict__ rh,
                                                            const float* __restrict__ w, const float* __restrict__ g,
                                                            float iw, int64_t n, float* __restrict__ grad,
                                                            float* __restrict__ x, float* __restrict__ r,
                                                            float* __restrict__ p, float* __restrict__ sv,
                                                            double* __restrict__ part) {
  double acc[1] = {0.0};
  ER_LOOP(i, n) {
    const float sv0 = s[i];
    const float two_s = 2.0f * sv0;
    float gr = two_s * dg[i] + two_s * rh[i];
    if (iw > 0.0f) gr = gr + two_s * ((iw * w[i]) * g[i]);
    grad[i] = gr;
    x[i] = 0.0f;
    r[i] = -gr;
    p[i] = -gr;
    sv[i] = sv0 * -gr;
    acc[0] += (double)(gr * gr);
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(kErThreads) k_er_gradient_finish(const double* __restrict__ part, int nb,
                                                                   double* sc, double tol) {
  double s[1];
  sum_partials<1>(part, nb, s);
  if (threadIdx.x == 0) {
    sc[DH_ER_GG] = s[0];
    sc[DH_ER_RS] = s[0];
    sc[DH_ER_RS_THR] = (tol * tol) * s[0];
    sc[DH_ER_STOP] = s[0] <= 1e-20 ? 1.0 : 0.0;
  }
}

// Ap = 4 s K^T(d K(sv)) + 4 s H^T(w H(sv)) [+ 4 iw s (w sv)]; sum p Ap.
__global__ void __launch_bounds__(kErThreads) k_er_cg_ap(const float* __restrict__ s, const float* __restrict__ kt,
                                                         const float* __restrict__ htu, const float* __restrict__ w,
                                                         const float* __restrict__ sv, float iw,
                                                         const float* __restrict__ p, float* __restrict__ ap, int64_t n,
                                                         double* __restrict__ part) {
  double acc[1] = {0.0};
  const float c4iw = 4.0f * iw;
  ER_LOOP(i, n) {
    const float s4 = 4.0f * s[i];
    float a = s4 * kt[i];
    float b = s4 * htu[i];
    if (iw > 0.0f) b = b + (c4iw * s[i]) * (w[i] * sv[i]);
    const float v = a + b;
    ap[i] = v;
    acc[0] += (double)(p[i] * v);
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(kErThreads) k_er_cg_ap_finish(const double* __restrict__ part, int nb, double* sc) {
  double s[1];
  sum_partials<1>(part, nb, s);
  if (threadIdx.x == 0) {
    sc[DH_ER_DENOM] = s[0];
    if (!isfinite(s[0]) || s[0] <= 1e-20) sc[DH_ER_STOP] = 1.0;  // non-positive curvature
  }
}

// x += a p, r -= a Ap with a = RS / DENOM; sum r^2.
__global__ void __launch_bounds__(kErThreads) k_er_cg_xr(float* __restrict__ x, float* __restrict__ r,
                                                         const float* __restrict__ p, const float* __restrict__ ap,
                                                         int64_t n, const double* __restrict__ sc,
                                                         double* __restrict__ part) {
  if (sc[DH_ER_STOP] != 0.0) return;  // uniform
  const float a = (float)(sc[DH_ER_RS] / sc[DH_ER_DENOM]);
  double acc[1] = {0.0};
  ER_LOOP(i, n) {
    x[i] = x[i] + a * p[i];
    const float rn = r[i] - a * ap[i];
    r[i] = rn;
    acc[0] += (double)(rn * rn);
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(kErThreads) k_er_cg_xr_finish(const double* __restrict__ part, int nb, double* sc) {
  if (sc[DH_ER_STOP] != 0.0) return;
  double s[1];
  sum_partials<1>(part, nb, s);
  if (threadIdx.x != 0) return;
  const double rs_new = s[0];
  if (!isfinite(rs_new) || rs_new < 0.0 || rs_new <= sc[DH_ER_RS_THR]) {
    sc[DH_ER_STOP] = 1.0;
    return;
  }
  sc[DH_ER_BETA] = rs_new / sc[DH_ER_RS];
  sc[DH_ER_RS] = rs_new;
}

// p = r + beta p, sv = s p.
__global__ void k_er_cg_p(const float* __restrict__ r, float* __restrict__ p, const float* __restrict__ s,
                          float* __restrict__ sv, int64_t n, const double* __restrict__ sc) {
  if (sc[DH_ER_STOP] != 0.0) return;
  const float b = (float)sc[DH_ER_BETA];
  ER_LOOP(i, n) {
    const float pn = r[i] + b * p[i];
    p[i] = pn;
    sv[i] = s[i] * pn;
  }
}

__global__ void __launch_bounds__(kErThreads) k_er_dot(const float* __restrict__ a, const float* __restrict__ b,
                                                       int64_t n, double* __restrict__ part) {
  double acc[1] = {0.0};
  ER_LOOP(i, n) acc[0] += (double)(a[i] * b[i]);
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(kErThreads) k_er_dot_finish(const double* __restrict__ part, int nb, double* sc,
                                                              int slot) {
  double s[1];
  sum_partials<1>(part, nb, s);
  if (threadIdx.x == 0) sc[slot] = s[0];
}

// out may alias a (in-place steps); each element is read before it is written by one thread.
__global__ void k_er_ew(int op, const float* a, const float* b, float alpha, float beta, float* out, float* out2,
                        int64_t n) {
  ER_LOOP(i, n) {
    const float av = a[i];
    switch (op) {
      case DH_ER_EW_DIV: out[i] = av / alpha; break;
      case DH_ER_EW_SQUARE_SCALE: out[i] = (av * av) * alpha; break;
      case DH_ER_EW_AXPY_SQUARE: {
        const float v = av + alpha * b[i];
        out[i] = v;
        out2[i] = v * v;
        break;
      }
      case DH_ER_EW_NEG: out[i] = -av; break;
      case DH_ER_EW_MUL: out[i] = av * b[i]; break;
      case DH_ER_EW_AFFINE: out[i] = alpha * av + beta; break;
      case DH_ER_EW_MAX_DIV: out[i] = fmaxf(av, alpha) / alpha; break;
      case DH_ER_EW_AXPY: out[i] = alpha * av + b[i]; break;
      default: out[i] = alpha * av; break;  // DH_ER_EW_SCALE
    }
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

int64_t er_partials_needed() { return (int64_t)kErBlocks * kErMaxSums; }

#define ER_LAUNCH(kern, nb, ...) hipLaunchKernelGGL(kern, dim3((unsigned)(nb)), dim3(kErThreads), 0, st, __VA_ARGS__)
#define ER_FINISH(kern, nb, ...) hipLaunchKernelGGL(kern, dim3(1), dim3(kErThreads), 0, st, part, nb, __VA_ARGS__)

static W6 w6_of(const float* w) {
  W6 r;
  for (int q = 0; q < 6; ++q) r.w[q] = w[q];
  return r;
}

hipError_t er_hess_weights(const float* g, int nz, int ny, int nx, const float* w6, float lam, float eps_reg,
                           float w_floor, float iw, float* w_out, float* u_out, double* sc, double* part,
                           hipStream_t st) {
  const int64_t n = (int64_t)nz * ny * nx;
  const int nb = er_blocks(n);
  ER_LAUNCH(k_er_hess_weights, nb, g, Dims{nz, ny, nx}, w6_of(w6), ErWeights{lam, eps_reg, w_floor, iw}, w_out, u_out,
            part);
  ER_FINISH(k_er_reg_finish, nb, sc);
  return hipGetLastError();
}

hipError_t er_weights_stacked(const float* hg, int nc, const float* g, int64_t n, float lam, float eps_reg,
                              float w_floor, float iw, float* w_out, float* u_out, double* sc, double* part,
                              hipStream_t st) {
  const int nb = er_blocks(n);
  ER_LAUNCH(k_er_weights_stacked, nb, hg, nc, g, n, ErWeights{lam, eps_reg, w_floor, iw}, w_out, u_out, part);
  ER_FINISH(k_er_reg_finish, nb, sc);
  return hipGetLastError();
}

hipError_t er_hess_apply_w(const float* v, const float* w, int nz, int ny, int nx, const float* w6, float* u_out,
                           hipStream_t st) {
  const int64_t n = (int64_t)nz * ny * nx;
  hipLaunchKernelGGL(k_er_hess_apply_w_tiled, tile_grid(nz, ny, nx), dim3(kTileThreads), 0, st, v, w,
                     Dims{nz, ny, nx}, w6_of(w6), u_out);
  (void)n;
  return hipGetLastError();
}

hipError_t er_scale_stacked(const float* x, const float* w, int nc, int64_t n, float* out, hipStream_t st) {
  ER_LAUNCH(k_er_scale_stacked, er_blocks(n), x, w, nc, n, out);
  return hipGetLastError();
}

hipError_t er_data(const float* m, const float* f, int64_t n, int poisson, float* score, float* hessw, double* sc,
                   double* part, hipStream_t st) {
  const int nb = er_blocks(n);
  ER_LAUNCH(k_er_data, nb, m, f, n, poisson, score, hessw, part);
  ER_FINISH(k_er_data_finish, nb, sc);
  return hipGetLastError();
}

hipError_t er_gradient(const float* s, const float* dg, const float* rh, const float* w, const float* g, float iw,
                       int64_t n, double cg_tol, float* grad, float* x, float* r, float* p, float* sv, double* sc,
                       double* part, hipStream_t st) {
  const int nb = er_blocks(n);
  ER_LAUNCH(k_er_gradient, nb, s, dg, rh, w, g, iw, n, grad, x, r, p, sv, part);
  ER_FINISH(k_er_gradient_finish, nb, sc, cg_tol);
  return hipGetLastError();
}

hipError_t er_cg_ap(const float* s, const float* kt, const float* htu, const float* w, const float* sv, float iw,
                    const float* p, float* ap, int64_t n, double* sc, double* part, hipStream_t st) {
  const int nb = er_blocks(n);
  ER_LAUNCH(k_er_cg_ap, nb, s, kt, htu, w, sv, iw, p, ap, n, part);
  ER_FINISH(k_er_cg_ap_finish, nb, sc);
  return hipGetLastError();
}

hipError_t er_cg_update(float* x, float* r, float* p, const float* ap, const float* s, float* sv, int64_t n,
                        double* sc, double* part, hipStream_t st) {
  const int nb = er_blocks(n);
  ER_LAUNCH(k_er_cg_xr, nb, x, r, p, ap, n, (const double*)sc, part);
  ER_FINISH(k_er_cg_xr_finish, nb, sc);
  ER_LAUNCH(k_er_cg_p, nb, r, p, s, sv, n, (const double*)sc);
  return hipGetLastError();
}

hipError_t er_dot(const float* a, const float* b, int64_t n, int slot, double* sc, double* part, hipStream_t st) {
  const int nb = er_blocks(n);
  ER_LAUNCH(k_er_dot, nb, a, b, n, part);
  ER_FINISH(k_er_dot_finish, nb, sc, slot);
  return hipGetLastError();
}

hipError_t er_ew(int op, const float* a, const float* b, float alpha, float beta, float* out, float* out2, int64_t n,
                 hipStream_t st) {
  ER_LAUNCH(k_er_ew, er_blocks(n), op, a, b, alpha, beta, out, out2, n);
  return hipGetLastError();
}

}  // namespace dh
