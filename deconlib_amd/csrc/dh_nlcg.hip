// NLCG (Schaefer, Schuster & Herz 2001 with the exact Newton-Raphson step
// length of COSM): the elementwise passes and float64 scalar logic around the
// three convolutions of an iteration (nlcg_mlx.py:84-529).
//
// Elementwise arithmetic is float32 in the reference's operation order; every
// sum is two-stage and deterministic (float64 per-workgroup partials in a
// fixed grid, then one workgroup sums them in a fixed order).  The finishing
// workgroup also runs the scalar step (Newton update of lambda, the accept
// test, Fletcher-Reeves gamma), so a whole line search enqueues without a
// host round trip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "dh_blocksum.hpp"
#include "dh_nlcg.hpp"

namespace dh {

constexpr int kNlThreads = kSumThreads;  // 4 waves; the block sums assume exactly this
constexpr int kNlBlocks = 1024;
constexpr int kNlMaxSums = 6;

#define NL_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * kNlThreads + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kNlThreads)

// Single-use streams: non-temporal, so they do not displace lines still wanted.
__device__ __forceinline__ float ldnt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stnt(float* p, float v) { __builtin_nontemporal_store(v, p); }

static int nl_blocks(int64_t n) {
  int64_t b = (n + kNlThreads - 1) / kNlThreads;
  if (b > kNlBlocks) b = kNlBlocks;
  return (int)(b < 1 ? 1 : b);
}

// Poisson I-divergence term (rl_mlx.py:58-80), eps 1e-6.
__device__ __forceinline__ float nl_idiv_term(float o, float mdl) {
  const float os = fmaxf(o, 1e-6f), ms = fmaxf(mdl, 1e-6f);
  return o * logf(os / ms) - (o - ms);
}

// Regularizer parts of the line search, exact polynomials in lambda of the
// dots A0..A5 = <Cf,Cf>, <Cf,Csd>, <Cf,Cdd>, <Csd,Csd>, <Csd,Cdd>, <Cdd,Cdd>,
// since C f(lambda) = Cf + 2 lambda Csd + lambda^2 Cdd:
//   sum u Csd_l = A1 + l (A2 + 2 A3) + 3 l^2 A4 + l^3 A5
//   sum Csd_l^2 = A3 + 2 l A4 + l^2 A5,   sum u Cdd = A2 + 2 l A4 + l^2 A5
//   sum u^2     = A0 + 4 l A1 + l^2 (2 A2 + 4 A3) + 4 l^3 A4 + l^4 A5
struct RegDots {
  double a[6];
};
__device__ __forceinline__ RegDots reg_dots_of(const double* sc) {
  RegDots r;
#pragma unroll
  for (int j = 0; j < 6; ++j) r.a[j] = sc[DH_NL_A0 + j];
  return r;
}

// ---------------------------------------------------------------------------
// set-up, model, line search
// ---------------------------------------------------------------------------

__global__ void k_nl_init(const float* f0, float* s, float* f, int64_t n, float floor_) {
  NL_LOOP(i, n) {
    const float v = sqrtf(fmaxf(f0[i], floor_));
    s[i] = v;
    f[i] = v * v;
  }
}

// w = 1 - g / max(m, eps); sums: m - g ln max(m, eps), I-div(g, m).
__global__ void __launch_bounds__(kNlThreads) k_nl_model(const float* __restrict__ m, const float* __restrict__ g,
                                                         float* __restrict__ w, int64_t n, float eps,
                                                         double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  NL_LOOP(i, n) {
    const float mv = ldnt(m + i), gv = ldnt(g + i);
    const float mc = fmaxf(mv, eps);
    if (w) stnt(w + i, 1.0f - gv / mc);
    acc[0] += (double)mv - (double)(gv * logf(mc));
    acc[1] += (double)nl_idiv_term(gv, mv);
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kNlThreads) k_nl_model_finish(const double* __restrict__ part, int nb,
                                                                double* sc) {
  double s[2];
  sum_partials<2>(part, nb, s);
  if (threadIdx.x == 0) {
    sc[DH_NL_MDATA] = s[0];
    sc[DH_NL_LOSS] = s[1];
  }
}

__global__ void k_nl_begin_step(const float* __restrict__ s, const float* __restrict__ d, float* __restrict__ sd,
                                float* __restrict__ dd, int64_t n, double* sc, int newton_iters) {
  // lambda = 0 for the new line search; nothing else in this launch reads it
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[DH_NL_LAM] = 0.0;
    sc[DH_NL_NEWTON_OK] = newton_iters > 0 ? 1.0 : 0.0;
  }
  NL_LOOP(i, n) {
    const float sv = s[i], dv = d[i];
    stnt(sd + i, sv * dv);
    stnt(dd + i, dv * dv);
  }
}

__global__ void __launch_bounds__(kNlThreads) k_nl_reg_dots(const float* __restrict__ cf,
                                                            const float* __restrict__ csd,
                                                            const float* __restrict__ cdd, int64_t n,
                                                            double* __restrict__ part) {
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  NL_LOOP(i, n) {
    const double a = ldnt(cf + i), b = ldnt(csd + i), c = ldnt(cdd + i);
    acc[0] += a * a;
    acc[1] += a * b;
    acc[2] += a * c;
    acc[3] += b * b;
    acc[4] += b * c;
    acc[5] += c * c;
  }
  block_partials<6>(acc, part);
}

__global__ void __launch_bounds__(kNlThreads) k_nl_reg_dots_finish(const double* __restrict__ part, int nb,
                                                                   double* sc) {
  double s[6];
  sum_partials<6>(part, nb, s);
  if (threadIdx.x == 0)
    for (int j = 0; j < 6; ++j) sc[DH_NL_A0 + j] = s[j];
}

// One Newton-Raphson pass at lambda = LAM:
//   m_l = max(Kss + 2l Ksd + l^2 Kdd + b, eps), Ksd_l = Ksd + l Kdd, w_l = 1 - g / m_l
//   sums: Ksd_l w_l, Kdd w_l, g Ksd_l^2 / m_l^2
__global__ void __launch_bounds__(kNlThreads) k_nl_newton(const float* __restrict__ m,
                                                          const float* __restrict__ ksd,
                                                          const float* __restrict__ kdd,
                                                          const float* __restrict__ g, int64_t n, float bg, float eps,
                                                          const double* __restrict__ sc, double* __restrict__ part) {
  if (sc[DH_NL_NEWTON_OK] == 0.0) return;  // uniform: the refinement has stopped
  const double lam = sc[DH_NL_LAM];
  const float c1 = (float)(2.0 * lam), c2 = (float)(lam * lam), cl = (float)lam;
  double acc[3] = {0.0, 0.0, 0.0};
  NL_LOOP(i, n) {
    const float kss = m[i] - bg;
    const float a = ksd[i], b = kdd[i], gv = g[i];
    const float ml = fmaxf(((kss + c1 * a) + c2 * b) + bg, eps);
    const float al = a + cl * b;
    const float wl = 1.0f - gv / ml;
    acc[0] += (double)(al * wl);
    acc[1] += (double)(b * wl);
    acc[2] += (double)(gv * (al * al) / (ml * ml));
  }
  block_partials<3>(acc, part);
}

__global__ void __launch_bounds__(kNlThreads) k_nl_newton_finish(const double* __restrict__ part, int nb,
                                                                 double* sc, double beta, double eps) {
  if (sc[DH_NL_NEWTON_OK] == 0.0) return;
  double s[3];
  sum_partials<3>(part, nb, s);
  if (threadIdx.x != 0) return;
  const double lam = sc[DH_NL_LAM];
  double d1 = 2.0 * s[0];
  double d2 = 2.0 * s[1] + 4.0 * s[2];
  if (beta != 0.0) {
    const RegDots r = reg_dots_of(sc);
    const double l = lam, l2 = lam * lam;
    const double u_csdl = r.a[1] + l * (r.a[2] + 2.0 * r.a[3]) + 3.0 * l2 * r.a[4] + l2 * l * r.a[5];
    const double csdl2 = r.a[3] + 2.0 * l * r.a[4] + l2 * r.a[5];
    const double u_cdd = r.a[2] + 2.0 * l * r.a[4] + l2 * r.a[5];
    d1 += 4.0 * beta * u_csdl;
    d2 += 8.0 * beta * csdl2 + 4.0 * beta * u_cdd;
  }
  sc[DH_NL_DPHI1] = d1;
  sc[DH_NL_DPHI2] = d2;
  if (!isfinite(d2) || fabs(d2) < eps) {
    sc[DH_NL_NEWTON_OK] = 0.0;
    return;
  }
  const double nl = lam - d1 / d2;
  sc[DH_NL_LAM] = nl;
  if (!isfinite(nl)) sc[DH_NL_NEWTON_OK] = 0.0;
}

// Trial model at the final lambda (not floored), its w, and the objective sums.
__global__ void __launch_bounds__(kNlThreads) k_nl_trial(const float* __restrict__ m, const float* __restrict__ ksd,
                                                         const float* __restrict__ kdd, const float* __restrict__ g,
                                                         int64_t n, float bg, float eps,
                                                         const double* __restrict__ sc, float* __restrict__ m_new,
                                                         float* __restrict__ w, double* __restrict__ part) {
  const double lam = sc[DH_NL_LAM];
  if (!isfinite(lam)) return;  // uniform: the step is rejected without a model
  const float c1 = (float)(2.0 * lam), c2 = (float)(lam * lam);
  double acc[2] = {0.0, 0.0};
  NL_LOOP(i, n) {
    const float kss = ldnt(m + i) - bg;
    const float gv = ldnt(g + i);
    const float mn = ((kss + c1 * ldnt(ksd + i)) + c2 * ldnt(kdd + i)) + bg;
    const float mc = fmaxf(mn, eps);
    stnt(m_new + i, mn);
    stnt(w + i, 1.0f - gv / mc);
    acc[0] += (double)mn - (double)(gv * logf(mc));
    acc[1] += (double)nl_idiv_term(gv, mn);
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kNlThreads) k_nl_trial_finish(const double* __restrict__ part, int nb,
                                                                double* sc, double beta) {
  const double lam = sc[DH_NL_LAM];
  const bool ok = isfinite(lam);
  double s[2] = {0.0, 0.0};
  if (ok) sum_partials<2>(part, nb, s);  // uniform branch
  if (threadIdx.x != 0) return;
  double phi0 = sc[DH_NL_MDATA];
  double phi = s[0];
  if (beta != 0.0) {
    const RegDots r = reg_dots_of(sc);
    const double l = lam, l2 = lam * lam;
    phi0 += beta * r.a[0];
    phi += beta * (r.a[0] + 4.0 * l * r.a[1] + l2 * (2.0 * r.a[2] + 4.0 * r.a[3]) + 4.0 * l2 * l * r.a[4] +
                   l2 * l2 * r.a[5]);
  }
  sc[DH_NL_PHI0] = phi0;
  if (!ok) {
    sc[DH_NL_PHI_TRIAL] = NAN;
    sc[DH_NL_ACCEPT] = 0.0;
    return;
  }
  sc[DH_NL_PHI_TRIAL] = phi;
  sc[DH_NL_TRIAL_DATA] = s[0];
  sc[DH_NL_TRIAL_LOSS] = s[1];
  // rejected only when phi >= phi0 (a NaN objective compares false, as in Python)
  sc[DH_NL_ACCEPT] = (phi >= phi0) ? 0.0 : 1.0;
}

// ---------------------------------------------------------------------------
// accepted step, gradient, direction
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kNlThreads) k_nl_update(float* __restrict__ s, const float* __restrict__ d,
                                                          float* __restrict__ f, int64_t n, float eps,
                                                          const double* __restrict__ sc, double* __restrict__ part) {
  const float cl = (float)sc[DH_NL_LAM];
  double acc[2] = {0.0, 0.0};
  NL_LOOP(i, n) {
    const float sn = s[i] + cl * ldnt(d + i);
    const float fn = sn * sn;
    const float fp = f[i];
    s[i] = sn;
    f[i] = fn;
    acc[0] += (double)(fp * (logf(fmaxf(fp, eps)) - logf(fmaxf(fn, eps))) - fp + fn);
    acc[1] += (double)fn;
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kNlThreads) k_nl_update_finish(const double* __restrict__ part, int nb,
                                                                 double* sc) {
  double s[2];
  sum_partials<2>(part, nb, s);
  if (threadIdx.x == 0) {
    sc[DH_NL_IDIV] = s[0];
    sc[DH_NL_SUMF] = s[1];
    sc[DH_NL_MDATA] = sc[DH_NL_TRIAL_DATA];
    sc[DH_NL_LOSS] = sc[DH_NL_TRIAL_LOSS];
    sc[DH_NL_LAM_ACCEPTED] = sc[DH_NL_LAM];
  }
}

// r = -(2 s KTw [+ 4 beta s CtCf]); sum r^2.
__global__ void __launch_bounds__(kNlThreads) k_nl_gradient(const float* __restrict__ s,
                                                            const float* __restrict__ ktw,
                                                            const float* __restrict__ ctcf, float c4b,
                                                            float* __restrict__ r, int64_t n,
                                                            double* __restrict__ part) {
  double acc[1] = {0.0};
  NL_LOOP(i, n) {
    const float sv = s[i];
    float gr = (2.0f * sv) * ldnt(ktw + i);
    if (ctcf) gr = gr + (c4b * sv) * ldnt(ctcf + i);
    const float rv = -gr;
    r[i] = rv;
    acc[0] += (double)(rv * rv);
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(kNlThreads) k_nl_gradient_finish(const double* __restrict__ part, int nb,
                                                                   double* sc, int mode, double eps) {
  double s[1];
  sum_partials<1>(part, nb, s);
  if (threadIdx.x != 0) return;
  const double rr_new = s[0];
  if (mode != 0) {
    const double rr = sc[DH_NL_RR];
    double gamma = rr > eps ? rr_new / rr : 0.0;
    if (mode == 2) gamma = 0.0;
    sc[DH_NL_GAMMA] = gamma;
  }
  sc[DH_NL_RR] = rr_new;
}

__global__ void k_nl_direction(const float* __restrict__ r, float* __restrict__ d, int64_t n,
                               const double* __restrict__ sc) {
  const float gm = (float)sc[DH_NL_GAMMA];
  NL_LOOP(i, n) d[i] = r[i] + gm * d[i];
}

// ---------------------------------------------------------------------------
// standalone sums (Hessian quadratic form, I-divergence between iterates)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kNlThreads) k_nl_quad(const float* __restrict__ g, const float* __restrict__ m,
                                                        const float* __restrict__ ksd, const float* __restrict__ v,
                                                        const float* __restrict__ d, int64_t n, float eps,
                                                        double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  NL_LOOP(i, n) {
    if (g) {
      const float mv = m[i], a = ksd[i];
      acc[0] += (double)((g[i] / fmaxf(mv * mv, eps)) * (a * a));
    }
    const float dv = d[i];
    acc[1] += (double)(v[i] * (dv * dv));
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kNlThreads) k_nl_idiv(const float* __restrict__ fp, const float* __restrict__ fc,
                                                        int64_t n, float eps, double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  NL_LOOP(i, n) {
    const float a = fp[i], b = fc[i];
    acc[0] += (double)(a * (logf(fmaxf(a, eps)) - logf(fmaxf(b, eps))) - a + b);
    acc[1] += (double)b;
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kNlThreads) k_nl_sum2_finish(const double* __restrict__ part, int nb,
                                                               double* out2) {
  double s[2];
  sum_partials<2>(part, nb, s);
  if (threadIdx.x == 0) {
    out2[0] = s[0];
    out2[1] = s[1];
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

int64_t nl_partials_needed() { return (int64_t)kNlBlocks * kNlMaxSums; }

#define NL_LAUNCH(kern, nb, ...) hipLaunchKernelGGL(kern, dim3((unsigned)(nb)), dim3(kNlThreads), 0, st, __VA_ARGS__)
#define NL_FINISH(kern, nb, ...) hipLaunchKernelGGL(kern, dim3(1), dim3(kNlThreads), 0, st, part, nb, __VA_ARGS__)

hipError_t nl_init(const float* f0, float* s, float* f, int64_t n, float floor_, hipStream_t st) {
  NL_LAUNCH(k_nl_init, nl_blocks(n), f0, s, f, n, floor_);
  return hipGetLastError();
}

hipError_t nl_model(const float* m, const float* g, float* w, int64_t n, float eps, double* sc, double* part,
                    hipStream_t st) {
  const int nb = nl_blocks(n);
  NL_LAUNCH(k_nl_model, nb, m, g, w, n, eps, part);
  NL_FINISH(k_nl_model_finish, nb, sc);
  return hipGetLastError();
}

hipError_t nl_begin_step(const float* s, const float* d, float* sd, float* dd, int64_t n, double* sc,
                         int newton_iters, hipStream_t st) {
  NL_LAUNCH(k_nl_begin_step, nl_blocks(n), s, d, sd, dd, n, sc, newton_iters);
  return hipGetLastError();
}

hipError_t nl_reg_dots(const float* cf, const float* csd, const float* cdd, int64_t n, double* sc, double* part,
                       hipStream_t st) {
  const int nb = nl_blocks(n);
  NL_LAUNCH(k_nl_reg_dots, nb, cf, csd, cdd, n, part);
  NL_FINISH(k_nl_reg_dots_finish, nb, sc);
  return hipGetLastError();
}

hipError_t nl_newton(const float* m, const float* ksd, const float* kdd, const float* g, int64_t n, float bg,
                     double beta, double eps, double* sc, double* part, hipStream_t st) {
  const int nb = nl_blocks(n);
  NL_LAUNCH(k_nl_newton, nb, m, ksd, kdd, g, n, bg, (float)eps, (const double*)sc, part);
  NL_FINISH(k_nl_newton_finish, nb, sc, beta, eps);
  return hipGetLastError();
}

hipError_t nl_trial(const float* m, const float* ksd, const float* kdd, const float* g, int64_t n, float bg,
                    double beta, double eps, double* sc, float* m_new, float* w, double* part, hipStream_t st) {
  const int nb = nl_blocks(n);
  NL_LAUNCH(k_nl_trial, nb, m, ksd, kdd, g, n, bg, (float)eps, (const double*)sc, m_new, w, part);
  NL_FINISH(k_nl_trial_finish, nb, sc, beta);
  return hipGetLastError();
}

hipError_t nl_update(float* s, const float* d, float* f, int64_t n, double eps, double* sc, double* part,
                     hipStream_t st) {
  const int nb = nl_blocks(n);
  NL_LAUNCH(k_nl_update, nb, s, d, f, n, (float)eps, (const double*)sc, part);
  NL_FINISH(k_nl_update_finish, nb, sc);
  return hipGetLastError();
}

hipError_t nl_gradient(const float* s, const float* ktw, const float* ctcf, float c4b, float* r, int64_t n, int mode,
                       double eps, double* sc, double* part, hipStream_t st) {
  const int nb = nl_blocks(n);
  NL_LAUNCH(k_nl_gradient, nb, s, ktw, ctcf, c4b, r, n, part);
  NL_FINISH(k_nl_gradient_finish, nb, sc, mode, eps);
  return hipGetLastError();
}

hipError_t nl_direction(const float* r, float* d, int64_t n, const double* sc, hipStream_t st) {
  NL_LAUNCH(k_nl_direction, nl_blocks(n), r, d, n, sc);
  return hipGetLastError();
}

hipError_t nl_quad(const float* g, const float* m, const float* ksd, const float* v, const float* d, int64_t n,
                   double eps, double* out2, double* part, hipStream_t st) {
  const int nb = nl_blocks(n);
  NL_LAUNCH(k_nl_quad, nb, g, m, ksd, v, d, n, (float)eps, part);
  NL_FINISH(k_nl_sum2_finish, nb, out2);
  return hipGetLastError();
}

hipError_t nl_idiv(const float* fp, const float* fc, int64_t n, double eps, double* out2, double* part,
                   hipStream_t st) {
  const int nb = nl_blocks(n);
  NL_LAUNCH(k_nl_idiv, nb, fp, fc, n, (float)eps, part);
  NL_FINISH(k_nl_sum2_finish, nb, out2);
  return hipGetLastError();
}

}  // namespace dh
