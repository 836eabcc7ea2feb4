// JetNewton: non-dimensional log-penalty deconvolution by two-metric
// active-set projected Newton with the exact Hessian (jetnewton_mlx.py:111-785).
// The kernels around the convolutions: the penalty's per-voxel r = H x, u =
// sum_c r_c^2, w = beta / (eta + u), c = -w^2 / beta from the Hessian stencil
// on the fly; the exact Hessian product's operand w q + 2 c <r, q> r with q =
// H v; the KKT residual and the Bertsekas active set with persistent
// freezing; the masked reduced CG with its absolute forcing tolerance; the
// two-metric direction; and the projected-arc trial point.  CG scalars and
// tests live in a float64 block on the device.
//
// Elementwise arithmetic is float32 in the reference's operation order; sums
// are two-stage and deterministic (dh_blocksum.hpp).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "dh_blocksum.hpp"
#include "dh_jetnewton.hpp"
#include "dh_stencil.hpp"

namespace dh {

constexpr int kJnThreads = kSumThreads;
constexpr int kJnBlocks = 1024;
constexpr int kJnMaxSums = 2;

#define JN_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * kJnThreads + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kJnThreads)

static int jn_blocks(int64_t n) {
  int64_t b = (n + kJnThreads - 1) / kJnThreads;
  if (b > kJnBlocks) b = kJnBlocks;
  return (int)(b < 1 ? 1 : b);
}

// Per-workgroup (sum of v0, max of v1) into part[2 * block + {0, 1}].
__device__ __forceinline__ void block_sum_max(double v0, double v1, double* __restrict__ part) {
  __shared__ double sh[2][kJnThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_down(v0, off, 64);
    v1 = fmax(v1, __shfl_down(v1, off, 64));
  }
  if (lane == 0) {
    sh[0][w] = v0;
    sh[1][w] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0, m = 0.0;
    for (int q = 0; q < kJnThreads / 64; ++q) {
      s += sh[0][q];
      m = fmax(m, sh[1][q]);
    }
    part[2 * (int64_t)blockIdx.x] = s;
    part[2 * (int64_t)blockIdx.x + 1] = m;
  }
}

// ---------------------------------------------------------------------------
// penalty weights and the Hessian product operand
// ---------------------------------------------------------------------------

struct JnPen {
  float beta, eta;
};

// r = H x (weights w6), u = sum r^2; sums ln(eta + u), u; optionally r, w, c, w r, u.
__global__ void __launch_bounds__(kJnThreads) k_jn_penalty(const float* __restrict__ x, Dims d, W6 w6, JnPen p,
                                                           float* __restrict__ r_out, float* __restrict__ w_out,
                                                           float* __restrict__ c_out, float* __restrict__ wr_out,
                                                           float* __restrict__ u_out, double* __restrict__ part) {
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  auto xv = [&](int z, int y, int xx) { return x[((int64_t)z * d.ny + y) * d.nx + xx]; };
  double acc[2] = {0.0, 0.0};
  JN_LOOP(i, n) {
    const int xx = (int)(i % d.nx);
    const int y = (int)((i / d.nx) % d.ny);
    const int z = (int)(i / ((int64_t)d.nx * d.ny));
    float h[6];
    hess6(xv, d, z, y, xx, w6.w, h);
    float u = h[0] * h[0];
#pragma unroll
    for (int c = 1; c < 6; ++c) u += h[c] * h[c];
    acc[0] += (double)logf(p.eta + u);
    acc[1] += (double)u;
    if (u_out) u_out[i] = u;
    if (r_out) {
      const float wv = p.beta / (p.eta + u);
      w_out[i] = wv;
      c_out[i] = -(wv * wv) / p.beta;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        r_out[c * n + i] = h[c];
        wr_out[c * n + i] = wv * h[c];
      }
    }
  }
  block_partials<2>(acc, part);
}

// The same from any operator's stacked responses r (nc, n).
__global__ void __launch_bounds__(kJnThreads) k_jn_penalty_stacked(const float* __restrict__ r, int nc, int64_t n,
                                                                   JnPen p, float* __restrict__ w_out,
                                                                   float* __restrict__ c_out,
                                                                   float* __restrict__ wr_out,
                                                                   float* __restrict__ u_out,
                                                                   double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  JN_LOOP(i, n) {
    float u = 0.0f;
    for (int c = 0; c < nc; ++c) {
      const float h = r[c * n + i];
      u = (c == 0) ? h * h : u + h * h;
    }
    acc[0] += (double)logf(p.eta + u);
    acc[1] += (double)u;
    if (u_out) u_out[i] = u;
    if (w_out) {
      const float wv = p.beta / (p.eta + u);
      w_out[i] = wv;
      c_out[i] = -(wv * wv) / p.beta;
      for (int c = 0; c < nc; ++c) wr_out[c * n + i] = wv * r[c * n + i];
    }
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kJnThreads) k_jn_penalty_finish(const double* __restrict__ part, int nb,
                                                                  double* sc) {
  double s[2];
  sum_partials<2>(part, nb, s);
  if (threadIdx.x == 0) {
    sc[DH_JN_LOGU] = s[0];
    sc[DH_JN_USUM] = s[1];
  }
}

// q = H v; out = w q + 2 (c <r, q>) r (the operand of H^T in the exact product), through
// the sliding LDS plane window of v.
__global__ void __launch_bounds__(kTileThreads) k_jn_hvp_operand_tiled(const float* __restrict__ v,
                                                                       const float* __restrict__ r,
                                                                       const float* __restrict__ w,
                                                                       const float* __restrict__ c, Dims d, W6 w6,
                                                                       float* __restrict__ out) {
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
      float q[6], rr[6];
      hess6_halo(f, z, t.y, t.x, w6.w, q);
      const int64_t i = ((int64_t)z * d.ny + t.y) * d.nx + t.x;
#pragma unroll
      for (int k = 0; k < 6; ++k) rr[k] = __builtin_nontemporal_load(r + k * n + i);
      float g = rr[0] * q[0];
#pragma unroll
      for (int k = 1; k < 6; ++k) g += rr[k] * q[k];
      const float m2 = 2.0f * (__builtin_nontemporal_load(c + i) * g), wv = __builtin_nontemporal_load(w + i);
#pragma unroll
      for (int k = 0; k < 6; ++k) __builtin_nontemporal_store(wv * q[k] + m2 * rr[k], out + k * n + i);
    }
    __syncthreads();
  }
}

// The same from the operator's stacked q (nc, n); out may alias q.
__global__ void k_jn_hvp_operand_stacked(const float* q, const float* __restrict__ r, const float* __restrict__ w,
                                         const float* __restrict__ c, int nc, int64_t n, float* out) {
  JN_LOOP(i, n) {
    float g = 0.0f;
    for (int k = 0; k < nc; ++k) g = (k == 0) ? r[i] * q[i] : g + r[k * n + i] * q[k * n + i];
    const float m2 = 2.0f * (c[i] * g), wv = w[i];
    for (int k = 0; k < nc; ++k) out[k * n + i] = wv * q[k * n + i] + m2 * r[k * n + i];
  }
}

// ---------------------------------------------------------------------------
// active set, reduced CG, direction, projected trial point
// ---------------------------------------------------------------------------

// diff = x - max(x - grad, 0): sum diff^2 and max |diff|.
__global__ void __launch_bounds__(kJnThreads) k_jn_kkt(const float* __restrict__ x, const float* __restrict__ g,
                                                       int64_t n, double* __restrict__ part) {
  double s = 0.0, m = 0.0;
  JN_LOOP(i, n) {
    const float xv = x[i];
    const float diff = xv - fmaxf(xv - g[i], 0.0f);
    s += (double)(diff * diff);
    m = fmax(m, (double)fabsf(diff));
  }
  block_sum_max(s, m, part);
}

__global__ void __launch_bounds__(kJnThreads) k_jn_kkt_finish(const double* __restrict__ part, int nb, double* sc,
                                                              double eps_bar) {
  __shared__ double sh[2][kJnThreads / 64];
  double s = 0.0, m = 0.0;
  for (int b = threadIdx.x; b < nb; b += kJnThreads) {
    s += part[2 * (int64_t)b];
    m = fmax(m, part[2 * (int64_t)b + 1]);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off, 64);
    m = fmax(m, __shfl_down(m, off, 64));
  }
  if (lane == 0) {
    sh[0][w] = s;
    sh[1][w] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = 0.0, mm = 0.0;
    for (int q = 0; q < kJnThreads / 64; ++q) {
      ss += sh[0][q];
      mm = fmax(mm, sh[1][q]);
    }
    sc[DH_JN_KKT2] = ss;
    sc[DH_JN_EPS] = fmin(eps_bar, sqrt(ss));
    sc[DH_JN_PG] = mm;
  }
}

// frozen |= (x < tau) & (g > delta); active = ((x <= eps) & (g > 0)) | frozen;
// the reduced CG starts at xcg = 0, r = p = where(free, -g, 0).
__global__ void __launch_bounds__(kJnThreads) k_jn_active(const float* x, const float* __restrict__ g,
                                                          int64_t n, float tau, float delta,
                                                          uint8_t* __restrict__ frozen, uint8_t* active,
                                                          float* __restrict__ r, float* __restrict__ p,
                                                          float* __restrict__ xcg, int use_frozen,
                                                          const double* __restrict__ sc, double* __restrict__ part) {
  const float eps = (float)sc[DH_JN_EPS];
  double acc[2] = {0.0, 0.0};
  JN_LOOP(i, n) {
    const float gv = g[i];
    if (!x) {  // a given active set: only the CG start
      const float rhs = active[i] ? 0.0f : -gv;
      r[i] = rhs;
      p[i] = rhs;
      xcg[i] = 0.0f;
      acc[0] += active[i] ? 1.0 : 0.0;
      acc[1] += (double)(rhs * rhs);
      continue;
    }
    const float xv = x[i];
    bool act = (xv <= eps) && (gv > 0.0f);
    if (use_frozen) {
      const bool fz = frozen[i] != 0 || ((xv < tau) && (gv > delta));
      frozen[i] = fz ? 1 : 0;
      act = act || fz;
    }
    active[i] = act ? 1 : 0;
    const float rhs = act ? 0.0f : -gv;
    if (r) {
      r[i] = rhs;
      p[i] = rhs;
      xcg[i] = 0.0f;
    }
    acc[0] += act ? 1.0 : 0.0;
    acc[1] += (double)(rhs * rhs);
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(kJnThreads) k_jn_active_finish(const double* __restrict__ part, int nb, double* sc,
                                                                 double eps) {
  double s[2];
  sum_partials<2>(part, nb, s);
  if (threadIdx.x == 0) {
    sc[DH_JN_ACTIVE] = s[0];
    const double r0 = sqrt(s[1]);
    const double tol_k = fmin(0.5, sqrt(r0)) * r0;
    sc[DH_JN_RZ] = s[1];
    sc[DH_JN_R0] = r0;
    sc[DH_JN_TOL_K] = tol_k;
    sc[DH_JN_STOP] = r0 <= fmax(tol_k, eps) ? 1.0 : 0.0;
    sc[DH_JN_NCG] = 0.0;
    sc[DH_JN_HITNEG] = 0.0;
    sc[DH_JN_USE_RHS] = 0.0;
  }
}

// ap = where(free, scale kt [+ htu], 0); sum p ap.
__global__ void __launch_bounds__(kJnThreads) k_jn_cg_ap(const float* __restrict__ kt, float scale,
                                                         const float* __restrict__ htu,
                                                         const uint8_t* __restrict__ active,
                                                         const float* __restrict__ p, float* __restrict__ ap,
                                                         int64_t n, double* __restrict__ part) {
  double acc[1] = {0.0};
  JN_LOOP(i, n) {
    float v = scale * kt[i];
    if (htu) v = v + htu[i];
    const float a = active[i] ? 0.0f : v;
    ap[i] = a;
    acc[0] += (double)(p[i] * a);
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(kJnThreads) k_jn_cg_ap_finish(const double* __restrict__ part, int nb, double* sc,
                                                                double eps) {
  double s[1];
  sum_partials<1>(part, nb, s);
  if (threadIdx.x == 0) {
    const double steps = sc[DH_JN_NCG] + 1.0;
    sc[DH_JN_NCG] = steps;
    sc[DH_JN_DENOM] = s[0];
    if (!isfinite(s[0]) || s[0] <= eps) {  // non-positive curvature
      sc[DH_JN_STOP] = 1.0;
      sc[DH_JN_HITNEG] = 1.0;
      if (steps == 1.0) sc[DH_JN_USE_RHS] = 1.0;
    }
  }
}

__global__ void __launch_bounds__(kJnThreads) k_jn_cg_xr(float* __restrict__ x, float* __restrict__ r,
                                                         const float* __restrict__ p, const float* __restrict__ ap,
                                                         int64_t n, const double* __restrict__ sc,
                                                         double* __restrict__ part) {
  if (sc[DH_JN_STOP] != 0.0) return;  // uniform
  const float a = (float)(sc[DH_JN_RZ] / sc[DH_JN_DENOM]);
  double acc[1] = {0.0};
  JN_LOOP(i, n) {
    x[i] = x[i] + a * p[i];
    const float rn = r[i] - a * ap[i];
    r[i] = rn;
    acc[0] += (double)(rn * rn);
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(kJnThreads) k_jn_cg_xr_finish(const double* __restrict__ part, int nb, double* sc) {
  if (sc[DH_JN_STOP] != 0.0) return;
  double s[1];
  sum_partials<1>(part, nb, s);
  if (threadIdx.x != 0) return;
  const double rn = sqrt(s[0]);
  if (!isfinite(rn) || rn <= sc[DH_JN_TOL_K]) {
    sc[DH_JN_STOP] = 1.0;
    return;
  }
  sc[DH_JN_BETA] = s[0] / sc[DH_JN_RZ];
  sc[DH_JN_RZ] = s[0];
}

__global__ void k_jn_cg_p(const float* __restrict__ r, float* __restrict__ p, int64_t n,
                          const double* __restrict__ sc) {
  if (sc[DH_JN_STOP] != 0.0) return;
  const float b = (float)sc[DH_JN_BETA];
  JN_LOOP(i, n) p[i] = r[i] + b * p[i];
}

// d = d_free + where(active, -g, 0), d_free = the CG solution (or where(free, -g, 0)
// after negative curvature on the first step); sum g d.
__global__ void __launch_bounds__(kJnThreads) k_jn_direction(const float* __restrict__ xcg,
                                                             const float* __restrict__ g,
                                                             const uint8_t* __restrict__ active,
                                                             float* __restrict__ d, int64_t n,
                                                             const double* __restrict__ sc,
                                                             double* __restrict__ part) {
  const bool use_rhs = sc[DH_JN_USE_RHS] != 0.0;
  double acc[1] = {0.0};
  JN_LOOP(i, n) {
    const float gv = g[i];
    const bool act = active[i] != 0;
    const float dfree = use_rhs ? (act ? 0.0f : -gv) : xcg[i];
    const float dv = dfree + (act ? -gv : 0.0f);
    d[i] = dv;
    acc[0] += (double)(gv * dv);
  }
  block_partials<1>(acc, part);
}

// x_t = max(x + t d, 0); sum g (x_t - x).
__global__ void __launch_bounds__(kJnThreads) k_jn_trial(const float* __restrict__ x, const float* __restrict__ d,
                                                         float t, const float* __restrict__ g,
                                                         float* __restrict__ xt, int64_t n,
                                                         double* __restrict__ part) {
  double acc[1] = {0.0};
  JN_LOOP(i, n) {
    const float xv = x[i];
    const float v = fmaxf(xv + t * d[i], 0.0f);
    xt[i] = v;
    acc[0] += (double)(g[i] * (v - xv));
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(kJnThreads) k_jn_sum1_finish(const double* __restrict__ part, int nb, double* sc,
                                                               int slot) {
  double s[1];
  sum_partials<1>(part, nb, s);
  if (threadIdx.x == 0) sc[slot] = s[0];
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

int64_t jn_partials_needed() { return (int64_t)kJnBlocks * kJnMaxSums; }

#define JN_LAUNCH(kern, nb, ...) hipLaunchKernelGGL(kern, dim3((unsigned)(nb)), dim3(kJnThreads), 0, st, __VA_ARGS__)
#define JN_FINISH(kern, nb, ...) hipLaunchKernelGGL(kern, dim3(1), dim3(kJnThreads), 0, st, part, nb, __VA_ARGS__)

static W6 jn_w6(const float* w) {
  W6 r;
  for (int q = 0; q < 6; ++q) r.w[q] = w[q];
  return r;
}

hipError_t jn_penalty(const float* x, int nz, int ny, int nx, const float* w6, float beta, float eta, float* r_out,
                      float* w_out, float* c_out, float* wr_out, float* u_out, double* sc, double* part,
                      hipStream_t st) {
  const int64_t n = (int64_t)nz * ny * nx;
  const int nb = jn_blocks(n);
  JN_LAUNCH(k_jn_penalty, nb, x, Dims{nz, ny, nx}, jn_w6(w6), JnPen{beta, eta}, r_out, w_out, c_out, wr_out, u_out,
            part);
  JN_FINISH(k_jn_penalty_finish, nb, sc);
  return hipGetLastError();
}

hipError_t jn_penalty_stacked(const float* r, int nc, int64_t n, float beta, float eta, float* w_out, float* c_out,
                              float* wr_out, float* u_out, double* sc, double* part, hipStream_t st) {
  const int nb = jn_blocks(n);
  JN_LAUNCH(k_jn_penalty_stacked, nb, r, nc, n, JnPen{beta, eta}, w_out, c_out, wr_out, u_out, part);
  JN_FINISH(k_jn_penalty_finish, nb, sc);
  return hipGetLastError();
}

hipError_t jn_hvp_operand(const float* v, const float* r, const float* w, const float* c, int nz, int ny, int nx,
                          const float* w6, float* out, hipStream_t st) {
  const int64_t n = (int64_t)nz * ny * nx;
  hipLaunchKernelGGL(k_jn_hvp_operand_tiled, tile_grid(nz, ny, nx), dim3(kTileThreads), 0, st, v, r, w, c,
                     Dims{nz, ny, nx}, jn_w6(w6), out);
  (void)n;
  return hipGetLastError();
}

hipError_t jn_hvp_operand_stacked(const float* q, const float* r, const float* w, const float* c, int nc, int64_t n,
                                  float* out, hipStream_t st) {
  JN_LAUNCH(k_jn_hvp_operand_stacked, jn_blocks(n), q, r, w, c, nc, n, out);
  return hipGetLastError();
}

hipError_t jn_kkt(const float* x, const float* g, int64_t n, double eps_bar, double* sc, double* part,
                  hipStream_t st) {
  const int nb = jn_blocks(n);
  JN_LAUNCH(k_jn_kkt, nb, x, g, n, part);
  JN_FINISH(k_jn_kkt_finish, nb, sc, eps_bar);
  return hipGetLastError();
}

hipError_t jn_active(const float* x, const float* g, int64_t n, float tau, float delta, uint8_t* frozen,
                     uint8_t* active, float* r, float* p, float* xcg, double eps, double* sc, double* part,
                     hipStream_t st) {
  const int nb = jn_blocks(n);
  JN_LAUNCH(k_jn_active, nb, x, g, n, tau, delta, frozen, active, r, p, xcg, frozen ? 1 : 0, (const double*)sc, part);
  JN_FINISH(k_jn_active_finish, nb, sc, eps);
  return hipGetLastError();
}

hipError_t jn_cg_ap(const float* kt, float scale, const float* htu, const uint8_t* active, const float* p, float* ap,
                    int64_t n, double eps, double* sc, double* part, hipStream_t st) {
  const int nb = jn_blocks(n);
  JN_LAUNCH(k_jn_cg_ap, nb, kt, scale, htu, active, p, ap, n, part);
  JN_FINISH(k_jn_cg_ap_finish, nb, sc, eps);
  return hipGetLastError();
}

hipError_t jn_cg_update(float* x, float* r, float* p, const float* ap, int64_t n, double* sc, double* part,
                        hipStream_t st) {
  const int nb = jn_blocks(n);
  JN_LAUNCH(k_jn_cg_xr, nb, x, r, p, ap, n, (const double*)sc, part);
  JN_FINISH(k_jn_cg_xr_finish, nb, sc);
  JN_LAUNCH(k_jn_cg_p, nb, r, p, n, (const double*)sc);
  return hipGetLastError();
}

hipError_t jn_direction(const float* xcg, const float* g, const uint8_t* active, float* d, int64_t n, double* sc,
                        double* part, hipStream_t st) {
  const int nb = jn_blocks(n);
  JN_LAUNCH(k_jn_direction, nb, xcg, g, active, d, n, (const double*)sc, part);
  JN_FINISH(k_jn_sum1_finish, nb, sc, (int)DH_JN_ND);
  return hipGetLastError();
}

hipError_t jn_trial(const float* x, const float* d, float t, const float* g, float* xt, int64_t n, double* sc,
                    double* part, hipStream_t st) {
  const int nb = jn_blocks(n);
  JN_LAUNCH(k_jn_trial, nb, x, d, t, g, xt, n, part);
  JN_FINISH(k_jn_sum1_finish, nb, sc, (int)DH_JN_DIRECTIONAL);
  return hipGetLastError();
}

}  // namespace dh
