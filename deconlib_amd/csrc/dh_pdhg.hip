// PDHG (Malitsky–Pock) with Hessian regularization: stencil, prox and FP64
// step-size kernels (pdhg_mlx.py:507-920, linops_mlx.py:182-262,
// linops_core_mlx.py:42-80).
//
// Scalars (tau, sigma, theta, the convergence counter and the histories) live
// in a float64 block on the device, so an iteration enqueues without any host
// round trip.  Reductions are two-stage and deterministic.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "dh_pdhg.hpp"
#include "dh_prox.hpp"
#include "dh_stencil.hpp"

namespace dh {

// Operands each thread reads once and outputs it writes once per launch
// (stencil inputs excepted: neighbouring tiles re-read their halos through L2)
// take the non-temporal hint, so the streams do not displace the halo lines.
__device__ __forceinline__ float ldnt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stnt(float* p, float v) { __builtin_nontemporal_store(v, p); }

#define PD_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

constexpr int kPdThreads = 256;
constexpr int kPdBlocks = 1024;

__device__ static double pd_block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r += sh[i];
  }
  __syncthreads();
  return r;
}

// Sums NV partial vectors laid out [block][NV] into out[0..NV).
__global__ void k_pd_finish(const double* __restrict__ part, int nblocks, int nv, double* out) {
  __shared__ double sh[16];
  for (int j = 0; j < nv; ++j) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += part[(int64_t)b * nv + j];
    double r = pd_block_sum(acc, sh);
    if (threadIdx.x == 0) out[j] = r;
  }
}

static hipError_t finish(const double* part, int nblocks, int nv, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pd_finish, dim3(1), dim3(1024), 0, st, part, nblocks, nv, out);
  return hipGetLastError();
}

// Sums two partial vectors [block][2] into *out0 and, if out1, *out1.
__global__ void k_pd_finish2(const double* __restrict__ part, int nblocks, double* out0, double* out1) {
  __shared__ double sh[16];
  for (int j = 0; j < 2; ++j) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += part[(int64_t)b * 2 + j];
    double r = pd_block_sum(acc, sh);
    double* o = j == 0 ? out0 : out1;
    if (threadIdx.x == 0 && o) o[0] = r;
  }
}

static hipError_t finish2(const double* part, int nblocks, double* out0, double* out1, hipStream_t st) {
  hipLaunchKernelGGL(k_pd_finish2, dim3(1), dim3(1024), 0, st, part, nblocks, out0, out1);
  return hipGetLastError();
}

hipError_t pd_move(double* sc, int from0, int to0, int from1, double* out1, hipStream_t st);

static int pd_blocks(int64_t n) {
  int64_t b = (n + kPdThreads - 1) / kPdThreads;
  if (b > kPdBlocks) b = kPdBlocks;
  return (int)(b < 1 ? 1 : b);
}

// ---------------------------------------------------------------------------
// y1: data-term dual prox (pdhg_mlx.py:115-189), fp32 as the reference
// ---------------------------------------------------------------------------

__global__ void k_pd_y1(float* __restrict__ y1, const float* __restrict__ ax, const float* __restrict__ axp,
                        const float* __restrict__ data, int64_t n, const double* __restrict__ sc, float bg,
                        int loss_type) {
  const PdY1Consts k = pd_y1_consts(sc, bg);
  PD_LOOP(i, n) stnt(y1 + i, pd_y1_prox(ldnt(y1 + i), ldnt(ax + i), ldnt(axp + i), ldnt(data + i), k, loss_type));
}

// The dual and primal updates write their multiply-adds as fmaf (xbar, y2 + sigma H xbar,
// the L1_2 norm's squares, x - tau grad), so that the fused sweep below and these
// kernels round identically wherever the compiler inlines them.
// y2 <- prox(y2 + sigma * H(xbar)), xbar = x + theta (x - xp)  (pdhg_mlx.py:729-734)
__global__ void __launch_bounds__(kTileThreads) k_pd_dual(float* __restrict__ y2, const float* __restrict__ x,
                                                          const float* __restrict__ xp, Dims d,
                                                          const double* __restrict__ sc, W6 w, float alpha, int l12) {
  __shared__ float pl[3 * kPlane];
  const Tile t = tile_of(d);
  const PlanePos pp(d, t);
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  const float sf = (float)sc[PD_SIGMA], th = (float)sc[PD_THETA];
  auto xbar = [&](float v, float vp) { return fmaf(th, v - vp, v); };
  float a0, a1, b0, b1;
  for (int zz = t.z0 - 1; zz <= t.z0; ++zz) {
    const int64_t bz = pp.base(d, zz);
    pp.load(x, bz, a0, a1);
    pp.load(xp, bz, b0, b1);
    pp.store(pl + slot_of(zz, t.z0) * kPlane, xbar(a0, b0), xbar(a1, b1));
  }
  {
    const int64_t bz = pp.base(d, t.z0 + 1);
    pp.load(x, bz, a0, a1);
    pp.load(xp, bz, b0, b1);
  }
  // this voxel's y2 components, loaded one plane ahead
  const int64_t pstride = (int64_t)d.ny * d.nx;
  float yc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (t.in) {
    const int64_t i0 = ((int64_t)t.z0 * d.ny + t.y) * d.nx + t.x;
#pragma unroll
    for (int q = 0; q < 6; ++q) yc[q] = ldnt(y2 + q * n + i0);
  }
  for (int z = t.z0; z < t.z1; ++z) {
    pp.store(pl + slot_of(z + 1, t.z0) * kPlane, xbar(a0, b0), xbar(a1, b1));
    __syncthreads();
    if (z + 1 < t.z1) {
      const int64_t bz = pp.base(d, z + 2);
      pp.load(x, bz, a0, a1);
      pp.load(xp, bz, b0, b1);
    }
    const int64_t i = ((int64_t)z * d.ny + t.y) * d.nx + t.x;
    float yn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (t.in && z + 1 < t.z1) {
#pragma unroll
      for (int q = 0; q < 6; ++q) yn[q] = ldnt(y2 + q * n + i + pstride);
    }
    if (t.in) {
      auto f = [&](int zz, int yy, int xx) { return lds_at(pl, slot_of(zz, t.z0), t, yy, xx); };
      float h[6];
      hess6_halo(f, z, t.y, t.x, w.w, h);
      float av[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) av[q] = fmaf(sf, h[q], yc[q]);
      if (!l12) {
#pragma unroll
        for (int q = 0; q < 6; ++q) stnt(y2 + q * n + i, fminf(fmaxf(av[q], -alpha), alpha));
      } else {
        float ss = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; ++q) ss = fmaf(av[q], av[q], ss);
        const float nrm = sqrtf(ss + 1e-12f);
        const float s = fmaxf(nrm / alpha, 1.0f);
#pragma unroll
        for (int q = 0; q < 6; ++q) stnt(y2 + q * n + i, av[q] / s);
      }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) yc[q] = yn[q];
    __syncthreads();
  }
}

// xn = max(0, x - tau (g + H^T y2)); partial sums of (xn - x)^2, x^2, xn^2
// (pdhg_mlx.py:741-742, 749-750, 805-806).  Components 0 (zz), 3 (yz) and
// 4 (xz) need z-neighbours (three-plane windows); 1, 2 and 5 only the current plane.
__global__ void __launch_bounds__(kTileThreads, 6) k_pd_primal(float* __restrict__ xn, const float* __restrict__ x,
                                                            const float* __restrict__ g,
                                                            const float* __restrict__ y2, Dims d,
                                                            const double* __restrict__ sc, W6 w,
                                                            double* __restrict__ part) {
  __shared__ float win[3][3 * kPlane];  // components 0, 3, 4
  __shared__ float cur[3][kPlane];      // components 1, 2, 5 at the current plane
  __shared__ double sh[16];
  const Tile t = tile_of(d);
  const PlanePos pp(d, t);
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  const float tf = (float)sc[PD_TAU];
  double s_dx = 0.0, s_x = 0.0, s_xn = 0.0;
  const float* wsrc[3] = {y2, y2 + 3 * n, y2 + 4 * n};
  const float* csrc[3] = {y2 + 1 * n, y2 + 2 * n, y2 + 5 * n};
  const float ww[3] = {w.w[0], w.w[3], w.w[4]}, cw[3] = {w.w[1], w.w[2], w.w[5]};
  // staged weights of the thread's two samples with the reflections hess_adj_halo
  // expects (exact sign flips): components 3 and 4 carry their y / x reflection, 5
  // both; the z reflection of 3 and 4 is applied per plane
  const HaloSign hs(d, t);
  const float wws[3][2] = {{ww[0], ww[0]}, {ww[1] * hs.y0, ww[1] * hs.y1}, {ww[2] * hs.x0, ww[2] * hs.x1}};
  const float cws[3][2] = {{cw[0], cw[0]}, {cw[1], cw[1]}, {cw[2] * hs.x0 * hs.y0, cw[2] * hs.x1 * hs.y1}};
  float wv[3][2], cv[3][2];
  auto wstore = [&](int j, int zz) {
    const float sz = j == 0 ? 1.0f : zsign(d, zz);
    pp.store(&win[j][slot_of(zz, t.z0) * kPlane], (wv[j][0] * wws[j][0]) * sz, (wv[j][1] * wws[j][1]) * sz);
  };
  // each component fetches only the halo hess_adj_halo reads of it: zz none, yz the
  // y halo, xz the x halo (windowed); yy y, xx x, xy both (current plane)
  auto wload = [&](int j, int64_t b) {
    if (j == 0) pp.load_need<false, false>(wsrc[0], b, wv[0][0], wv[0][1]);
    else if (j == 1) pp.load_need<false, true>(wsrc[1], b, wv[1][0], wv[1][1]);
    else pp.load_need<true, false>(wsrc[2], b, wv[2][0], wv[2][1]);
  };
  auto cload = [&](int j, int64_t b) {
    if (j == 0) pp.load_need<false, true>(csrc[0], b, cv[0][0], cv[0][1]);
    else if (j == 1) pp.load_need<true, false>(csrc[1], b, cv[1][0], cv[1][1]);
    else pp.load_need<true, true>(csrc[2], b, cv[2][0], cv[2][1]);
  };
  for (int zz = t.z0 - 1; zz <= t.z0; ++zz) {
    const int64_t bz = pp.base(d, zz);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      wload(j, bz);
      wstore(j, zz);
    }
  }
  {
    const int64_t bn = pp.base(d, t.z0 + 1), bc = pp.base(d, t.z0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      wload(j, bn);
      cload(j, bc);
    }
  }
  // this voxel's gradient and iterate, loaded one plane ahead
  float gc = 0.f, xc = 0.f;
  if (t.in) {
    const int64_t i0 = ((int64_t)t.z0 * d.ny + t.y) * d.nx + t.x;
    gc = ldnt(g + i0);
    xc = ldnt(x + i0);
  }
  for (int z = t.z0; z < t.z1; ++z) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      wstore(j, z + 1);
      pp.store(cur[j], cv[j][0] * cws[j][0], cv[j][1] * cws[j][1]);
    }
    __syncthreads();
    if (z + 1 < t.z1) {
      const int64_t bn = pp.base(d, z + 2), bc = pp.base(d, z + 1);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        wload(j, bn);
        cload(j, bc);
      }
    }
    float gn = 0.f, xnext = 0.f;
    if (t.in && z + 1 < t.z1) {
      const int64_t i1 = ((int64_t)(z + 1) * d.ny + t.y) * d.nx + t.x;
      gn = ldnt(g + i1);
      xnext = ldnt(x + i1);
    }
    if (t.in) {
      auto hw = [&](int q, int zz, int yy, int xx) {
        const int o = (yy - t.y0 + 1) * kPW + (xx - t.x0 + 1);
        switch (q) {
          case 0: return win[0][slot_of(zz, t.z0) * kPlane + o];
          case 3: return win[1][slot_of(zz, t.z0) * kPlane + o];
          case 4: return win[2][slot_of(zz, t.z0) * kPlane + o];
          case 1: return cur[0][o];
          case 2: return cur[1][o];
          default: return cur[2][o];
        }
      };
      const float lt = hess_adj_halo(hw, z, t.y, t.x);
      const int64_t i = ((int64_t)z * d.ny + t.y) * d.nx + t.x;
      const float grad = gc + lt;
      const float xv = xc;
      const float v = fmaxf(fmaf(-tf, grad, xv), 0.0f);
      stnt(xn + i, v);
      const float dv = v - xv;
      s_dx += (double)(dv * dv);
      s_x += (double)(xv * xv);
      s_xn += (double)(v * v);
    }
    gc = gn;
    xc = xnext;
    __syncthreads();
  }
  const int64_t b = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  double r0 = pd_block_sum(s_dx, sh);
  double r1 = pd_block_sum(s_x, sh);
  double r2 = pd_block_sum(s_xn, sh);
  if (threadIdx.x == 0) {
    part[b * 3 + 0] = r0;
    part[b * 3 + 1] = r1;
    part[b * 3 + 2] = r2;
  }
}

// ---------------------------------------------------------------------------
// Fused dual and primal steps (pdhg_mlx.py:729-750).  g = A^T y1_new does not
// depend on y2, so the adjoint convolution runs before this kernel; one z sweep
// per tile then computes y2_new = prox(y2 + sigma H xbar) on the tile and on the
// one-voxel ring around it that H^T reads (each ring recomputed by the
// neighbouring tiles from y2 and xbar, which no tile overwrites: y2_new and x_new
// go to buffers of their own), and from the staged y2_new the primal step
// x_new = max(0, x - tau (g + H^T y2_new)) with the partial sums of
// (x_new - x)^2, x^2, x_new^2.  y2_new is written once and never read back: per
// voxel x, xp, y2, g are read and y2_new, x_new written once (64 B, against 92 B
// for k_pd_dual followed by k_pd_primal), with the same arithmetic in the same
// order, so the same values.
//
// Thread t owns tile voxel (t % kTX, t / kTX) in every role (xbar, y2_new, x_new;
// coalesced, and its x stays in registers from staging to the primal step); the
// rings are spread over the first threads: the xbar ring (two voxels wide, 304
// points) and the y2_new ring, where H^T reads only components 1, 3, 5 on the y
// sides, 2, 4, 5 on the x sides and 5 at the corners (all six under the L1_2
// projection, whose norm needs them).
// ---------------------------------------------------------------------------

constexpr int kFW = kTX + 4, kFH = kTY + 4, kFPlane = kFW * kFH;  // xbar window: two-voxel ring
constexpr int kXRing = kFPlane - kTX * kTY;                         // 304
// planes per workgroup: the z sweep restarts its pipeline (four xbar planes, two
// y2_new planes) at every chunk, so chunks are longer than the other stencils'
constexpr int kZCF = 64;
constexpr int kYRing = kPlane - kTX * kTY;                          // 148
static_assert(kXRing <= kTileThreads && kYRing <= kTileThreads, "one ring point per thread");

// Weighted component q of Hessian3D.forward at (z, y, x): hess6_halo's expression
// for that component alone (the same operations, so the same value).
template <class F>
__device__ __forceinline__ float hess_comp(const F& f, int q, int z, int y, int x, const float* w) {
  const int zm = z - 1, zp = z + 1, ym = y - 1, yp = y + 1, xm = x - 1, xp = x + 1;
  float h;
  switch (q) {
    case 0: h = (f(zm, y, x) - 2.0f * f(z, y, x)) + f(zp, y, x); break;
    case 1: h = (f(z, ym, x) - 2.0f * f(z, y, x)) + f(z, yp, x); break;
    case 2: h = (f(z, y, xm) - 2.0f * f(z, y, x)) + f(z, y, xp); break;
    case 3: {
      const float dz_yp = (f(zp, yp, x) - f(zm, yp, x)) / 2.0f;
      const float dz_ym = (f(zp, ym, x) - f(zm, ym, x)) / 2.0f;
      h = (dz_yp - dz_ym) / 2.0f;
      break;
    }
    case 4: {
      const float dz_xp = (f(zp, y, xp) - f(zm, y, xp)) / 2.0f;
      const float dz_xm = (f(zp, y, xm) - f(zm, y, xm)) / 2.0f;
      h = (dz_xp - dz_xm) / 2.0f;
      break;
    }
    default: {
      const float dy_xp = (f(z, yp, xp) - f(z, ym, xp)) / 2.0f;
      const float dy_xm = (f(z, yp, xm) - f(z, ym, xm)) / 2.0f;
      h = (dy_xp - dy_xm) / 2.0f;
      break;
    }
  }
  return h * w[q];
}

template <int L12>
__global__ void __launch_bounds__(kTileThreads, 4) k_pd_dual_primal(float* __restrict__ y2o, float* __restrict__ xn,
                                                                    const float* __restrict__ x,
                                                                    const float* __restrict__ xp,
                                                                    const float* __restrict__ y2i,
                                                                    const float* __restrict__ g, Dims d,
                                                                    const double* __restrict__ sc, W6 w,
                                                                    float alpha, double* __restrict__ part) {
  // four plane slots each (plane z in slot z & 3): one barrier per step suffices
  // when step z stages xbar(z + 4), computes y2_new(z + 2) and x_new(z)
  __shared__ float xb[4][kFPlane];    // xbar (window rows of kFW)
  __shared__ float yw[3][4][kPlane];  // weighted y2_new components 0, 3, 4 (rows of kPW)
  __shared__ float yc[3][4][kPlane];  // components 1, 2, 5
  __shared__ double sh[16];
  const Tile t = tile_of(d, kZCF);
  const int z0 = t.z0, z1 = t.z1;
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  const int64_t pstride = (int64_t)d.ny * d.nx;
  const float sf = (float)sc[PD_SIGMA], th = (float)sc[PD_THETA], tf = (float)sc[PD_TAU];
  const int thr = threadIdx.x;
  auto plane = [&](int z) { return (int64_t)clampi(z, d.nz) * pstride; };

  // the thread's voxel: clamped in-plane offset (also its y2_new / x_new offset when t.in;
  // in-plane offsets fit 32 bits, host-checked), its xbar- and y2_new-window slots; a
  // voxel past the volume's x or y end (a partial tile) is ring for its in-volume
  // neighbours: computed at the clamped position (xbar-window centre ocy, ocx) and
  // staged with the reflection signs osy, osx
  const int ocyg = clampi(t.y, d.ny), ocxg = clampi(t.x, d.nx);
  const int voff = ocyg * d.nx + ocxg;
  const int xwin = (t.ty + 2) * kFW + (t.tx + 2);
  const int ywin = (t.ty + 1) * kPW + (t.tx + 1);
  const int ocy = ocyg - t.y0 + 2, ocx = ocxg - t.x0 + 2;
  const float osy = t.y < d.ny ? 1.0f : -1.0f, osx = t.x < d.nx ? 1.0f : -1.0f;
  // its xbar ring point (thr < kXRing): rows 0, 1 and kFH-2, kFH-1, then the side columns
  int xr_ly, xr_lx;
  {
    const int h = thr;
    if (h < 4 * kFW) {
      const int r = h / kFW;
      xr_ly = r < 2 ? r : kFH - 4 + r;
      xr_lx = h - r * kFW;
    } else {
      const int k = h - 4 * kFW;  // 4 columns x kTY rows
      xr_ly = 2 + (k >> 2);
      const int c = k & 3;
      xr_lx = c < 2 ? c : kFW - 4 + c;
    }
  }
  const bool xr_has = thr < kXRing;
  const int xr_off = clampi(t.y0 + xr_ly - 2, d.ny) * d.nx + clampi(t.x0 + xr_lx - 2, d.nx);
  const int xr_win = xr_ly * kFW + xr_lx;
  // its y2_new ring point (thr < kYRing): y sides, x sides, corners; computed at the
  // clamped position (xbar-window centre yr_cy, yr_cx); reflection signs outside the volume
  int yr_ly, yr_lx, yr_kind;  // kind 0: y side (components 1, 3, 5), 1: x side (2, 4, 5), 2: corner (5)
  {
    const int h = thr;
    if (h < 2 * kTX) {
      yr_kind = 0;
      yr_ly = h < kTX ? 0 : kPH - 1;
      yr_lx = 1 + (h & (kTX - 1));
    } else if (h < 2 * kTX + 2 * kTY) {
      yr_kind = 1;
      const int k = h - 2 * kTX;
      yr_ly = 1 + (k % kTY);
      yr_lx = k < kTY ? 0 : kPW - 1;
    } else {
      yr_kind = 2;
      const int k = (h - 2 * kTX - 2 * kTY) & 3;
      yr_ly = (k & 2) ? kPH - 1 : 0;
      yr_lx = (k & 1) ? kPW - 1 : 0;
    }
  }
  const bool yr_has = thr < kYRing;
  const int yr_gy = t.y0 + yr_ly - 1, yr_gx = t.x0 + yr_lx - 1;
  const int yr_cyg = clampi(yr_gy, d.ny), yr_cxg = clampi(yr_gx, d.nx);
  const int yr_off = yr_cyg * d.nx + yr_cxg;
  const int yr_cy = yr_cyg - t.y0 + 2, yr_cx = yr_cxg - t.x0 + 2;
  const int yr_win = yr_ly * kPW + yr_lx;
  const float yr_sy = (yr_gy < 0 || yr_gy >= d.ny) ? -1.0f : 1.0f;
  const float yr_sx = (yr_gx < 0 || yr_gx >= d.nx) ? -1.0f : 1.0f;
  // components of the ring point (L1: the three H^T reads there; L1_2: all six)
  constexpr int NR = L12 ? 6 : 3;
  int rq[NR];
  if (L12) {
#pragma unroll
    for (int q = 0; q < NR; ++q) rq[q] = q;
  } else {
    rq[0] = yr_kind == 0 ? 1 : (yr_kind == 1 ? 2 : 5);
    rq[1] = yr_kind == 0 ? 3 : (yr_kind == 1 ? 4 : 5);
    rq[2] = 5;
  }

  float ax, axp, rx, rxp;          // x, xp of the next xbar plane: own voxel, ring point
  float yv[6], yr[NR];             // y2 of the next y2_new plane: own voxel, ring point
  float yv1[6], yr1[NR], yv2[6], yr2[NR];  // the same one and two planes further (requested ahead)
  float xq0 = 0.f, xq1 = 0.f, xq2 = 0.f, xq3 = 0.f;  // own x on planes z .. z+3 (the primal's x)
  auto xload = [&](int z) {
    const int64_t b = plane(z);
    ax = x[b + voff];
    axp = xp[b + voff];
    rx = xr_has ? x[b + xr_off] : 0.0f;
    rxp = xr_has ? xp[b + xr_off] : 0.0f;
  };
  auto xbar = [&](float v, float vp) { return fmaf(th, v - vp, v); };
  auto xstage = [&](int z) {
    xb[z & 3][xwin] = xbar(ax, axp);
    if (xr_has) xb[z & 3][xr_win] = xbar(rx, rxp);
  };
  auto yload_to = [&](int p, float* ov, float* orr) {
    const int64_t b = plane(p);
#pragma unroll
    for (int q = 0; q < 6; ++q) ov[q] = y2i[q * n + b + voff];
#pragma unroll
    for (int q = 0; q < NR; ++q) orr[q] = yr_has ? y2i[rq[q] * n + b + yr_off] : 0.0f;
  };
  auto yload = [&](int p) { yload_to(p, yv, yr); };
  auto prox = [&](float* o, int m) {
    if (!L12) {
      for (int q = 0; q < m; ++q) o[q] = fminf(fmaxf(o[q], -alpha), alpha);
    } else {
      float ss = 0.0f;
#pragma unroll
      for (int q = 0; q < 6; ++q) ss = fmaf(o[q], o[q], ss);
      const float nrm = sqrtf(ss + 1e-12f);
      const float s = fmaxf(nrm / alpha, 1.0f);
#pragma unroll
      for (int q = 0; q < 6; ++q) o[q] = o[q] / s;
    }
  };
  // y2_new of plane p (from yv / yr): own voxel (written out if `write`) and ring point,
  // staged with the weights and reflections k_pd_primal stages y2 with
  auto ycompute = [&](int p, bool write) {
    const int a = p & 3;
    auto f = [&](int zz, int yy, int xx) { return xb[zz & 3][yy * kFW + xx]; };
    {
      float h[6];
      hess6_halo(f, p, ocy, ocx, w.w, h);
      float o[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) o[q] = fmaf(sf, h[q], yv[q]);
      prox(o, 6);
      if (write && t.in) {
        const int64_t b = plane(p);
#pragma unroll
        for (int q = 0; q < 6; ++q) stnt(y2o + q * n + b + voff, o[q]);
      }
      yw[0][a][ywin] = o[0] * w.w[0];
      yw[1][a][ywin] = o[3] * (w.w[3] * osy);
      yw[2][a][ywin] = o[4] * (w.w[4] * osx);
      yc[0][a][ywin] = o[1] * w.w[1];
      yc[1][a][ywin] = o[2] * w.w[2];
      yc[2][a][ywin] = o[5] * ((w.w[5] * osx) * osy);
    }
    if (yr_has) {
      float o[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) o[q] = fmaf(sf, hess_comp(f, rq[q], p, yr_cy, yr_cx, w.w), yr[q]);
      prox(o, NR);
      // (ring points: the y sides carry the y reflection, the x sides the x one)
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        switch (rq[q]) {
          case 0: yw[0][a][yr_win] = o[q] * w.w[0]; break;
          case 1: yc[0][a][yr_win] = o[q] * w.w[1]; break;
          case 2: yc[1][a][yr_win] = o[q] * w.w[2]; break;
          case 3: yw[1][a][yr_win] = o[q] * (w.w[3] * yr_sy); break;
          case 4: yw[2][a][yr_win] = o[q] * (w.w[4] * yr_sx); break;
          default: yc[2][a][yr_win] = o[q] * ((w.w[5] * yr_sx) * yr_sy); break;
        }
      }
    }
  };
  double s_dx = 0.0, s_x = 0.0, s_xn = 0.0;
  // x_new of plane z at the own voxel; y2_new of planes z +- 1 beyond the volume's z
  // ends is that of the end plane, components 3 and 4 reflected (k_pd_primal's staging)
  auto xcompute = [&](int z, float gz, float xv) {
    if (!t.in) return;
    const int cm = (z > 0 ? z - 1 : z) & 3, c0 = z & 3, cp = (z + 1 < d.nz ? z + 1 : z) & 3;
    const float szm = z > 0 ? 1.0f : -1.0f, szp = z + 1 < d.nz ? 1.0f : -1.0f;
    auto hw = [&](int q, int zz, int yy, int xx) {
      const int o = (yy - t.y0 + 1) * kPW + (xx - t.x0 + 1);
      if (q == 1) return yc[0][c0][o];
      if (q == 2) return yc[1][c0][o];
      if (q == 5) return yc[2][c0][o];
      const int j = q == 0 ? 0 : (q == 3 ? 1 : 2);
      const float v = yw[j][zz < z ? cm : (zz > z ? cp : c0)][o];
      if (q == 0) return v;
      return v * (zz < z ? szm : (zz > z ? szp : 1.0f));
    };
    const float lt = hess_adj_halo(hw, z, t.y, t.x);
    const int64_t i = (int64_t)z * pstride + voff;
    const float grad = gz + lt;
    const float v = fmaxf(fmaf(-tf, grad, xv), 0.0f);
    stnt(xn + i, v);
    const float dv = v - xv;
    s_dx += (double)(dv * dv);
    s_x += (double)(xv * xv);
    s_xn += (double)(v * v);
  };
  auto gload = [&](int z) { return t.in ? ldnt(g + (int64_t)z * pstride + voff) : 0.0f; };
  auto keep = [&](int q) {
    if (q == z0) xq0 = ax;
    else if (q == z0 + 1) xq1 = ax;
    else if (q == z0 + 2) xq2 = ax;
    else if (q == z0 + 3) xq3 = ax;
  };

  // Prologue: y2_new is needed on planes [pa, pb] (z0 - 1 .. z1 inside the volume); before
  // step z0, planes pa .. z0 + 1 of it and xbar planes up to z0 + 3.  Plane q of xbar may
  // replace plane q - 4 once every y2_new plane that reads q - 4 is done.
  const int pa = z0 > 0 ? z0 - 1 : 0;
  const int pb = z1 < d.nz ? z1 : d.nz - 1;
  const int ylast = min(z0 + 1, pb);
  int nx_ = pa - 1, ny_ = pa;  // next xbar plane to stage, next y2_new plane to compute
  for (int k = 0; k < 4; ++k, ++nx_) {
    xload(nx_);
    xstage(nx_);
    keep(nx_);
  }
  __syncthreads();
  for (; ny_ <= ylast && ny_ + 1 < nx_; ++ny_) {
    yload(ny_);
    ycompute(ny_, ny_ >= z0 && ny_ < z1);
  }
  while (nx_ <= z0 + 3 || ny_ <= ylast) {
    __syncthreads();
    xload(nx_);
    xstage(nx_);
    keep(nx_);
    ++nx_;
    __syncthreads();
    for (; ny_ <= ylast && ny_ + 1 < nx_; ++ny_) {
      yload(ny_);
      ycompute(ny_, ny_ >= z0 && ny_ < z1);
    }
  }
  // operands of step z0 (xbar one step ahead, y2 two steps ahead from here on)
  if (z0 + 3 <= pb) xload(z0 + 4);
  if (z0 + 2 <= pb) yload(z0 + 2);
  if (z0 + 3 <= pb) yload_to(z0 + 3, yv1, yr1);
  float gz = gload(z0);
  __syncthreads();
  for (int z = z0; z < z1; ++z) {
    // xbar(z + 4) replaces xbar(z), whose last reader (y2_new(z + 1)) ran before the barrier
    const bool st4 = z + 3 <= pb;
    if (st4) xstage(z + 4);
    const float xq4 = ax;
    // later steps' operands, requested before this step's work
    if (z + 4 <= pb) {
      xload(z + 5);
      yload_to(z + 4, yv2, yr2);
    }
    const float gnext = z + 1 < z1 ? gload(z + 1) : 0.0f;
    // y2_new(z + 2) replaces y2_new(z - 2), last read by x_new(z - 1)
    if (z + 2 <= pb) ycompute(z + 2, z + 2 < z1);
    xcompute(z, gz, xq0);
    gz = gnext;
    xq0 = xq1;
    xq1 = xq2;
    xq2 = xq3;
    xq3 = xq4;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      yv[q] = yv1[q];
      yv1[q] = yv2[q];
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      yr[q] = yr1[q];
      yr1[q] = yr2[q];
    }
    __syncthreads();
  }
  const int64_t b = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const double r0 = pd_block_sum(s_dx, sh);
  const double r1 = pd_block_sum(s_x, sh);
  const double r2 = pd_block_sum(s_xn, sh);
  if (threadIdx.x == 0) {
    part[b * 3 + 0] = r0;
    part[b * 3 + 1] = r1;
    part[b * 3 + 2] = r2;
  }
}

// sum |H(a - b)|^2 (dL of the step-size rule; H is linear, so this is the
// reference's |L x_new - L x|^2 up to rounding) and, if want_abs, the
// regularizer loss of a: sum |H a| (1, "L1") or sum sqrt(sum_q (H a)_q^2 + 1e-12) (2, "L1_2")
// Hessian3D.adjoint, out = H^T y2, through the same sliding LDS window as k_pd_primal:
// each workgroup stages the weighted components (0, 3, 4 over three planes; 1, 2, 5 at
// the current plane) with their halo once, instead of every voxel gathering ~19
// neighbours across six component planes (the same arithmetic as hess_adj on
// global memory, so the results are bit-identical).
__global__ void __launch_bounds__(kTileThreads, 6) k_hess_adj_tiled(const float* __restrict__ y2,
                                                                 float* __restrict__ out, Dims d, W6 w) {
  __shared__ float win[3][3 * kPlane];  // components 0, 3, 4
  __shared__ float cur[3][kPlane];      // components 1, 2, 5 at the current plane
  const Tile t = tile_of(d);
  const PlanePos pp(d, t);
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  const float* wsrc[3] = {y2, y2 + 3 * n, y2 + 4 * n};
  const float* csrc[3] = {y2 + 1 * n, y2 + 2 * n, y2 + 5 * n};
  const float ww[3] = {w.w[0], w.w[3], w.w[4]}, cw[3] = {w.w[1], w.w[2], w.w[5]};
  // staged weights of the thread's two samples with the reflections hess_adj_halo
  // expects (exact sign flips): components 3 and 4 carry their y / x reflection, 5
  // both; the z reflection of 3 and 4 is applied per plane
  const HaloSign hs(d, t);
  const float wws[3][2] = {{ww[0], ww[0]}, {ww[1] * hs.y0, ww[1] * hs.y1}, {ww[2] * hs.x0, ww[2] * hs.x1}};
  const float cws[3][2] = {{cw[0], cw[0]}, {cw[1], cw[1]}, {cw[2] * hs.x0 * hs.y0, cw[2] * hs.x1 * hs.y1}};
  float wv[3][2], cv[3][2];
  auto wstore = [&](int j, int zz) {
    const float sz = j == 0 ? 1.0f : zsign(d, zz);
    pp.store(&win[j][slot_of(zz, t.z0) * kPlane], (wv[j][0] * wws[j][0]) * sz, (wv[j][1] * wws[j][1]) * sz);
  };
  // each component fetches only the halo hess_adj_halo reads of it: zz none, yz the
  // y halo, xz the x halo (windowed); yy y, xx x, xy both (current plane)
  auto wload = [&](int j, int64_t b) {
    if (j == 0) pp.load_need<false, false>(wsrc[0], b, wv[0][0], wv[0][1]);
    else if (j == 1) pp.load_need<false, true>(wsrc[1], b, wv[1][0], wv[1][1]);
    else pp.load_need<true, false>(wsrc[2], b, wv[2][0], wv[2][1]);
  };
  auto cload = [&](int j, int64_t b) {
    if (j == 0) pp.load_need<false, true>(csrc[0], b, cv[0][0], cv[0][1]);
    else if (j == 1) pp.load_need<true, false>(csrc[1], b, cv[1][0], cv[1][1]);
    else pp.load_need<true, true>(csrc[2], b, cv[2][0], cv[2][1]);
  };
  for (int zz = t.z0 - 1; zz <= t.z0; ++zz) {
    const int64_t bz = pp.base(d, zz);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      wload(j, bz);
      wstore(j, zz);
    }
  }
  {
    const int64_t bn = pp.base(d, t.z0 + 1), bc = pp.base(d, t.z0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      wload(j, bn);
      cload(j, bc);
    }
  }
  for (int z = t.z0; z < t.z1; ++z) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      wstore(j, z + 1);
      pp.store(cur[j], cv[j][0] * cws[j][0], cv[j][1] * cws[j][1]);
    }
    __syncthreads();
    if (z + 1 < t.z1) {
      const int64_t bn = pp.base(d, z + 2), bc = pp.base(d, z + 1);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        wload(j, bn);
        cload(j, bc);
      }
    }
    if (t.in) {
      auto hw = [&](int q, int zz, int yy, int xx) {
        const int o = (yy - t.y0 + 1) * kPW + (xx - t.x0 + 1);
        switch (q) {
          case 0: return win[0][slot_of(zz, t.z0) * kPlane + o];
          case 3: return win[1][slot_of(zz, t.z0) * kPlane + o];
          case 4: return win[2][slot_of(zz, t.z0) * kPlane + o];
          case 1: return cur[0][o];
          case 2: return cur[1][o];
          default: return cur[2][o];
        }
      };
      stnt(out + ((int64_t)z * d.ny + t.y) * d.nx + t.x, hess_adj_halo(hw, z, t.y, t.x));
    }
    __syncthreads();
  }
}

constexpr int kZCS = 32;  // planes per workgroup of the sums
// Planes rotate through four LDS slots, so one barrier per plane suffices: the slot
// a plane's store reuses was last read two planes earlier, before the previous
// barrier.  Plane loads run two planes ahead of the computed one.
__global__ void __launch_bounds__(kTileThreads) k_pd_hess_sums(const float* __restrict__ a,
                                                               const float* __restrict__ b, Dims d, W6 w,
                                                               int want_abs, double* __restrict__ part) {
  __shared__ float pdl[4 * kPlane], pa[4 * kPlane];
  __shared__ double sh[16];
  const Tile t = tile_of(d, kZCS);
  const PlanePos pp(d, t);
  auto slot = [&](int z) { return (z - t.z0 + 1) & 3; };
  double s_sq = 0.0, s_abs = 0.0;
  float a0, a1, b0, b1;
  for (int zz = t.z0 - 1; zz <= t.z0; ++zz) {
    const int64_t bz = pp.base(d, zz);
    pp.load(a, bz, a0, a1);
    pp.load(b, bz, b0, b1);
    pp.store(pdl + slot(zz) * kPlane, a0 - b0, a1 - b1);
    if (want_abs) pp.store(pa + slot(zz) * kPlane, a0, a1);
  }
  // planes z + 1 (r1) and z + 2 (r2) in registers
  float r1a0, r1a1, r1b0, r1b1, r2a0 = 0.f, r2a1 = 0.f, r2b0 = 0.f, r2b1 = 0.f;
  {
    const int64_t bz = pp.base(d, t.z0 + 1);
    pp.load(a, bz, r1a0, r1a1);
    pp.load(b, bz, r1b0, r1b1);
  }
  if (t.z0 + 2 <= t.z1) {
    const int64_t bz = pp.base(d, t.z0 + 2);
    pp.load(a, bz, r2a0, r2a1);
    pp.load(b, bz, r2b0, r2b1);
  }
  for (int z = t.z0; z < t.z1; ++z) {
    pp.store(pdl + slot(z + 1) * kPlane, r1a0 - r1b0, r1a1 - r1b1);
    if (want_abs) pp.store(pa + slot(z + 1) * kPlane, r1a0, r1a1);
    r1a0 = r2a0;
    r1a1 = r2a1;
    r1b0 = r2b0;
    r1b1 = r2b1;
    if (z + 3 <= t.z1) {
      const int64_t bz = pp.base(d, z + 3);
      pp.load(a, bz, r2a0, r2a1);
      pp.load(b, bz, r2b0, r2b1);
    }
    __syncthreads();
    if (t.in) {
      auto fd = [&](int zz, int yy, int xx) { return lds_at(pdl, slot(zz), t, yy, xx); };
      float h[6];
      hess6_halo(fd, z, t.y, t.x, w.w, h);
      // the voxel's six squares summed in float, then accumulated in double
      float hh = 0.0f;
#pragma unroll
      for (int q = 0; q < 6; ++q) hh = fmaf(h[q], h[q], hh);
      s_sq += (double)hh;
      if (want_abs) {
        auto fa = [&](int zz, int yy, int xx) { return lds_at(pa, slot(zz), t, yy, xx); };
        hess6_halo(fa, z, t.y, t.x, w.w, h);
        float ss = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          if (want_abs == 1) s_abs += (double)fabsf(h[q]);
          ss += h[q] * h[q];
        }
        if (want_abs == 2) s_abs += (double)sqrtf(ss + 1e-12f);
      }
    }
  }
  const int64_t blk = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  double r0 = pd_block_sum(s_sq, sh);
  double r1 = pd_block_sum(s_abs, sh);
  if (threadIdx.x == 0) {
    part[blk * 2 + 0] = r0;
    part[blk * 2 + 1] = r1;
  }
}

// Hessian3D.forward into out (6, nz, ny, nx)
__global__ void k_hess_fwd(const float* __restrict__ f, float* __restrict__ out, Dims d, W6 w) {
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  auto fv = [&](int z, int y, int xx) { return f[((int64_t)z * d.ny + y) * d.nx + xx]; };
  PD_LOOP(i, n) {
    const int xx = (int)(i % d.nx);
    const int y = (int)((i / d.nx) % d.ny);
    const int z = (int)(i / ((int64_t)d.nx * d.ny));
    float h[6];
    hess6(fv, d, z, y, xx, w.w, h);
#pragma unroll
    for (int q = 0; q < 6; ++q) out[q * n + i] = h[q];
  }
}

// sum (a - b)^2 and, if want_loss, the Poisson (or Gaussian) data loss of a + bg
// against data (pdhg_mlx.py:752-759, 860-872)
__global__ void k_pd_data_sums(const float* __restrict__ a, const float* __restrict__ b,
                               const float* __restrict__ data, int64_t n, float bg, int want_loss, int loss_type,
                               double* __restrict__ part) {
  __shared__ double sh[16];
  double s_sq = 0.0, s_loss = 0.0;
  auto one = [&](float av, float bv, const float* dp) {
    const float dd = av - bv;
    s_sq += (double)(dd * dd);
    if (want_loss) {
      const float fm = av + bg;
      const float dv = *dp;
      if (loss_type == 0) {
        const float fms = fmaxf(fm, 1e-6f);
        s_loss += (double)(dv * logf((dv + 1e-6f) / fms) - (dv - fm));
      } else {
        const float r = dv - fm;
        s_loss += 0.5 * (double)(r * r);
      }
    }
  };
  // 16-byte loads of a and b where both are 16-byte aligned (four samples per
  // thread and iteration in flight), then the remainder one sample at a time
  int64_t done = 0;
  if ((((uintptr_t)a | (uintptr_t)b) & 15) == 0) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int64_t n4 = n / 4;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += (int64_t)gridDim.x * blockDim.x) {
      const f4 av = __builtin_nontemporal_load(reinterpret_cast<const f4*>(a) + q);
      const f4 bv = __builtin_nontemporal_load(reinterpret_cast<const f4*>(b) + q);
      one(av.x, bv.x, data + 4 * q);
      one(av.y, bv.y, data + 4 * q + 1);
      one(av.z, bv.z, data + 4 * q + 2);
      one(av.w, bv.w, data + 4 * q + 3);
    }
    done = 4 * n4;
  }
  for (int64_t i = done + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    one(ldnt(a + i), ldnt(b + i), data + i);
  double r0 = pd_block_sum(s_sq, sh);
  double r1 = pd_block_sum(s_loss, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2 + 0] = r0;
    part[blockIdx.x * 2 + 1] = r1;
  }
}

// Malitsky–Pock step-size update and convergence test in FP64 (pdhg_mlx.py:761-835).
__global__ void k_pd_scalars(double* sc, double* hist, int k, PdParams pp) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double tau = sc[PD_TAU], sigma = sc[PD_SIGMA], theta = sc[PD_THETA], step = sc[PD_STEP];
  const double dx2 = sc[PD_DX2], dK2 = sc[PD_DA2] + sc[PD_DL2];
  const double ratio = dx2 / (2.0 * sigma * fmax(dK2, 1e-12));
  const double tc = fmin(sqrt(1.0 + theta) * tau, pp.delta * sqrt(ratio));
  double tn = (tc * sigma * dK2 > pp.delta * dx2) ? pp.eta * tc : tc;
  tn = fmax(tn, step * 1e-4);
  const double thn = tn / tau;
  double sn = sigma / fmax(thn, 0.1);
  sn = fmin(fmax(sn, step * 1e-4), step * 1e4);
  const double rel = sqrt(dx2 / (fmax(sc[PD_XX], sc[PD_XNXN]) + 1e-12));
  hist[3 * k + 0] = tn;
  hist[3 * k + 1] = sn;
  hist[3 * k + 2] = rel;
  if (k >= pp.min_iter) {
    if (rel < pp.tol) {
      sc[PD_COUNT] += 1.0;
      if (sc[PD_COUNT] >= pp.patience) {
        sc[PD_DONE] = 1.0;
        sc[PD_FINAL] = k + 1;
        return;
      }
    } else {
      sc[PD_COUNT] = 0.0;
    }
  }
  sc[PD_TAU] = tn;
  sc[PD_SIGMA] = sn;
  sc[PD_THETA] = thn;
}

// ---------------------------------------------------------------------------
// First-order (Gradient1D/2D/3D, linops_mlx.py:46-127) and identity
// (pdhg_mlx.py:241-306) regularizers.  Components [z, y, x] of the forward
// difference d1_fwd (linops_core_mlx.py:25-32) with weights w3; the identity
// has one component, x itself.  Volumes are small next to the FFT passes'
// traffic, so the neighbours come through the caches (no LDS tiling).
// ---------------------------------------------------------------------------

struct W3 {
  float w[3];
};

struct Vox {
  int z, y, x;
};

__device__ __forceinline__ Vox vox_of(int64_t i, const Dims& d) {
  Vox v;
  v.x = (int)(i % d.nx);
  v.y = (int)((i / d.nx) % d.ny);
  v.z = (int)(i / ((int64_t)d.nx * d.ny));
  return v;
}

// weighted forward differences of f at voxel i: w_a (f[next along a] - f[i]), 0 on the last plane
template <class F>
__device__ __forceinline__ void grad3(const F& f, int64_t i, const Vox& v, const Dims& d, const W3& w, float* g) {
  const float c = f(i);
  const int64_t sz = (int64_t)d.ny * d.nx;
  g[0] = w.w[0] * ((v.z + 1 < d.nz ? f(i + sz) : c) - c);
  g[1] = w.w[1] * ((v.y + 1 < d.ny ? f(i + d.nx) : c) - c);
  g[2] = w.w[2] * ((v.x + 1 < d.nx ? f(i + 1) : c) - c);
}

// d1_fwd_adj of the weighted component q along one axis of n samples at index k
// (stride s): [-g0, g[k-1] - g[k], g[n-2]] (linops_core_mlx.py:35-39; n == 1 gives -g0)
__device__ __forceinline__ float fwd_adj_at(const float* __restrict__ g, int64_t i, int k, int n, int64_t s, float w) {
  if (k == 0) return -(w * g[i]);
  if (k == n - 1) return w * g[i - s];
  return w * g[i - s] - w * g[i];
}

// adjoint of grad3 at voxel i: adj_z + adj_y + adj_x (Gradient3D.adjoint order)
__device__ __forceinline__ float grad3_adj(const float* __restrict__ y, int64_t n, int64_t i, const Vox& v,
                                           const Dims& d, const W3& w) {
  const int64_t sz = (int64_t)d.ny * d.nx;
  const float az = fwd_adj_at(y, i, v.z, d.nz, sz, w.w[0]);
  const float ay = fwd_adj_at(y + n, i, v.y, d.ny, d.nx, w.w[1]);
  const float ax = fwd_adj_at(y + 2 * n, i, v.x, d.nx, 1, w.w[2]);
  return (az + ay) + ax;
}

template <int NC>
__device__ __forceinline__ void reg_prox_store(float* __restrict__ y2, int64_t n, int64_t i, const float* av,
                                               float alpha, int l12) {
  if (!l12) {
#pragma unroll
    for (int q = 0; q < NC; ++q) stnt(y2 + q * n + i, fminf(fmaxf(av[q], -alpha), alpha));
  } else {
    float ss = 0.0f;
#pragma unroll
    for (int q = 0; q < NC; ++q) ss += av[q] * av[q];
    const float s = fmaxf(sqrtf(ss + 1e-12f) / alpha, 1.0f);
#pragma unroll
    for (int q = 0; q < NC; ++q) stnt(y2 + q * n + i, av[q] / s);
  }
}

// y2 <- prox(y2 + sigma L(xbar)), xbar = x + theta (x - xp)  (pdhg_mlx.py:729-734)
template <int KIND>
__global__ void k_pr_dual(float* __restrict__ y2, const float* __restrict__ x, const float* __restrict__ xp, Dims d,
                          const double* __restrict__ sc, W3 w, float alpha, int l12) {
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  const float sf = (float)sc[PD_SIGMA], th = (float)sc[PD_THETA];
  auto xbar = [&](int64_t j) {
    const float v = x[j];
    return v + th * (v - xp[j]);
  };
  PD_LOOP(i, n) {
    if (KIND == DH_REG_GRADIENT) {
      float g[3];
      grad3(xbar, i, vox_of(i, d), d, w, g);
      float av[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) av[q] = ldnt(y2 + q * n + i) + sf * g[q];
      reg_prox_store<3>(y2, n, i, av, alpha, l12);
    } else {
      float av[1] = {ldnt(y2 + i) + sf * xbar(i)};
      reg_prox_store<1>(y2, n, i, av, alpha, l12);
    }
  }
}

// xn = max(0, x - tau (g + L^T y2)); partial sums of (xn - x)^2, x^2, xn^2
template <int KIND>
__global__ void k_pr_primal(float* __restrict__ xn, const float* __restrict__ x, const float* __restrict__ g,
                            const float* __restrict__ y2, Dims d, const double* __restrict__ sc, W3 w,
                            double* __restrict__ part) {
  __shared__ double sh[16];
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  const float tf = (float)sc[PD_TAU];
  double s_dx = 0.0, s_x = 0.0, s_xn = 0.0;
  PD_LOOP(i, n) {
    const float lt = KIND == DH_REG_GRADIENT ? grad3_adj(y2, n, i, vox_of(i, d), d, w) : y2[i];
    const float xv = ldnt(x + i);
    const float v = fmaxf(xv - tf * (ldnt(g + i) + lt), 0.0f);
    stnt(xn + i, v);
    const float dv = v - xv;
    s_dx += (double)(dv * dv);
    s_x += (double)(xv * xv);
    s_xn += (double)(v * v);
  }
  const double r0 = pd_block_sum(s_dx, sh);
  const double r1 = pd_block_sum(s_x, sh);
  const double r2 = pd_block_sum(s_xn, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 3 + 0] = r0;
    part[blockIdx.x * 3 + 1] = r1;
    part[blockIdx.x * 3 + 2] = r2;
  }
}

// sum |L(a - b)|^2 and, if want_abs, sum |L a| (1) or sum sqrt(sum_q (L a)_q^2 + 1e-12) (2)
template <int KIND>
__global__ void k_pr_sums(const float* __restrict__ a, const float* __restrict__ b, Dims d, W3 w, int want_abs,
                          double* __restrict__ part) {
  __shared__ double sh[16];
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  double s_sq = 0.0, s_abs = 0.0;
  auto fd = [&](int64_t j) { return a[j] - b[j]; };
  auto fa = [&](int64_t j) { return a[j]; };
  PD_LOOP(i, n) {
    constexpr int NC = KIND == DH_REG_GRADIENT ? 3 : 1;
    float h[NC];
    const Vox v = vox_of(i, d);
    if (KIND == DH_REG_GRADIENT) grad3(fd, i, v, d, w, h);
    else h[0] = fd(i);
#pragma unroll
    for (int q = 0; q < NC; ++q) s_sq += (double)(h[q] * h[q]);
    if (want_abs) {
      if (KIND == DH_REG_GRADIENT) grad3(fa, i, v, d, w, h);
      else h[0] = fa(i);
      float ss = 0.0f;
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        if (want_abs == 1) s_abs += (double)fabsf(h[q]);
        ss += h[q] * h[q];
      }
      if (want_abs == 2) s_abs += (double)sqrtf(ss + 1e-12f);
    }
  }
  const double r0 = pd_block_sum(s_sq, sh);
  const double r1 = pd_block_sum(s_abs, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x * 2 + 0] = r0;
    part[blockIdx.x * 2 + 1] = r1;
  }
}

__global__ void k_grad_fwd(const float* __restrict__ f, float* __restrict__ out, Dims d, W3 w) {
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  auto fv = [&](int64_t j) { return f[j]; };
  PD_LOOP(i, n) {
    float g[3];
    grad3(fv, i, vox_of(i, d), d, w, g);
#pragma unroll
    for (int q = 0; q < 3; ++q) out[q * n + i] = g[q];
  }
}

__global__ void k_grad_adj(const float* __restrict__ y, float* __restrict__ out, Dims d, W3 w) {
  const int64_t n = (int64_t)d.nz * d.ny * d.nx;
  PD_LOOP(i, n) out[i] = grad3_adj(y, n, i, vox_of(i, d), d, w);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t pd_y1(float* y1, const float* ax, const float* axp, const float* data, int64_t n, const double* sc,
                 float bg, int loss_type, hipStream_t st) {
  hipLaunchKernelGGL(k_pd_y1, dim3(pd_blocks(n)), dim3(kPdThreads), 0, st, y1, ax, axp, data, n, sc, bg, loss_type);
  return hipGetLastError();
}


hipError_t pd_dual(float* y2, const float* x, const float* xp, int nz, int ny, int nx, const double* sc,
                   const float* w, float alpha, int l12, hipStream_t st) {
  Dims d{nz, ny, nx};
  W6 ww;
  for (int q = 0; q < 6; ++q) ww.w[q] = w[q];
  hipLaunchKernelGGL(k_pd_dual, tile_grid(nz, ny, nx), dim3(kTileThreads), 0, st, y2, x, xp, d, sc, ww, alpha, l12);
  return hipGetLastError();
}

hipError_t pd_primal(float* xn, const float* x, const float* g, const float* y2, int nz, int ny, int nx,
                     double* sc, const float* w, double* part, hipStream_t st) {
  Dims d{nz, ny, nx};
  W6 ww;
  for (int q = 0; q < 6; ++q) ww.w[q] = w[q];
  hipLaunchKernelGGL(k_pd_primal, tile_grid(nz, ny, nx), dim3(kTileThreads), 0, st, xn, x, g, y2, d, sc, ww, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return finish(part, tile_blocks(nz, ny, nx), 3, sc + PD_DX2, st);  // DX2, XX, XNXN are consecutive
}

hipError_t pd_dual_primal(float* y2_new, float* xn, const float* x, const float* xp, const float* y2, const float* g,
                          int nz, int ny, int nx, double* sc, const float* w, float alpha, int l12, double* part,
                          hipStream_t st) {
  Dims d{nz, ny, nx};
  W6 ww;
  for (int q = 0; q < 6; ++q) ww.w[q] = w[q];
  const dim3 grid = tile_grid(nz, ny, nx, kZCF);
  if (l12)
    hipLaunchKernelGGL(k_pd_dual_primal<1>, grid, dim3(kTileThreads), 0, st, y2_new, xn, x, xp, y2, g, d, sc, ww,
                       alpha, part);
  else
    hipLaunchKernelGGL(k_pd_dual_primal<0>, grid, dim3(kTileThreads), 0, st, y2_new, xn, x, xp, y2, g, d, sc, ww,
                       alpha, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return finish(part, (int)(grid.x * grid.y * grid.z), 3, sc + PD_DX2, st);  // DX2, XX, XNXN are consecutive
}

hipError_t pd_hess_sums(const float* a, const float* b, int nz, int ny, int nx, const float* w, int want_abs,
                        double* sc, double* loss_out, double* part, hipStream_t st) {
  Dims d{nz, ny, nx};
  W6 ww;
  for (int q = 0; q < 6; ++q) ww.w[q] = w[q];
  const dim3 grid = tile_grid(nz, ny, nx, kZCS);
  hipLaunchKernelGGL(k_pd_hess_sums, grid, dim3(kTileThreads), 0, st, a, b, d, ww, want_abs, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return finish2(part, (int)(grid.x * grid.y * grid.z), sc + PD_DL2, loss_out, st);
}

hipError_t pd_data_sums(const float* a, const float* b, const float* data, int64_t n, float bg, int want_loss,
                        int loss_type, double* sc, double* loss_out, double* part, hipStream_t st) {
  const int nb = pd_blocks(n);
  hipLaunchKernelGGL(k_pd_data_sums, dim3(nb), dim3(kPdThreads), 0, st, a, b, data, n, bg, want_loss, loss_type,
                     part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return finish2(part, nb, sc + PD_DA2, loss_out, st);
}

__global__ void k_pd_move(double* sc, int from0, int to0, int from1, double* out1) {
  if (threadIdx.x == 0) {
    sc[to0] = sc[from0];
    if (out1) out1[0] = sc[from1];
  }
}

hipError_t pd_move(double* sc, int from0, int to0, int from1, double* out1, hipStream_t st) {
  hipLaunchKernelGGL(k_pd_move, dim3(1), dim3(64), 0, st, sc, from0, to0, from1, out1);
  return hipGetLastError();
}

// tau = sigma = step = 0.99 / sqrt(|A|^2 + |L|^2), theta = 1, every other slot 0 (pdhg_mlx.py:693-701)
__global__ void k_pd_init(double* sc, const double* blur_norm_sq, double reg_norm_sq) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i < DH_PD_NSLOTS; ++i) sc[i] = 0.0;
  const double s = 0.99 / sqrt(blur_norm_sq[0] + reg_norm_sq);
  sc[PD_TAU] = sc[PD_SIGMA] = sc[PD_STEP] = s;
  sc[PD_THETA] = 1.0;
}

hipError_t pd_init(double* sc, const double* blur_norm_sq, double reg_norm_sq, hipStream_t st) {
  hipLaunchKernelGGL(k_pd_init, dim3(1), dim3(64), 0, st, sc, blur_norm_sq, reg_norm_sq);
  return hipGetLastError();
}

hipError_t pd_scalars(double* sc, double* hist, int k, const PdParams& pp, hipStream_t st) {
  hipLaunchKernelGGL(k_pd_scalars, dim3(1), dim3(64), 0, st, sc, hist, k, pp);
  return hipGetLastError();
}

static W3 w3_of(const float* w) {
  W3 r;
  for (int q = 0; q < 3; ++q) r.w[q] = w ? w[q] : 1.0f;
  return r;
}

hipError_t pr_dual(int kind, float* y2, const float* x, const float* xp, int nz, int ny, int nx, const double* sc,
                   const float* w, float alpha, int l12, hipStream_t st) {
  const Dims d{nz, ny, nx};
  const int64_t n = (int64_t)nz * ny * nx;
  if (kind == DH_REG_GRADIENT)
    hipLaunchKernelGGL(k_pr_dual<DH_REG_GRADIENT>, dim3(pd_blocks(n)), dim3(kPdThreads), 0, st, y2, x, xp, d, sc,
                       w3_of(w), alpha, l12);
  else
    hipLaunchKernelGGL(k_pr_dual<DH_REG_IDENTITY>, dim3(pd_blocks(n)), dim3(kPdThreads), 0, st, y2, x, xp, d, sc,
                       w3_of(w), alpha, l12);
  return hipGetLastError();
}

hipError_t pr_primal(int kind, float* xn, const float* x, const float* g, const float* y2, int nz, int ny, int nx,
                     double* sc, const float* w, double* part, hipStream_t st) {
  const Dims d{nz, ny, nx};
  const int64_t n = (int64_t)nz * ny * nx;
  const int nb = pd_blocks(n);
  if (kind == DH_REG_GRADIENT)
    hipLaunchKernelGGL(k_pr_primal<DH_REG_GRADIENT>, dim3(nb), dim3(kPdThreads), 0, st, xn, x, g, y2, d, sc, w3_of(w),
                       part);
  else
    hipLaunchKernelGGL(k_pr_primal<DH_REG_IDENTITY>, dim3(nb), dim3(kPdThreads), 0, st, xn, x, g, y2, d, sc, w3_of(w),
                       part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return finish(part, nb, 3, sc + PD_DX2, st);
}

hipError_t pr_sums(int kind, const float* a, const float* b, int nz, int ny, int nx, const float* w, int want_abs,
                   double* sc, double* loss_out, double* part, hipStream_t st) {
  const Dims d{nz, ny, nx};
  const int64_t n = (int64_t)nz * ny * nx;
  const int nb = pd_blocks(n);
  if (kind == DH_REG_GRADIENT)
    hipLaunchKernelGGL(k_pr_sums<DH_REG_GRADIENT>, dim3(nb), dim3(kPdThreads), 0, st, a, b, d, w3_of(w), want_abs,
                       part);
  else
    hipLaunchKernelGGL(k_pr_sums<DH_REG_IDENTITY>, dim3(nb), dim3(kPdThreads), 0, st, a, b, d, w3_of(w), want_abs,
                       part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return finish2(part, nb, sc + PD_DL2, loss_out, st);
}

hipError_t grad_forward(const float* f, float* out, int nz, int ny, int nx, const float* w, hipStream_t st) {
  hipLaunchKernelGGL(k_grad_fwd, dim3(pd_blocks((int64_t)nz * ny * nx)), dim3(kPdThreads), 0, st, f, out,
                     Dims{nz, ny, nx}, w3_of(w));
  return hipGetLastError();
}

hipError_t grad_adjoint(const float* y, float* out, int nz, int ny, int nx, const float* w, hipStream_t st) {
  hipLaunchKernelGGL(k_grad_adj, dim3(pd_blocks((int64_t)nz * ny * nx)), dim3(kPdThreads), 0, st, y, out,
                     Dims{nz, ny, nx}, w3_of(w));
  return hipGetLastError();
}

int64_t pd_partials_needed(int nz, int ny, int nx) {
  const int64_t n = (int64_t)nz * ny * nx;
  return 3 * std::max<int64_t>(pd_blocks(n), tile_blocks(nz, ny, nx)) + 16;
}

static dim3 ew_blocks(int64_t n) {
  int64_t nb = (n + kPdThreads - 1) / kPdThreads;
  if (nb > 65535LL * 32) nb = 65535LL * 32;
  return dim3((unsigned)nb);
}

hipError_t hess_forward(const float* f, float* out, int nz, int ny, int nx, const float* w, hipStream_t st) {
  Dims d{nz, ny, nx};
  W6 ww;
  for (int q = 0; q < 6; ++q) ww.w[q] = w[q];
  hipLaunchKernelGGL(k_hess_fwd, ew_blocks((int64_t)nz * ny * nx), dim3(kPdThreads), 0, st, f, out, d, ww);
  return hipGetLastError();
}

hipError_t hess_adjoint(const float* y2, float* out, int nz, int ny, int nx, const float* w, hipStream_t st) {
  Dims d{nz, ny, nx};
  W6 ww;
  for (int q = 0; q < 6; ++q) ww.w[q] = w[q];
  hipLaunchKernelGGL(k_hess_adj_tiled, tile_grid(nz, ny, nx), dim3(kTileThreads), 0, st, y2, out, d, ww);
  return hipGetLastError();
}

}  // namespace dh
