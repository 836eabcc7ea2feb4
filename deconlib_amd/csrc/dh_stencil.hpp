// Hessian3D stencils with Neumann (replicate) boundaries (linops_mlx.py:182-262,
// linops_core_mlx.py:42-80): device helpers shared by the PDHG and ER-Decon kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dh {

struct Dims {
  int nz, ny, nx;
};

__device__ __forceinline__ int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// Six weighted components [zz, yy, xx, yz, xz, xy] of Hessian3D.forward at
// (z, y, x) for f given by an accessor taking clamped coordinates.
template <class F>
__device__ __forceinline__ void hess6(const F& f, const Dims& d, int z, int y, int x, const float* w, float* h) {
  const int zm = clampi(z - 1, d.nz), zp = clampi(z + 1, d.nz);
  const int ym = clampi(y - 1, d.ny), yp = clampi(y + 1, d.ny);
  const int xm = clampi(x - 1, d.nx), xp = clampi(x + 1, d.nx);
  const float c = f(z, y, x);
  // d2: fpad[:-2] - 2 * fpad[1:-1] + fpad[2:]
  h[0] = (f(zm, y, x) - 2.0f * c) + f(zp, y, x);
  h[1] = (f(z, ym, x) - 2.0f * c) + f(z, yp, x);
  h[2] = (f(z, y, xm) - 2.0f * c) + f(z, y, xp);
  // d1_cen: (fpad[2:] - fpad[:-2]) / 2, applied twice
  const float dz_yp = (f(zp, yp, x) - f(zm, yp, x)) / 2.0f;
  const float dz_ym = (f(zp, ym, x) - f(zm, ym, x)) / 2.0f;
  h[3] = (dz_yp - dz_ym) / 2.0f;  // H_yz = d1_cen(Dz, y)
  const float dz_xp = (f(zp, y, xp) - f(zm, y, xp)) / 2.0f;
  const float dz_xm = (f(zp, y, xm) - f(zm, y, xm)) / 2.0f;
  h[4] = (dz_xp - dz_xm) / 2.0f;  // H_xz = d1_cen(Dz, x)
  const float dy_xp = (f(z, yp, xp) - f(z, ym, xp)) / 2.0f;
  const float dy_xm = (f(z, yp, xm) - f(z, ym, xm)) / 2.0f;
  h[5] = (dy_xp - dy_xm) / 2.0f;  // H_xy = d1_cen(Dy, x)
#pragma unroll
  for (int q = 0; q < 6; ++q) h[q] *= w[q];
}

// d1_cen_adj of g along one axis at index i (n samples):
//   fpad = [-g0, g..., -g_{n-1}];  out[i] = (fpad[i] - fpad[i + 2]) / 2
template <class G>
__device__ __forceinline__ float d1adj(const G& g, int i, int n) {
  const float l = (i == 0) ? -g(0) : g(i - 1);
  const float r = (i == n - 1) ? -g(n - 1) : g(i + 1);
  return (l - r) / 2.0f;
}

// d2 (self-adjoint) along one axis at index i with replicate padding.
template <class G>
__device__ __forceinline__ float d2at(const G& g, int i, int n) {
  return (g(clampi(i - 1, n)) - 2.0f * g(i)) + g(clampi(i + 1, n));
}

// Hessian3D.adjoint at (z, y, x): hw(q, z, y, x) returns H[q] * w[q].
template <class HW>
__device__ __forceinline__ float hess_adj(const HW& hw, const Dims& d, int z, int y, int x) {
  const float a0 = d2at([&](int i) { return hw(0, i, y, x); }, z, d.nz);
  const float a1 = d2at([&](int i) { return hw(1, z, i, x); }, y, d.ny);
  const float a2 = d2at([&](int i) { return hw(2, z, y, i); }, x, d.nx);
  // d1_cen_adj(d1_cen_adj(Hw3, y), z)
  const float a3 = d1adj([&](int iz) { return d1adj([&](int iy) { return hw(3, iz, iy, x); }, y, d.ny); }, z, d.nz);
  // d1_cen_adj(d1_cen_adj(Hw4, x), z)
  const float a4 = d1adj([&](int iz) { return d1adj([&](int ix) { return hw(4, iz, y, ix); }, x, d.nx); }, z, d.nz);
  // d1_cen_adj(d1_cen_adj(Hw5, x), y)
  const float a5 = d1adj([&](int iy) { return d1adj([&](int ix) { return hw(5, z, iy, ix); }, x, d.nx); }, y, d.ny);
  return ((((a0 + a1) + a2) + a3) + a4) + a5;
}

struct W6 {
  float w[6];
};

}  // namespace dh
