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

// hess6 on an accessor whose neighbours one step outside the volume already hold the
// replicated boundary values (an LDS window filled from clamped coordinates, as
// the tiled kernels below stage it): the same reads in the same order, so the
// same values, without the index clamps.
template <class F>
__device__ __forceinline__ void hess6_halo(const F& f, int z, int y, int x, const float* w, float* h) {
  const int zm = z - 1, zp = z + 1, ym = y - 1, yp = y + 1, xm = x - 1, xp = x + 1;
  const float c = f(z, y, x);
  h[0] = (f(zm, y, x) - 2.0f * c) + f(zp, y, x);
  h[1] = (f(z, ym, x) - 2.0f * c) + f(z, yp, x);
  h[2] = (f(z, y, xm) - 2.0f * c) + f(z, y, xp);
  const float dz_yp = (f(zp, yp, x) - f(zm, yp, x)) / 2.0f;
  const float dz_ym = (f(zp, ym, x) - f(zm, ym, x)) / 2.0f;
  h[3] = (dz_yp - dz_ym) / 2.0f;
  const float dz_xp = (f(zp, y, xp) - f(zm, y, xp)) / 2.0f;
  const float dz_xm = (f(zp, y, xm) - f(zm, y, xm)) / 2.0f;
  h[4] = (dz_xp - dz_xm) / 2.0f;
  const float dy_xp = (f(z, yp, xp) - f(z, ym, xp)) / 2.0f;
  const float dy_xm = (f(z, yp, xm) - f(z, ym, xm)) / 2.0f;
  h[5] = (dy_xp - dy_xm) / 2.0f;
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

// Hessian3D.adjoint at (z, y, x) on an accessor whose neighbours one step outside
// the volume are staged so that plain differences reproduce hess_adj's boundary
// rules: the d2 components (0, 1, 2) hold the replicated boundary value there, the
// mixed components (3: y and z, 4: x and z, 5: x and y) the boundary value negated
// once for each of their two axes crossed (d1_cen_adj's fpad = [-g0, g, -g_{n-1}]).
// Same values, same order as hess_adj, without the boundary branches.
template <class HW>
__device__ __forceinline__ float hess_adj_halo(const HW& hw, int z, int y, int x) {
  const float a0 = (hw(0, z - 1, y, x) - 2.0f * hw(0, z, y, x)) + hw(0, z + 1, y, x);
  const float a1 = (hw(1, z, y - 1, x) - 2.0f * hw(1, z, y, x)) + hw(1, z, y + 1, x);
  const float a2 = (hw(2, z, y, x - 1) - 2.0f * hw(2, z, y, x)) + hw(2, z, y, x + 1);
  auto dy3 = [&](int iz) { return (hw(3, iz, y - 1, x) - hw(3, iz, y + 1, x)) / 2.0f; };
  const float a3 = (dy3(z - 1) - dy3(z + 1)) / 2.0f;
  auto dx4 = [&](int iz) { return (hw(4, iz, y, x - 1) - hw(4, iz, y, x + 1)) / 2.0f; };
  const float a4 = (dx4(z - 1) - dx4(z + 1)) / 2.0f;
  auto dx5 = [&](int iy) { return (hw(5, z, iy, x - 1) - hw(5, z, iy, x + 1)) / 2.0f; };
  const float a5 = (dx5(y - 1) - dx5(y + 1)) / 2.0f;
  return ((((a0 + a1) + a2) + a3) + a4) + a5;
}

struct W6 {
  float w[6];
};

// ---------------------------------------------------------------------------
// LDS-tiled stencils: a workgroup owns a TY x TX column of the volume over ZC
// consecutive planes and slides a three-plane window (with a one-voxel halo,
// filled from clamped coordinates = the Neumann boundary) down z, so each
// voxel of x, x_prev and y2 is read from HBM about once.
// ---------------------------------------------------------------------------

constexpr int kTX = 64, kTY = 8, kZC = 16;
constexpr int kPW = kTX + 2, kPH = kTY + 2, kPlane = kPW * kPH;
constexpr int kTileThreads = kTX * kTY;

struct Tile {
  int x0, y0, z0, z1;  // z1 = one past the last plane of this workgroup
  int tx, ty, x, y;
  bool in;             // (y, x) inside the volume
};

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each XCD
// (and its L2) runs a contiguous x-then-y range of tiles, whose halos overlap.
__device__ __forceinline__ void tile_coords(int& bx, int& by, int& bz) {
  const int64_t nb = (int64_t)gridDim.x * gridDim.y * gridDim.z;
  const int64_t b = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const int64_t q = nb / 8, r = nb % 8, x = b % 8, i = b / 8;
  const int64_t l = x * q + (x < r ? x : r) + i;
  bx = (int)(l % gridDim.x);
  by = (int)((l / gridDim.x) % gridDim.y);
  bz = (int)(l / ((int64_t)gridDim.x * gridDim.y));
}

__device__ __forceinline__ Tile tile_of(const Dims& d, int zc = kZC) {
  Tile t;
  int bx, by, bz;
  tile_coords(bx, by, bz);
  t.x0 = bx * kTX;
  t.y0 = by * kTY;
  t.z0 = bz * zc;
  t.z1 = min(t.z0 + zc, d.nz);
  t.tx = threadIdx.x % kTX;
  t.ty = threadIdx.x / kTX;
  t.x = t.x0 + t.tx;
  t.y = t.y0 + t.ty;
  t.in = t.x < d.nx && t.y < d.ny;
  return t;
}

// Fills one halo'd plane: dst[ly * kPW + lx] = f(global index at clamped (z, y, x)).
template <class F>
__device__ __forceinline__ void fill_plane(float* dst, const Dims& d, const Tile& t, int z, F f) {
  const int zc = clampi(z, d.nz);
  for (int i = threadIdx.x; i < kPlane; i += kTileThreads) {
    const int ly = i / kPW, lx = i - ly * kPW;
    const int gy = clampi(t.y0 + ly - 1, d.ny), gx = clampi(t.x0 + lx - 1, d.nx);
    dst[i] = f(((int64_t)zc * d.ny + gy) * d.nx + gx);
  }
}

__device__ __forceinline__ int slot_of(int z, int z0) { return (z - z0 + 1) % 3; }

// Each thread stages at most two samples of every halo'd plane (kPlane <=
// 2 * kTileThreads).  Their in-plane offsets are fixed, so the next plane's
// loads are issued into registers before the current plane is computed and
// land in LDS after it (software pipelining of the z sweep).
struct PlanePos {
  int64_t off0, off1;
  bool has1;
  bool hx0, hy0, hx1, hy1;  // sample in the x (y) halo column (row)
  __device__ PlanePos(const Dims& d, const Tile& t) {
    const int i0 = threadIdx.x, i1 = threadIdx.x + kTileThreads;
    const int ly0 = i0 / kPW, lx0 = i0 - ly0 * kPW;
    off0 = (int64_t)clampi(t.y0 + ly0 - 1, d.ny) * d.nx + clampi(t.x0 + lx0 - 1, d.nx);
    has1 = i1 < kPlane;
    const int ly1 = has1 ? i1 / kPW : 0, lx1 = has1 ? i1 - ly1 * kPW : 0;
    off1 = (int64_t)clampi(t.y0 + ly1 - 1, d.ny) * d.nx + clampi(t.x0 + lx1 - 1, d.nx);
    hx0 = lx0 == 0 || lx0 == kPW - 1;
    hy0 = ly0 == 0 || ly0 == kPH - 1;
    hx1 = lx1 == 0 || lx1 == kPW - 1;
    hy1 = ly1 == 0 || ly1 == kPH - 1;
  }
  // load for a stencil that reads the x halo only if NX and the y halo only if NY:
  // samples in a halo it never reads are not fetched (their LDS slots stay unread)
  template <bool NX, bool NY>
  __device__ __forceinline__ void load_need(const float* __restrict__ src, int64_t b, float& v0, float& v1) const {
    const bool l0 = (NX || !hx0) && (NY || !hy0);
    const bool l1 = has1 && (NX || !hx1) && (NY || !hy1);
    v0 = l0 ? src[b + off0] : 0.0f;
    v1 = l1 ? src[b + off1] : 0.0f;
  }
  __device__ __forceinline__ int64_t base(const Dims& d, int z) const {
    return (int64_t)clampi(z, d.nz) * d.ny * d.nx;
  }
  __device__ __forceinline__ void load(const float* __restrict__ src, int64_t b, float& v0, float& v1) const {
    v0 = src[b + off0];
    v1 = has1 ? src[b + off1] : 0.0f;
  }
  __device__ __forceinline__ void store(float* dst, float v0, float v1) const {
    dst[threadIdx.x] = v0;
    if (has1) dst[threadIdx.x + kTileThreads] = v1;
  }
};

// Reflection signs of a thread's two staged samples (PlanePos order) for
// hess_adj_halo: -1 where the sample's y (x) lies outside the volume, else +1;
// zsign gives the same for a staged plane z.
struct HaloSign {
  float y0, x0, y1, x1;
  __device__ HaloSign(const Dims& d, const Tile& t) {
    auto sg = [](int v, int n) { return (v < 0 || v >= n) ? -1.0f : 1.0f; };
    const int i0 = threadIdx.x, i1 = threadIdx.x + kTileThreads;
    const int ly0 = i0 / kPW, lx0 = i0 - ly0 * kPW;
    const int ly1 = i1 / kPW, lx1 = i1 - ly1 * kPW;
    y0 = sg(t.y0 + ly0 - 1, d.ny);
    x0 = sg(t.x0 + lx0 - 1, d.nx);
    y1 = sg(t.y0 + ly1 - 1, d.ny);
    x1 = sg(t.x0 + lx1 - 1, d.nx);
  }
};
__device__ __forceinline__ float zsign(const Dims& d, int z) { return (z < 0 || z >= d.nz) ? -1.0f : 1.0f; }

__device__ __forceinline__ float lds_at(const float* planes, int slot, const Tile& t, int yy, int xx) {
  return planes[slot * kPlane + (yy - t.y0 + 1) * kPW + (xx - t.x0 + 1)];
}

static inline dim3 tile_grid(int nz, int ny, int nx, int zc = kZC) {
  return dim3((unsigned)((nx + kTX - 1) / kTX), (unsigned)((ny + kTY - 1) / kTY), (unsigned)((nz + zc - 1) / zc));
}

static inline int tile_blocks(int nz, int ny, int nx) {
  const dim3 g = tile_grid(nz, ny, nx);
  return (int)(g.x * g.y * g.z);
}

}  // namespace dh
