// Small in-register DFTs for the Stockham passes (gfx950, wave64).
//
// dft<R, DIR>(u) overwrites u[0..R) with y_q = sum_m u_m exp(DIR*2*pi*i*m*q/R),
// DIR = -1 forward, +1 inverse (unnormalised).  Radices 2, 3, 4, 5 are written
// out; 8 and 16 are two register Stockham passes over the leaves, with their
// twiddles folded to constants after unrolling.
#pragma once
#include <hip/hip_runtime.h>

namespace dh {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by DIR*i
template <int DIR>
__device__ __forceinline__ float2 mul_dir_i(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// cos / sin of 2*pi*m/16 (folded to constants once the callers unroll)
__device__ __forceinline__ constexpr float cos16(int m) {
  switch (m & 15) {
    case 0: return 1.0f;
    case 1: case 15: return 0.92387953251128674f;
    case 2: case 14: return 0.70710678118654752f;
    case 3: case 13: return 0.38268343236508977f;
    case 4: case 12: return 0.0f;
    case 5: case 11: return -0.38268343236508977f;
    case 6: case 10: return -0.70710678118654752f;
    case 7: case 9: return -0.92387953251128674f;
    default: return -1.0f;
  }
}
__device__ __forceinline__ constexpr float sin16(int m) { return cos16(m - 4); }

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<1, DIR> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2* u) {
    float2 a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
  }
};

template <int DIR>
struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(float2* u) {
    const float h = 0.86602540378443865f;  // sqrt(3)/2
    float2 t = cadd(u[1], u[2]);
    float2 m = make_float2(u[0].x - 0.5f * t.x, u[0].y - 0.5f * t.y);
    float2 d = csub(u[1], u[2]);
    float2 s = mul_dir_i<DIR>(cscale(d, h));
    u[0] = cadd(u[0], t);
    u[1] = cadd(m, s);
    u[2] = csub(m, s);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2* u) {
    float2 a = cadd(u[0], u[2]), b = csub(u[0], u[2]);
    float2 c = cadd(u[1], u[3]), d = mul_dir_i<DIR>(csub(u[1], u[3]));
    u[0] = cadd(a, c);
    u[2] = csub(a, c);
    u[1] = cadd(b, d);
    u[3] = csub(b, d);
  }
};

template <int DIR>
struct Dft<5, DIR> {
  __device__ __forceinline__ static void run(float2* u) {
    const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
    const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
    float2 a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
    float2 a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
    float2 u0 = u[0];
    float2 r1 = make_float2(u0.x + c1 * a1.x + c2 * a2.x, u0.y + c1 * a1.y + c2 * a2.y);
    float2 r2 = make_float2(u0.x + c2 * a1.x + c1 * a2.x, u0.y + c2 * a1.y + c1 * a2.y);
    float2 i1 = mul_dir_i<DIR>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    float2 i2 = mul_dir_i<DIR>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    u[0] = make_float2(u0.x + a1.x + a2.x, u0.y + a1.y + a2.y);
    u[1] = cadd(r1, i1);
    u[4] = csub(r1, i1);
    u[2] = cadd(r2, i2);
    u[3] = csub(r2, i2);
  }
};

// One register Stockham pass of radix R and span P over an N-point array.
template <int N, int R, int P, int DIR>
__device__ __forceinline__ void reg_pass(const float2* in, float2* out) {
  constexpr int NB = N / R;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int k = i % P;
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = in[i + r * NB];
    if (k != 0) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int m = (r * k * (N / (P * R))) * (16 / N);  // index into the 16-point table
        const float2 w = make_float2(cos16(m), DIR * sin16(m));
        u[r] = cmul(u[r], w);
      }
    }
    Dft<R, DIR>::run(u);
    const int j = (i / P) * P * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[j + r * P] = u[r];
  }
}

template <int DIR>
struct Dft<8, DIR> {
  __device__ __forceinline__ static void run(float2* u) {
    float2 t[8];
    reg_pass<8, 2, 1, DIR>(u, t);
    reg_pass<8, 4, 2, DIR>(t, u);
  }
};

template <int DIR>
struct Dft<16, DIR> {
  __device__ __forceinline__ static void run(float2* u) {
    float2 t[16];
    reg_pass<16, 4, 1, DIR>(u, t);
    reg_pass<16, 4, 4, DIR>(t, u);
  }
};

}  // namespace dh
