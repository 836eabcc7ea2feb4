// Compile-time specialised FFT passes: one instantiation per transform length.
//
// For a length N the radix schedule, the span of every pass, the threads per
// line T = N / R_max and each thread's butterfly count are constants, so every
// register array is indexed statically and the kernel carries exactly the
// registers its passes need (the runtime-length kernels in dh_fft.hip keep
// every radix's code live at once).  Bodies are shared by the single-volume
// and batched entry points.
#pragma once
#include "dh_dft.hpp"
#include "dh_fft.hpp"

namespace dh {
namespace ct {

struct Sched {
  int np;
  int r[12];
};

// 16s first, then 8/4/2, then 5s and 3s (the same rule as the runtime plans).
constexpr Sched default_sched(int n) {
  Sched s{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = n, c2 = 0, c3 = 0, c5 = 0;
  while (m % 2 == 0) m /= 2, ++c2;
  while (m % 3 == 0) m /= 3, ++c3;
  while (m % 5 == 0) m /= 5, ++c5;
  int n16 = c2 / 4, rem = c2 % 4;
  if (rem == 1 && n16 > 0) {
    --n16;
    for (int i = 0; i < n16; ++i) s.r[s.np++] = 16;
    s.r[s.np++] = 8;
    s.r[s.np++] = 4;
  } else {
    for (int i = 0; i < n16; ++i) s.r[s.np++] = 16;
    if (rem == 3) s.r[s.np++] = 8;
    if (rem == 2) s.r[s.np++] = 4;
    if (rem == 1) s.r[s.np++] = 2;
  }
  for (int i = 0; i < c5; ++i) s.r[s.np++] = 5;
  for (int i = 0; i < c3; ++i) s.r[s.np++] = 3;
  return s;
}

constexpr Sched reversed(Sched s) {
  Sched o{s.np, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  for (int i = 0; i < s.np; ++i) o.r[i] = s.r[s.np - 1 - i];
  return o;
}

constexpr int span_of(Sched s, int p) {
  int v = 1;
  for (int i = 0; i < p; ++i) v *= s.r[i];
  return v;
}

constexpr int rmax(Sched s) {
  int v = 1;
  for (int i = 0; i < s.np; ++i) v = s.r[i] > v ? s.r[i] : v;
  return v;
}

// Measured exceptions to the rule (one MI355X; others measured no better).
// Columns (Y and Z passes): 1024 points as 16 x 8 x 8 (Y passes at 64x512x512 82.3 ->
// 79.5 us and 85.3 -> 82.9 us; 8 x 8 x 16 no gain, 8 x 8 x 8 x 2 slower); 64 as 8 x 8
// (Z pass of a 64x128x128 canvas 9.6 -> 6.8 us; 4 x 16 8.1 us); 128 stays 16 x 8
// (8 x 16 and 8 x 8 x 2 slower), 256 stays 16 x 16 (8 x 8 x 4 slower).
constexpr Sched make_sched(int n) {
  if (n == 1024) return Sched{3, {16, 8, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  if (n == 64) return Sched{2, {8, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  return default_sched(n);
}

// Rows (the X passes) up to 2048 points: 8s first, then 4 or 2, then 5s and 3s.  A
// row of 512 points as 8 x 8 x 8 runs 64 threads of 8 points instead of 32 of 16, so
// an X workgroup holds half the registers per thread and the CU more workgroups (X
// round trip at 64x512x512 103 -> 90 us, X forward 63 -> 52 us; 16 x 16 x 2 was slower,
// 147 us); measured the same way 128 points as 8 x 8 x 2 (X forward 14.2 -> 11.8 us at
// 64x256x128), 256 as 8 x 8 x 4 (27.7 -> 16.6 us) and 1024 as 8 x 8 x 8 x 2 (103 ->
// 67 us at 32x512x1024), 2048 unchanged.  A column or Z pass of 512 points keeps
// 16 x 8 x 4: as 8 x 8 x 8 the 256x2048x2048 Z pass took 32 ms instead of 17.
constexpr Sched make_row_sched(int n) {
  if (n > 2048) return default_sched(n);
  Sched s{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int m = n, c2 = 0, c3 = 0, c5 = 0;
  while (m % 2 == 0) m /= 2, ++c2;
  while (m % 3 == 0) m /= 3, ++c3;
  while (m % 5 == 0) m /= 5, ++c5;
  for (int i = 0; i < c2 / 3; ++i) s.r[s.np++] = 8;
  if (c2 % 3 == 2) s.r[s.np++] = 4;
  if (c2 % 3 == 1) s.r[s.np++] = 2;
  for (int i = 0; i < c5; ++i) s.r[s.np++] = 5;
  for (int i = 0; i < c3; ++i) s.r[s.np++] = 3;
  return s;
}

// ROW: the row (X) schedule, else the column (Y, Z) one.
template <int N, bool ROW = false>
struct Geo {
  static constexpr Sched F = ROW ? make_row_sched(N) : make_sched(N);
  static constexpr Sched I = reversed(F);
  static constexpr int T = N / rmax(F);  // threads per line
};

// ---- per-pass twiddle tables ----
// Pass p of a schedule (radix R, span P > 1) owns tw_pows(R) rows of P entries:
// row j, entry k = w(k)^(2^j), w(k) = exp(-2*pi*i*k / (P*R)).  A butterfly loads
// w, w^2, w^4, w^8 (as its radix needs) at index k, so lanes with consecutive k
// read consecutive entries (no LDS bank conflicts when the table sits in LDS),
// and forms the other powers with at most three complex products.
constexpr int tw_pows(int R) { return R == 16 ? 4 : (R == 8 || R == 5) ? 3 : (R == 4 || R == 3) ? 2 : 1; }

// Entries before pass p's rows (schedule order).
constexpr int tw_pass_offset(Sched s, int p) {
  int off = 0, span = 1;
  for (int q = 0; q < p; ++q) {
    if (span > 1) off += tw_pows(s.r[q]) * span;
    span *= s.r[q];
  }
  return off;
}
constexpr int tw_table_size(Sched s) { return tw_pass_offset(s, s.np); }

// u[r] *= w^r for r = 1..R-1 (w conjugated when DIR > 0) from pass rows tw, index k.
template <int R, int DIR, int P>
__device__ __forceinline__ void twiddle(float2 (&u)[R], const float2* __restrict__ tw, int k) {
  auto ld = [&](int j) {
    float2 w = tw[j * P + k];
    if (DIR > 0) w.y = -w.y;
    return w;
  };
  if constexpr (R == 2) {
    u[1] = cmul(u[1], ld(0));
  } else if constexpr (R == 3) {
    u[1] = cmul(u[1], ld(0));
    u[2] = cmul(u[2], ld(1));
  } else if constexpr (R == 4 || R == 5) {
    const float2 w1 = ld(0), w2 = ld(1);
    u[1] = cmul(u[1], w1);
    u[2] = cmul(u[2], w2);
    u[3] = cmul(u[3], cmul(w1, w2));
    if constexpr (R == 5) u[4] = cmul(u[4], ld(2));
  } else {
    static_assert(R == 8 || R == 16, "radix");
    const float2 w1 = ld(0), w2 = ld(1), w4 = ld(2);
    const float2 w3 = cmul(w1, w2), w5 = cmul(w1, w4), w6 = cmul(w2, w4), w7 = cmul(w3, w4);
    u[1] = cmul(u[1], w1);
    u[2] = cmul(u[2], w2);
    u[3] = cmul(u[3], w3);
    u[4] = cmul(u[4], w4);
    u[5] = cmul(u[5], w5);
    u[6] = cmul(u[6], w6);
    u[7] = cmul(u[7], w7);
    if constexpr (R == 16) {
      const float2 w8 = ld(3);
      u[8] = cmul(u[8], w8);
      u[9] = cmul(u[9], cmul(w1, w8));
      u[10] = cmul(u[10], cmul(w2, w8));
      u[11] = cmul(u[11], cmul(w3, w8));
      u[12] = cmul(u[12], cmul(w4, w8));
      u[13] = cmul(u[13], cmul(w5, w8));
      u[14] = cmul(u[14], cmul(w6, w8));
      u[15] = cmul(u[15], cmul(w7, w8));
    }
  }
}

// Butterflies per thread of a radix-R pass over N points with T threads.
template <int N, int T, int R>
struct Bfly {
  static constexpr int M = (N / R + T - 1) / T;
};

// Register operands of one Stockham pass: u[m][r] = element i + r*N/R, i = t + m*T.
template <int N, int T, int R, class Ld>
__device__ __forceinline__ void pass_load(float2 (&u)[(N / R + T - 1) / T][R], int t, Ld ld) {
  constexpr int NB = N / R;
  constexpr int M = Bfly<N, T, R>::M;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if ((M * T == NB) || i < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) u[m][r] = ld(i + r * NB);
    }
  }
}

// Twiddles, butterflies and the Stockham-ordered stores of one pass; tw = the
// pass's table rows (unused when P == 1).
template <int N, int T, int R, int P, int DIR, class St>
__device__ __forceinline__ void pass_store(float2 (&u)[(N / R + T - 1) / T][R], int t, const float2* __restrict__ tw,
                                           St st) {
  constexpr int NB = N / R;
  constexpr int M = Bfly<N, T, R>::M;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if ((M * T == NB) || i < NB) {
      // the last pass of a line (P == NB) has i < P: j = i, written so that a store's
      // index stays the thread's index plus a constant
      const int k = (P == NB) ? i : i % P;
      if constexpr (P > 1) twiddle<R, DIR, P>(u[m], tw, k);
      Dft<R, DIR>::run(u[m]);
      const int j = (P == NB) ? i : (i / P) * P * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) st(j + r * P, u[m][r]);
    }
  }
}

// One Stockham pass with everything compile-time; t < T.
template <int N, int T, int R, int P, int DIR, class Ld, class St>
__device__ __forceinline__ void pass(int t, const float2* __restrict__ tw, Ld ld, St st, bool sync_mid) {
  float2 u[Bfly<N, T, R>::M][R];
  pass_load<N, T, R>(u, t, ld);
  if (sync_mid) __syncthreads();
  pass_store<N, T, R, P, DIR>(u, t, tw, st);
}

// Passes [p, S.np) of schedule S; the first uses gld, the last gst, the rest
// LDS.  LAST_LDS: gst writes the same LDS the last pass reads, so that pass
// needs the barrier between its loads and stores.
// tw = the schedule's tables (pass rows in schedule order).
template <int N, bool INV, int DIR, int p, bool LAST_LDS = false, bool ROW = false, class GLd, class GSt, class LLd,
          class LSt>
__device__ __forceinline__ void passes(int t, const float2* __restrict__ tw, GLd gld, GSt gst, LLd lld, LSt lst,
                                       bool first_from_global) {
  constexpr Sched S = INV ? Geo<N, ROW>::I : Geo<N, ROW>::F;
  if constexpr (p < S.np) {
    constexpr int R = S.r[p];
    constexpr int P = span_of(S, p);
    constexpr int T = Geo<N, ROW>::T;
    constexpr bool last = (p == S.np - 1);
    const bool first = (p == 0) && first_from_global;
    constexpr int off = tw_pass_offset(S, p);
    const float2* twp = tw + off;
    if (first && last) pass<N, T, R, P, DIR>(t, twp, gld, gst, false);
    else if (first) pass<N, T, R, P, DIR>(t, twp, gld, lst, false);
    else if (last) pass<N, T, R, P, DIR>(t, twp, lld, gst, LAST_LDS);
    else pass<N, T, R, P, DIR>(t, twp, lld, lst, true);
    if (!last) __syncthreads();
    passes<N, INV, DIR, p + 1, LAST_LDS, ROW>(t, tw, gld, gst, lld, lst, first_from_global);
  }
}

// Full line transform: global -> (LDS ...) -> global.
template <int N, int DIR, class GLd, class GSt, class LLd, class LSt>
__device__ __forceinline__ void line(int t, const float2* __restrict__ tw, GLd gld, GSt gst, LLd lld, LSt lst) {
  if constexpr (N == 1) {
    if (t == 0) gst(0, gld(0));
  } else {
    passes<N, false, DIR, 0>(t, tw, gld, gst, lld, lst, true);
  }
}

}  // namespace ct
}  // namespace dh
