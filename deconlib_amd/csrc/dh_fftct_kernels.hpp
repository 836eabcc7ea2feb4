// Kernels and host launchers for the compile-time specialised passes.
//
// Each launch covers up to kPack volumes of the same transform length (one
// "shape class"), their descriptors passed by value in the kernel arguments,
// so a single volume needs no descriptor upload and a batch of volumes of
// one length is a single launch per pass.
#pragma once
#include "dh_fftct.hpp"

namespace dh {
namespace ct {

// Every specialised launch uses at most this many threads per workgroup.
constexpr int kMaxThreads = 512;

constexpr int kNumXcd = 8;
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb) {
  const int64_t q = nb / kNumXcd, r = nb % kNumXcd;
  const int64_t x = b % kNumXcd, i = b / kNumXcd;
  return x * q + (x < r ? x : r) + i;
}

template <class A>
__device__ __forceinline__ int pick(const Pack<A>& pk, int64_t b) {
  int v = 0;
  for (int i = 1; i < pk.nvol; ++i)
    if (pk.prefix[i] <= b) v = i;
  return v;
}

__device__ __forceinline__ float pro_apply(const Pro& p, const float* src, int64_t idx) {
  switch (p.kind) {
    case PRO_ONES: return 1.0f;
    case PRO_SUB_FLOOR: return fmaxf(src[idx] - p.a, p.b);
    case PRO_DIV_DEV: return src[idx] / (float)(*p.dptr);
    default: return src[idx];
  }
}

__device__ __forceinline__ float idiv(float o, float mdl) {
  const float os = fmaxf(o, 1e-6f), ms = fmaxf(mdl, 1e-6f);
  return o * logf(os / ms) - (o - ms);
}

__device__ __forceinline__ float epi_apply(const Epi& e, float v, int64_t idx, double& acc) {
  switch (e.kind) {
    case EPI_FLOOR: return fmaxf(v, e.a);
    case EPI_RL_RATIO: {
      const float mdl = v + e.a;
      const float o = e.aux0[idx];
      if (e.partials) acc += (double)idiv(o, mdl);
      return o / fmaxf(mdl, e.b);
    }
    case EPI_RL_UPDATE: return fmaxf(e.aux0[idx] * (v / e.aux1[idx]), e.b);
    case EPI_ADD: {
      const float mdl = v + e.a;
      if (e.partials && e.aux0) acc += (double)idiv(e.aux0[idx], mdl);
      return mdl;
    }
    default: return v;
  }
}

// epi_apply on the sample pair (idx, idx + 1) with float2 operand loads;
// the pair is accumulated in the same order as two epi_apply calls.
__device__ __forceinline__ float2 epi_apply2(const Epi& e, float2 v, int64_t idx, double& acc) {
  switch (e.kind) {
    case EPI_FLOOR: return make_float2(fmaxf(v.x, e.a), fmaxf(v.y, e.a));
    case EPI_RL_RATIO: {
      const float2 o = *reinterpret_cast<const float2*>(e.aux0 + idx);
      const float m0 = v.x + e.a, m1 = v.y + e.a;
      if (e.partials) {
        acc += (double)idiv(o.x, m0);
        acc += (double)idiv(o.y, m1);
      }
      return make_float2(o.x / fmaxf(m0, e.b), o.y / fmaxf(m1, e.b));
    }
    case EPI_RL_UPDATE: {
      const float2 xv = *reinterpret_cast<const float2*>(e.aux0 + idx);
      const float2 sv = *reinterpret_cast<const float2*>(e.aux1 + idx);
      return make_float2(fmaxf(xv.x * (v.x / sv.x), e.b), fmaxf(xv.y * (v.y / sv.y), e.b));
    }
    case EPI_ADD: {
      const float m0 = v.x + e.a, m1 = v.y + e.a;
      if (e.partials && e.aux0) {
        const float2 o = *reinterpret_cast<const float2*>(e.aux0 + idx);
        acc += (double)idiv(o.x, m0);
        acc += (double)idiv(o.y, m1);
      }
      return make_float2(m0, m1);
    }
    default: return v;
  }
}

__device__ __forceinline__ double bsum(double v, double* scratch) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r += scratch[i];
  }
  return r;
}

// ---- column pass (Y forward / inverse, OTF Z) ----
template <int N, int DIR>
__device__ __forceinline__ void col_body(const ColArgs& a, int64_t lb, int L, float2* lds) {
  constexpr int T = Geo<N>::T;
  const int ln = threadIdx.x % L;
  const int t = threadIdx.x / L;
  const int64_t tile = lb % a.ntiles, batch = lb / a.ntiles;
  const int64_t col = tile * L + ln;
  const bool ok = col < a.ncols;
  const float2* sb = a.src + batch * a.src_b + col;
  float2* db = a.dst + batch * a.dst_b + col;
  const int lo = a.lo, hi = a.hi, shift = N - a.lo - a.hi;
  auto gld = [&](int e) -> float2 {
    int r = -1;
    if (e < lo) r = e;
    else if (e >= N - hi) r = e - shift;
    if (!ok || r < 0) return make_float2(0.f, 0.f);
    return sb[(int64_t)r * a.src_rs];
  };
  auto gst = [&](int e, float2 v) {
    if (ok && e < a.n_out) db[(int64_t)e * a.dst_rs] = make_float2(v.x * a.scale, v.y * a.scale);
  };
  auto lld = [&](int e) { return lds[e * L + ln]; };
  auto lst = [&](int e, float2 v) { lds[e * L + ln] = v; };
  (void)T;
  line<N, DIR>(t, a.tw, gld, gst, lld, lst);
}

template <int N, int DIR>
__global__ void __launch_bounds__(kMaxThreads) k_col(Pack<ColArgs> pk, int L) {
  extern __shared__ float2 lds[];
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int v = pick(pk, b);
  col_body<N, DIR>(pk.a[v], b - pk.prefix[v], L, lds);
}

// ---- fused Z pass: forward FFT, x OTF, inverse FFT, in place ----
template <int N, int R, int P>
__device__ __forceinline__ void z_mid(int t, const float2* __restrict__ tw, float2* lds, int L, int ln,
                                      const float2* cb, int64_t rs, int lo, bool ok, const float2* __restrict__ otf,
                                      int64_t col, int conj, float2* gdst, int n_out) {
  constexpr int T = Geo<N>::T;
  constexpr int NB = N / R;
  constexpr int M = (NB + T - 1) / T;
  constexpr bool from_lds = Geo<N>::F.np > 1;
  float2 u[M][R];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if ((M * T == NB) || i < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = i + r * NB;
        if (from_lds) u[m][r] = lds[e * L + ln];
        else u[m][r] = (ok && e < lo) ? cb[(int64_t)e * rs] : make_float2(0.f, 0.f);
      }
    }
  }
  if (from_lds) __syncthreads();
  constexpr int step = N / (P * R);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if ((M * T == NB) || i < NB) {
      const int k = i % P;
      if (P > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) u[m][r] = cmul(u[m][r], tw[r * k * step]);
      }
      Dft<R, -1>::run(u[m]);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float2 o = ok ? otf[(int64_t)(i + r * NB) * rs + col] : make_float2(0.f, 0.f);
        u[m][r] = conj ? cmulc(u[m][r], o) : cmul(u[m][r], o);
      }
      Dft<R, 1>::run(u[m]);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = i * R + r;
        if (from_lds) lds[e * L + ln] = u[m][r];
        else if (ok && e < n_out) gdst[(int64_t)e * rs] = u[m][r];
      }
    }
  }
}

template <int N, int p, class GLd, class LLd, class LSt>
__device__ __forceinline__ void z_fwd_passes(int t, const float2* __restrict__ tw, GLd gld, LLd lld, LSt lst) {
  constexpr Sched S = Geo<N>::F;
  if constexpr (p < S.np - 1) {
    constexpr int R = S.r[p];
    constexpr int P = span_of(S, p);
    if (p == 0) pass<N, Geo<N>::T, R, P, -1>(t, tw, gld, lst, false);
    else pass<N, Geo<N>::T, R, P, -1>(t, tw, lld, lst, true);
    __syncthreads();
    z_fwd_passes<N, p + 1>(t, tw, gld, lld, lst);
  }
}

template <int N>
__device__ __forceinline__ void z_body(const ZArgs& a, int64_t lb, int L, float2* lds) {
  const int ln = threadIdx.x % L;
  const int t = threadIdx.x / L;
  const int64_t col = lb * L + ln;
  const bool ok = col < a.ncols;
  float2* cb = a.buf + col;
  const int64_t rs = a.ncols;
  const int lo = a.lo, n_out = a.n_out;
  if constexpr (N == 1) {
    if (t == 0 && ok && lo > 0) {
      const float2 v = cb[0], o = a.otf[col];
      cb[0] = a.conj ? cmulc(v, o) : cmul(v, o);
    }
  } else {
    auto gld = [&](int e) -> float2 { return (ok && e < lo) ? cb[(int64_t)e * rs] : make_float2(0.f, 0.f); };
    auto gst = [&](int e, float2 v) {
      if (ok && e < n_out) cb[(int64_t)e * rs] = v;
    };
    auto lld = [&](int e) { return lds[e * L + ln]; };
    auto lst = [&](int e, float2 v) { lds[e * L + ln] = v; };
    constexpr Sched F = Geo<N>::F;
    z_fwd_passes<N, 0>(t, a.tw, gld, lld, lst);
    z_mid<N, F.r[F.np - 1], span_of(F, F.np - 1)>(t, a.tw, lds, L, ln, cb, rs, lo, ok, a.otf, col, a.conj, cb, n_out);
    if constexpr (F.np > 1) {
      __syncthreads();
      passes<N, true, 1, 1>(t, a.tw, gld, gst, lld, lst, false);
    }
  }
}

template <int N>
__global__ void __launch_bounds__(kMaxThreads) k_z(Pack<ZArgs> pk, int L) {
  extern __shared__ float2 lds[];
  const int64_t b = xcd_remap(blockIdx.x, gridDim.x);
  const int v = pick(pk, b);
  z_body<N>(pk.a[v], b - pk.prefix[v], L, lds);
}

// ---- X forward: real rows -> half spectrum (M-point complex FFT) ----
template <int M>
__device__ __forceinline__ void xf_body(const XfArgs& a, int64_t lb, int RPB, float2* lds) {
  constexpr int T = Geo<M>::T;
  const int rl = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int64_t row = lb * RPB + rl;
  const bool ok = row < a.nrows;
  constexpr int P2 = 2 * M;
  const int lo = a.lo, hi = a.hi, shift = P2 - a.lo - a.hi;
  const float* sr = a.src + row * a.rowlen;
  float2* lrow = lds + rl * (M + 1);
  auto sample = [&](int q) -> float {
    int c = -1;
    if (q < lo) c = q;
    else if (q >= P2 - hi) c = q - shift;
    if (!ok || c < 0) return 0.0f;
    return pro_apply(a.pro, sr, c);
  };
  const bool vec2 = a.vec2 != 0;
  auto gld = [&](int e) -> float2 {
    if (vec2 && ok && 2 * e + 1 < lo) {
      const float2 v = *reinterpret_cast<const float2*>(sr + 2 * e);
      switch (a.pro.kind) {
        case PRO_ONES: return make_float2(1.0f, 1.0f);
        case PRO_SUB_FLOOR: return make_float2(fmaxf(v.x - a.pro.a, a.pro.b), fmaxf(v.y - a.pro.a, a.pro.b));
        case PRO_DIV_DEV: {
          const float s = (float)(*a.pro.dptr);
          return make_float2(v.x / s, v.y / s);
        }
        default: return v;
      }
    }
    return make_float2(sample(2 * e), sample(2 * e + 1));
  };
  auto lld = [&](int e) { return lrow[e]; };
  auto lst = [&](int e, float2 v) { lrow[e] = v; };
  // all passes end in LDS (natural order), then the real-to-half-spectrum step
  if constexpr (M == 1) {
    if (t == 0) lrow[0] = gld(0);
  } else {
    passes<M, false, -1, 0, true>(t, a.twm, gld, lst, lld, lst, true);
  }
  __syncthreads();
  if (ok) {
    float2* dr = a.dst + row * a.H;
    for (int k = t; k <= M; k += T) {
      const float2 zk = lrow[k == M ? 0 : k];
      const float2 zm = lrow[k == 0 ? 0 : M - k];
      const float2 zmc = make_float2(zm.x, -zm.y);
      const float2 E = make_float2(0.5f * (zk.x + zmc.x), 0.5f * (zk.y + zmc.y));
      const float2 D = make_float2(zk.x - zmc.x, zk.y - zmc.y);
      const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);  // -i/2 * D
      dr[k] = cadd(E, cmul(a.twp[k], O));
    }
  }
}

template <int M>
__global__ void __launch_bounds__(kMaxThreads) k_xf(Pack<XfArgs> pk, int RPB) {
  extern __shared__ float2 lds[];
  const int64_t b = blockIdx.x;
  const int v = pick(pk, b);
  xf_body<M>(pk.a[v], b - pk.prefix[v], RPB, lds);
}

// ---- X inverse: half spectrum -> cropped real rows with the fused epilogue ----
template <int M>
__device__ __forceinline__ void xi_body(const XiArgs& a, int64_t lb, int RPB, float2* lds) {
  constexpr int T = Geo<M>::T;
  const int rl = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int64_t row = lb * RPB + rl;
  const bool ok = row < a.nrows;
  const int H = a.H, s2 = a.s2;
  float2* lrow = lds + rl * (M + 1);
  if (ok) {
    const float2* sr = a.src + row * H;
    for (int k = t; k < H; k += T) lrow[k] = sr[k];
  } else {
    for (int k = t; k < H; k += T) lrow[k] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* twp = a.twp;
  auto pre = [&](int q) -> float2 {
    float2 xq = lrow[q];
    float2 xm = lrow[M - q];
    if (q == 0) {
      xq.y = 0.f;
      xm.y = 0.f;
    }
    const float2 xmc = make_float2(xm.x, -xm.y);
    const float2 A = cadd(xq, xmc);
    const float2 B = csub(xq, xmc);
    const float2 w = make_float2(twp[q].x, -twp[q].y);
    const float2 wb = cmul(w, B);
    return make_float2(A.x - wb.y, A.y + wb.x);
  };
  double acc = 0.0;
  float* drow = a.dst + row * s2;
  const int64_t base = row * s2;
  const bool vec2 = a.vec2 != 0;
  auto gst = [&](int e, float2 v) {
    if (!ok) return;
    const int q0 = 2 * e, q1 = 2 * e + 1;
    if (vec2 && q1 < s2) {
      *reinterpret_cast<float2*>(drow + q0) = epi_apply2(a.epi, v, base + q0, acc);
      return;
    }
    if (q0 < s2) drow[q0] = epi_apply(a.epi, v.x, base + q0, acc);
    if (q1 < s2) drow[q1] = epi_apply(a.epi, v.y, base + q1, acc);
  };
  auto lld = [&](int e) { return lrow[e]; };
  auto lst = [&](int e, float2 v) { lrow[e] = v; };
  if constexpr (M == 1) {
    float2 v = make_float2(0.f, 0.f);
    if (t == 0) v = pre(0);
    __syncthreads();
    if (t == 0) gst(0, v);
  } else {
    // the first pass reads the LDS row and overwrites it: full barrier between
    constexpr Sched S = Geo<M>::F;
    constexpr int R0 = S.r[0];
    if constexpr (S.np == 1) {
      pass<M, T, R0, 1, 1>(t, a.twm, pre, gst, true);
    } else {
      pass<M, T, R0, 1, 1>(t, a.twm, pre, lst, true);
      __syncthreads();
      passes<M, false, 1, 1>(t, a.twm, pre, gst, lld, lst, false);
    }
  }
  if (a.epi.partials) {
    __syncthreads();
    const double tot = bsum(acc, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) a.epi.partials[lb] = tot;
  }
}

template <int M>
__global__ void __launch_bounds__(kMaxThreads) k_xi(Pack<XiArgs> pk, int RPB) {
  extern __shared__ float2 lds[];
  const int64_t b = blockIdx.x;
  const int v = pick(pk, b);
  xi_body<M>(pk.a[v], b - pk.prefix[v], RPB, lds);
}

// ---- host launchers (explicitly instantiated per length) ----
template <int N>
struct Launch {
  static constexpr int T = Geo<N>::T;
  static hipError_t col(int dir, const Pack<ColArgs>& pk, int L, int64_t nblocks, hipStream_t st) {
    const size_t shm = (size_t)N * L * sizeof(float2);
    if (dir < 0) {
      (void)hipFuncSetAttribute((const void*)k_col<N, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      hipLaunchKernelGGL((k_col<N, -1>), dim3((unsigned)nblocks), dim3(L * T), shm, st, pk, L);
    } else {
      (void)hipFuncSetAttribute((const void*)k_col<N, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      hipLaunchKernelGGL((k_col<N, 1>), dim3((unsigned)nblocks), dim3(L * T), shm, st, pk, L);
    }
    return hipGetLastError();
  }
  static hipError_t z(const Pack<ZArgs>& pk, int L, int64_t nblocks, hipStream_t st) {
    const size_t shm = (size_t)N * L * sizeof(float2);
    (void)hipFuncSetAttribute((const void*)k_z<N>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_z<N>), dim3((unsigned)nblocks), dim3(L * T), shm, st, pk, L);
    return hipGetLastError();
  }
  static hipError_t xf(const Pack<XfArgs>& pk, int R, int64_t nblocks, hipStream_t st) {
    const size_t shm = (size_t)R * (N + 1) * sizeof(float2);
    (void)hipFuncSetAttribute((const void*)k_xf<N>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_xf<N>), dim3((unsigned)nblocks), dim3(R * T), shm, st, pk, R);
    return hipGetLastError();
  }
  static hipError_t xi(const Pack<XiArgs>& pk, int R, int64_t nblocks, hipStream_t st) {
    size_t shm = (size_t)R * (N + 1) * sizeof(float2);
    if (shm < 16 * sizeof(double)) shm = 16 * sizeof(double);
    (void)hipFuncSetAttribute((const void*)k_xi<N>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_xi<N>), dim3((unsigned)nblocks), dim3(R * T), shm, st, pk, R);
    return hipGetLastError();
  }
};

}  // namespace ct
}  // namespace dh
