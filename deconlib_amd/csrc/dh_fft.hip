// Pruned, fused 3-D FFT convolution passes for gfx950.
//
// One LinearFFTConvolver application (linops_mlx.py:520-530) is five kernels:
//   X  r2c  : rows of the signal volume S (zero padding never materialised)
//   Y  fwd  : columns over (s0, ., H), input rows [0, s1) only
//   Z  conv : forward FFT over z, x OTF (or conj OTF), inverse FFT, rows [0, s0) kept
//   Y  inv  : columns, output rows [0, s1) only
//   X  c2r  : rows, output samples [0, s2) only, with the solver's fused epilogue
//
// Every pass is a mixed-radix Stockham FFT (radices 16, 8, 5, 4, 3, 2) whose
// first pass reads global memory and whose last pass writes global memory
// directly (both coalesced: lanes walk consecutive columns or consecutive row
// samples); the intermediate passes exchange through LDS.
//
// These runtime-length kernels cover canvases without a compile-time
// specialised length (dh_fftct*.hpp); each pass is a __device__ body over one
// tile (a group of columns or rows of one volume).
#include "dh_dft.hpp"
#include "dh_fft.hpp"

namespace dh {

template <int R>
struct MBv {
  static constexpr int v = (kEPT + R - 1) / R + 1;
};

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Map the hardware block id to a logical one so that XCD x runs a
// contiguous range of logical tiles: a bijection on [0, nb).
constexpr int kNumXcd = 8;
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb) {
  const int64_t q = nb / kNumXcd, r = nb % kNumXcd;
  const int64_t x = b % kNumXcd, i = b / kNumXcd;
  return x * q + (x < r ? x : r) + i;
}

// One Stockham pass of radix R and span P over an n-point line.
//   ld(e)    -> input element e
//   st(e, v) <- output element e
// If sync_mid, all loads complete (block-wide) before any store.
template <int R, int DIR, class Ld, class St>
__device__ __forceinline__ void stockham_pass(int n, int P, int t, int T, const float2* __restrict__ tw,
                                              Ld ld, St st, bool sync_mid) {
  constexpr int M = MBv<R>::v;
  const int nb = n / R;
  float2 u[M][R];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if (i < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) u[m][r] = ld(i + r * nb);
    }
  }
  if (sync_mid) __syncthreads();
  const int step = n / (P * R);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if (i < nb) {
      const int k = i % P;
      if (P > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * step];
          if (DIR > 0) w.y = -w.y;
          u[m][r] = cmul(u[m][r], w);
        }
      }
      Dft<R, DIR>::run(u[m]);
      const int j = (i / P) * P * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) st(j + r * P, u[m][r]);
    }
  }
}

template <int DIR, class Ld, class St>
__device__ __forceinline__ void pass_any(int R, int n, int P, int t, int T, const float2* __restrict__ tw,
                                         Ld ld, St st, bool sync_mid) {
  switch (R) {
    case 16: stockham_pass<16, DIR>(n, P, t, T, tw, ld, st, sync_mid); break;
    case 8: stockham_pass<8, DIR>(n, P, t, T, tw, ld, st, sync_mid); break;
    case 5: stockham_pass<5, DIR>(n, P, t, T, tw, ld, st, sync_mid); break;
    case 4: stockham_pass<4, DIR>(n, P, t, T, tw, ld, st, sync_mid); break;
    case 3: stockham_pass<3, DIR>(n, P, t, T, tw, ld, st, sync_mid); break;
    default: stockham_pass<2, DIR>(n, P, t, T, tw, ld, st, sync_mid); break;
  }
}

// Runs all passes of schedule s: first pass loads with gld, last pass stores
// with gst, the rest go through the LDS line accessor.  Threads with t >= n
// do no butterflies but take every barrier.
template <int DIR, class GLd, class GSt, class LIdx>
__device__ __forceinline__ void fft_line(const AxisSched& s, int t, int T, const float2* __restrict__ tw,
                                         float2* lds, LIdx lidx, GLd gld, GSt gst) {
  const int n = s.n;
  if (s.np == 0) {  // 1-point transform: copy
    if (t == 0) gst(0, gld(0));
    return;
  }
  auto lld = [&](int e) { return lds[lidx(e)]; };
  auto lst = [&](int e, float2 v) { lds[lidx(e)] = v; };
  for (int p = 0; p < s.np; ++p) {
    const bool first = (p == 0), last = (p == s.np - 1);
    const int R = s.radix[p], P = s.span[p];
    if (first && last) {
      pass_any<DIR>(R, n, P, t, T, tw, gld, gst, false);
    } else if (first) {
      pass_any<DIR>(R, n, P, t, T, tw, gld, lst, false);
      __syncthreads();
    } else if (last) {
      pass_any<DIR>(R, n, P, t, T, tw, lld, gst, false);
    } else {
      pass_any<DIR>(R, n, P, t, T, tw, lld, lst, true);
      __syncthreads();
    }
  }
}

__device__ __forceinline__ int threads_per_line_d(int n) { return (n + kEPT - 1) / kEPT; }

// ---------------------------------------------------------------------------
// Column FFT body (Y passes, OTF Z pass): L columns of one batch plane
// ---------------------------------------------------------------------------

template <int DIR>
__device__ __forceinline__ void col_fft_body(const AxisSched& s, const float2* __restrict__ tw,
                                             const float2* __restrict__ src, float2* __restrict__ dst,
                                             int64_t ncols, int64_t src_rs, int64_t dst_rs, int lo, int hi,
                                             int n_out, float scale, int64_t tile, int L, float2* lds) {
  const int n = s.n;
  const int T = threads_per_line_d(n);
  const int ln = threadIdx.x % L;
  int t = threadIdx.x / L;
  if (t >= T) t = n;  // idle thread
  const int64_t col = tile * L + ln;
  const bool ok = col < ncols;
  const float2* sb = src + col;
  float2* db = dst + col;
  const int shift = n - lo - hi;
  auto gld = [&](int e) -> float2 {
    int r = -1;
    if (e < lo) r = e;
    else if (e >= n - hi) r = e - shift;
    if (!ok || r < 0) return make_float2(0.f, 0.f);
    return sb[(int64_t)r * src_rs];
  };
  auto gst = [&](int e, float2 v) {
    if (ok && e < n_out) db[(int64_t)e * dst_rs] = make_float2(v.x * scale, v.y * scale);
  };
  auto lidx = [&](int e) { return e * L + ln; };
  fft_line<DIR>(s, t, T, tw, lds, lidx, gld, gst);
}

// ---------------------------------------------------------------------------
// Fused Z body: forward FFT -> x OTF -> inverse FFT, in place, L columns.
// ---------------------------------------------------------------------------

template <int R>
__device__ __forceinline__ void z_mid_pass(int n, int P, int t, int T, const float2* __restrict__ tw,
                                           float2* lds, int L, int ln, bool from_lds, bool to_lds,
                                           const float2* gsrc, int64_t rs, int lo, bool ok,
                                           const float2* __restrict__ otf, int64_t ocol, int64_t ors,
                                           int conj, float2* gdst, int n_out) {
  constexpr int M = MBv<R>::v;
  const int nb = n / R;
  float2 u[M][R];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if (i < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = i + r * nb;
        if (from_lds) u[m][r] = lds[e * L + ln];
        else u[m][r] = (ok && e < lo) ? gsrc[(int64_t)e * rs] : make_float2(0.f, 0.f);
      }
    }
  }
  if (from_lds && to_lds) __syncthreads();
  const int step = n / (P * R);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if (i < nb) {
      const int k = i % P;
      if (P > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) u[m][r] = cmul(u[m][r], tw[r * k * step]);
      }
      Dft<R, -1>::run(u[m]);
      // forward output element e = i + r*nb (i < P here: last forward pass)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 o = ok ? otf[(int64_t)(i + r * nb) * ors + ocol] : make_float2(0.f, 0.f);
        u[m][r] = conj ? cmulc(u[m][r], o) : cmul(u[m][r], o);
      }
      // first inverse pass: same radix, span 1, no twiddle
      Dft<R, 1>::run(u[m]);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = i * R + r;
        if (to_lds) lds[e * L + ln] = u[m][r];
        else if (ok && e < n_out) gdst[(int64_t)e * rs] = u[m][r];
      }
    }
  }
}

__device__ __forceinline__ void z_conv_body(const AxisSched& f, const AxisSched& iv, const float2* __restrict__ tw,
                                            float2* buf, const float2* __restrict__ otf, int64_t ncols, int lo,
                                            int n_out, int conj, int64_t tile, int L, float2* lds) {
  const int n = f.n;
  const int T = threads_per_line_d(n);
  const int ln = threadIdx.x % L;
  int t = threadIdx.x / L;
  if (t >= T) t = n;  // idle thread
  const int64_t col = tile * L + ln;
  const bool ok = col < ncols;
  float2* cb = buf + col;
  const int64_t rs = ncols;
  auto lidx = [&](int e) { return e * L + ln; };
  auto lld = [&](int e) { return lds[lidx(e)]; };
  auto lst = [&](int e, float2 v) { lds[lidx(e)] = v; };
  auto gld = [&](int e) -> float2 { return (ok && e < lo) ? cb[(int64_t)e * rs] : make_float2(0.f, 0.f); };
  auto gst = [&](int e, float2 v) {
    if (ok && e < n_out) cb[(int64_t)e * rs] = v;
  };
  if (n == 1) {  // the Z "FFT" is the identity: multiply by the OTF only
    if (t == 0 && ok && lo > 0) {
      const float2 v = cb[0], o = otf[col];
      cb[0] = conj ? cmulc(v, o) : cmul(v, o);
    }
    return;
  }
  const int np = f.np;
  for (int p = 0; p < np - 1; ++p) {  // forward passes except the last
    const int R = f.radix[p], P = f.span[p];
    if (p == 0) pass_any<-1>(R, n, P, t, T, tw, gld, lst, false);
    else pass_any<-1>(R, n, P, t, T, tw, lld, lst, true);
    __syncthreads();
  }
  {  // merged: last forward pass, x OTF, first inverse pass
    const int R = f.radix[np - 1], P = f.span[np - 1];
    const bool from_lds = np > 1, to_lds = np > 1;
    switch (R) {
#define DH_ZMID(RR)                                                                              \
  case RR:                                                                                       \
    z_mid_pass<RR>(n, P, t, T, tw, lds, L, ln, from_lds, to_lds, cb, rs, lo, ok, otf, col, ncols, \
                   conj, cb, n_out);                                                             \
    break;
      DH_ZMID(16) DH_ZMID(8) DH_ZMID(5) DH_ZMID(4) DH_ZMID(3) DH_ZMID(2)
#undef DH_ZMID
    }
    if (to_lds) __syncthreads();
  }
  for (int p = 1; p < iv.np; ++p) {  // remaining inverse passes
    const int R = iv.radix[p], P = iv.span[p];
    if (p == iv.np - 1) pass_any<1>(R, n, P, t, T, tw, lld, gst, false);
    else {
      pass_any<1>(R, n, P, t, T, tw, lld, lst, true);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// X bodies (rows): real <-> half spectrum through an m = P2/2 complex FFT.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float apply_pro(const Pro& p, const float* src, int64_t idx) {
  switch (p.kind) {
    case PRO_ONES: return 1.0f;
    case PRO_SUB_FLOOR: return fmaxf(src[idx] - p.a, p.b);
    case PRO_DIV_DEV: return src[idx] / (float)(*p.dptr);
    default: return src[idx];
  }
}

// rows [row0, row0 + RPB) of the volume; tid -> (row, t)
__device__ __forceinline__ void x_r2c_body(const AxisSched& s, const float2* __restrict__ twm,
                                           const float2* __restrict__ twp, const float* __restrict__ src,
                                           float2* __restrict__ dst, int64_t nrows, int rowlen, int lo, int hi,
                                           int H, const Pro& pro, int64_t row0, int RPB, float2* lds) {
  const int m = s.n;
  const int T = threads_per_line_d(m);
  int rl = threadIdx.x / T;
  int t = threadIdx.x % T;
  const bool active = rl < RPB;
  if (!active) {
    rl = 0;
    t = m;
  }
  const int64_t row = row0 + rl;
  const bool ok = active && row < nrows;
  const int P2 = 2 * m;
  const int shift = P2 - lo - hi;
  const float* sr = pro_row(pro, src, ok ? row : 0, rowlen);
  float2* lrow = lds + rl * (m + 1);
  auto sample = [&](int q) -> float {
    int c = -1;
    if (q < lo) c = q;
    else if (q >= P2 - hi) c = q - shift;
    if (!ok || c < 0) return 0.0f;
    return apply_pro(pro, sr, pro_col(pro, c));
  };
  auto gld = [&](int e) -> float2 { return make_float2(sample(2 * e), sample(2 * e + 1)); };
  auto lld = [&](int e) { return lrow[e]; };
  auto lst = [&](int e, float2 v) { lrow[e] = v; };
  if (s.np == 0) {
    if (active)
      for (int e = t; e < m; e += T) lrow[e] = gld(e);
  } else {
    for (int p = 0; p < s.np; ++p) {
      const int R = s.radix[p], P = s.span[p];
      if (p == 0) pass_any<-1>(R, m, P, t, T, twm, gld, lst, false);
      else pass_any<-1>(R, m, P, t, T, twm, lld, lst, true);
      if (p != s.np - 1) __syncthreads();
    }
  }
  __syncthreads();
  if (ok) {
    float2* dr = dst + row * H;
    for (int k = t; k <= m; k += T) {
      const float2 zk = lrow[k == m ? 0 : k];
      const float2 zm = lrow[k == 0 ? 0 : m - k];
      const float2 zmc = make_float2(zm.x, -zm.y);
      const float2 E = make_float2(0.5f * (zk.x + zmc.x), 0.5f * (zk.y + zmc.y));
      const float2 D = make_float2(zk.x - zmc.x, zk.y - zmc.y);
      const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);  // -i/2 * D
      dr[k] = cadd(E, cmul(twp[k], O));
    }
  }
}

__device__ __forceinline__ float idiv_term(float o, float mdl) {
  const float os = fmaxf(o, 1e-6f), ms = fmaxf(mdl, 1e-6f);
  return o * logf(os / ms) - (o - ms);
}

__device__ __forceinline__ float apply_epi(const Epi& e, float v, int64_t idx, double& acc) {
  switch (e.kind) {
    case EPI_FLOOR: return fmaxf(v, e.a);
    case EPI_RL_RATIO: {
      const float mdl = v + e.a;
      const float o = e.aux0[idx];
      if (e.partials) acc += (double)idiv_term(o, mdl);
      return o / fmaxf(mdl, e.b);
    }
    case EPI_RL_UPDATE: return fmaxf(e.aux0[idx] * (v / e.aux1[idx]), e.b);
    case EPI_ADD: {
      const float mdl = v + e.a;
      if (e.partials && e.aux0) acc += (double)idiv_term(e.aux0[idx], mdl);
      return mdl;
    }
    case EPI_SQDIFF: {
      const float dv = v - e.aux0[idx];
      if (e.partials) acc += (double)(dv * dv);
      return v;
    }
    default: return v;
  }
}

// Block-wide deterministic double sum; result valid in thread 0.
__device__ double block_sum(double v, double* scratch) {
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

// pidx: index of this workgroup's partial sum within the volume
__device__ __forceinline__ void x_c2r_body(const AxisSched& s, const float2* __restrict__ twm,
                                           const float2* __restrict__ twp, const float2* __restrict__ src,
                                           float* dst, int64_t nrows, int s2, int H, const Epi& epi, int64_t row0,
                                           int RPB, int64_t pidx, float2* lds) {
  const int m = s.n;
  const int T = threads_per_line_d(m);
  int rl = threadIdx.x / T;
  int t = threadIdx.x % T;
  const bool active = rl < RPB;
  if (!active) {
    rl = 0;
    t = m;
  }
  const int64_t row = row0 + rl;
  const bool ok = active && row < nrows;
  float2* lrow = lds + rl * (m + 1);
  if (active) {
    if (ok) {
      const float2* sr = src + row * H;
      for (int k = t; k < H; k += T) lrow[k] = sr[k];
    } else {
      for (int k = t; k < H; k += T) lrow[k] = make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  // Z'_q = (X_q + conj X_{m-q}) + i exp(+2 pi i q / P2) (X_q - conj X_{m-q});
  // DC and Nyquist bins are taken as real (the imaginary parts an irfft ignores).
  auto pre = [&](int q) -> float2 {
    float2 xq = lrow[q];
    float2 xm = lrow[m - q];
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
  float* drow = dst + row * s2;
  const int64_t base = row * s2;
  auto gst = [&](int e, float2 v) {
    if (!ok) return;
    const int q0 = 2 * e, q1 = 2 * e + 1;
    if (q0 < s2) drow[q0] = apply_epi(epi, v.x, base + q0, acc);
    if (q1 < s2) drow[q1] = apply_epi(epi, v.y, base + q1, acc);
  };
  auto lld = [&](int e) { return lrow[e]; };
  auto lst = [&](int e, float2 v) { lrow[e] = v; };
  if (s.np == 0) {
    float2 v = make_float2(0.f, 0.f);
    if (t == 0) v = pre(0);
    __syncthreads();
    if (t == 0) gst(0, v);
  } else {
    for (int p = 0; p < s.np; ++p) {
      const int R = s.radix[p], P = s.span[p];
      const bool last = p == s.np - 1;
      if (p == 0 && last) pass_any<1>(R, m, P, t, T, twm, pre, gst, true);
      else if (p == 0) pass_any<1>(R, m, P, t, T, twm, pre, lst, true);
      else if (last) pass_any<1>(R, m, P, t, T, twm, lld, gst, false);
      else pass_any<1>(R, m, P, t, T, twm, lld, lst, true);
      if (!last) __syncthreads();
    }
  }
  if (epi.partials) {
    __syncthreads();
    double tot = block_sum(acc, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) epi.partials[pidx] = tot;
  }
}

// ---------------------------------------------------------------------------
// Single-volume kernels
// ---------------------------------------------------------------------------

template <int DIR>
__global__ void k_col_fft(AxisSched s, const float2* __restrict__ tw, const float2* __restrict__ src,
                          float2* __restrict__ dst, int64_t ncols, int64_t src_b, int64_t dst_b, int64_t src_rs,
                          int64_t dst_rs, int lo, int hi, int n_out, float scale, int L) {
  extern __shared__ float2 lds[];
  // XCD-aware tile order: consecutive column tiles (which share cache lines)
  // go to the same XCD and its L2.
  const int64_t nb = (int64_t)gridDim.x * gridDim.y;
  const int64_t lin = xcd_remap((int64_t)blockIdx.y * gridDim.x + blockIdx.x, nb);
  const int64_t tile = lin % gridDim.x, batch = lin / gridDim.x;
  col_fft_body<DIR>(s, tw, src + batch * src_b, dst + batch * dst_b, ncols, src_rs, dst_rs, lo, hi, n_out, scale,
                    tile, L, lds);
}

__global__ void k_z_conv(AxisSched f, AxisSched iv, const float2* __restrict__ tw, float2* buf,
                         const float2* __restrict__ otf, int64_t ncols, int lo, int n_out, int conj, int L) {
  extern __shared__ float2 lds[];
  z_conv_body(f, iv, tw, buf, otf, ncols, lo, n_out, conj, xcd_remap(blockIdx.x, gridDim.x), L, lds);
}

__global__ void k_x_r2c(AxisSched s, const float2* __restrict__ twm, const float2* __restrict__ twp,
                        const float* __restrict__ src, float2* __restrict__ dst, int64_t nrows, int rowlen, int lo,
                        int hi, int H, int RPB, Pro pro) {
  extern __shared__ float2 lds[];
  x_r2c_body(s, twm, twp, src, dst, nrows, rowlen, lo, hi, H, pro, (int64_t)blockIdx.x * RPB, RPB, lds);
}

__global__ void k_x_c2r(AxisSched s, const float2* __restrict__ twm, const float2* __restrict__ twp,
                        const float2* __restrict__ src, float* dst, int64_t nrows, int s2, int H, int RPB, Epi epi) {
  extern __shared__ float2 lds[];
  x_c2r_body(s, twm, twp, src, dst, nrows, s2, H, epi, (int64_t)blockIdx.x * RPB, RPB, blockIdx.x, lds);
}

__global__ void k_sum_partials(const double* __restrict__ p, int64_t n, double* out) {
  __shared__ double scratch[16];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += p[i];
  double r = block_sum(acc, scratch);
  if (threadIdx.x == 0) out[0] = r;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

int threads_per_line(int n) { return (n + kEPT - 1) / kEPT; }

// Largest LDS a workgroup may take on gfx950 (160 KiB per CU).
constexpr int64_t kMaxLds = 160 * 1024;

static void allow_big_lds() {
  static bool done = false;
  if (done) return;
  done = true;
  (void)hipFuncSetAttribute((const void*)k_col_fft<-1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
  (void)hipFuncSetAttribute((const void*)k_col_fft<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
  (void)hipFuncSetAttribute((const void*)k_z_conv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
  (void)hipFuncSetAttribute((const void*)k_x_r2c, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
  (void)hipFuncSetAttribute((const void*)k_x_c2r, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
}

bool col_tile_ok(int n, int L) {
  const int T = threads_per_line(n);
  return L >= 1 && L * T <= 1024 && (int64_t)L * n * 8 <= kMaxLds;
}

bool row_tile_ok(int m, int R) {
  const int T = threads_per_line(m);
  return R >= 1 && R * T <= 1024 && (int64_t)R * (m + 1) * 8 <= kMaxLds;
}

ColLaunch col_launch_for(int n, int64_t ncols, int tile) {
  ColLaunch c;
  c.T = threads_per_line(n);
  if (tile > 0 && col_tile_ok(n, tile)) {
    c.L = tile;
    return c;
  }
  int L = 1;
  // widest power-of-two column tile with <= 512 threads and <= 64 KiB of LDS
  while (L < 64 && (2 * L) * c.T <= 512 && (int64_t)(2 * L) * n * 8 <= 65536) L *= 2;
  while (L > 1 && L / 2 >= ncols) L /= 2;
  c.L = L;
  return c;
}

RowLaunch row_launch_for(int m, int tile) {
  RowLaunch r;
  r.T = threads_per_line(m);
  if (tile > 0 && row_tile_ok(m, tile)) {
    r.R = tile;
    return r;
  }
  int R = 1;
  while (R < 64 && (2 * R) * r.T <= 256 && (int64_t)(2 * R) * (m + 1) * 8 <= 65536) R *= 2;
  r.R = R;
  return r;
}

hipError_t launch_col_fft(int dir, const AxisSched& s, const float2* tw, const float2* src, float2* dst,
                          int64_t ncols, int nbatch, int64_t src_b, int64_t dst_b, int64_t src_rs,
                          int64_t dst_rs, int lo, int hi, int n_out, float scale, int tile, hipStream_t st) {
  allow_big_lds();
  ColLaunch c = col_launch_for(s.n, ncols, tile);
  dim3 grid((unsigned)((ncols + c.L - 1) / c.L), (unsigned)nbatch);
  dim3 block(c.L * c.T);
  size_t shm = (size_t)s.n * c.L * sizeof(float2);
  if (dir < 0)
    hipLaunchKernelGGL(k_col_fft<-1>, grid, block, shm, st, s, tw, src, dst, ncols, src_b, dst_b, src_rs,
                       dst_rs, lo, hi, n_out, scale, c.L);
  else
    hipLaunchKernelGGL(k_col_fft<1>, grid, block, shm, st, s, tw, src, dst, ncols, src_b, dst_b, src_rs,
                       dst_rs, lo, hi, n_out, scale, c.L);
  return hipGetLastError();
}

hipError_t launch_z_conv(const AxisSched& f, const AxisSched& iv, const float2* tw, float2* buf,
                         const float2* otf, int64_t ncols, int lo, int n_out, int conj, int tile, hipStream_t st) {
  allow_big_lds();
  ColLaunch c = f.n == 1 ? ColLaunch{256, 1} : col_launch_for(f.n, ncols, tile);
  dim3 grid((unsigned)((ncols + c.L - 1) / c.L));
  dim3 block(c.L * c.T);
  size_t shm = (size_t)f.n * c.L * sizeof(float2);
  hipLaunchKernelGGL(k_z_conv, grid, block, shm, st, f, iv, tw, buf, otf, ncols, lo, n_out, conj, c.L);
  return hipGetLastError();
}

hipError_t launch_x_r2c(const AxisSched& s, const float2* twm, const float2* twp, const float* src,
                        float2* dst, int64_t nrows, int rowlen, int lo, int hi, int H, const Pro& pro, int tile,
                        hipStream_t st) {
  allow_big_lds();
  RowLaunch r = row_launch_for(s.n, tile);
  dim3 grid((unsigned)((nrows + r.R - 1) / r.R));
  dim3 block(r.R * r.T);
  size_t shm = (size_t)r.R * (s.n + 1) * sizeof(float2);
  hipLaunchKernelGGL(k_x_r2c, grid, block, shm, st, s, twm, twp, src, dst, nrows, rowlen, lo, hi, H, r.R,
                     pro);
  return hipGetLastError();
}

hipError_t launch_x_c2r(const AxisSched& s, const float2* twm, const float2* twp, const float2* src,
                        float* dst, int64_t nrows, int s2, int H, const Epi& epi, int tile, int64_t* nblocks,
                        hipStream_t st) {
  allow_big_lds();
  RowLaunch r = row_launch_for(s.n, tile);
  dim3 grid((unsigned)((nrows + r.R - 1) / r.R));
  dim3 block(r.R * r.T);
  size_t shm = (size_t)r.R * (s.n + 1) * sizeof(float2);
  if (shm < 16 * sizeof(double)) shm = 16 * sizeof(double);
  if (nblocks) *nblocks = grid.x;
  hipLaunchKernelGGL(k_x_c2r, grid, block, shm, st, s, twm, twp, src, dst, nrows, s2, H, r.R, epi);
  return hipGetLastError();
}

hipError_t launch_sum_partials(const double* partials, int64_t n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, st, partials, n, out);
  return hipGetLastError();
}

}  // namespace dh
