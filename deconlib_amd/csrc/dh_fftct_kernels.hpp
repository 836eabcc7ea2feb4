// Kernels and host launchers for the compile-time specialised passes.
//
// Each launch covers up to kPack volumes of the same transform length (one
// "shape class"), their descriptors passed by value in the kernel arguments,
// so a single volume needs no descriptor upload and a batch of volumes of
// one length is a single launch per pass.
#pragma once
#include <mutex>
#include <type_traits>

#include "dh_fftct.hpp"
#include "dh_prox.hpp"

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

__device__ __forceinline__ float idiv(float o, float mdl) {
  const float os = fmaxf(o, 1e-6f), ms = fmaxf(mdl, 1e-6f);
  return o * logf(os / ms) - (o - ms);
}

// Every sample counts (no reduction box).
struct AllSamples {
  __device__ __forceinline__ bool operator()(int) const { return true; }
};

// Samples q0 + j of a row inside the reduction box: in_row = the row's (z, y)
// lies in it (a per-lane flag), [x0, x1) its uniform x range.
struct InBox {
  bool in_row;
  int x0, x1, q0;
  __device__ __forceinline__ bool operator()(int j) const { return in_row && q0 + j >= x0 && q0 + j < x1; }
};

// cnt(j): sample j (0, or 1 for the second of a pair) enters the partial sums;
// evaluated only when the epilogue reduces.
template <class Cnt = AllSamples>
__device__ __forceinline__ float epi_apply(const Epi& e, float v, int64_t idx, double& acc, Cnt cnt = Cnt()) {
  switch (e.kind) {
    case EPI_FLOOR: return fmaxf(v, e.a);
    case EPI_RL_RATIO: {
      const float mdl = v + e.a;
      const float o = e.aux0[idx];
      if (e.partials && cnt(0)) acc += (double)idiv(o, mdl);
      return o / fmaxf(mdl, e.b);
    }
    case EPI_RL_UPDATE: return fmaxf(e.aux0[idx] * (v / e.aux1[idx]), e.b);
    case EPI_ADD: {
      const float mdl = v + e.a;
      if (e.partials && e.aux0 && cnt(0)) acc += (double)idiv(e.aux0[idx], mdl);
      return mdl;
    }
    case EPI_SQDIFF: {
      const float dv = v - e.aux0[idx];
      if (e.partials && cnt(0)) acc += (double)(dv * dv);
      return v;
    }
    default: return v;
  }
}

// epi_apply on the sample pair (idx, idx + 1) with float2 operand loads;
// the pair is accumulated in the same order as two epi_apply calls.  K >= 0
// fixes the kind at compile time (e.kind is then not read).
template <int K = -1, class Cnt = AllSamples>
__device__ __forceinline__ float2 epi_apply2(const Epi& e, float2 v, int64_t idx, double& acc, Cnt cnt = Cnt()) {
  switch (K >= 0 ? K : e.kind) {
    case EPI_FLOOR: return make_float2(fmaxf(v.x, e.a), fmaxf(v.y, e.a));
    case EPI_RL_RATIO: {
      const float2 o = *reinterpret_cast<const float2*>(e.aux0 + idx);
      const float m0 = v.x + e.a, m1 = v.y + e.a;
      if (e.partials) {
        if (cnt(0)) acc += (double)idiv(o.x, m0);
        if (cnt(1)) acc += (double)idiv(o.y, m1);
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
        if (cnt(0)) acc += (double)idiv(o.x, m0);
        if (cnt(1)) acc += (double)idiv(o.y, m1);
      }
      return make_float2(m0, m1);
    }
    case EPI_SQDIFF: {
      if (e.partials) {
        const float2 o = *reinterpret_cast<const float2*>(e.aux0 + idx);
        const float d0 = v.x - o.x, d1 = v.y - o.y;
        if (cnt(0)) acc += (double)(d0 * d0);
        if (cnt(1)) acc += (double)(d1 * d1);
      }
      return v;
    }
    default: return v;
  }
}

// Workgroup sum of v, returned by thread 0.  A workgroup smaller than a wave
// (or a partial last wave) adds only its existing lanes; for whole waves lane
// 0's reduction tree is the plain one.
__device__ __forceinline__ double bsum(double v, double* scratch) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int valid = min(64, (int)blockDim.x - (w << 6));
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(v, off, 64);
    if (lane + off < valid) v += o;
  }
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

// ---- streaming global accesses ----
// Every pass reads each of its inputs once and writes each output once per
// launch: non-temporal loads and stores keep these streams from displacing
// reusable lines (twiddles, the other operands of the pass) in the caches.
typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_nt(const float2* p) {
  const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
  const f2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
}
// Non-temporal (NT) or ordinary access: a column pass whose workgroups each take
// part of a block's 64-byte row runs uses ordinary accesses, so that the cache
// can join the pieces the workgroups load and store.
template <bool NT>
__device__ __forceinline__ float2 ld_s(const float2* p) {
  if constexpr (NT) return ld_nt(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st_s(float2* p, float2 v) {
  if constexpr (NT) st_nt(p, v);
  else *p = v;
}

// ---- column pass (Y forward / inverse, OTF Z) ----
// LDS slot of line element e, column ln: 8 extra slots every 16 elements spread
// a wave's (rows x columns) accesses over the banks.
__device__ __forceinline__ int col_lds(int e, int ln, int L) { return e * L + ln + (L < 8 ? L : 8) * (e >> 4); }

// Hides the lane coordinates from loop-invariant code motion in the persistent
// loops below: otherwise every pass's LDS and twiddle addresses are hoisted out
// of the tile loop and stay live, which doubles the register count.  Emits no
// instruction.
__device__ __forceinline__ void opaque(int& a, int& b) { asm volatile("" : "+v"(a), "+v"(b)); }
__device__ __forceinline__ void opaque(int& a) { asm volatile("" : "+v"(a)); }

// Addressing of one lane's column in one tile (logical block lb of a volume):
// the tile's plane bases sb/db are workgroup-uniform (scalar registers), the
// lane's column offset and row stride are in bytes of type Off.  With Off =
// uint32_t (every offset inside one plane below 2^32 bytes, col_narrow) an
// access is a scalar base plus one 32-bit vector offset; uint64_t serves
// planes larger than that.
template <class Off>
struct ColLane {
  const char* sb;
  char* db;
  Off so, dof, sr, dr;
  bool ok;
};

template <class Off>
__device__ __forceinline__ ColLane<Off> col_lane(const ColArgs& a, int64_t lb, int L, int ln) {
  const int64_t tile = lb % a.ntiles, batch = lb / a.ntiles;
  const int64_t col = tile * L + ln;
  int64_t soff = col, doff = col, srs = a.src_rs, drs = a.dst_rs;
  if (a.blk > 0) {
    const int64_t b = col / kBlk, j = col - b * kBlk;
    const int64_t w = a.ncols - b * kBlk < kBlk ? a.ncols - b * kBlk : kBlk;
    soff = b * kBlk * a.src_rs + j;
    doff = b * kBlk * a.dst_rs + j;
    srs = drs = w;
  }
  return ColLane<Off>{reinterpret_cast<const char*>(a.src + batch * a.src_b),
                      reinterpret_cast<char*>(a.dst + batch * a.dst_b),
                      (Off)(soff * 8),
                      (Off)(doff * 8),
                      (Off)(srs * 8),
                      (Off)(drs * 8),
                      col < a.ncols};
}

// True when every element offset of a volume's plane fits 32 bits in bytes:
// blocked planes span ncols x rows elements, plain planes rows x row stride.
__host__ __device__ inline bool col_narrow(const ColArgs& a, int N) {
  constexpr int64_t kLim = int64_t(1) << 31;
  auto span = [&](int64_t rs) {
    return a.blk > 0 ? (a.ncols + kBlk) * (rs + N) : a.ncols + (int64_t)(N + 1) * rs;
  };
  return 8 * span(a.src_rs) < kLim && 8 * span(a.dst_rs) < kLim;
}

// Line element e of the zero-padded canvas: rows [0, lo) and [N-hi, N) hold
// the source rows 0..lo+hi-1, the rest is zero.
template <int N, bool NT = true, class Off>
__device__ __forceinline__ float2 col_ld(const ColArgs& a, const ColLane<Off>& c, int e) {
  int r = -1;
  if (e < a.lo) r = e;
  else if (e >= N - a.hi) r = e - (N - a.lo - a.hi);
  if (!c.ok || r < 0) return make_float2(0.f, 0.f);
  return ld_s<NT>(reinterpret_cast<const float2*>(c.sb + (c.so + (Off)r * c.sr)));
}

// col_ld for hi == 0 (no wrapped rows: the Y and Z passes of an application).
template <bool NT = true, class Off>
__device__ __forceinline__ float2 col_ld_lo(int lo, const ColLane<Off>& c, int e) {
  if (!c.ok || e >= lo) return make_float2(0.f, 0.f);
  return ld_s<NT>(reinterpret_cast<const float2*>(c.sb + (c.so + (Off)e * c.sr)));
}

// Persistent column pass: each workgroup walks tiles b = blockIdx.x,
// blockIdx.x + gridDim.x, ... and loads the next tile's first-pass operands
// into registers while the current tile's remaining passes run.
// Workgroups are dealt to the 8 XCDs in turn (workgroup b to XCD b % 8).  Numbering
// the tiles XCD by XCD instead gives each XCD a contiguous run of them, so tiles that
// share a block's row runs also share an L2.  A bijection of [0, G): for every XCD x
// the b = x + 8 i < G map to consecutive numbers from x q + min(x, r).
__device__ __forceinline__ int64_t xcd_major(int64_t b, int64_t G) {
  constexpr int64_t kXcd = 8;
  const int64_t q = G / kXcd, r = G % kXcd, x = b % kXcd;
  return x * q + (x < r ? x : r) + b / kXcd;
}

template <int N, int DIR, bool NT, class Off>
__device__ __forceinline__ void col_run(const Pack<ColArgs>& pk, int L, int64_t b, float2* lds) {
  constexpr int T = Geo<N>::T;
  constexpr Sched S = Geo<N>::F;
  constexpr int R0 = S.np > 0 ? S.r[0] : 1;
  // Blocked layout: a wave takes kBlk columns x consecutive line elements (one
  // contiguous run per access); plain layout: L consecutive columns.
  const int G = (pk.a[0].blk > 0 && L > kBlk) ? kBlk : L;
  const int thr = threadIdx.x;
  const int ln0 = thr % G + G * (thr / (G * T));
  const int t0 = (thr / G) % T;
  const int64_t total = pk.prefix[pk.nvol];
  if constexpr (N == 1) {  // length-1 transform: scaled copy
    for (; b < total; b += gridDim.x) {
      const int v = pick(pk, b);
      const ColArgs& a = pk.a[v];
      const ColLane<Off> c = col_lane<Off>(a, b - pk.prefix[v], L, ln0);
      const float2 x = col_ld<N, NT>(a, c, 0);
      if (c.ok && a.n_out > 0)
        *reinterpret_cast<float2*>(c.db + c.dof) = make_float2(x.x * a.scale, x.y * a.scale);
    }
    return;
  }
  float2 u[Bfly<N, T, R0>::M][R0];
  auto fetch = [&](int64_t bb) {
    int ln = ln0, t = t0;
    opaque(ln, t);
    const int v = pick(pk, bb);
    const ColArgs& a = pk.a[v];
    const ColLane<Off> c = col_lane<Off>(a, bb - pk.prefix[v], L, ln);
    if (a.hi == 0) {
      const int lo = a.lo;
      pass_load<N, T, R0>(u, t, [&](int e) { return col_ld_lo<NT>(lo, c, e); });
    } else {
      pass_load<N, T, R0>(u, t, [&](int e) { return col_ld<N, NT>(a, c, e); });
    }
  };
  fetch(b);
  for (;;) {
    int ln = ln0, t = t0;
    opaque(ln, t);
    auto lld = [&](int e) { return lds[col_lds(e, ln, L)]; };
    auto lst = [&](int e, float2 v) { lds[col_lds(e, ln, L)] = v; };
    const int v = pick(pk, b);
    const ColArgs& a = pk.a[v];
    const ColLane<Off> c = col_lane<Off>(a, b - pk.prefix[v], L, ln);
    const int n_out = a.n_out;
    const float scale = a.scale;
    auto gst = [&](int e, float2 val) {
      if (c.ok && e < n_out)
        st_s<NT>(reinterpret_cast<float2*>(c.db + (c.dof + (Off)e * c.dr)), make_float2(val.x * scale, val.y * scale));
    };
    const int64_t bn = b + gridDim.x;
    if constexpr (S.np == 1) {
      pass_store<N, T, R0, 1, DIR>(u, t, a.tw, gst);
      if (bn < total) fetch(bn);
    } else {
      pass_store<N, T, R0, 1, DIR>(u, t, a.tw, lst);
      if (bn < total) fetch(bn);
      __syncthreads();
      passes<N, false, DIR, 1>(t, a.tw, lld, gst, lld, lst, false);
    }
    if (bn >= total) break;
    b = bn;
    __syncthreads();  // the next tile's first pass overwrites the LDS lines
  }
}

// WIDE: 64-bit lane offsets, for volumes that are not col_narrow (the launcher
// picks the instantiation; a separate kernel, so that the 32-bit one keeps its
// register count).
template <int N, int DIR, bool NT = true, bool WIDE = false>
__global__ void __launch_bounds__(kMaxThreads) k_col(Pack<ColArgs> pk, int L) {
  extern __shared__ float2 lds[];
  const int64_t total = pk.prefix[pk.nvol];
  const int64_t b = NT ? (int64_t)blockIdx.x : xcd_major(blockIdx.x, gridDim.x);
  if (b >= total) return;
  col_run<N, DIR, NT, std::conditional_t<WIDE, uint64_t, uint32_t>>(pk, L, b, lds);
}

// ---- fused Z pass: forward FFT, x OTF, inverse FFT, in place ----

// Entries of the forward-schedule tables; the inverse schedule's rows follow them.
template <int N>
constexpr int kTwF = tw_table_size(Geo<N>::F);

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
        else u[m][r] = (ok && e < lo) ? ld_nt(cb + (int64_t)e * rs) : make_float2(0.f, 0.f);
      }
    }
  }
  if (from_lds) __syncthreads();
  constexpr int off = tw_pass_offset(Geo<N>::F, Geo<N>::F.np - 1);
  const float2* twp = tw + off;
  const float2* ot = otf + col + (int64_t)t * rs;  // OTF element t + c at c rows' uniform offset
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int i = t + m * T;
    if ((M * T == NB) || i < NB) {
      const int k = i % P;
      if constexpr (P > 1) twiddle<R, -1, P>(u[m], twp, k);
      Dft<R, -1>::run(u[m]);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float2 o = ok ? ld_nt(ot + (int64_t)(m * T + r * NB) * rs) : make_float2(0.f, 0.f);
        u[m][r] = conj ? cmulc(u[m][r], o) : cmul(u[m][r], o);
      }
      Dft<R, 1>::run(u[m]);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int e = i * R + r;
        if (from_lds) lds[e * L + ln] = u[m][r];
        else if (ok && e < n_out) st_nt(gdst + (int64_t)e * rs, u[m][r]);
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
    constexpr int off = tw_pass_offset(S, p);
    const float2* twp = tw + off;
    if (p == 0) pass<N, Geo<N>::T, R, P, -1>(t, twp, gld, lst, false);
    else pass<N, Geo<N>::T, R, P, -1>(t, twp, lld, lst, true);
    __syncthreads();
    z_fwd_passes<N, p + 1>(t, tw, gld, lld, lst);
  }
}

template <int N>
__device__ __forceinline__ void z_body(const ZArgs& a, int64_t lb, int L, float2* lds, int thr) {
  const int ln = thr % L;
  const int t = thr / L;
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
    auto gld = [&](int e) -> float2 { return (ok && e < lo) ? ld_nt(cb + (int64_t)e * rs) : make_float2(0.f, 0.f); };
    auto gst = [&](int e, float2 v) {
      if (ok && e < n_out) st_nt(cb + (int64_t)e * rs, v);
    };
    auto lld = [&](int e) { return lds[e * L + ln]; };
    auto lst = [&](int e, float2 v) { lds[e * L + ln] = v; };
    constexpr Sched F = Geo<N>::F;
    z_fwd_passes<N, 0>(t, a.tw, gld, lld, lst);
    z_mid<N, F.r[F.np - 1], span_of(F, F.np - 1)>(t, a.tw, lds, L, ln, cb, rs, lo, ok, a.otf, col, a.conj, cb, n_out);
    if constexpr (F.np > 1) {
      __syncthreads();
      passes<N, true, 1, 1>(t, a.tw + kTwF<N>, gld, gst, lld, lst, false);
    }
  }
}

// Persistent fused Z pass (same tile walk and next-tile prefetch as k_col);
// lengths with a single radix pass work straight from global memory.
template <int N>
__global__ void __launch_bounds__(kMaxThreads) k_z(Pack<ZArgs> pk, int L) {
  extern __shared__ float2 lds[];
  constexpr Sched F = Geo<N>::F;
  const int64_t total = pk.prefix[pk.nvol];
  int64_t b = blockIdx.x;
  if constexpr (F.np < 2) {
    for (; b < total; b += gridDim.x) {
      int thr = threadIdx.x;
      opaque(thr);
      const int v = pick(pk, b);
      z_body<N>(pk.a[v], b - pk.prefix[v], L, lds, thr);
    }
  } else {
    // The next tile's operands are prefetched unless the middle pass (radix
    // 16: 16 data + 16 OTF operands per butterfly) leaves no registers for them.
    constexpr bool kPrefetch = F.r[F.np - 1] <= 8;
    constexpr int T = Geo<N>::T;
    constexpr int R0 = F.r[0];
    const int ln0 = threadIdx.x % L;
    const int t0 = threadIdx.x / L;
    if (b >= total) return;
    float2 u[Bfly<N, T, R0>::M][R0];
    auto fetch = [&](int64_t bb) {
      int ln = ln0, t = t0;
      opaque(ln, t);
      const int v = pick(pk, bb);
      const ZArgs& a = pk.a[v];
      const int64_t col = (bb - pk.prefix[v]) * L + ln;
      const bool ok = col < a.ncols;
      // element e = t + c sits c rows after the thread's first one: c * ncols is
      // uniform, so no per-element 64-bit multiply
      const float2* cbt = a.buf + col + (int64_t)t * a.ncols;
      pass_load<N, T, R0>(u, t, [&](int e) {
        return (ok && e < a.lo) ? ld_nt(cbt + (int64_t)(e - t) * a.ncols) : make_float2(0.f, 0.f);
      });
    };
    fetch(b);
    for (;;) {
      int ln = ln0, t = t0;
      opaque(ln, t);
      auto lld = [&](int e) { return lds[e * L + ln]; };
      auto lst = [&](int e, float2 v) { lds[e * L + ln] = v; };
      const int v = pick(pk, b);
      const ZArgs& a = pk.a[v];
      const int64_t col = (b - pk.prefix[v]) * L + ln;
      const bool ok = col < a.ncols;
      float2* cb = a.buf + col;
      const int64_t rs = a.ncols;
      float2* cbt = cb + (int64_t)t * rs;  // as in fetch: the last pass stores at t + constant
      auto gst = [&](int e, float2 val) {
        if (ok && e < a.n_out) st_nt(cbt + (int64_t)(e - t) * rs, val);
      };
      pass_store<N, T, R0, 1, -1>(u, t, a.tw, lst);
      const int64_t bn = b + gridDim.x;
      if (kPrefetch && bn < total) fetch(bn);
      __syncthreads();
      z_fwd_passes<N, 1>(t, a.tw, lld, lld, lst);
      z_mid<N, F.r[F.np - 1], span_of(F, F.np - 1)>(t, a.tw, lds, L, ln, cb, rs, a.lo, ok, a.otf, col, a.conj, cb,
                                                    a.n_out);
      __syncthreads();
      passes<N, true, 1, 1>(t, a.tw + kTwF<N>, lld, gst, lld, lst, false);
      if (bn >= total) break;
      b = bn;
      if (!kPrefetch) fetch(b);
      __syncthreads();
    }
  }
}

// ---- X rows: LDS layout and the blocked spectrum row walk ----
// LDS slot of row element e: one spare slot every 32 elements, so Stockham
// stores at stride R do not fall on one bank.
__device__ __forceinline__ int xrow_lds(int e) { return e + (e >> 5); }

// Visits every (row rl, column k < H) of a workgroup's RPB rows (RPB a power of
// two) in blocked-layout order: a wave covers kBlk columns x consecutive rows,
// one contiguous run.  fn(rl, k, global offset of (row, k), row valid).
template <class Fn>
__device__ __forceinline__ void blocked_rows(int64_t row0, int RPB, int64_t nrows, int H, int prow, Fn fn) {
  const int lg = __builtin_ctz(RPB);
  const int nbk = (H + kBlk - 1) / kBlk;
  const int total = nbk * RPB * kBlk;
  const int64_t z0 = row0 / prow;
  const int y0 = (int)(row0 - z0 * prow);
  for (int f = threadIdx.x; f < total; f += blockDim.x) {
    const int j = f & (kBlk - 1);
    const int rl = (f / kBlk) & (RPB - 1);
    const int b = f >> (lg + 3);
    const int k = b * kBlk + j;
    if (k >= H) continue;
    int64_t z = z0;
    int y = y0 + rl;
    if (y >= prow) {
      z += y / prow;
      y %= prow;
    }
    const int w = (b == nbk - 1) ? H - b * kBlk : kBlk;
    const int64_t off = z * prow * (int64_t)H + (int64_t)b * kBlk * prow + (int64_t)y * w + j;
    fn(rl, k, off, row0 + rl < nrows);
  }
}

// blocked_rows for a workgroup of RPB x T threads (T threads per row, a
// multiple of kBlk): thread thr then always has column j = thr % kBlk of row
// rl = (thr / kBlk) % RPB, and its blocks advance by T / kBlk, so the row split
// and the 64-bit offset are formed once and stepped (the same visits, in the
// same order, as blocked_rows).
// thr = threadIdx.x, passed through opaque() by the caller so that the set-up is
// formed here rather than hoisted (and held in registers) across earlier passes.
template <int T, class Fn>
__device__ __forceinline__ void blocked_rows_t(int thr, int64_t row0, int RPB, int64_t nrows, int H, int prow,
                                               Fn fn) {
  if constexpr (T % kBlk != 0) {
    blocked_rows(row0, RPB, nrows, H, prow, fn);
  } else {
    const int nbk = (H + kBlk - 1) / kBlk;
    const int j = thr & (kBlk - 1);
    const int rl = (thr / kBlk) & (RPB - 1);
    int64_t z = row0 / prow;
    int y = (int)(row0 - z * prow) + rl;
    if (y >= prow) {
      z += y / prow;
      y %= prow;
    }
    const bool valid = row0 + rl < nrows;
    const int64_t zb = z * prow * (int64_t)H;
    const int64_t step = (int64_t)(T / kBlk) * kBlk * prow;
    int b = thr >> (__builtin_ctz(RPB) + 3);
    int64_t off = zb + (int64_t)b * kBlk * prow + (int64_t)y * kBlk + j;  // a full-width block
    for (; b < nbk - 1; b += T / kBlk, off += step) fn(rl, b * kBlk + j, off, valid);
    if (b == nbk - 1) {
      const int w = H - b * kBlk;
      if (j < w) fn(rl, b * kBlk + j, zb + (int64_t)b * kBlk * prow + (int64_t)y * w + j, valid);
    }
  }
}

// Both twiddle tables of an X workgroup in LDS after its rows: the pass tables
// of the length-M transform at [0, tw_table_size), the M+1 split twiddles at
// [M, 2M].  One transform length per launch, so one copy serves every volume.
template <int M>
__device__ __forceinline__ void stage_x_tables(float2* ltw, const float2* twm, const float2* twp) {
  constexpr int TWM = tw_table_size(Geo<M, true>::F);
  static_assert(TWM <= M, "the pass tables fit the LDS slot reserved for them");
  for (int i = threadIdx.x; i < 2 * M + 1; i += blockDim.x) {
    if (i < TWM) ltw[i] = twm[i];
    else if (i >= M) ltw[i] = twp[i - M];
  }
}

// ---- X forward: real rows -> half spectrum (M-point complex FFT) ----
// Complex sample e (real samples 2e, 2e+1 of the zero-padded row of length 2M,
// prologue applied) of row `row` of volume a.
// PRO_DIV_DEV (the OTF build's kernel / sum) is applied afterwards by xf_div.
// EXT: 0 the plain prologues; 1 also PRO_SPREAD; 2 also PRO_PD_Y1.  A launch takes
// the smallest instantiation that covers its pack, so the others keep lean code.
// Row `row` of volume a: ok = row < nrows, rb = row * rowlen (its offset in
// signal-shaped arrays), sr = its first source sample (xf_row).
template <int EXT>
__device__ __forceinline__ const float* xf_row(const XfArgs& a, int64_t row, bool ok) {
  if constexpr (EXT > 0) return pro_row(a.pro, a.src, ok ? row : 0, a.rowlen);
  return a.src + row * a.rowlen;
}

template <int M, int EXT>
__device__ __forceinline__ float2 xf_ld(const XfArgs& a, const float* sr, int64_t rb, bool ok, int e) {
  constexpr int P2 = 2 * M;
  const int kind = a.pro.kind;
  // PRO_ONES never reads the source (which may then be null)
  const bool read = kind != PRO_ONES;
  float2 raw = make_float2(0.f, 0.f);
  bool ok0, ok1;
  if constexpr (EXT > 0) {
    if (kind == PRO_SPREAD) {  // coarse rows (hi == 0), each sample replicated f2 (1 or 2) times
      ok0 = ok && 2 * e < a.lo;
      ok1 = ok && 2 * e + 1 < a.lo;
      if (a.pro.f2 == 2) {
        if (ok0) raw.x = raw.y = __builtin_nontemporal_load(sr + e);
      } else {
        if (ok0) raw.x = __builtin_nontemporal_load(sr + 2 * e);
        if (ok1) raw.y = __builtin_nontemporal_load(sr + 2 * e + 1);
      }
      return raw;
    }
    if (EXT > 1 && kind == PRO_PD_Y1) {  // convolution input rows (hi == 0)
      const PdY1Consts kc = pd_y1_consts(a.pro.sc, a.pro.a);
      const int64_t i0 = rb + 2 * e;
      ok0 = ok && 2 * e < a.lo;
      ok1 = ok && 2 * e + 1 < a.lo;
      if (a.vec2) {  // every operand row 8-byte aligned, even row length: ok0 == ok1
        if (ok0) {
          const float2 y = ld_nt(reinterpret_cast<const float2*>(a.src + i0));
          const float2 ax = ld_nt(reinterpret_cast<const float2*>(a.pro.ax + i0));
          const float2 axp = ld_nt(reinterpret_cast<const float2*>(a.pro.axp + i0));
          const float2 d = ld_nt(reinterpret_cast<const float2*>(a.pro.obs + i0));
          raw = make_float2(pd_y1_prox(y.x, ax.x, axp.x, d.x, kc, a.pro.loss),
                            pd_y1_prox(y.y, ax.y, axp.y, d.y, kc, a.pro.loss));
          st_nt(reinterpret_cast<float2*>(a.pro.y1w + i0), raw);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (j == 0 ? ok0 : ok1) {
            const float v = pd_y1_prox(__builtin_nontemporal_load(a.src + i0 + j),
                                       __builtin_nontemporal_load(a.pro.ax + i0 + j),
                                       __builtin_nontemporal_load(a.pro.axp + i0 + j),
                                       __builtin_nontemporal_load(a.pro.obs + i0 + j), kc, a.pro.loss);
            __builtin_nontemporal_store(v, a.pro.y1w + i0 + j);
            if (j == 0) raw.x = v;
            else raw.y = v;
          }
        }
      }
      return raw;
    }
  }
  if (a.vec2) {  // hi == 0, even row length, 8-byte aligned rows
    ok0 = ok1 = ok && 2 * e + 1 < a.lo;
    if (ok0 && read) raw = ld_nt(reinterpret_cast<const float2*>(sr + 2 * e));
  } else {
    auto src_of = [&](int q) {
      if (q < a.lo) return q;
      if (q >= P2 - a.hi) return q - (P2 - a.lo - a.hi);
      return -1;
    };
    const int c0 = src_of(2 * e), c1 = src_of(2 * e + 1);
    ok0 = ok && c0 >= 0;
    ok1 = ok && c1 >= 0;
    if (ok0 && read) raw.x = __builtin_nontemporal_load(sr + c0);
    if (ok1 && read) raw.y = __builtin_nontemporal_load(sr + c1);
  }
  auto pro = [&](float x, bool valid) {
    const float y = kind == PRO_ONES ? 1.0f : (kind == PRO_SUB_FLOOR ? fmaxf(x - a.pro.a, a.pro.b) : x);
    return valid ? y : 0.0f;
  };
  return make_float2(pro(raw.x, ok0), pro(raw.y, ok1));
}

// PRO_DIV_DEV on a thread's first-pass operands (zero padding stays zero).
template <int N, int T, int R>
__device__ __forceinline__ void xf_div(float2 (&u)[(N / R + T - 1) / T][R], const Pro& p) {
  if (p.kind != PRO_DIV_DEV) return;
  const float s = (float)(*p.dptr);
#pragma unroll
  for (int m = 0; m < (N / R + T - 1) / T; ++m)
#pragma unroll
    for (int r = 0; r < R; ++r) u[m][r] = make_float2(u[m][r].x / s, u[m][r].y / s);
}

// Spectrum sample k of a row whose M-point transform z sits in LDS at lr.
template <int M>
__device__ __forceinline__ float2 xf_post(const float2* lr, const float2* twp, int k) {
  const float2 zk = lr[xrow_lds(k == M ? 0 : k)];
  const float2 zm = lr[xrow_lds(k == 0 ? 0 : M - k)];
  const float2 zmc = make_float2(zm.x, -zm.y);
  const float2 E = make_float2(0.5f * (zk.x + zmc.x), 0.5f * (zk.y + zmc.y));
  const float2 D = make_float2(zk.x - zmc.x, zk.y - zmc.y);
  const float2 O = make_float2(0.5f * D.y, -0.5f * D.x);  // -i/2 * D
  return cadd(E, cmul(twp[k], O));
}

// Persistent: workgroups walk tiles of RPB rows, prefetching the next tile's
// samples while the current tile's passes and stores run.  Both twiddle tables
// sit in LDS after the rows (one length per launch, so one copy per workgroup).
template <int M, int EXT>
__global__ void __launch_bounds__(kMaxThreads) k_xf(Pack<XfArgs> pk, int RPB) {
  extern __shared__ float2 lds[];
  constexpr int T = Geo<M, true>::T;
  constexpr Sched S = Geo<M, true>::F;
  constexpr int R0 = S.np > 0 ? S.r[0] : 1;
  constexpr int RS = xrow_stride(M);
  const int rl0 = threadIdx.x / T;
  const int t0 = threadIdx.x % T;
  const int64_t total = pk.prefix[pk.nvol];
  int64_t b = blockIdx.x;
  if (b >= total) return;
  const float2* twm = pk.a[0].twm;
  const float2* twp = pk.a[0].twp;
  if constexpr (!xrow_tw_global(M)) {
    float2* ltw = lds + RPB * RS;
    stage_x_tables<M>(ltw, pk.a[0].twm, pk.a[0].twp);
    twm = ltw;
    twp = ltw + M;
  }
  float2 u[Bfly<M, T, R0>::M][R0];
  auto fetch = [&](int64_t bb) {
    int rl = rl0, t = t0;
    opaque(rl, t);
    const int v = pick(pk, bb);
    const XfArgs& a = pk.a[v];
    const int64_t row = (bb - pk.prefix[v]) * RPB + rl;
    const bool ok = row < a.nrows;
    const int64_t rb = row * a.rowlen;
    const float* sr = xf_row<EXT>(a, row, ok);
    if constexpr (M == 1) {
      u[0][0] = xf_ld<M, EXT>(a, sr, rb, ok, 0);
    } else {
      pass_load<M, T, R0>(u, t, [&](int e) { return xf_ld<M, EXT>(a, sr, rb, ok, e); });
    }
    xf_div<M, T, R0>(u, a.pro);
  };
  fetch(b);
  for (;;) {
    int rl = rl0, t = t0;
    opaque(rl, t);
    float2* lrow = lds + rl * RS;
    auto lld = [&](int e) { return lrow[xrow_lds(e)]; };
    auto lst = [&](int e, float2 v) { lrow[xrow_lds(e)] = v; };
    const int v = pick(pk, b);
    const XfArgs& a = pk.a[v];
    const int64_t lb = b - pk.prefix[v];
    const int64_t row = lb * RPB + rl;
    if constexpr (M == 1) {
      if (t == 0) lrow[0] = u[0][0];
    } else {
      pass_store<M, T, R0, 1, -1>(u, t, twm, lst);
    }
    const int64_t bn = b + gridDim.x;
    if (bn < total) fetch(bn);
    __syncthreads();
    // remaining passes, in LDS (the last one reads and writes the row: LAST_LDS)
    passes<M, false, -1, 1, true, true>(t, twm, lld, lst, lld, lst, false);
    __syncthreads();
    if (a.blk > 0) {
      int thr = threadIdx.x;
      opaque(thr);
      blocked_rows_t<T>(thr, lb * RPB, RPB, a.nrows, a.H, a.prow, [&](int r, int k, int64_t off, bool valid) {
        if (valid) st_nt(a.dst + off, xf_post<M>(lds + r * RS, twp, k));
      });
    } else if (row < a.nrows) {
      float2* dr = a.dst + row * a.H;
      for (int k = t; k <= M; k += T) st_nt(dr + k, xf_post<M>(lrow, twp, k));
    }
    if (bn >= total) break;
    b = bn;
    __syncthreads();
  }
}

// ---- X inverse: half spectrum -> cropped real rows with the fused epilogue ----
// Elements of the H = M+1 spectrum samples of a tile's rows one thread loads:
// ceil((M + kBlk) / T) covers both the blocked walk (at most ceil(nbk*kBlk/T)
// visits) and plain rows (ceil(H/T)).
template <int M>
struct XiEpt {
  static constexpr int value = (M + kBlk + Geo<M, true>::T - 1) / Geo<M, true>::T;
};

// Loads (q) or places into LDS (!load) this thread's spectrum samples of the
// tile whose rows start at row0: visit i of the blocked walk, or sample
// k = t + i*T of the thread's own row for the plain layout.
// STEP: the stepped blocked walk (fewer instructions per load; the kernels whose
// registers allow it use it, see k_xi).
template <int M, bool STEP>
// nthr: threads that walk these rows (the workgroup, or half of it in k_xi_bs).
__device__ __forceinline__ void xi_tile(const XiArgs& a, int64_t row0, int RPB,
                                        float2 (&q)[(M + kBlk + Geo<M, true>::T - 1) / Geo<M, true>::T], float2* lds, bool load,
                                        int thr, int nthr) {
  constexpr int T = Geo<M, true>::T;
  constexpr int RS = xrow_stride(M);
  const int H = a.H;
  if (STEP && T % kBlk == 0 && a.blk > 0) {
    // A workgroup of RPB x T threads: visit i of thread thr is column j = thr % kBlk
    // of row r = (thr / kBlk) % RPB in block bk0 + i * T / kBlk (the same visits as
    // the general walk below), so the row's plane and offsets are formed once.
    const int lg = __builtin_ctz(RPB);
    const int nbk = (H + kBlk - 1) / kBlk;
    const int j = thr & (kBlk - 1);
    const int r = (thr >> 3) & (RPB - 1);
    const int bk0 = thr >> (lg + 3);
    constexpr int BS = T / kBlk;
    if (!load) {
      float2* lr = lds + r * RS;
#pragma unroll
      for (int i = 0; i < XiEpt<M>::value; ++i) {
        const int bk = bk0 + i * BS, k = bk * kBlk + j;
        if (bk < nbk && k < H) lr[xrow_lds(k)] = q[i];
      }
      return;
    }
    const bool rv = row0 + r < a.nrows;
    int64_t z = row0 / a.prow;
    int y = (int)(row0 - z * a.prow) + r;
    if (y >= a.prow) {
      z += y / a.prow;
      y %= a.prow;
    }
    // the row's plane; offsets inside one plane (H x prow samples) fit 32 bits
    const float2* zb = a.src + z * a.prow * (int64_t)H;
    const int rowk = kBlk * a.prow;
#pragma unroll
    for (int i = 0; i < XiEpt<M>::value; ++i) {
      const int bk = bk0 + i * BS, k = bk * kBlk + j;
      float2 val = make_float2(0.f, 0.f);
      if (rv && bk < nbk && k < H) {
        const int w = (bk == nbk - 1) ? H - bk * kBlk : kBlk;
        val = ld_nt(zb + (uint32_t)(bk * rowk + y * w + j));
      }
      q[i] = val;
    }
  } else if (a.blk > 0) {
    const int lg = __builtin_ctz(RPB);
    const int nbk = (H + kBlk - 1) / kBlk;
    const int total = nbk * RPB * kBlk;
    const int64_t z0 = row0 / a.prow;
    const int y0 = (int)(row0 - z0 * a.prow);
#pragma unroll
    for (int i = 0; i < XiEpt<M>::value; ++i) {
      const int f = thr + i * nthr;
      const int j = f & (kBlk - 1);
      const int r = (f / kBlk) & (RPB - 1);
      const int bk = f >> (lg + 3);
      const int k = bk * kBlk + j;
      if (f >= total || k >= H) continue;
      if (!load) {
        lds[r * RS + xrow_lds(k)] = q[i];
        continue;
      }
      float2 val = make_float2(0.f, 0.f);
      if (row0 + r < a.nrows) {
        int64_t z = z0;
        int y = y0 + r;
        if (y >= a.prow) {
          z += y / a.prow;
          y %= a.prow;
        }
        const int w = (bk == nbk - 1) ? H - bk * kBlk : kBlk;
        val = ld_nt(a.src + (z * a.prow * (int64_t)H + (int64_t)bk * kBlk * a.prow + (int64_t)y * w + j));
      }
      q[i] = val;
    }
  } else {
    const int rl = thr / T, t = thr % T;
    const int64_t row = row0 + rl;
#pragma unroll
    for (int i = 0; i < XiEpt<M>::value; ++i) {
      const int k = t + i * T;
      if (k >= H) continue;
      if (!load) lds[rl * RS + xrow_lds(k)] = q[i];
      else q[i] = row < a.nrows ? ld_nt(a.src + row * H + k) : make_float2(0.f, 0.f);
    }
  }
}

// RT: round trip.  The epilogue's output rows (zero beyond s2) are written to the
// LDS row instead of (or besides) a.dst, transformed by the forward X pass in
// LDS, and the spectrum stored to a.spec; bit-identical to an X inverse pass
// writing the rows followed by an X forward pass reading them.
// BOX: the epilogue's reduction honours a.epi.box (a separate instantiation, so
// that launches without a box carry none of its registers).
template <int M, bool RT, bool BOX = false, bool BS = false>
__device__ __forceinline__ void xi_body(const XiArgs& a, int64_t lb, int RPB, float2* lds, const float2* twm,
                                        const float2* twp, int thr) {
  constexpr int T = Geo<M, true>::T;
  const int rl = thr / T;
  const int t = thr % T;
  const int64_t row = lb * RPB + rl;
  const bool ok = row < a.nrows;
  const int s2 = a.s2;
  constexpr int RS = xrow_stride(M);
  float2* lrow = lds + rl * RS;
  auto pre = [&](int q) -> float2 {
    float2 xq = lrow[xrow_lds(q)];
    float2 xm = lrow[xrow_lds(M - q)];
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
  float* drow = a.dst ? a.dst + row * s2 : nullptr;
  const int64_t base = row * s2;
  const bool vec2 = a.vec2 != 0;
  // reduction box: whether this row's (z, y) lies in it, and its x range
  bool in_row = true;
  int bx0 = 0, bx1 = s2;
  if constexpr (BOX) {
    const int* bx = a.epi.box;
    if (bx[1] > bx[0]) {
      const int64_t z = row / a.prow;
      const int y = (int)(row - z * a.prow);
      in_row = z >= bx[0] && z < bx[1] && y >= bx[2] && y < bx[3];
      bx0 = bx[4];
      bx1 = bx[5];
    }
  }
  auto lld = [&](int e) { return lrow[xrow_lds(e)]; };
  auto lst = [&](int e, float2 v) { lrow[xrow_lds(e)] = v; };
  // the inverse transform, its output row left in LDS in natural order
  if constexpr (M == 1) {
    float2 v = make_float2(0.f, 0.f);
    if (t == 0) v = pre(0);
    __syncthreads();
    if (t == 0) lst(0, v);
  } else {
    // the first pass reads the LDS row and overwrites it: full barrier between;
    // so does the last one (LAST_LDS)
    constexpr Sched S = Geo<M, true>::F;
    constexpr int R0 = S.r[0];
    pass<M, T, R0, 1, 1>(t, twm, pre, lst, true);
    if constexpr (S.np > 1) {
      __syncthreads();
      passes<M, false, 1, 1, true, true>(t, twm, pre, lst, lld, lst, false);
    }
  }
  __syncthreads();
  if constexpr (BS) {
    // k_xi_bs: LDS rows 0..3 hold fine rows (2z', 2y'), (2z', 2y'+1), (2z'+1, 2y'),
    // (2z'+1, 2y'+1) of coarse row lb; output sample x' sums fine x = 2x', 2x'+1 (the
    // .x and .y of LDS pair x') in k_block_sum's order: z innermost, then y, then x.
    const int m2 = s2 / 2;
    float* orow = a.dst + lb * m2;
    for (int x = thr; x < m2; x += RPB * T) {
      const int sl = xrow_lds(x);
      const float2 v00 = lds[sl], v01 = lds[RS + sl], v10 = lds[2 * RS + sl], v11 = lds[3 * RS + sl];
      float r3 = 0.0f;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        float r2 = 0.0f;
        float r1 = 0.0f;
        r1 += c ? v00.y : v00.x;
        r1 += c ? v10.y : v10.x;
        r2 += r1;
        r1 = 0.0f;
        r1 += c ? v01.y : v01.x;
        r1 += c ? v11.y : v11.x;
        r2 += r1;
        r3 += r2;
      }
      __builtin_nontemporal_store(r3, orow + x);
    }
    return;
  }
  // The epilogue over this thread's sample pairs e = t, t + T, ... of its row:
  // one loop per epilogue kind, so that each kind's code appears once (not once
  // per pair of the unrolled transform) and four pairs' operand loads are in
  // flight together.  Output pairs beyond s2 (and rows past the volume) are
  // zero; a round trip writes every pair back to the LDS row.
  constexpr int XE = (M + T - 1) / T;
  auto cnt_of = [&](int q) {
    using Cnt = std::conditional_t<BOX, InBox, AllSamples>;
    Cnt c;
    if constexpr (BOX) c = InBox{in_row, bx0, bx1, q};
    return c;
  };
  auto sweep2 = [&](auto kind) {
    constexpr int K = decltype(kind)::value;
#pragma unroll 4
    for (int i = 0; i < XE; ++i) {
      const int e = t + i * T;
      if (XE * T != M && e >= M) break;
      const int q0 = 2 * e;
      float2 o = make_float2(0.f, 0.f);
      if (ok && q0 < s2) {
        o = epi_apply2<K>(a.epi, lrow[xrow_lds(e)], base + q0, acc, cnt_of(q0));
        if (!RT || drow) st_nt(reinterpret_cast<float2*>(drow + q0), o);
      }
      if constexpr (RT) lrow[xrow_lds(e)] = o;
    }
  };
  if (vec2) {
    // (s2 even: a pair lies wholly inside the row or wholly beyond it)
    switch (a.epi.kind) {
      case EPI_FLOOR: sweep2(std::integral_constant<int, EPI_FLOOR>()); break;
      case EPI_RL_RATIO: sweep2(std::integral_constant<int, EPI_RL_RATIO>()); break;
      case EPI_RL_UPDATE: sweep2(std::integral_constant<int, EPI_RL_UPDATE>()); break;
      case EPI_ADD: sweep2(std::integral_constant<int, EPI_ADD>()); break;
      case EPI_SQDIFF: sweep2(std::integral_constant<int, EPI_SQDIFF>()); break;
      default: sweep2(std::integral_constant<int, EPI_NONE>()); break;
    }
  } else {
    // odd row lengths or unaligned operands: one sample at a time
#pragma unroll 1
    for (int e = t; e < M; e += T) {
      const float2 v = lrow[xrow_lds(e)];
      const int q0 = 2 * e, q1 = 2 * e + 1;
      float2 o = make_float2(0.f, 0.f);
      if (ok) {
        if (q0 < s2) {
          o.x = epi_apply(a.epi, v.x, base + q0, acc, cnt_of(q0));
          if (!RT || drow) __builtin_nontemporal_store(o.x, drow + q0);
        }
        if (q1 < s2) {
          o.y = epi_apply(a.epi, v.y, base + q1, acc, cnt_of(q1));
          if (!RT || drow) __builtin_nontemporal_store(o.y, drow + q1);
        }
      }
      if constexpr (RT) lrow[xrow_lds(e)] = o;
    }
  }
  if constexpr (RT) {
    // fresh lane coordinates: keeps the forward phase's addresses from being
    // computed (and held in registers) during the inverse phase
    int thr2 = thr;
    opaque(thr2);
    const int t2 = thr2 % T;
    float2* lrow2 = lds + (thr2 / T) * RS;
    auto lld2 = [&](int e) { return lrow2[xrow_lds(e)]; };
    auto lst2 = [&](int e, float2 v) { lrow2[xrow_lds(e)] = v; };
    __syncthreads();
    passes<M, false, -1, 0, true, true>(t2, twm, lld2, lst2, lld2, lst2, false);
    __syncthreads();
    if (a.blk > 0) {
      int thr3 = thr;
      opaque(thr3);
      blocked_rows_t<T>(thr3, lb * RPB, RPB, a.nrows, a.H, a.prow, [&](int r, int k, int64_t off, bool valid) {
        if (valid) st_nt(a.spec + off, xf_post<M>(lds + r * RS, twp, k));
      });
    } else if (ok) {
      float2* sr = a.spec + row * a.H;
      for (int k = t2; k <= M; k += T) st_nt(sr + k, xf_post<M>(lrow2, twp, k));
    }
  }
  if (a.epi.partials) {
    __syncthreads();
    const double tot = bsum(acc, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) a.epi.partials[lb] = tot;
  }
}

// Persistent: the next tile's spectrum samples are loaded into registers while
// the current tile is transformed; twiddles in LDS as for k_xf.
template <int M, bool RT, bool BOX = false>
__global__ void __launch_bounds__(kMaxThreads) k_xi(Pack<XiArgs> pk, int RPB) {
  extern __shared__ float2 lds[];
  constexpr int RS = xrow_stride(M);
  const int64_t total = pk.prefix[pk.nvol];
  int64_t b = blockIdx.x;
  if (b >= total) return;
  const float2* twm = pk.a[0].twm;
  const float2* twp = pk.a[0].twp;
  if constexpr (!xrow_tw_global(M)) {
    float2* ltw = lds + RPB * RS;
    stage_x_tables<M>(ltw, pk.a[0].twm, pk.a[0].twp);
    twm = ltw;
    twp = ltw + M;
  }
  float2 q[XiEpt<M>::value];
  // the stepped load walk costs registers where a prefetched tile stays live
  // across the body of a long row: there the general walk keeps the occupancy
  constexpr bool kStep = RT || M <= 768;
  if constexpr (RT) {
    // the round trip holds two transforms' registers: load each tile at its start
    // rather than keeping a prefetched tile live across the body
    for (; b < total; b += gridDim.x) {
      int thr = threadIdx.x;
      opaque(thr);
      const int v = pick(pk, b);
      const XiArgs& a = pk.a[v];
      const int64_t lb = b - pk.prefix[v];
      xi_tile<M, kStep>(a, lb * RPB, RPB, q, lds, true, thr, blockDim.x);
      xi_tile<M, kStep>(a, lb * RPB, RPB, q, lds, false, thr, blockDim.x);
      __syncthreads();
      xi_body<M, RT, BOX>(a, lb, RPB, lds, twm, twp, thr);
      __syncthreads();
    }
    return;
  }
  {
    int thr = threadIdx.x;
    opaque(thr);
    const int v = pick(pk, b);
    xi_tile<M, kStep>(pk.a[v], (b - pk.prefix[v]) * RPB, RPB, q, lds, true, thr, blockDim.x);
  }
  for (;;) {
    int thr = threadIdx.x;
    opaque(thr);
    const int v = pick(pk, b);
    const XiArgs& a = pk.a[v];
    const int64_t lb = b - pk.prefix[v];
    xi_tile<M, kStep>(a, lb * RPB, RPB, q, lds, false, thr, blockDim.x);
    const int64_t bn = b + gridDim.x;
    if (bn < total) {
      int thr2 = threadIdx.x;
      opaque(thr2);
      const int vn = pick(pk, bn);
      xi_tile<M, kStep>(pk.a[vn], (bn - pk.prefix[vn]) * RPB, RPB, q, lds, true, thr2, blockDim.x);
    }
    __syncthreads();
    xi_body<M, RT>(a, lb, RPB, lds, twm, twp, thr);
    if (bn >= total) break;
    b = bn;
    __syncthreads();
  }
}

// ---- X inverse with the 2x2x2 block sum (FractionalAreaDownsample(2) after the
// convolution, its block sum core_operators.py:205-219 fused): tile lb is coarse row
// (z', y') = (lb / (prow/2), lb % (prow/2)); the first half of the workgroup loads fine
// rows (2z', 2y'), (2z', 2y'+1) into LDS rows 0-1, the second half (2z'+1, 2y'),
// (2z'+1, 2y'+1) into rows 2-3 (prow and the plane count even).  Persistent, with the
// next tile's spectrum prefetched into registers as in k_xi.
template <int M>
__global__ void __launch_bounds__(kMaxThreads) k_xi_bs(Pack<XiArgs> pk, int RPB) {
  extern __shared__ float2 lds[];
  constexpr int T = Geo<M, true>::T;
  constexpr int RS = xrow_stride(M);
  const int64_t total = pk.prefix[pk.nvol];
  int64_t b = blockIdx.x;
  if (b >= total) return;
  const float2* twm = pk.a[0].twm;
  const float2* twp = pk.a[0].twp;
  if constexpr (!xrow_tw_global(M)) {
    float2* ltw = lds + RPB * RS;
    stage_x_tables<M>(ltw, pk.a[0].twm, pk.a[0].twp);
    twm = ltw;
    twp = ltw + M;
  }
  float2 q[XiEpt<M>::value];
  constexpr bool kStep = M <= 768;
  const int half = threadIdx.x >= 2 * T ? 1 : 0;
  auto tile = [&](int64_t bb, bool load) {
    int thr = threadIdx.x - half * 2 * T;
    opaque(thr);
    const int v = pick(pk, bb);
    const XiArgs& a = pk.a[v];
    const int64_t lb = bb - pk.prefix[v];
    const int64_t cy = a.prow / 2;
    const int64_t zc = lb / cy, yc = lb - zc * cy;
    xi_tile<M, kStep>(a, (2 * zc + half) * a.prow + 2 * yc, 2, q, lds + half * 2 * RS, load, thr, 2 * T);
  };
  tile(b, true);
  for (;;) {
    tile(b, false);
    const int64_t bn = b + gridDim.x;
    if (bn < total) tile(bn, true);
    __syncthreads();
    int thr = threadIdx.x;
    opaque(thr);
    const int v = pick(pk, b);
    xi_body<M, false, false, true>(pk.a[v], b - pk.prefix[v], 4, lds, twm, twp, thr);
    if (bn >= total) break;
    b = bn;
    __syncthreads();
  }
}

// ---- host launchers (explicitly instantiated per length) ----
// Compute units of the current device.
inline int device_cus() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (cache[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 1;
    cache[dev] = n;
  }
  return cache[dev];
}

// Workgroups of a persistent launch: as many as the device holds at once
// (occupancy x compute units), at most nblocks.
inline unsigned persistent_grid(const void* kernel, int threads, size_t shm, int64_t nblocks) {
  struct Entry {
    const void* k;
    int threads;
    size_t shm;
    int per;
  };
  static std::mutex mu;
  static Entry cache[128];
  static int n = 0;
  int per = 0;
  {
    std::lock_guard<std::mutex> g(mu);
    for (int i = 0; i < n && !per; ++i)
      if (cache[i].k == kernel && cache[i].threads == threads && cache[i].shm == shm) per = cache[i].per;
    if (!per) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, shm) != hipSuccess || per < 1) per = 1;
      if (n < 128) cache[n++] = Entry{kernel, threads, shm, per};
    }
  }
  const int64_t cap = (int64_t)per * device_cus();
  return (unsigned)(nblocks < cap ? nblocks : cap);
}

// ---- PSF synthesis: |IFFT2(P exp(2 pi i kz z))|^2 per z-plane ----
// Y pass: L consecutive columns of one z-plane per workgroup.  The first pass
// loads the pupil and applies the defocus phase, reduced to [0, 1) turns in
// float64 before sincospi; the inverse transform along y is stored to work.
template <int N>
__global__ void __launch_bounds__(kMaxThreads) k_psf_y(Pack<PsfYArgs> pk, int L) {
  extern __shared__ float2 lds[];
  constexpr int T = Geo<N>::T;
  const int ln = threadIdx.x % L;
  const int t = threadIdx.x / L;
  const int64_t total = pk.prefix[pk.nvol];
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int v = pick(pk, b);
    const PsfYArgs& a = pk.a[v];
    const int64_t lb = b - pk.prefix[v];
    const int64_t zi = lb / a.ntiles;
    const int64_t c = (lb - zi * a.ntiles) * L + ln;
    const bool ok = c < a.nx;
    const double zz = a.z[zi];
    auto gld = [&](int e) -> float2 {
      if (!ok) return make_float2(0.f, 0.f);
      const int64_t i = (int64_t)e * a.nx + c;
      const double turns = a.kz[i] * zz;
      float sn, cs;
      sincospif(2.0f * (float)(turns - floor(turns)), &sn, &cs);
      return cmul(a.pupil[i], make_float2(cs, sn));
    };
    auto gst = [&](int e, float2 val) {
      if (ok) st_nt(a.work + (zi * N + e) * a.nx + c, val);
    };
    auto lld = [&](int e) { return lds[e * L + ln]; };
    auto lst = [&](int e, float2 val) { lds[e * L + ln] = val; };
    if constexpr (N == 1) {
      if (t == 0) gst(0, gld(0));
    } else {
      passes<N, false, 1, 0>(t, a.tw, gld, gst, lld, lst, true);
    }
    __syncthreads();
  }
}

// X pass: R rows per workgroup; the last pass stores (or, for a later channel
// of a vectorial PSF, adds) |v|^2 * scale and the workgroup's float64 sum of
// them goes to partials[tile].
template <int N>
__global__ void __launch_bounds__(kMaxThreads) k_psf_x(Pack<PsfXArgs> pk, int R) {
  extern __shared__ float2 lds[];
  constexpr int T = Geo<N>::T;
  constexpr int RS = xrow_stride(N);
  const int rl = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int64_t total = pk.prefix[pk.nvol];
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int v = pick(pk, b);
    const PsfXArgs& a = pk.a[v];
    const int64_t lb = b - pk.prefix[v];
    const int64_t row = lb * R + rl;
    const bool ok = row < a.nrows;
    const float2* src = a.work + row * N;
    float* dst = a.psf + row * N;
    float2* lrow = lds + rl * RS;
    double acc = 0.0;
    auto gld = [&](int e) { return ok ? ld_nt(src + e) : make_float2(0.f, 0.f); };
    auto gst = [&](int e, float2 val) {
      if (ok) {
        const float p = (val.x * val.x + val.y * val.y) * a.scale;
        if (a.accumulate) dst[e] += p;
        else __builtin_nontemporal_store(p, dst + e);
        acc += (double)p;
      }
    };
    auto lld = [&](int e) { return lrow[xrow_lds(e)]; };
    auto lst = [&](int e, float2 val) { lrow[xrow_lds(e)] = val; };
    if constexpr (N == 1) {
      if (t == 0) gst(0, gld(0));
    } else {
      passes<N, false, 1, 0>(t, a.tw, gld, gst, lld, lst, true);
    }
    __syncthreads();
    const double tot = bsum(acc, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) a.partials[lb] = tot;
    __syncthreads();
  }
}

template <int N>
struct Launch {
  static constexpr int T = Geo<N>::T;         // column / Z passes
  static constexpr int TX = Geo<N, true>::T;  // X passes (row schedule)
  // one workgroup per tile, or a persistent grid
  static unsigned grid(const void* k, int threads, size_t shm, int64_t nblocks, bool persist) {
    return persist ? persistent_grid(k, threads, shm, nblocks) : (unsigned)nblocks;
  }
  template <int DIR, bool NT, bool WIDE = false>
  static void col_as(const Pack<ColArgs>& pk, int L, int64_t nblocks, bool persist, size_t shm, hipStream_t st) {
    const void* k = (const void*)k_col<N, DIR, NT, WIDE>;
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_col<N, DIR, NT, WIDE>), dim3(grid(k, L * T, shm, nblocks, persist)), dim3(L * T), shm, st,
                       pk, L);
  }
  // blocked columns in tiles narrower than a block: ordinary accesses (ld_s/st_s).
  // Planes too large for 32-bit offsets (the OTF build's Z transform of a very large
  // canvas: a forward transform of plain columns) take the 64-bit kernel; other
  // shapes that large are not supported.
  static hipError_t col(int dir, const Pack<ColArgs>& pk, int L, int64_t nblocks, bool persist, hipStream_t st) {
    const size_t shm = (size_t)col_lds_elems(N, L) * sizeof(float2);
    const bool split = pk.a[0].blk > 0 && L < kBlk;
    bool narrow = true;
    for (int i = 0; i < pk.nvol; ++i) narrow = narrow && col_narrow(pk.a[i], N);
    if (!narrow) {
      if (dir >= 0 || split) return hipErrorInvalidValue;
      col_as<-1, true, true>(pk, L, nblocks, persist, shm, st);
      return hipGetLastError();
    }
    if (dir < 0) {
      if (split) col_as<-1, false>(pk, L, nblocks, persist, shm, st);
      else col_as<-1, true>(pk, L, nblocks, persist, shm, st);
    } else {
      if (split) col_as<1, false>(pk, L, nblocks, persist, shm, st);
      else col_as<1, true>(pk, L, nblocks, persist, shm, st);
    }
    return hipGetLastError();
  }
  static hipError_t z(const Pack<ZArgs>& pk, int L, int64_t nblocks, bool persist, hipStream_t st) {
    const size_t shm = (size_t)N * L * sizeof(float2);
    const void* k = (const void*)k_z<N>;
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_z<N>), dim3(grid(k, L * T, shm, nblocks, persist)), dim3(L * T), shm, st, pk, L);
    return hipGetLastError();
  }
  static hipError_t xf(const Pack<XfArgs>& pk, int R, int64_t nblocks, bool persist, hipStream_t st) {
    const size_t shm = (size_t)xrow_lds_elems(N, R) * sizeof(float2);
    int ext = 0;
    for (int i = 0; i < pk.nvol; ++i)
      ext = std::max(ext, pk.a[i].pro.kind == PRO_PD_Y1 ? 2 : (pk.a[i].pro.kind == PRO_SPREAD ? 1 : 0));
    const void* k = ext == 2 ? (const void*)k_xf<N, 2> : ext == 1 ? (const void*)k_xf<N, 1> : (const void*)k_xf<N, 0>;
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    const dim3 g(grid(k, R * TX, shm, nblocks, persist)), blk(R * TX);
    if (ext == 2) hipLaunchKernelGGL((k_xf<N, 2>), g, blk, shm, st, pk, R);
    else if (ext == 1) hipLaunchKernelGGL((k_xf<N, 1>), g, blk, shm, st, pk, R);
    else hipLaunchKernelGGL((k_xf<N, 0>), g, blk, shm, st, pk, R);
    return hipGetLastError();
  }
  static hipError_t xi(const Pack<XiArgs>& pk, int R, int64_t nblocks, bool persist, hipStream_t st) {
    size_t shm = (size_t)xrow_lds_elems(N, R) * sizeof(float2);
    if (shm < 16 * sizeof(double)) shm = 16 * sizeof(double);
    if (pk.a[0].epi.kind == EPI_BLOCKSUM2) {  // one volume per launch, four rows per tile
      if (R != 4) return hipErrorInvalidValue;
      const void* kb = (const void*)k_xi_bs<N>;
      (void)hipFuncSetAttribute(kb, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      hipLaunchKernelGGL((k_xi_bs<N>), dim3(grid(kb, R * TX, shm, nblocks, persist)), dim3(R * TX), shm, st, pk, R);
      return hipGetLastError();
    }
    const void* k = (const void*)k_xi<N, false>;
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_xi<N, false>), dim3(grid(k, R * TX, shm, nblocks, persist)), dim3(R * TX), shm, st, pk, R);
    return hipGetLastError();
  }
  static hipError_t psfy(const Pack<PsfYArgs>& pk, int L, int64_t nblocks, hipStream_t st) {
    const size_t shm = (size_t)N * L * sizeof(float2);
    const void* k = (const void*)k_psf_y<N>;
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_psf_y<N>), dim3((unsigned)nblocks), dim3(L * T), shm, st, pk, L);
    return hipGetLastError();
  }
  static hipError_t psfx(const Pack<PsfXArgs>& pk, int R, int64_t nblocks, hipStream_t st) {
    size_t shm = (size_t)R * xrow_stride(N) * sizeof(float2);
    if (shm < 16 * sizeof(double)) shm = 16 * sizeof(double);
    const void* k = (const void*)k_psf_x<N>;
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_psf_x<N>), dim3((unsigned)nblocks), dim3(R * T), shm, st, pk, R);
    return hipGetLastError();
  }
  // the boxed instantiation when any volume of the pack has a reduction box
  static hipError_t xr(const Pack<XiArgs>& pk, int R, int64_t nblocks, bool persist, hipStream_t st) {
    size_t shm = (size_t)xrow_lds_elems(N, R) * sizeof(float2);
    if (shm < 16 * sizeof(double)) shm = 16 * sizeof(double);
    bool boxed = false;
    for (int v = 0; v < pk.nvol; ++v) boxed = boxed || pk.a[v].epi.box[1] > pk.a[v].epi.box[0];
    const void* k = boxed ? (const void*)k_xi<N, true, true> : (const void*)k_xi<N, true, false>;
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (boxed)
      hipLaunchKernelGGL((k_xi<N, true, true>), dim3(grid(k, R * TX, shm, nblocks, persist)), dim3(R * TX), shm, st,
                         pk, R);
    else
      hipLaunchKernelGGL((k_xi<N, true, false>), dim3(grid(k, R * TX, shm, nblocks, persist)), dim3(R * TX), shm, st,
                         pk, R);
    return hipGetLastError();
  }
};

}  // namespace ct
}  // namespace dh
