// C ABI of libdeconhip.so: plans, OTF construction, one convolution
// application, and the RL elementwise / reduction entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/deconhip.h"
#include "dh_fft.hpp"
#include "dh_fftct.hpp"
#include "dh_ops.hpp"
#include "dh_erdecon.hpp"
#include "dh_jetnewton.hpp"
#include "dh_nlcg.hpp"
#include "dh_pdhg.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char* where) {
  if (e != hipSuccess) return fail(DH_EHIP, std::string(where) + ": " + hipGetErrorString(e));
  return DH_OK;
}

#define DH_TRY(expr)            \
  do {                          \
    int _rc = (expr);           \
    if (_rc != DH_OK) return _rc; \
  } while (0)

bool is_smooth(int n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

int next_smooth(int n) {
  while (!is_smooth(n)) ++n;
  return n;
}

// Radix schedule for a 5-smooth n: 16s first, then 8/4/2, then 5s and 3s.
bool make_sched(int n, dh::AxisSched& f, dh::AxisSched& iv) {
  if (!is_smooth(n)) return false;
  int m = n, c2 = 0, c3 = 0, c5 = 0;
  while (m % 2 == 0) m /= 2, ++c2;
  while (m % 3 == 0) m /= 3, ++c3;
  while (m % 5 == 0) m /= 5, ++c5;
  std::vector<int> r;
  int n16 = c2 / 4, rem = c2 % 4;
  if (rem == 1 && n16 > 0) {
    --n16;
    for (int i = 0; i < n16; ++i) r.push_back(16);
    r.push_back(8);
    r.push_back(4);
  } else {
    for (int i = 0; i < n16; ++i) r.push_back(16);
    if (rem == 3) r.push_back(8);
    if (rem == 2) r.push_back(4);
    if (rem == 1) r.push_back(2);
  }
  for (int i = 0; i < c5; ++i) r.push_back(5);
  for (int i = 0; i < c3; ++i) r.push_back(3);
  if ((int)r.size() > dh::kMaxPasses) return false;
  f.n = iv.n = n;
  f.np = iv.np = (int)r.size();
  int sp = 1;
  for (size_t p = 0; p < r.size(); ++p) {
    f.radix[p] = r[p];
    f.span[p] = sp;
    sp *= r[p];
  }
  sp = 1;
  for (size_t p = 0; p < r.size(); ++p) {
    iv.radix[p] = r[r.size() - 1 - p];
    iv.span[p] = sp;
    sp *= iv.radix[p];
  }
  return true;
}

void push_table(std::vector<float2>& tw, int n, int count) {
  for (int m = 0; m < count; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)n;
    tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
  }
}

// Per-pass twiddle rows of the specialised kernels for schedule s: pass q
// (radix R, span P > 1) gets tw_pows(R) rows of P entries, row j entry k =
// exp(-2*pi*i * 2^j * k / (P*R)) (the layout dh::ct::twiddle reads).
void push_pass_tables(std::vector<float2>& tw, const dh::ct::Sched& s) {
  int span = 1;
  for (int q = 0; q < s.np; ++q) {
    const int R = s.r[q];
    if (span > 1)
      for (int j = 0; j < dh::ct::tw_pows(R); ++j)
        for (int k = 0; k < span; ++k) {
          const double a = -2.0 * M_PI * (double)((1 << j) * k) / (double)(span * R);
          tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
        }
    span *= R;
  }
}

// Per-pass event timing (measurement only; off unless dh_timing_enable(1)).
struct TimedPass {
  int pass;
  hipEvent_t a, b;
};
bool g_timing = false;
int g_timing_mask = (1 << DH_PASS_COUNT) - 1;
std::vector<TimedPass> g_pending;
std::vector<hipEvent_t> g_free;  // read events, reused
double g_ms[DH_PASS_COUNT] = {0};
int64_t g_cnt[DH_PASS_COUNT] = {0};

// Timing events record without the system-scope fence: a default event's record
// writes back and invalidates the caches, which stalls the stream between the
// passes it brackets and slows the next pass (the elapsed time between two events
// needs no fence once the host has synchronised on the later one).
int timing_event(hipEvent_t* e) {
  if (!g_free.empty()) {
    *e = g_free.back();
    g_free.pop_back();
    return DH_OK;
  }
  return hip_check(hipEventCreateWithFlags(e, hipEventDisableSystemFence), "event create");
}

bool g_open = false;  // timing_begin recorded the opening event of the current pass

int timing_begin(int pass, hipStream_t st) {
  g_open = g_timing && (g_timing_mask >> pass & 1);
  if (!g_open) return DH_OK;
  TimedPass t{pass, nullptr, nullptr};
  DH_TRY(timing_event(&t.a));
  DH_TRY(timing_event(&t.b));
  DH_TRY(hip_check(hipEventRecord(t.a, st), "event record"));
  g_pending.push_back(t);
  return DH_OK;
}

int timing_end(hipStream_t st) {
  if (!g_open) return DH_OK;
  g_open = false;
  return hip_check(hipEventRecord(g_pending.back().b, st), "event record");
}

}  // namespace

namespace dh {
int set_last_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace dh

struct dh_conv_plan {
  int ndim = 3;
  int circular = 0;
  int s[3], k[3], fft[3], P[3];
  int H = 0, M = 0;
  int klo[3], khi[3];
  dh::AxisSched zf, zi, yf, yi, xf, xi;
  int64_t tw_z = 0, tw_y = 0, tw_xm = 0, tw_xp = 0;
  std::vector<float2> tw_host;
  const float2* tw_dev = nullptr;
  int64_t b1_off = 0, b2_off = 0, part_off = 0, scal_off = 0, ws_bytes = 0;
  int64_t n_partials = 0;
  int tile[DH_PASS_COUNT] = {0, 0, 0, 0, 0};
  bool ct = false;  // every axis length has compile-time specialised kernels
  hipEvent_t bound = nullptr;  // completion of the twiddle upload (tw_host must outlive it)
};

namespace {

// ---- launch shapes of the specialised kernels ----
constexpr int kCtMaxThreads = 512;
constexpr int64_t kCtDefaultLds = 64 * 1024;
constexpr int64_t kCtRowDefaultLds = 96 * 1024;
constexpr int64_t kCtMaxLds = 160 * 1024;

// Column kernels: the fused Z pass (LDS n*L), a plain-layout column pass and a
// blocked-layout (Y) column pass, whose tile is at most one block or whole blocks.
enum ColKind { COL_Z = 0, COL_PLAIN = 1, COL_BLOCKED = 2 };

int64_t col_lds_bytes(int n, int L, int kind) {
  return 8 * (kind == COL_Z ? (int64_t)n * L : dh::col_lds_elems(n, L));
}
bool ct_col_tile_ok(int n, int L, int kind) {
  if (L < 1 || L * dh::ct_threads(n) > kCtMaxThreads || col_lds_bytes(n, L, kind) > kCtMaxLds) return false;
  return kind != COL_BLOCKED || L <= dh::kBlk || L % dh::kBlk == 0;
}
// rows per workgroup: a power of two (the blocked row walk relies on it)
bool ct_row_tile_ok(int m, int R) {
  return R >= 1 && (R & (R - 1)) == 0 && R * dh::ct_row_threads(m) <= kCtMaxThreads &&
         dh::xrow_lds_elems(m, R) * 8 <= kCtMaxLds;
}
// Default launch shapes (what the autotuner settles on for the 64x512x512
// workload, canvas 128x1024x1024): columns per workgroup = the widest power of
// two within 64 KiB LDS and 512 threads (256 for the fused Z pass unless that
// leaves it fewer than 16 columns, as for 512-point columns), one whole
// block for the blocked Y layout whenever it fits 160 KiB and 512 threads; rows per workgroup = the widest power of two within 512
// threads and 96 KiB LDS, halved (down to one wave) while a volume has fewer
// than kMinRowTiles tiles: an X pass keeps a tile's load, transforms and stores
// in sequence, so small volumes need many workgroups in flight to hide latency.
int ct_col_tile(int n, int64_t ncols, int over, int kind) {
  if (over > 0 && ct_col_tile_ok(n, over, kind)) return over;
  const int T = dh::ct_threads(n);
  const int cap = kind == COL_BLOCKED ? dh::kBlk : 64;
  const int max_threads = kind == COL_Z ? 256 : kCtMaxThreads;
  // a blocked pass takes a whole block when it fits at all: narrower tiles
  // split each block's 64-byte row runs between workgroups
  const int64_t lds_cap = kind == COL_BLOCKED ? kCtMaxLds : kCtDefaultLds;
  int L = 1;
  while (L < cap && 2 * L * T <= max_threads && col_lds_bytes(n, 2 * L, kind) <= lds_cap) L *= 2;
  // the fused Z pass reads each plane's L columns as one run: below 16 columns
  // (128-byte runs, long columns) it takes up to the full thread limit
  if (kind == COL_Z)
    while (L < 16 && 2 * L * T <= kCtMaxThreads && col_lds_bytes(n, 2 * L, kind) <= lds_cap) L *= 2;
  while (L > 1 && L / 2 >= ncols) L /= 2;
  return L;
}
constexpr int64_t kMinRowTiles = 2048;  // 8 workgroups per CU
int ct_row_tile(int m, int over, int64_t nrows) {
  if (over > 0 && ct_row_tile_ok(m, over)) return over;
  const int T = dh::ct_row_threads(m);
  int R = 1;
  while (R < 64 && 2 * R * T <= kCtMaxThreads && dh::xrow_lds_elems(m, 2 * R) * 8 <= kCtRowDefaultLds) R *= 2;
  const int r_min = T >= 64 ? 1 : 64 / T;
  while (R > r_min && (nrows + R - 1) / R < kMinRowTiles) R /= 2;
  return R;
}

static_assert(DH_TILE_PERSISTENT == dh::kTilePersistent, "tile flag mismatch");

// Tile override of a pass (0 = heuristic) and whether it launches persistently
// (by default the X forward: one-time per-workgroup set-up, the twiddle tables in
// LDS.  The X inverse and round trip launch one workgroup per tile: measured 6 %
// faster at 512-point rows and 11 % at 2048-point rows than the persistent grid).
int tile_override(const dh_conv_plan* p, int pass) { return p->tile[pass] & dh::kTileMask; }
int persist_flag(const dh_conv_plan* p, int pass) {
  const int v = p->tile[pass];
  const bool on = v > 0 ? (v & dh::kTilePersistent) != 0 : pass == DH_PASS_X_FWD;
  return on ? dh::kTilePersistent : 0;
}

bool aligned8(const void* p) { return ((uintptr_t)p & 7) == 0; }
// float2 row access is valid when every row starts on an 8-byte boundary
int xf_vec2(const float* src, int rowlen) { return (src && rowlen % 2 == 0 && aligned8(src)) ? 1 : 0; }
int xi_vec2(const float* dst, int s2, const dh::Epi& e) {
  return (s2 % 2 == 0 && aligned8(dst) && aligned8(e.aux0) && aligned8(e.aux1)) ? 1 : 0;
}

template <class A>
void pack_one(dh::Pack<A>& pk, const A& a, int64_t nblk) {
  pk.a[0] = a;
  pk.prefix[0] = 0;
  pk.prefix[1] = nblk;
  pk.nvol = 1;
}

// The five pass descriptors of one application with the specialised kernels,
// with the launch tile and workgroup count of each.  The X/Y spectra in the
// workspace use the blocked layout.
struct CtPasses {
  dh::XfArgs xf;
  dh::ColArgs yf, yi;
  dh::ZArgs z;
  dh::XiArgs xi;
  int t_xf = 0, t_yf = 0, t_z = 0, t_yi = 0, t_xi = 0;
  int64_t n_xf = 0, n_yf = 0, n_z = 0, n_yi = 0, n_xi = 0;
  bool has_y = false;
};

CtPasses ct_passes(const dh_conv_plan* p, int adjoint, const float* in, float* out, const void* otf, char* ws,
                   const dh::Pro& pro, const dh::Epi& epi) {
  CtPasses c;
  float2* B1 = (float2*)(ws + p->b1_off);
  float2* B2 = (float2*)(ws + p->b2_off);
  const int s0 = p->s[0], s1 = p->s[1], s2 = p->s[2], P1 = p->P[1];
  const int64_t H = p->H, nrows = (int64_t)s0 * s1;
  const float2* tw = p->tw_dev;
  c.t_xf = ct_row_tile(p->M, tile_override(p, DH_PASS_X_FWD), nrows);
  c.n_xf = (nrows + c.t_xf - 1) / c.t_xf;
  c.t_xf |= persist_flag(p, DH_PASS_X_FWD);
  int vec2 = xf_vec2(in, s2);
  if (pro.kind == dh::PRO_PD_Y1)
    vec2 = vec2 && aligned8(pro.ax) && aligned8(pro.axp) && aligned8(pro.obs) && aligned8(pro.y1w);
  c.xf = dh::XfArgs{in, B1, tw + p->tw_xm, tw + p->tw_xp, nrows, s2, s2, 0, (int)H, pro, vec2, dh::kBlk, s1};
  c.has_y = P1 > 1;
  if (c.has_y) {
    c.t_yf = ct_col_tile(P1, H, tile_override(p, DH_PASS_Y_FWD), COL_BLOCKED);
    const int64_t nt = (H + c.t_yf - 1) / c.t_yf;
    c.n_yf = nt * s0;
    c.t_yf |= persist_flag(p, DH_PASS_Y_FWD);
    c.yf = dh::ColArgs{B1, B2, tw + p->tw_y, H, (int64_t)s1 * H, (int64_t)P1 * H, s1, P1, nt, s1, 0, P1, 1.0f,
                       dh::kBlk};
    c.t_yi = ct_col_tile(P1, H, tile_override(p, DH_PASS_Y_INV), COL_BLOCKED);
    const int64_t nti = (H + c.t_yi - 1) / c.t_yi;
    c.n_yi = nti * s0;
    c.t_yi |= persist_flag(p, DH_PASS_Y_INV);
    c.yi = dh::ColArgs{B2, B1, tw + p->tw_y, H, (int64_t)P1 * H, (int64_t)s1 * H, P1, s1, nti, P1, 0, s1, 1.0f,
                       dh::kBlk};
  }
  const int64_t ncols = (int64_t)P1 * H;
  c.t_z = p->P[0] == 1 ? 256 : ct_col_tile(p->P[0], ncols, tile_override(p, DH_PASS_Z), COL_Z);
  c.n_z = (ncols + c.t_z - 1) / c.t_z;
  c.t_z |= persist_flag(p, DH_PASS_Z);
  c.z = dh::ZArgs{c.has_y ? B2 : B1, (const float2*)otf, tw + p->tw_z, ncols, s0, s0, adjoint ? 1 : 0};
  if (epi.kind == dh::EPI_BLOCKSUM2) {  // k_xi_bs: four fine rows per coarse row, persistent
    c.t_xi = 4;
    c.n_xi = nrows / 4;
    c.t_xi |= dh::kTilePersistent;
  } else {
    c.t_xi = ct_row_tile(p->M, tile_override(p, DH_PASS_X_INV), nrows);
    c.n_xi = (nrows + c.t_xi - 1) / c.t_xi;
    c.t_xi |= persist_flag(p, DH_PASS_X_INV);
  }
  c.xi = dh::XiArgs{B1, out, nullptr, tw + p->tw_xm, tw + p->tw_xp, nrows, s2, (int)H, epi, xi_vec2(out, s2, epi),
                    dh::kBlk, s1};
  return c;
}

template <class A>
int launch_one(int kind, int n, int dir, const A& a, int tile, int64_t nb, int pass, const char* what,
               hipStream_t st) {
  dh::Pack<A> pk;
  pack_one(pk, a, nb);
  DH_TRY(timing_begin(pass, st));
  DH_TRY(hip_check(dh::ct_launch(kind, n, dir, &pk, tile, nb, st), what));
  return timing_end(st);
}

// One application with the specialised kernels.
int ct_apply(const dh_conv_plan* p, int adjoint, const float* in, float* out, const void* otf, char* ws,
             const dh::Pro& pro, const dh::Epi& epi, int64_t* xi_blocks, hipStream_t st) {
  const CtPasses c = ct_passes(p, adjoint, in, out, otf, ws, pro, epi);
  DH_TRY(launch_one(dh::CT_XF, p->M, -1, c.xf, c.t_xf, c.n_xf, DH_PASS_X_FWD, "x forward pass", st));
  if (c.has_y) DH_TRY(launch_one(dh::CT_COL, p->P[1], -1, c.yf, c.t_yf, c.n_yf, DH_PASS_Y_FWD, "y forward pass", st));
  DH_TRY(launch_one(dh::CT_Z, p->P[0], 0, c.z, c.t_z, c.n_z, DH_PASS_Z, "z pass", st));
  if (c.has_y) DH_TRY(launch_one(dh::CT_COL, p->P[1], 1, c.yi, c.t_yi, c.n_yi, DH_PASS_Y_INV, "y inverse pass", st));
  DH_TRY(launch_one(dh::CT_XI, p->M, 1, c.xi, c.t_xi, c.n_xi, DH_PASS_X_INV, "x inverse pass", st));
  *xi_blocks = c.n_xi;
  return DH_OK;
}

// One fused RL iteration with the specialised kernels (see dh_rl_iteration).
int ct_rl_iteration(const dh_conv_plan* p, const void* otf, char* ws, float* x, const float* obs, const float* sens,
                    float bg, float eps, double* reduce_out, int spectrum_ready, int keep_spectrum, hipStream_t st) {
  float2* B1 = (float2*)(ws + p->b1_off);
  const dh::Pro none{dh::PRO_NONE, 0.f, 0.f, nullptr};
  dh::Epi ratio{dh::EPI_RL_RATIO, bg, eps, obs, nullptr, nullptr};
  if (reduce_out) ratio.partials = (double*)(ws + p->part_off);
  const dh::Epi update{dh::EPI_RL_UPDATE, 0.f, eps, x, sens, nullptr};
  // forward convolution of x; its inverse X pass computes the ratio and runs the
  // adjoint's forward X pass on it (the ratio never reaches HBM)
  CtPasses f = ct_passes(p, 0, x, nullptr, otf, ws, none, ratio);
  f.xi.spec = B1;
  if (!spectrum_ready) DH_TRY(launch_one(dh::CT_XF, p->M, -1, f.xf, f.t_xf, f.n_xf, DH_PASS_X_FWD, "x forward pass", st));
  if (f.has_y) DH_TRY(launch_one(dh::CT_COL, p->P[1], -1, f.yf, f.t_yf, f.n_yf, DH_PASS_Y_FWD, "y forward pass", st));
  DH_TRY(launch_one(dh::CT_Z, p->P[0], 0, f.z, f.t_z, f.n_z, DH_PASS_Z, "z pass", st));
  if (f.has_y) DH_TRY(launch_one(dh::CT_COL, p->P[1], 1, f.yi, f.t_yi, f.n_yi, DH_PASS_Y_INV, "y inverse pass", st));
  if (reduce_out && f.n_xi > p->n_partials) return fail(DH_EWORKSPACE, "partials area too small");
  DH_TRY(launch_one(dh::CT_XR, p->M, 1, f.xi, f.t_xi, f.n_xi, DH_PASS_X_INV, "x round trip (ratio)", st));
  // adjoint (correlation) of the ratio; its inverse X pass applies the update in
  // place and, when asked, leaves the forward X spectrum of the new x in B1
  CtPasses a = ct_passes(p, 1, nullptr, x, otf, ws, none, update);
  if (a.has_y) DH_TRY(launch_one(dh::CT_COL, p->P[1], -1, a.yf, a.t_yf, a.n_yf, DH_PASS_Y_FWD, "y forward pass", st));
  DH_TRY(launch_one(dh::CT_Z, p->P[0], 0, a.z, a.t_z, a.n_z, DH_PASS_Z, "z pass", st));
  if (a.has_y) DH_TRY(launch_one(dh::CT_COL, p->P[1], 1, a.yi, a.t_yi, a.n_yi, DH_PASS_Y_INV, "y inverse pass", st));
  if (keep_spectrum) {
    a.xi.spec = B1;
    DH_TRY(launch_one(dh::CT_XR, p->M, 1, a.xi, a.t_xi, a.n_xi, DH_PASS_X_INV, "x round trip (update)", st));
  } else {
    DH_TRY(launch_one(dh::CT_XI, p->M, 1, a.xi, a.t_xi, a.n_xi, DH_PASS_X_INV, "x inverse pass", st));
  }
  if (reduce_out)
    DH_TRY(hip_check(dh::launch_sum_partials(ratio.partials, f.n_xi, reduce_out, st), "epilogue reduction"));
  return DH_OK;
}

}  // namespace

extern "C" {

int dh_abi_version(void) { return DH_ABI_VERSION; }

int dh_rl_iteration(const dh_conv_plan* p, const void* otf_dev, void* work_dev, int64_t work_bytes, float* x,
                    const float* obs, const float* sens, float background, float eps, double* reduce_out,
                    int spectrum_ready, int keep_spectrum, void* stream) {
  if (!p || !otf_dev || !work_dev || !x || !obs || !sens) return fail(DH_EINVAL, "null argument");
  if (!p->tw_dev) return fail(DH_EINVAL, "plan is not bound (dh_conv_plan_bind)");
  if (work_bytes < p->ws_bytes) return fail(DH_EWORKSPACE, "workspace too small");
  if (!p->ct) return fail(DH_EUNSUPPORTED, "the fused RL iteration needs a canvas with specialised kernels");
  return ct_rl_iteration(p, otf_dev, (char*)work_dev, x, obs, sens, background, eps, reduce_out, spectrum_ready,
                         keep_spectrum, (hipStream_t)stream);
}

const char* dh_last_error(void) { return g_err.c_str(); }

int dh_conv_plan_create(int ndim, const int64_t* signal_shape, const int64_t* kernel_shape,
                        const int64_t* fft_shape, int circular, dh_conv_plan** out) {
  if (!out || !signal_shape || !kernel_shape) return fail(DH_EINVAL, "null argument");
  if (ndim < 1 || ndim > 3) return fail(DH_EINVAL, "ndim must be 1, 2 or 3");
  auto* p = new dh_conv_plan();
  p->ndim = ndim;
  p->circular = circular ? 1 : 0;
  const int off = 3 - ndim;
  for (int a = 0; a < 3; ++a) {
    p->s[a] = p->k[a] = p->fft[a] = 1;
  }
  for (int a = 0; a < ndim; ++a) {
    if (signal_shape[a] < 1 || kernel_shape[a] < 1 || signal_shape[a] > (1 << 20) || kernel_shape[a] > (1 << 20)) {
      delete p;
      return fail(DH_EINVAL, "shape entries must be in [1, 2^20]");
    }
    p->s[off + a] = (int)signal_shape[a];
    p->k[off + a] = (int)kernel_shape[a];
  }
  for (int a = 0; a < 3; ++a) {
    int minimum = p->s[a] + p->k[a] - 1;
    int f;
    if (fft_shape && a >= off) {
      f = (int)fft_shape[a - off];
    } else if (circular) {
      f = p->s[a];
    } else {
      f = next_smooth(std::max(p->s[a] + p->k[a] - 1, p->k[a]));
    }
    if (circular) {
      if (p->s[a] != p->k[a] || f != p->s[a]) {
        delete p;
        return fail(DH_EINVAL, "circular plan needs equal signal, kernel and fft shapes");
      }
    } else if (f < minimum) {
      delete p;
      return fail(DH_EINVAL, "fft_shape is too small for linear convolution; minimum is n + m - 1 per axis");
    }
    p->fft[a] = f;
  }
  // Canvas this plan computes on: any canvas >= n + m - 1 gives the same
  // linear convolution.  Prefer a length with compile-time specialised
  // kernels (at most 1.5x the 5-smooth minimum), else the next 5-smooth
  // length; the x extent must be even (real-to-complex pass).
  for (int a = 0; a < 3; ++a) {
    int c = p->fft[a];
    if (!circular) {
      int sm = next_smooth(c);
      if (a == 2) {
        while (sm % 2 != 0 || !is_smooth(sm)) ++sm;
      }
      int best = sm;
      for (int cand = c; cand <= sm + sm / 2; ++cand) {
        const bool ok = (a == 2) ? (cand % 2 == 0 && dh::ct_supported(cand / 2)) : dh::ct_supported(cand);
        if (ok) {
          best = cand;
          break;
        }
      }
      c = best;
    } else if (!is_smooth(c) || (a == 2 && c % 2 != 0 && c != 1)) {
      delete p;
      return fail(DH_EUNSUPPORTED, "circular plan needs 5-smooth extents and an even last axis");
    }
    p->P[a] = c;
  }
  if (p->P[2] == 1) {
    delete p;
    return fail(DH_EUNSUPPORTED, "last axis of length 1 is not supported");
  }
  for (int a = 0; a < 3; ++a) {
    // one line per workgroup holds at most 16 samples per thread, 1024 threads
    const int line = (a == 2) ? p->P[2] / 2 : p->P[a];
    if (line > 16384) {
      delete p;
      return fail(DH_EUNSUPPORTED, "canvas extent above 16384 (32768 on the last axis) is not supported");
    }
  }
  for (int a = 0; a < 3; ++a) {
    if (circular) {
      p->klo[a] = p->k[a];
      p->khi[a] = 0;
    } else {
      p->klo[a] = (p->k[a] + 1) / 2;
      p->khi[a] = p->k[a] - p->klo[a];
    }
  }
  p->M = p->P[2] / 2;
  p->H = p->M + 1;
  p->ct = dh::ct_supported(p->P[0]) && dh::ct_supported(p->P[1]) && dh::ct_supported(p->M);
  if (!make_sched(p->P[0], p->zf, p->zi) || !make_sched(p->P[1], p->yf, p->yi) ||
      !make_sched(p->M, p->xf, p->xi)) {
    delete p;
    return fail(DH_EUNSUPPORTED, "canvas extents must be 5-smooth");
  }
  // twiddle tables
  // specialised kernels: per-pass rows (Z: forward then inverse schedule);
  // runtime-length kernels: one table of n entries per axis
  p->tw_z = 0;
  if (p->ct) {
    push_pass_tables(p->tw_host, dh::ct::make_sched(p->P[0]));
    push_pass_tables(p->tw_host, dh::ct::reversed(dh::ct::make_sched(p->P[0])));
  } else {
    push_table(p->tw_host, p->P[0], p->P[0]);
  }
  p->tw_y = (int64_t)p->tw_host.size();
  if (p->ct) push_pass_tables(p->tw_host, dh::ct::make_sched(p->P[1]));
  else push_table(p->tw_host, p->P[1], p->P[1]);
  p->tw_xm = (int64_t)p->tw_host.size();
  if (p->ct) push_pass_tables(p->tw_host, dh::ct::make_row_sched(p->M));
  else push_table(p->tw_host, p->M, p->M);
  p->tw_xp = (int64_t)p->tw_host.size();
  push_table(p->tw_host, p->P[2], p->M + 1);
  if (p->tw_host.empty()) p->tw_host.push_back(make_float2(1.f, 0.f));
  // workspace layout
  const int64_t H = p->H;
  const int64_t rows1 = std::max<int64_t>((int64_t)p->s[0] * p->s[1], (int64_t)p->k[0] * p->k[1]);
  const int64_t b1 = rows1 * H * 8;
  const int64_t b2 = (int64_t)std::max(p->s[0], p->k[0]) * p->P[1] * H * 8;
  p->n_partials = (int64_t)p->s[0] * p->s[1] + 2048;
  auto align = [](int64_t v) { return (v + 255) / 256 * 256; };
  p->b1_off = 0;
  p->b2_off = align(b1);
  p->part_off = p->b2_off + align(b2);
  p->scal_off = p->part_off + align(p->n_partials * 8);
  p->ws_bytes = p->scal_off + 256;
  *out = p;
  return DH_OK;
}

int dh_conv_plan_info(const dh_conv_plan* p, dh_plan_info* info) {
  if (!p || !info) return fail(DH_EINVAL, "null argument");
  info->ndim = p->ndim;
  for (int a = 0; a < 3; ++a) {
    info->signal_shape[a] = p->s[a];
    info->kernel_shape[a] = p->k[a];
    info->fft_shape[a] = p->fft[a];
    info->canvas_shape[a] = p->P[a];
  }
  info->otf_bytes = (int64_t)p->P[0] * p->P[1] * p->H * 8;
  info->work_bytes = p->ws_bytes;
  info->twiddle_bytes = (int64_t)p->tw_host.size() * 8;
  info->spectrum_block = p->ct ? dh::kBlk : 0;
  return DH_OK;
}

int dh_conv_plan_bind(dh_conv_plan* p, void* twiddles_dev, void* stream) {
  if (!p || !twiddles_dev) return fail(DH_EINVAL, "null argument");
  hipStream_t st = (hipStream_t)stream;
  DH_TRY(hip_check(hipMemcpyAsync(twiddles_dev, p->tw_host.data(), p->tw_host.size() * 8,
                                  hipMemcpyHostToDevice, st),
                   "twiddle upload"));
  // no stream synchronisation: the host table lives in the plan, and destroying the
  // plan first waits for this event
  if (!p->bound) DH_TRY(hip_check(hipEventCreateWithFlags(&p->bound, hipEventDisableTiming), "event create"));
  DH_TRY(hip_check(hipEventRecord(p->bound, st), "event record"));
  p->tw_dev = (const float2*)twiddles_dev;
  return DH_OK;
}

int dh_conv_plan_set_tile(dh_conv_plan* p, int pass, int tile) {
  if (!p || pass < 0 || pass >= DH_PASS_COUNT || tile < 0) return fail(DH_EINVAL, "bad set_tile arguments");
  if (tile > 0) {
    const bool row = (pass == DH_PASS_X_FWD || pass == DH_PASS_X_INV);
    const int n = row ? p->M : (pass == DH_PASS_Z ? p->P[0] : p->P[1]);
    const int kind = pass == DH_PASS_Z ? COL_Z : COL_BLOCKED;
    const int t = tile & dh::kTileMask;
    if (tile & ~(dh::kTileMask | dh::kTilePersistent)) return fail(DH_EINVAL, "unknown tile flags");
    if ((tile & dh::kTilePersistent) && !p->ct)
      return fail(DH_EINVAL, "persistent launches need a canvas with specialised kernels");
    const bool ok = p->ct ? (row ? ct_row_tile_ok(n, t) : ct_col_tile_ok(n, t, kind))
                          : (row ? dh::row_tile_ok(n, t) : dh::col_tile_ok(n, t));
    if (!ok) return fail(DH_EINVAL, "tile does not fit the workgroup limits (threads, 160 KiB LDS)");
  }
  p->tile[pass] = tile;
  return DH_OK;
}

int dh_conv_plan_get_tile(const dh_conv_plan* p, int pass) {
  if (!p || pass < 0 || pass >= DH_PASS_COUNT) return -1;
  const bool row = (pass == DH_PASS_X_FWD || pass == DH_PASS_X_INV);
  const int n = row ? p->M : (pass == DH_PASS_Z ? p->P[0] : p->P[1]);
  const int64_t ncols = (pass == DH_PASS_Z) ? (int64_t)p->P[1] * p->H : p->H;
  if (p->ct)
    return persist_flag(p, pass) |
           (row ? ct_row_tile(n, tile_override(p, pass), (int64_t)p->s[0] * p->s[1])
                : ct_col_tile(n, ncols, tile_override(p, pass), pass == DH_PASS_Z ? COL_Z : COL_BLOCKED));
  if (row) return dh::row_launch_for(n, p->tile[pass]).R;
  return dh::col_launch_for(n, ncols, p->tile[pass]).L;
}

// ---- PSF synthesis (psf/widefield.py:18-73) ----
struct dh_psf_plan {
  hipEvent_t bound = nullptr;  // completion of the twiddle upload
  int ny = 0, nx = 0;
  int L = 1, R = 1;  // Y columns / X rows per workgroup
  std::vector<float2> tw_host;
  int64_t tw_y = 0, tw_x = 0;
  const float2* tw_dev = nullptr;
};

namespace {
int64_t psf_row_tiles(const dh_psf_plan* p, int nz) { return ((int64_t)nz * p->ny + p->R - 1) / p->R; }
int64_t psf_work_elems(const dh_psf_plan* p, int nz) { return (int64_t)nz * p->ny * p->nx; }
int64_t psf_part_off(const dh_psf_plan* p, int nz) { return (psf_work_elems(p, nz) * 8 + 255) / 256 * 256; }
}  // namespace

int dh_psf_plan_create(int ny, int nx, dh_psf_plan** out) {
  if (!out) return fail(DH_EINVAL, "null argument");
  if (ny < 1 || nx < 1) return fail(DH_EINVAL, "pupil shape must be positive");
  if (!dh::ct_supported(ny) || !dh::ct_supported(nx))
    return fail(DH_EUNSUPPORTED, "GPU PSF synthesis needs pupil sides of the form 2^a * {1, 3, 5} <= 4096");
  auto* p = new dh_psf_plan;
  p->ny = ny;
  p->nx = nx;
  p->tw_y = 0;
  push_pass_tables(p->tw_host, dh::ct::make_sched(ny));
  p->tw_x = (int64_t)p->tw_host.size();
  push_pass_tables(p->tw_host, dh::ct::make_sched(nx));
  if (p->tw_host.empty()) p->tw_host.push_back(make_float2(1.f, 0.f));
  // Y: up to 16 consecutive columns (128-byte row runs) within 512 threads and 160 KiB
  const int Ty = dh::ct_threads(ny);
  while (p->L < 16 && 2 * p->L * Ty <= kCtMaxThreads && (int64_t)ny * 2 * p->L * 8 <= kCtMaxLds) p->L *= 2;
  // X: up to 16 rows within 512 threads and 96 KiB
  const int Tx = dh::ct_threads(nx);
  while (p->R < 16 && 2 * p->R * Tx <= kCtMaxThreads &&
         (int64_t)2 * p->R * dh::xrow_stride(nx) * 8 <= kCtRowDefaultLds)
    p->R *= 2;
  *out = p;
  return DH_OK;
}

int dh_psf_plan_info(const dh_psf_plan* p, int nz, int nchan, int64_t* twiddle_bytes, int64_t* work_bytes) {
  if (!p || nz < 1 || nchan < 1) return fail(DH_EINVAL, "bad PSF plan info arguments");
  if (twiddle_bytes) *twiddle_bytes = (int64_t)p->tw_host.size() * 8;
  if (work_bytes) *work_bytes = psf_part_off(p, nz) + ((int64_t)nchan * psf_row_tiles(p, nz) + 1) * 8;
  return DH_OK;
}

int dh_psf_plan_bind(dh_psf_plan* p, void* twiddles_dev, void* stream) {
  if (!p || !twiddles_dev) return fail(DH_EINVAL, "null argument");
  hipStream_t st = (hipStream_t)stream;
  DH_TRY(hip_check(hipMemcpyAsync(twiddles_dev, p->tw_host.data(), p->tw_host.size() * 8, hipMemcpyHostToDevice, st),
                   "twiddle upload"));
  if (!p->bound) DH_TRY(hip_check(hipEventCreateWithFlags(&p->bound, hipEventDisableTiming), "event create"));
  DH_TRY(hip_check(hipEventRecord(p->bound, st), "event record"));
  p->tw_dev = (const float2*)twiddles_dev;
  return DH_OK;
}

int dh_psf_synth(const dh_psf_plan* p, int nchan, const void* pupils, const double* weights, const double* kz,
                 const double* z, int nz, int normalize, float* psf, void* work, int64_t work_bytes, double* total_out,
                 void* stream) {
  if (!p || !pupils || !weights || !kz || !z || !psf || !work) return fail(DH_EINVAL, "null argument");
  if (nz < 1 || nchan < 1) return fail(DH_EINVAL, "nz and nchan must be positive");
  if (!p->tw_dev) return fail(DH_EINVAL, "PSF plan is not bound (dh_psf_plan_bind)");
  int64_t need = 0;
  dh_psf_plan_info(p, nz, nchan, nullptr, &need);
  if (work_bytes < need) return fail(DH_EWORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)work;
  float2* buf = (float2*)ws;
  double* part = (double*)(ws + psf_part_off(p, nz));
  const int64_t nrt = psf_row_tiles(p, nz);
  double* total = total_out ? total_out : part + nchan * nrt;
  const int64_t ntiles = (p->nx + p->L - 1) / p->L;
  const double s = 1.0 / ((double)p->ny * (double)p->nx);
  const int64_t plane = (int64_t)p->ny * p->nx;
  // channel by channel through one work buffer: y pass, then x pass storing
  // (first channel) or adding w_c |.|^2; per-channel partials, summed in order
  for (int c = 0; c < nchan; ++c) {
    const float2* pc = (const float2*)pupils + c * plane;
    dh::Pack<dh::PsfYArgs> py;
    pack_one(py, dh::PsfYArgs{pc, kz, z, buf, p->tw_dev + p->tw_y, p->nx, ntiles}, (int64_t)nz * ntiles);
    DH_TRY(hip_check(dh::ct_launch(dh::CT_PSF_Y, p->ny, 1, &py, p->L, (int64_t)nz * ntiles, st), "psf y pass"));
    dh::Pack<dh::PsfXArgs> px;
    pack_one(px, dh::PsfXArgs{buf, psf, p->tw_dev + p->tw_x, (int64_t)nz * p->ny, (float)(weights[c] * s * s),
                              c > 0 ? 1 : 0, part + c * nrt},
             nrt);
    DH_TRY(hip_check(dh::ct_launch(dh::CT_PSF_X, p->nx, 1, &px, p->R, nrt, st), "psf x pass"));
  }
  DH_TRY(hip_check(dh::launch_sum_partials(part, nchan * nrt, total, st), "psf sum"));
  if (normalize) DH_TRY(hip_check(dh::launch_div_total(psf, psf_work_elems(p, nz), total, st), "psf normalisation"));
  return DH_OK;
}

int dh_psf_plan_destroy(dh_psf_plan* p) {
  if (p && p->bound) {
    hipEventSynchronize(p->bound);
    hipEventDestroy(p->bound);
  }
  delete p;
  return DH_OK;
}

int dh_conv_plan_destroy(dh_conv_plan* p) {
  if (p && p->bound) {
    hipEventSynchronize(p->bound);
    hipEventDestroy(p->bound);
  }
  delete p;
  return DH_OK;
}

static int otf_build(const dh_conv_plan* p, const float* kernel_dev, int normalize, void* otf_dev, void* work_dev,
                     int64_t work_bytes, double* norm_sq_host, double* norm_sq_dev, void* stream) {
  if (!p || !kernel_dev || !otf_dev || !work_dev) return fail(DH_EINVAL, "null argument");
  if (!p->tw_dev) return fail(DH_EINVAL, "plan is not bound (dh_conv_plan_bind)");
  if (work_bytes < p->ws_bytes) return fail(DH_EWORKSPACE, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)work_dev;
  float2* B1 = (float2*)(ws + p->b1_off);
  float2* B2 = (float2*)(ws + p->b2_off);
  double* part = (double*)(ws + p->part_off);
  double* scal = (double*)(ws + p->scal_off);
  const int64_t nk = (int64_t)p->k[0] * p->k[1] * p->k[2];
  dh::Pro pro{dh::PRO_NONE, 0.f, 0.f, nullptr};
  if (normalize) {
    DH_TRY(hip_check(dh::launch_reduce(dh::RED_SUM, kernel_dev, nk, part, scal, st), "kernel sum"));
    pro.kind = dh::PRO_DIV_DEV;
    pro.dptr = scal;
  }
  const int64_t H = p->H;
  const double total = (double)p->P[0] * p->P[1] * p->P[2];
  const int64_t ncols = (int64_t)p->P[1] * H;
  const int64_t krows = (int64_t)p->k[0] * p->k[1];
  float2* zsrc = B1;
  if (p->ct) {
    const float2* tw = p->tw_dev;
    const int R = ct_row_tile(p->M, 0, krows);
    const int64_t nbx = (krows + R - 1) / R;
    dh::Pack<dh::XfArgs> px;
    pack_one(px, dh::XfArgs{kernel_dev, B1, tw + p->tw_xm, tw + p->tw_xp, krows, p->k[2], p->klo[2], p->khi[2],
                            (int)H, pro, 0, dh::kBlk, p->k[1]},
             nbx);
    DH_TRY(hip_check(dh::ct_launch(dh::CT_XF, p->M, -1, &px, R, nbx, st), "otf x pass"));
    if (p->P[1] > 1) {
      const int L = ct_col_tile(p->P[1], H, 0, COL_BLOCKED);
      const int64_t nt = (H + L - 1) / L;
      dh::Pack<dh::ColArgs> py;
      pack_one(py, dh::ColArgs{B1, B2, tw + p->tw_y, H, (int64_t)p->k[1] * H, (int64_t)p->P[1] * H, p->k[1],
                               p->P[1], nt, p->klo[1], p->khi[1], p->P[1], 1.0f, dh::kBlk},
               nt * p->k[0]);
      DH_TRY(hip_check(dh::ct_launch(dh::CT_COL, p->P[1], -1, &py, L, nt * p->k[0], st), "otf y pass"));
      zsrc = B2;
    }
    // the Z transform treats each (blocked) plane as a flat row of ncols columns
    const int L = ct_col_tile(p->P[0], ncols, 0, COL_PLAIN);
    const int64_t nt = (ncols + L - 1) / L;
    dh::Pack<dh::ColArgs> pz;
    pack_one(pz, dh::ColArgs{zsrc, (float2*)otf_dev, tw + p->tw_z, ncols, 0, 0, ncols, ncols, nt, p->klo[0],
                             p->khi[0], p->P[0], (float)(1.0 / total), 0},
             nt);
    DH_TRY(hip_check(dh::ct_launch(dh::CT_COL, p->P[0], -1, &pz, L, nt, st), "otf z pass"));
  } else {
    DH_TRY(hip_check(dh::launch_x_r2c(p->xf, p->tw_dev + p->tw_xm, p->tw_dev + p->tw_xp, kernel_dev, B1, krows,
                                      p->k[2], p->klo[2], p->khi[2], p->H, pro, 0, st),
                     "otf x pass"));
    if (p->P[1] > 1) {
      DH_TRY(hip_check(dh::launch_col_fft(-1, p->yf, p->tw_dev + p->tw_y, B1, B2, H, p->k[0], (int64_t)p->k[1] * H,
                                          (int64_t)p->P[1] * H, H, H, p->klo[1], p->khi[1], p->P[1], 1.0f, 0, st),
                       "otf y pass"));
      zsrc = B2;
    }
    DH_TRY(hip_check(dh::launch_col_fft(-1, p->zf, p->tw_dev + p->tw_z, zsrc, (float2*)otf_dev, ncols, 1, 0, 0,
                                        ncols, ncols, p->klo[0], p->khi[0], p->P[0], (float)(1.0 / total), 0, st),
                     "otf z pass"));
  }
  if (norm_sq_host || norm_sq_dev) {
    DH_TRY(hip_check(dh::launch_cmax_abs2((const float2*)otf_dev, (int64_t)p->P[0] * ncols, part, scal + 1, st),
                     "otf norm"));
  }
  if (norm_sq_dev)
    DH_TRY(hip_check(dh::launch_scale_double(scal + 1, total * total, norm_sq_dev, st), "otf norm scale"));
  if (norm_sq_host) {
    double v = 0.0;
    DH_TRY(hip_check(hipMemcpyAsync(&v, scal + 1, sizeof(double), hipMemcpyDeviceToHost, st), "otf norm copy"));
    DH_TRY(hip_check(hipStreamSynchronize(st), "otf norm sync"));
    *norm_sq_host = v * total * total;
  }
  return DH_OK;
}

int dh_otf_build(const dh_conv_plan* p, const float* kernel_dev, int normalize, void* otf_dev, void* work_dev,
                 int64_t work_bytes, double* norm_sq_host, void* stream) {
  return otf_build(p, kernel_dev, normalize, otf_dev, work_dev, work_bytes, norm_sq_host, nullptr, stream);
}

int dh_otf_build_dev(const dh_conv_plan* p, const float* kernel_dev, int normalize, void* otf_dev, void* work_dev,
                     int64_t work_bytes, double* norm_sq_dev, void* stream) {
  return otf_build(p, kernel_dev, normalize, otf_dev, work_dev, work_bytes, nullptr, norm_sq_dev, stream);
}

}  // extern "C"

namespace {

// DH_PRO_SPREAD set-up: the factors must divide the plan's signal extents (the
// input is then the coarse volume), the x factor is 1 or 2, and the rows of one
// signal fit 32 bits.
// DH_PRO_PD_Y1 set-up: operands present, a specialised canvas.
int prologue_setup(const dh_conv_plan* p, const dh_fuse_args* fuse, dh::Pro& pro) {
  if (pro.kind == DH_PRO_PD_Y1) {
    if (!p->ct) return fail(DH_EINVAL, "the y1 prologue needs a canvas with specialised kernels");
    if (!fuse->pro_aux[0] || !fuse->pro_aux[1] || !fuse->pro_aux[2] || !fuse->pro_out || !fuse->pro_scalars)
      return fail(DH_EINVAL, "the y1 prologue needs pro_aux[0..2], pro_out and pro_scalars");
    if (fuse->pro_loss != 0 && fuse->pro_loss != 1) return fail(DH_EINVAL, "pro_loss must be 0 or 1");
    pro.ax = fuse->pro_aux[0];
    pro.axp = fuse->pro_aux[1];
    pro.obs = fuse->pro_aux[2];
    pro.y1w = fuse->pro_out;
    pro.sc = fuse->pro_scalars;
    pro.loss = fuse->pro_loss;
    return DH_OK;
  }
  if (pro.kind != DH_PRO_SPREAD) return DH_OK;
  const int f0 = fuse->spread[0], f1 = fuse->spread[1], f2 = fuse->spread[2];
  if (f0 < 1 || f1 < 1 || f2 < 1 || p->s[0] % f0 || p->s[1] % f1 || p->s[2] % f2)
    return fail(DH_EINVAL, "spread factors must divide the signal extents");
  if (f2 > 2) return fail(DH_EINVAL, "the spread prologue takes an x factor of 1 or 2");
  if ((int64_t)p->s[0] * p->s[1] > INT32_MAX) return fail(DH_EINVAL, "spread prologue: too many rows");
  pro.f0 = f0;
  pro.f1 = f1;
  pro.f2 = f2;
  pro.rows1 = p->s[1];
  return DH_OK;
}

}  // namespace

extern "C" {

int dh_conv_apply(const dh_conv_plan* p, int adjoint, const float* in, float* out, const void* otf_dev,
                  void* work_dev, int64_t work_bytes, const dh_fuse_args* fuse, void* stream) {
  if (!p || !out || !otf_dev || !work_dev) return fail(DH_EINVAL, "null argument");
  if (!p->tw_dev) return fail(DH_EINVAL, "plan is not bound (dh_conv_plan_bind)");
  if (work_bytes < p->ws_bytes) return fail(DH_EWORKSPACE, "workspace too small");
  dh::Pro pro{dh::PRO_NONE, 0.f, 0.f, nullptr};
  dh::Epi epi{dh::EPI_NONE, 0.f, 0.f, nullptr, nullptr, nullptr};
  double* reduce_out = nullptr;
  char* ws = (char*)work_dev;
  if (fuse) {
    pro.kind = fuse->prologue;
    pro.a = fuse->pro_a;
    pro.b = fuse->pro_b;
    epi.kind = fuse->epilogue;
    epi.a = fuse->epi_a;
    epi.b = fuse->epi_b;
    epi.aux0 = fuse->aux0;
    epi.aux1 = fuse->aux1;
    if (pro.kind < DH_PRO_NONE || pro.kind > DH_PRO_PD_Y1) return fail(DH_EINVAL, "unknown prologue");
    DH_TRY(prologue_setup(p, fuse, pro));
    if (epi.kind < DH_EPI_NONE || epi.kind > DH_EPI_BLOCKSUM2) return fail(DH_EINVAL, "unknown epilogue");
    if (epi.kind == DH_EPI_BLOCKSUM2) {
      if (!p->ct) return fail(DH_EINVAL, "the block-sum epilogue needs a canvas with specialised kernels");
      if (p->s[0] % 2 || p->s[1] % 2 || p->s[2] % 2)
        return fail(DH_EINVAL, "the block-sum epilogue needs even signal extents");
      if (!ct_row_tile_ok(p->M, 4)) return fail(DH_EINVAL, "the block-sum epilogue's four-row tile does not fit");
    }
    if ((epi.kind == DH_EPI_RL_RATIO || epi.kind == DH_EPI_RL_UPDATE || epi.kind == DH_EPI_SQDIFF) && !epi.aux0)
      return fail(DH_EINVAL, "epilogue needs aux0");
    if (epi.kind == DH_EPI_RL_UPDATE && !epi.aux1) return fail(DH_EINVAL, "epilogue needs aux1");
    if (fuse->reduce_out) {
      if (!(epi.kind == DH_EPI_RL_RATIO || epi.kind == DH_EPI_SQDIFF || (epi.kind == DH_EPI_ADD && epi.aux0)))
        return fail(DH_EINVAL, "this epilogue has no reduction");
      reduce_out = fuse->reduce_out;
      epi.partials = (double*)(ws + p->part_off);
    }
  }
  if (!in && pro.kind != DH_PRO_ONES) return fail(DH_EINVAL, "null input");
  hipStream_t st = (hipStream_t)stream;
  if (p->ct) {
    int64_t nb = 0;
    DH_TRY(ct_apply(p, adjoint, in, out, otf_dev, ws, pro, epi, &nb, st));
    if (reduce_out) {
      if (nb > p->n_partials) return fail(DH_EWORKSPACE, "partials area too small");
      DH_TRY(hip_check(dh::launch_sum_partials(epi.partials, nb, reduce_out, st), "epilogue reduction"));
    }
    return DH_OK;
  }
  float2* B1 = (float2*)(ws + p->b1_off);
  float2* B2 = (float2*)(ws + p->b2_off);
  const int s0 = p->s[0], s1 = p->s[1], s2 = p->s[2];
  const int64_t H = p->H;
  const int64_t nrows = (int64_t)s0 * s1;
  DH_TRY(timing_begin(DH_PASS_X_FWD, st));
  DH_TRY(hip_check(dh::launch_x_r2c(p->xf, p->tw_dev + p->tw_xm, p->tw_dev + p->tw_xp, in, B1, nrows, s2, s2, 0,
                                    p->H, pro, p->tile[DH_PASS_X_FWD], st),
                   "x forward pass"));
  DH_TRY(timing_end(st));
  float2* zb = B1;
  if (p->P[1] > 1) {
    DH_TRY(timing_begin(DH_PASS_Y_FWD, st));
    DH_TRY(hip_check(dh::launch_col_fft(-1, p->yf, p->tw_dev + p->tw_y, B1, B2, H, s0, (int64_t)s1 * H,
                                        (int64_t)p->P[1] * H, H, H, s1, 0, p->P[1], 1.0f, p->tile[DH_PASS_Y_FWD], st),
                     "y forward pass"));
    DH_TRY(timing_end(st));
    zb = B2;
  }
  DH_TRY(timing_begin(DH_PASS_Z, st));
  DH_TRY(hip_check(dh::launch_z_conv(p->zf, p->zi, p->tw_dev + p->tw_z, zb, (const float2*)otf_dev,
                                     (int64_t)p->P[1] * H, s0, s0, adjoint ? 1 : 0, p->tile[DH_PASS_Z], st),
                   "z pass"));
  DH_TRY(timing_end(st));
  if (p->P[1] > 1) {
    DH_TRY(timing_begin(DH_PASS_Y_INV, st));
    DH_TRY(hip_check(dh::launch_col_fft(1, p->yi, p->tw_dev + p->tw_y, B2, B1, H, s0, (int64_t)p->P[1] * H,
                                        (int64_t)s1 * H, H, H, p->P[1], 0, s1, 1.0f, p->tile[DH_PASS_Y_INV], st),
                     "y inverse pass"));
    DH_TRY(timing_end(st));
  }
  int64_t nblocks = 0;
  DH_TRY(timing_begin(DH_PASS_X_INV, st));
  DH_TRY(hip_check(dh::launch_x_c2r(p->xi, p->tw_dev + p->tw_xm, p->tw_dev + p->tw_xp, B1, out, nrows, s2, p->H,
                                    epi, p->tile[DH_PASS_X_INV], &nblocks, st),
                   "x inverse pass"));
  DH_TRY(timing_end(st));
  if (reduce_out) {
    if (nblocks > p->n_partials) return fail(DH_EWORKSPACE, "partials area too small");
    DH_TRY(hip_check(dh::launch_sum_partials(epi.partials, nblocks, reduce_out, st), "epilogue reduction"));
  }
  return DH_OK;
}

int64_t dh_pd_partials(int nz, int ny, int nx) {
  if (nz < 1 || ny < 1 || nx < 1) return 0;
  return dh::pd_partials_needed(nz, ny, nx);
}

int dh_pd_y1(float* y1, const float* ax, const float* axp, const float* data, int64_t n, const double* scalars,
             float background, int loss_type, void* stream) {
  if (!y1 || !ax || !axp || !data || !scalars || n < 1) return fail(DH_EINVAL, "bad pd_y1 arguments");
  if (loss_type != 0 && loss_type != 1) return fail(DH_EINVAL, "loss_type must be 0 (poisson) or 1 (gaussian)");
  return hip_check(dh::pd_y1(y1, ax, axp, data, n, scalars, background, loss_type, (hipStream_t)stream), "pd_y1");
}

static bool dims_ok(int nz, int ny, int nx) { return nz >= 1 && ny >= 1 && nx >= 1; }

int dh_pd_dual(float* y2, const float* x, const float* xp, int nz, int ny, int nx, const double* scalars,
               const float* w, float alpha, int l12, void* stream) {
  if (!y2 || !x || !xp || !scalars || !w || !dims_ok(nz, ny, nx)) return fail(DH_EINVAL, "bad pd_dual arguments");
  return hip_check(dh::pd_dual(y2, x, xp, nz, ny, nx, scalars, w, alpha, l12, (hipStream_t)stream), "pd_dual");
}

int dh_pd_primal(float* x_new, const float* x, const float* g, const float* y2, int nz, int ny, int nx,
                 double* scalars, const float* w, double* partials, void* stream) {
  if (!x_new || !x || !g || !y2 || !scalars || !w || !partials || !dims_ok(nz, ny, nx))
    return fail(DH_EINVAL, "bad pd_primal arguments");
  return hip_check(dh::pd_primal(x_new, x, g, y2, nz, ny, nx, scalars, w, partials, (hipStream_t)stream),
                   "pd_primal");
}

int dh_pd_dual_primal(float* y2_new, float* x_new, const float* x, const float* xp, const float* y2,
                      const float* g, int nz, int ny, int nx, double* scalars, const float* w, float alpha,
                      int l12, double* partials, void* stream) {
  if (!y2_new || !x_new || !x || !xp || !y2 || !g || !scalars || !w || !partials || !dims_ok(nz, ny, nx))
    return fail(DH_EINVAL, "bad pd_dual_primal arguments");
  if ((int64_t)ny * nx > INT32_MAX) return fail(DH_EINVAL, "pd_dual_primal: a plane holds more than 2^31 voxels");
  const int64_t n = (int64_t)nz * ny * nx;
  auto overlaps = [](const void* a, int64_t an, const void* b, int64_t bn) {
    const char* pa = (const char*)a;
    const char* pb = (const char*)b;
    return pa < pb + bn && pb < pa + an;
  };
  const int64_t vb = 4 * n, yb = 24 * n;
  if (overlaps(y2_new, yb, y2, yb) || overlaps(y2_new, yb, x, vb) || overlaps(y2_new, yb, xp, vb) ||
      overlaps(y2_new, yb, g, vb) || overlaps(x_new, vb, x, vb) || overlaps(x_new, vb, xp, vb) ||
      overlaps(x_new, vb, y2, yb) || overlaps(x_new, vb, g, vb) || overlaps(x_new, vb, y2_new, yb))
    return fail(DH_EINVAL, "pd_dual_primal outputs must not overlap its inputs or each other");
  return hip_check(dh::pd_dual_primal(y2_new, x_new, x, xp, y2, g, nz, ny, nx, scalars, w, alpha, l12, partials,
                                      (hipStream_t)stream),
                   "pd_dual_primal");
}

int dh_pd_hess_sums(const float* a, const float* b, int nz, int ny, int nx, const float* w, double* scalars,
                    double* reg_loss_out, int reg_norm, double* partials, void* stream) {
  if (!a || !b || !w || !scalars || !partials || !dims_ok(nz, ny, nx))
    return fail(DH_EINVAL, "bad pd_hess_sums arguments");
  if (reg_loss_out && reg_norm != 1 && reg_norm != 2) return fail(DH_EINVAL, "reg_norm must be 1 or 2");
  return hip_check(dh::pd_hess_sums(a, b, nz, ny, nx, w, reg_loss_out ? reg_norm : 0, scalars, reg_loss_out, partials,
                                    (hipStream_t)stream),
                   "pd_hess_sums");
}

int dh_pd_data_sums(const float* a, const float* b, const float* data, int64_t n, float background, int loss_type,
                    double* scalars, double* data_loss_out, double* partials, void* stream) {
  if (!a || !b || !scalars || !partials || n < 1 || (data_loss_out && !data))
    return fail(DH_EINVAL, "bad pd_data_sums arguments");
  return hip_check(dh::pd_data_sums(a, b, data, n, background, data_loss_out ? 1 : 0, loss_type, scalars,
                                    data_loss_out, partials, (hipStream_t)stream),
                   "pd_data_sums");
}

int dh_pd_init_scalars(double* scalars, const double* blur_norm_sq_dev, double reg_norm_sq, void* stream) {
  if (!scalars || !blur_norm_sq_dev) return fail(DH_EINVAL, "null argument");
  return hip_check(dh::pd_init(scalars, blur_norm_sq_dev, reg_norm_sq, (hipStream_t)stream), "pd_init_scalars");
}

int dh_pd_scalars(double* scalars, double* hist, int k, const dh_pd_params* params, void* stream) {
  if (!scalars || !hist || !params || k < 0) return fail(DH_EINVAL, "bad pd_scalars arguments");
  return hip_check(dh::pd_scalars(scalars, hist, k, *params, (hipStream_t)stream), "pd_scalars");
}

int dh_hessian_forward(const float* f, float* out, int nz, int ny, int nx, const float* w, void* stream) {
  if (!f || !out || !w || !dims_ok(nz, ny, nx)) return fail(DH_EINVAL, "bad hessian_forward arguments");
  return hip_check(dh::hess_forward(f, out, nz, ny, nx, w, (hipStream_t)stream), "hessian_forward");
}

int dh_hessian_adjoint(const float* y, float* out, int nz, int ny, int nx, const float* w, void* stream) {
  if (!y || !out || !w || !dims_ok(nz, ny, nx)) return fail(DH_EINVAL, "bad hessian_adjoint arguments");
  return hip_check(dh::hess_adjoint(y, out, nz, ny, nx, w, (hipStream_t)stream), "hessian_adjoint");
}

int dh_reg_components(int reg) {
  return reg == DH_REG_HESSIAN ? 6 : (reg == DH_REG_GRADIENT ? 3 : (reg == DH_REG_IDENTITY ? 1 : 0));
}

static int reg_check(int reg, const float* w, int nz, int ny, int nx) {
  if (reg != DH_REG_HESSIAN && reg != DH_REG_GRADIENT && reg != DH_REG_IDENTITY)
    return fail(DH_EINVAL, "unknown regularizer");
  if (reg != DH_REG_IDENTITY && !w) return fail(DH_EINVAL, "null weights");
  if (!dims_ok(nz, ny, nx)) return fail(DH_EINVAL, "bad volume shape");
  return DH_OK;
}

int dh_pd_reg_dual(int reg, float* y2, const float* x, const float* xp, int nz, int ny, int nx,
                   const double* scalars, const float* w, float alpha, int l12, void* stream) {
  if (reg == DH_REG_HESSIAN) return dh_pd_dual(y2, x, xp, nz, ny, nx, scalars, w, alpha, l12, stream);
  if (!y2 || !x || !xp || !scalars) return fail(DH_EINVAL, "null argument");
  DH_TRY(reg_check(reg, w, nz, ny, nx));
  return hip_check(dh::pr_dual(reg, y2, x, xp, nz, ny, nx, scalars, w, alpha, l12, (hipStream_t)stream), "pd_reg_dual");
}

int dh_pd_reg_primal(int reg, float* x_new, const float* x, const float* g, const float* y2, int nz, int ny,
                     int nx, double* scalars, const float* w, double* partials, void* stream) {
  if (reg == DH_REG_HESSIAN) return dh_pd_primal(x_new, x, g, y2, nz, ny, nx, scalars, w, partials, stream);
  if (!x_new || !x || !g || !y2 || !scalars || !partials) return fail(DH_EINVAL, "null argument");
  DH_TRY(reg_check(reg, w, nz, ny, nx));
  return hip_check(dh::pr_primal(reg, x_new, x, g, y2, nz, ny, nx, scalars, w, partials, (hipStream_t)stream),
                   "pd_reg_primal");
}

int dh_pd_reg_sums(int reg, const float* a, const float* b, int nz, int ny, int nx, const float* w,
                   double* scalars, double* reg_loss_out, int reg_norm, double* partials, void* stream) {
  if (reg == DH_REG_HESSIAN)
    return dh_pd_hess_sums(a, b, nz, ny, nx, w, scalars, reg_loss_out, reg_norm, partials, stream);
  if (!a || !b || !scalars || !partials) return fail(DH_EINVAL, "null argument");
  DH_TRY(reg_check(reg, w, nz, ny, nx));
  if (reg_loss_out && reg_norm != 1 && reg_norm != 2) return fail(DH_EINVAL, "reg_norm must be 1 or 2");
  return hip_check(dh::pr_sums(reg, a, b, nz, ny, nx, w, reg_loss_out ? reg_norm : 0, scalars, reg_loss_out,
                               partials, (hipStream_t)stream),
                   "pd_reg_sums");
}

int dh_reg_forward(int reg, const float* f, float* out, int nz, int ny, int nx, const float* w, void* stream) {
  if (reg == DH_REG_HESSIAN) return dh_hessian_forward(f, out, nz, ny, nx, w, stream);
  if (!f || !out) return fail(DH_EINVAL, "null argument");
  DH_TRY(reg_check(reg, w, nz, ny, nx));
  hipStream_t st = (hipStream_t)stream;
  if (reg == DH_REG_IDENTITY)
    return hip_check(hipMemcpyAsync(out, f, (size_t)nz * ny * nx * 4, hipMemcpyDeviceToDevice, st), "identity copy");
  return hip_check(dh::grad_forward(f, out, nz, ny, nx, w, st), "gradient forward");
}

int dh_reg_adjoint(int reg, const float* y, float* out, int nz, int ny, int nx, const float* w, void* stream) {
  if (reg == DH_REG_HESSIAN) return dh_hessian_adjoint(y, out, nz, ny, nx, w, stream);
  if (!y || !out) return fail(DH_EINVAL, "null argument");
  DH_TRY(reg_check(reg, w, nz, ny, nx));
  hipStream_t st = (hipStream_t)stream;
  if (reg == DH_REG_IDENTITY)
    return hip_check(hipMemcpyAsync(out, y, (size_t)nz * ny * nx * 4, hipMemcpyDeviceToDevice, st), "identity copy");
  return hip_check(dh::grad_adjoint(y, out, nz, ny, nx, w, st), "gradient adjoint");
}

// ---- NLCG ----
int64_t dh_nlcg_partials(void) { return dh::nl_partials_needed(); }

int dh_nlcg_init(const float* f0, float* s, float* f, int64_t n, float floor_value, void* stream) {
  if (!f0 || !s || !f || n < 1) return fail(DH_EINVAL, "bad nlcg_init arguments");
  return hip_check(dh::nl_init(f0, s, f, n, floor_value, (hipStream_t)stream), "nlcg_init");
}

int dh_nlcg_model(const float* m, const float* g, float* w, int64_t n, float eps, double* scalars,
                  double* partials, void* stream) {
  if (!m || !g || !scalars || !partials || n < 1) return fail(DH_EINVAL, "bad nlcg_model arguments");
  return hip_check(dh::nl_model(m, g, w, n, eps, scalars, partials, (hipStream_t)stream), "nlcg_model");
}

int dh_nlcg_begin_step(const float* s, const float* d, float* sd, float* dd, int64_t n, double* scalars,
                       int newton_iters, void* stream) {
  if (!s || !d || !sd || !dd || !scalars || n < 1 || newton_iters < 0)
    return fail(DH_EINVAL, "bad nlcg_begin_step arguments");
  return hip_check(dh::nl_begin_step(s, d, sd, dd, n, scalars, newton_iters, (hipStream_t)stream),
                   "nlcg_begin_step");
}

int dh_nlcg_reg_dots(const float* cf, const float* csd, const float* cdd, int64_t n, double* scalars,
                     double* partials, void* stream) {
  if (!cf || !csd || !cdd || !scalars || !partials || n < 1) return fail(DH_EINVAL, "bad nlcg_reg_dots arguments");
  return hip_check(dh::nl_reg_dots(cf, csd, cdd, n, scalars, partials, (hipStream_t)stream), "nlcg_reg_dots");
}

int dh_nlcg_newton(const float* m, const float* ksd, const float* kdd, const float* g, int64_t n, float background,
                   double beta, double eps, double* scalars, double* partials, void* stream) {
  if (!m || !ksd || !kdd || !g || !scalars || !partials || n < 1) return fail(DH_EINVAL, "bad nlcg_newton arguments");
  return hip_check(dh::nl_newton(m, ksd, kdd, g, n, background, beta, eps, scalars, partials, (hipStream_t)stream),
                   "nlcg_newton");
}

int dh_nlcg_trial(const float* m, const float* ksd, const float* kdd, const float* g, int64_t n, float background,
                  double beta, double eps, double* scalars, float* m_new, float* w, double* partials, void* stream) {
  if (!m || !ksd || !kdd || !g || !scalars || !m_new || !w || !partials || n < 1)
    return fail(DH_EINVAL, "bad nlcg_trial arguments");
  return hip_check(
      dh::nl_trial(m, ksd, kdd, g, n, background, beta, eps, scalars, m_new, w, partials, (hipStream_t)stream),
      "nlcg_trial");
}

int dh_nlcg_update(float* s, const float* d, float* f, int64_t n, double eps, double* scalars, double* partials,
                   void* stream) {
  if (!s || !d || !f || !scalars || !partials || n < 1) return fail(DH_EINVAL, "bad nlcg_update arguments");
  return hip_check(dh::nl_update(s, d, f, n, eps, scalars, partials, (hipStream_t)stream), "nlcg_update");
}

int dh_nlcg_gradient(const float* s, const float* ktw, const float* ctcf, float c4b, float* r, int64_t n, int mode,
                     double eps, double* scalars, double* partials, void* stream) {
  if (!s || !ktw || !r || !scalars || !partials || n < 1 || mode < 0 || mode > 2)
    return fail(DH_EINVAL, "bad nlcg_gradient arguments");
  return hip_check(dh::nl_gradient(s, ktw, ctcf, c4b, r, n, mode, eps, scalars, partials, (hipStream_t)stream),
                   "nlcg_gradient");
}

int dh_nlcg_direction(const float* r, float* d, int64_t n, const double* scalars, void* stream) {
  if (!r || !d || !scalars || n < 1) return fail(DH_EINVAL, "bad nlcg_direction arguments");
  return hip_check(dh::nl_direction(r, d, n, scalars, (hipStream_t)stream), "nlcg_direction");
}

int dh_nlcg_quad(const float* g, const float* m, const float* ksd, const float* v, const float* d, int64_t n,
                 double eps, double* out2, double* partials, void* stream) {
  if ((g && (!m || !ksd)) || !v || !d || !out2 || !partials || n < 1) return fail(DH_EINVAL, "bad nlcg_quad arguments");
  return hip_check(dh::nl_quad(g, m, ksd, v, d, n, eps, out2, partials, (hipStream_t)stream), "nlcg_quad");
}

int dh_nlcg_idiv(const float* fp, const float* fc, int64_t n, double eps, double* out2, double* partials,
                 void* stream) {
  if (!fp || !fc || !out2 || !partials || n < 1) return fail(DH_EINVAL, "bad nlcg_idiv arguments");
  return hip_check(dh::nl_idiv(fp, fc, n, eps, out2, partials, (hipStream_t)stream), "nlcg_idiv");
}

// ---- ER-Decon ----
int64_t dh_er_partials(void) { return dh::er_partials_needed(); }

int dh_er_hess_weights(const float* g, int nz, int ny, int nx, const float* w6, float lam, float eps_reg,
                       float w_floor, float intensity_weight, float* w_out, float* u_out, double* scalars,
                       double* partials, void* stream) {
  if (!g || !w6 || !scalars || !partials || !dims_ok(nz, ny, nx) || (w_out && !u_out) || !(eps_reg > 0.0f))
    return fail(DH_EINVAL, "bad er_hess_weights arguments");
  return hip_check(dh::er_hess_weights(g, nz, ny, nx, w6, lam, eps_reg, w_floor, intensity_weight, w_out, u_out,
                                       scalars, partials, (hipStream_t)stream),
                   "er_hess_weights");
}

int dh_er_weights_stacked(const float* hg, int nc, const float* g, int64_t n, float lam, float eps_reg,
                          float w_floor, float intensity_weight, float* w_out, float* u_out, double* scalars,
                          double* partials, void* stream) {
  if (!hg || !g || nc < 1 || n < 1 || !scalars || !partials || (w_out && !u_out) || !(eps_reg > 0.0f))
    return fail(DH_EINVAL, "bad er_weights_stacked arguments");
  return hip_check(dh::er_weights_stacked(hg, nc, g, n, lam, eps_reg, w_floor, intensity_weight, w_out, u_out,
                                          scalars, partials, (hipStream_t)stream),
                   "er_weights_stacked");
}

int dh_er_hess_apply_w(const float* v, const float* w, int nz, int ny, int nx, const float* w6, float* u_out,
                       void* stream) {
  if (!v || !w || !w6 || !u_out || !dims_ok(nz, ny, nx)) return fail(DH_EINVAL, "bad er_hess_apply_w arguments");
  return hip_check(dh::er_hess_apply_w(v, w, nz, ny, nx, w6, u_out, (hipStream_t)stream), "er_hess_apply_w");
}

int dh_er_scale_stacked(const float* x, const float* w, int nc, int64_t n, float* out, void* stream) {
  if (!x || !w || !out || nc < 1 || n < 1) return fail(DH_EINVAL, "bad er_scale_stacked arguments");
  return hip_check(dh::er_scale_stacked(x, w, nc, n, out, (hipStream_t)stream), "er_scale_stacked");
}

int dh_er_data(const float* m, const float* f, int64_t n, int poisson, float* score, float* hessw,
               double* scalars, double* partials, void* stream) {
  if (!m || !f || n < 1 || !scalars || !partials || (score && poisson && !hessw))
    return fail(DH_EINVAL, "bad er_data arguments");
  return hip_check(dh::er_data(m, f, n, poisson ? 1 : 0, score, hessw, scalars, partials, (hipStream_t)stream),
                   "er_data");
}

int dh_er_gradient(const float* s, const float* dg, const float* rh, const float* w, const float* g,
                   float intensity_weight, int64_t n, double cg_tol, float* grad, float* x, float* r, float* p,
                   float* sv, double* scalars, double* partials, void* stream) {
  if (!s || !dg || !rh || !grad || !x || !r || !p || !sv || !scalars || !partials || n < 1 ||
      (intensity_weight > 0.0f && (!w || !g)))
    return fail(DH_EINVAL, "bad er_gradient arguments");
  return hip_check(dh::er_gradient(s, dg, rh, w, g, intensity_weight, n, cg_tol, grad, x, r, p, sv, scalars, partials,
                                   (hipStream_t)stream),
                   "er_gradient");
}

int dh_er_cg_ap(const float* s, const float* kt, const float* htu, const float* w, const float* sv,
                float intensity_weight, const float* p, float* ap, int64_t n, double* scalars, double* partials,
                void* stream) {
  if (!s || !kt || !htu || !p || !ap || !scalars || !partials || n < 1 || (intensity_weight > 0.0f && (!w || !sv)))
    return fail(DH_EINVAL, "bad er_cg_ap arguments");
  return hip_check(
      dh::er_cg_ap(s, kt, htu, w, sv, intensity_weight, p, ap, n, scalars, partials, (hipStream_t)stream),
      "er_cg_ap");
}

int dh_er_cg_update(float* x, float* r, float* p, const float* ap, const float* s, float* sv, int64_t n,
                    double* scalars, double* partials, void* stream) {
  if (!x || !r || !p || !ap || !s || !sv || !scalars || !partials || n < 1)
    return fail(DH_EINVAL, "bad er_cg_update arguments");
  return hip_check(dh::er_cg_update(x, r, p, ap, s, sv, n, scalars, partials, (hipStream_t)stream), "er_cg_update");
}

int dh_er_dot(const float* a, const float* b, int64_t n, int slot, double* scalars, double* partials, void* stream) {
  if (!a || !b || n < 1 || slot < 0 || slot >= DH_ER_NSLOTS || !scalars || !partials)
    return fail(DH_EINVAL, "bad er_dot arguments");
  return hip_check(dh::er_dot(a, b, n, slot, scalars, partials, (hipStream_t)stream), "er_dot");
}

int dh_er_ew(int op, const float* a, const float* b, float alpha, float beta, float* out, float* out2, int64_t n,
             void* stream) {
  const bool needs_b = op == DH_ER_EW_AXPY_SQUARE || op == DH_ER_EW_MUL || op == DH_ER_EW_AXPY;
  if (op < DH_ER_EW_DIV || op > DH_ER_EW_AXPY || !a || !out || n < 1 || (needs_b && !b) ||
      (op == DH_ER_EW_AXPY_SQUARE && !out2))
    return fail(DH_EINVAL, "bad er_ew arguments");
  return hip_check(dh::er_ew(op, a, b, alpha, beta, out, out2, n, (hipStream_t)stream), "er_ew");
}

// ---- JetNewton ----
int64_t dh_jn_partials(void) { return dh::jn_partials_needed(); }

int dh_jn_penalty(const float* x, int nz, int ny, int nx, const float* w6, float beta, float eta, float* r_out,
                  float* w_out, float* c_out, float* wr_out, float* u_out, double* scalars, double* partials,
                  void* stream) {
  if (!x || !w6 || !scalars || !partials || !dims_ok(nz, ny, nx) || (r_out && (!w_out || !c_out || !wr_out)) ||
      !(beta > 0.0f))
    return fail(DH_EINVAL, "bad jn_penalty arguments");
  return hip_check(dh::jn_penalty(x, nz, ny, nx, w6, beta, eta, r_out, w_out, c_out, wr_out, u_out, scalars, partials,
                                  (hipStream_t)stream),
                   "jn_penalty");
}

int dh_jn_penalty_stacked(const float* r, int nc, int64_t n, float beta, float eta, float* w_out, float* c_out,
                          float* wr_out, float* u_out, double* scalars, double* partials, void* stream) {
  if (!r || nc < 1 || n < 1 || !scalars || !partials || (w_out && (!c_out || !wr_out)) || !(beta > 0.0f))
    return fail(DH_EINVAL, "bad jn_penalty_stacked arguments");
  return hip_check(dh::jn_penalty_stacked(r, nc, n, beta, eta, w_out, c_out, wr_out, u_out, scalars, partials,
                                          (hipStream_t)stream),
                   "jn_penalty_stacked");
}

int dh_jn_hvp_operand(const float* v, const float* r, const float* w, const float* c, int nz, int ny, int nx,
                      const float* w6, float* out, void* stream) {
  if (!v || !r || !w || !c || !w6 || !out || !dims_ok(nz, ny, nx)) return fail(DH_EINVAL, "bad jn_hvp_operand arguments");
  return hip_check(dh::jn_hvp_operand(v, r, w, c, nz, ny, nx, w6, out, (hipStream_t)stream), "jn_hvp_operand");
}

int dh_jn_hvp_operand_stacked(const float* q, const float* r, const float* w, const float* c, int nc, int64_t n,
                              float* out, void* stream) {
  if (!q || !r || !w || !c || !out || nc < 1 || n < 1) return fail(DH_EINVAL, "bad jn_hvp_operand_stacked arguments");
  return hip_check(dh::jn_hvp_operand_stacked(q, r, w, c, nc, n, out, (hipStream_t)stream), "jn_hvp_operand_stacked");
}

int dh_jn_kkt(const float* x, const float* g, int64_t n, double eps_bar, double* scalars, double* partials,
              void* stream) {
  if (!x || !g || n < 1 || !scalars || !partials) return fail(DH_EINVAL, "bad jn_kkt arguments");
  return hip_check(dh::jn_kkt(x, g, n, eps_bar, scalars, partials, (hipStream_t)stream), "jn_kkt");
}

int dh_jn_active(const float* x, const float* g, int64_t n, float freeze_tau, float freeze_delta, uint8_t* frozen,
                 uint8_t* active, float* r, float* p, float* xcg, double eps, double* scalars, double* partials,
                 void* stream) {
  if (!g || !active || n < 1 || !scalars || !partials || (r && (!p || !xcg)) || (!x && (!r || frozen)))
    return fail(DH_EINVAL, "bad jn_active arguments");
  return hip_check(dh::jn_active(x, g, n, freeze_tau, freeze_delta, frozen, active, r, p, xcg, eps, scalars, partials,
                                 (hipStream_t)stream),
                   "jn_active");
}

int dh_jn_cg_ap(const float* kt, float scale, const float* htu, const uint8_t* active, const float* p, float* ap,
                int64_t n, double eps, double* scalars, double* partials, void* stream) {
  if (!kt || !active || !p || !ap || n < 1 || !scalars || !partials) return fail(DH_EINVAL, "bad jn_cg_ap arguments");
  return hip_check(dh::jn_cg_ap(kt, scale, htu, active, p, ap, n, eps, scalars, partials, (hipStream_t)stream),
                   "jn_cg_ap");
}

int dh_jn_cg_update(float* x, float* r, float* p, const float* ap, int64_t n, double* scalars, double* partials,
                    void* stream) {
  if (!x || !r || !p || !ap || n < 1 || !scalars || !partials) return fail(DH_EINVAL, "bad jn_cg_update arguments");
  return hip_check(dh::jn_cg_update(x, r, p, ap, n, scalars, partials, (hipStream_t)stream), "jn_cg_update");
}

int dh_jn_direction(const float* xcg, const float* g, const uint8_t* active, float* d, int64_t n, double* scalars,
                    double* partials, void* stream) {
  if (!xcg || !g || !active || !d || n < 1 || !scalars || !partials) return fail(DH_EINVAL, "bad jn_direction arguments");
  return hip_check(dh::jn_direction(xcg, g, active, d, n, scalars, partials, (hipStream_t)stream), "jn_direction");
}

int dh_jn_trial(const float* x, const float* d, float t, const float* g, float* x_t, int64_t n, double* scalars,
                double* partials, void* stream) {
  if (!x || !d || !g || !x_t || n < 1 || !scalars || !partials) return fail(DH_EINVAL, "bad jn_trial arguments");
  return hip_check(dh::jn_trial(x, d, t, g, x_t, n, scalars, partials, (hipStream_t)stream), "jn_trial");
}

int dh_timing_enable(int on) {
  g_timing = on != 0;
  return DH_OK;
}

int dh_timing_select(int mask) {
  g_timing_mask = mask & ((1 << DH_PASS_COUNT) - 1);
  return DH_OK;
}

int dh_timing_read(double* ms, int64_t* counts, int reset) {
  for (auto& t : g_pending) {
    DH_TRY(hip_check(hipEventSynchronize(t.b), "event sync"));
    float e = 0.f;
    DH_TRY(hip_check(hipEventElapsedTime(&e, t.a, t.b), "event elapsed"));
    g_ms[t.pass] += e;
    g_cnt[t.pass] += 1;
    g_free.push_back(t.a);
    g_free.push_back(t.b);
  }
  g_pending.clear();
  for (int i = 0; i < DH_PASS_COUNT; ++i) {
    if (ms) ms[i] = g_ms[i];
    if (counts) counts[i] = g_cnt[i];
    if (reset) {
      g_ms[i] = 0.0;
      g_cnt[i] = 0;
    }
  }
  return DH_OK;
}

}  // extern "C"

namespace {

int fill_fuse(const dh_fuse_args* fuse, dh::Pro& pro, dh::Epi& epi, double** reduce_out) {
  pro = dh::Pro{dh::PRO_NONE, 0.f, 0.f, nullptr};
  epi = dh::Epi{dh::EPI_NONE, 0.f, 0.f, nullptr, nullptr, nullptr};
  *reduce_out = nullptr;
  if (!fuse) return DH_OK;
  pro.kind = fuse->prologue;
  pro.a = fuse->pro_a;
  pro.b = fuse->pro_b;
  epi.kind = fuse->epilogue;
  epi.a = fuse->epi_a;
  epi.b = fuse->epi_b;
  epi.aux0 = fuse->aux0;
  epi.aux1 = fuse->aux1;
  if (pro.kind < DH_PRO_NONE || pro.kind > DH_PRO_PD_Y1) return fail(DH_EINVAL, "unknown prologue");
  if (epi.kind == DH_EPI_BLOCKSUM2) return fail(DH_EINVAL, "the block-sum epilogue is not available in batches");
  if (epi.kind < DH_EPI_NONE || epi.kind > DH_EPI_SQDIFF) return fail(DH_EINVAL, "unknown epilogue");
  if ((epi.kind == DH_EPI_RL_RATIO || epi.kind == DH_EPI_RL_UPDATE || epi.kind == DH_EPI_SQDIFF) && !epi.aux0)
    return fail(DH_EINVAL, "epilogue needs aux0");
  if (epi.kind == DH_EPI_RL_UPDATE && !epi.aux1) return fail(DH_EINVAL, "epilogue needs aux1");
  if (fuse->reduce_out) {
    if (!(epi.kind == DH_EPI_RL_RATIO || epi.kind == DH_EPI_SQDIFF || (epi.kind == DH_EPI_ADD && epi.aux0)))
      return fail(DH_EINVAL, "this epilogue has no reduction");
    *reduce_out = fuse->reduce_out;
  }
  return DH_OK;
}

// Batched application with the specialised kernels: per pass, the volumes are
// grouped by transform length and launch shape, and each group of up to kPack
// volumes is one launch with its descriptors in the kernel arguments.
template <class A>
struct Group {
  int n = 0, tile = 0;
  std::vector<A> args;
  std::vector<int64_t> nblk;
};

template <class A>
int launch_groups(int kind, int dir, std::vector<Group<A>>& groups, int pass, hipStream_t st) {
  for (auto& g : groups) {
    for (size_t s = 0; s < g.args.size(); s += dh::kPack) {
      dh::Pack<A> pk;
      pk.nvol = (int)std::min<size_t>(dh::kPack, g.args.size() - s);
      pk.prefix[0] = 0;
      for (int v = 0; v < pk.nvol; ++v) {
        pk.a[v] = g.args[s + v];
        pk.prefix[v + 1] = pk.prefix[v] + g.nblk[s + v];
      }
      if (pk.prefix[pk.nvol] == 0) continue;
      DH_TRY(timing_begin(pass, st));
      DH_TRY(hip_check(dh::ct_launch(kind, g.n, dir, &pk, g.tile, pk.prefix[pk.nvol], st), "batched pass"));
      DH_TRY(timing_end(st));
    }
  }
  return DH_OK;
}

template <class A>
void add_to_group(std::vector<Group<A>>& groups, int n, int tile, const A& a, int64_t nblk) {
  for (auto& g : groups) {
    if (g.n == n && g.tile == tile) {
      g.args.push_back(a);
      g.nblk.push_back(nblk);
      return;
    }
  }
  Group<A> g;
  g.n = n;
  g.tile = tile;
  g.args.push_back(a);
  g.nblk.push_back(nblk);
  groups.push_back(std::move(g));
}

int ct_batch_apply(int nitems, const dh_batch_item* items, int adjoint, hipStream_t st) {
  std::vector<Group<dh::XfArgs>> gxf;
  std::vector<Group<dh::ColArgs>> gyf, gyi;
  std::vector<Group<dh::ZArgs>> gz;
  std::vector<Group<dh::XiArgs>> gxi;
  struct Red {
    double* partials;
    int64_t n;
    double* out;
  };
  std::vector<Red> reds;
  for (int i = 0; i < nitems; ++i) {
    const dh_batch_item& it = items[i];
    const dh_conv_plan* p = it.plan;
    if (!it.otf || !it.work || !it.out) return fail(DH_EINVAL, "null batch item field");
    if (!p->tw_dev) return fail(DH_EINVAL, "plan is not bound (dh_conv_plan_bind)");
    if (it.work_bytes < p->ws_bytes) return fail(DH_EWORKSPACE, "workspace too small");
    dh::Pro pro;
    dh::Epi epi;
    double* red = nullptr;
    DH_TRY(fill_fuse(&it.fuse, pro, epi, &red));
    DH_TRY(prologue_setup(p, &it.fuse, pro));
    if (!it.in && pro.kind != DH_PRO_ONES) return fail(DH_EINVAL, "null input");
    char* ws = (char*)it.work;
    if (red) epi.partials = (double*)(ws + p->part_off);
    const CtPasses c = ct_passes(p, adjoint, it.in, it.out, it.otf, ws, pro, epi);
    if (red) {
      if (c.n_xi > p->n_partials) return fail(DH_EWORKSPACE, "partials area too small");
      reds.push_back(Red{epi.partials, c.n_xi, red});
    }
    add_to_group(gxf, p->M, c.t_xf, c.xf, c.n_xf);
    if (c.has_y) {
      add_to_group(gyf, p->P[1], c.t_yf, c.yf, c.n_yf);
      add_to_group(gyi, p->P[1], c.t_yi, c.yi, c.n_yi);
    }
    add_to_group(gz, p->P[0], c.t_z, c.z, c.n_z);
    add_to_group(gxi, p->M, c.t_xi, c.xi, c.n_xi);
  }
  DH_TRY(launch_groups(dh::CT_XF, -1, gxf, DH_PASS_X_FWD, st));
  DH_TRY(launch_groups(dh::CT_COL, -1, gyf, DH_PASS_Y_FWD, st));
  DH_TRY(launch_groups(dh::CT_Z, 0, gz, DH_PASS_Z, st));
  DH_TRY(launch_groups(dh::CT_COL, 1, gyi, DH_PASS_Y_INV, st));
  DH_TRY(launch_groups(dh::CT_XI, 1, gxi, DH_PASS_X_INV, st));
  for (const Red& r : reds)
    DH_TRY(hip_check(dh::launch_sum_partials(r.partials, r.n, r.out, st), "batched epilogue reduction"));
  return DH_OK;
}

// Batched fused RL iterations: the stages of dh_rl_iteration, each stage one
// grouped launch per transform length over every item.
int ct_batch_rl(int nitems, const dh_rl_item* items, hipStream_t st) {
  std::vector<Group<dh::XfArgs>> gxf;
  std::vector<Group<dh::ColArgs>> gyf1, gyi1, gyf2, gyi2;
  std::vector<Group<dh::ZArgs>> gz1, gz2;
  std::vector<Group<dh::XiArgs>> gr1, gr2, gi2;
  struct Red {
    double* partials;
    int64_t n;
    double* out;
  };
  std::vector<Red> reds;
  for (int i = 0; i < nitems; ++i) {
    const dh_rl_item& it = items[i];
    const dh_conv_plan* p = it.plan;
    char* ws = (char*)it.work;
    float2* B1 = (float2*)(ws + p->b1_off);
    const dh::Pro none{dh::PRO_NONE, 0.f, 0.f, nullptr};
    dh::Epi ratio{dh::EPI_RL_RATIO, it.background, it.eps, it.obs, nullptr, nullptr};
    if (it.reduce_out) ratio.partials = (double*)(ws + p->part_off);
    for (int d = 0; d < 6; ++d) ratio.box[d] = it.data_box[d];
    const dh::Epi update{dh::EPI_RL_UPDATE, 0.f, it.eps, it.x, it.sens, nullptr};
    CtPasses f = ct_passes(p, 0, it.x, nullptr, it.otf, ws, none, ratio);
    f.xi.spec = B1;
    CtPasses a = ct_passes(p, 1, nullptr, it.x, it.otf, ws, none, update);
    if (it.reduce_out) {
      if (f.n_xi > p->n_partials) return fail(DH_EWORKSPACE, "partials area too small");
      reds.push_back(Red{ratio.partials, f.n_xi, it.reduce_out});
    }
    if (!it.spectrum_ready) add_to_group(gxf, p->M, f.t_xf, f.xf, f.n_xf);
    if (f.has_y) {
      add_to_group(gyf1, p->P[1], f.t_yf, f.yf, f.n_yf);
      add_to_group(gyi1, p->P[1], f.t_yi, f.yi, f.n_yi);
      add_to_group(gyf2, p->P[1], a.t_yf, a.yf, a.n_yf);
      add_to_group(gyi2, p->P[1], a.t_yi, a.yi, a.n_yi);
    }
    add_to_group(gz1, p->P[0], f.t_z, f.z, f.n_z);
    add_to_group(gz2, p->P[0], a.t_z, a.z, a.n_z);
    add_to_group(gr1, p->M, f.t_xi, f.xi, f.n_xi);
    if (it.keep_spectrum) {
      a.xi.spec = B1;
      add_to_group(gr2, p->M, a.t_xi, a.xi, a.n_xi);
    } else {
      add_to_group(gi2, p->M, a.t_xi, a.xi, a.n_xi);
    }
  }
  DH_TRY(launch_groups(dh::CT_XF, -1, gxf, DH_PASS_X_FWD, st));
  DH_TRY(launch_groups(dh::CT_COL, -1, gyf1, DH_PASS_Y_FWD, st));
  DH_TRY(launch_groups(dh::CT_Z, 0, gz1, DH_PASS_Z, st));
  DH_TRY(launch_groups(dh::CT_COL, 1, gyi1, DH_PASS_Y_INV, st));
  DH_TRY(launch_groups(dh::CT_XR, 1, gr1, DH_PASS_X_INV, st));
  DH_TRY(launch_groups(dh::CT_COL, -1, gyf2, DH_PASS_Y_FWD, st));
  DH_TRY(launch_groups(dh::CT_Z, 0, gz2, DH_PASS_Z, st));
  DH_TRY(launch_groups(dh::CT_COL, 1, gyi2, DH_PASS_Y_INV, st));
  DH_TRY(launch_groups(dh::CT_XR, 1, gr2, DH_PASS_X_INV, st));
  DH_TRY(launch_groups(dh::CT_XI, 1, gi2, DH_PASS_X_INV, st));
  for (const Red& r : reds)
    DH_TRY(hip_check(dh::launch_sum_partials(r.partials, r.n, r.out, st), "batched epilogue reduction"));
  return DH_OK;
}

}  // namespace

extern "C" {

int dh_batch_rl_iteration(int nitems, const dh_rl_item* items, void* stream) {
  if (!items || nitems < 1) return fail(DH_EINVAL, "bad batch arguments");
  for (int i = 0; i < nitems; ++i) {
    const dh_rl_item& it = items[i];
    if (!it.plan || !it.otf || !it.work || !it.x || !it.obs || !it.sens)
      return fail(DH_EINVAL, "null field in RL batch item");
    if (!it.plan->tw_dev) return fail(DH_EINVAL, "plan is not bound (dh_conv_plan_bind)");
    if (it.work_bytes < it.plan->ws_bytes) return fail(DH_EWORKSPACE, "workspace too small");
    if (!it.plan->ct) return fail(DH_EUNSUPPORTED, "the fused RL iteration needs a canvas with specialised kernels");
    const int32_t* b = it.data_box;
    if (b[1] > b[0]) {
      const int* s = it.plan->s;
      for (int d = 0; d < 3; ++d)
        if (b[2 * d] < 0 || b[2 * d + 1] <= b[2 * d] || b[2 * d + 1] > s[d])
          return fail(DH_EINVAL, "data_box outside the signal grid");
    }
  }
  return ct_batch_rl(nitems, items, (hipStream_t)stream);
}

int dh_batch_conv_apply(int nitems, const dh_batch_item* items, int adjoint, void* stream) {
  if (!items || nitems < 1) return fail(DH_EINVAL, "bad batch arguments");
  for (int i = 0; i < nitems; ++i)
    if (!items[i].plan) return fail(DH_EINVAL, "null plan in batch item");
  hipStream_t st = (hipStream_t)stream;
  // specialised lengths: grouped launches; the rest: one application each
  std::vector<dh_batch_item> cts;
  for (int i = 0; i < nitems; ++i) {
    const dh_batch_item& it = items[i];
    if (it.plan->ct) {
      cts.push_back(it);
    } else {
      DH_TRY(dh_conv_apply(it.plan, adjoint, it.in, it.out, it.otf, it.work, it.work_bytes, &it.fuse, stream));
    }
  }
  if (!cts.empty()) DH_TRY(ct_batch_apply((int)cts.size(), cts.data(), adjoint, st));
  return DH_OK;
}

int64_t dh_reduce_partials(int64_t n) { return dh::reduce_nblocks(n); }

int dh_reduce(int op, const float* x, int64_t n, double* partials_dev, double* result_dev, void* stream) {
  if (!x || !partials_dev || !result_dev || n < 1) return fail(DH_EINVAL, "bad reduce arguments");
  if (op < DH_RED_SUM || op > DH_RED_SUMSQ) return fail(DH_EINVAL, "unknown reduction");
  return hip_check(dh::launch_reduce(op, x, n, partials_dev, result_dev, (hipStream_t)stream), "reduce");
}

int dh_rl_support(float* sens, float* x, int64_t n, const double* maxval_dev, void* stream) {
  if (!sens || !maxval_dev) return fail(DH_EINVAL, "null argument");
  return hip_check(dh::launch_rl_support(sens, x, n, maxval_dev, (hipStream_t)stream), "rl_support");
}

int dh_clamp_min(const float* in, float* out, int64_t n, float lo, void* stream) {
  if (!in || !out) return fail(DH_EINVAL, "null argument");
  return hip_check(dh::launch_clamp_min(in, out, n, lo, (hipStream_t)stream), "clamp_min");
}

int dh_sub_floor(const float* in, float* out, int64_t n, float a, float b, void* stream) {
  if (!in || !out) return fail(DH_EINVAL, "null argument");
  return hip_check(dh::launch_sub_floor(in, out, n, a, b, (hipStream_t)stream), "sub_floor");
}

int dh_rl_ratio(const float* model, const float* obs, float* out, int64_t n, float bg, float eps,
                double* partials_dev, double* reduce_out, void* stream) {
  if (!model || !obs || !out) return fail(DH_EINVAL, "null argument");
  if (reduce_out && !partials_dev) return fail(DH_EINVAL, "reduction needs partials");
  return hip_check(dh::launch_rl_ratio(model, obs, out, n, bg, eps, partials_dev, reduce_out, (hipStream_t)stream),
                   "rl_ratio");
}

int dh_rl_update(float* x, const float* corr, const float* sens, int64_t n, float eps, void* stream) {
  if (!x || !corr || !sens) return fail(DH_EINVAL, "null argument");
  return hip_check(dh::launch_rl_update(x, corr, sens, n, eps, (hipStream_t)stream), "rl_update");
}

int dh_poisson_idiv(const float* obs, const float* model, int64_t n, float eps, double* partials_dev,
                    double* result_dev, void* stream) {
  if (!obs || !model || !partials_dev || !result_dev || n < 1) return fail(DH_EINVAL, "bad poisson_idiv arguments");
  if (!(eps >= 0.0f)) return fail(DH_EINVAL, "eps must be >= 0");
  return hip_check(dh::launch_poisson_idiv(obs, model, n, eps, partials_dev, result_dev, (hipStream_t)stream),
                   "poisson_idiv");
}

int dh_add_scalar(const float* in, const float* obs, float* out, int64_t n, float a, double* partials_dev,
                  double* reduce_out, void* stream) {
  if (!in || !out) return fail(DH_EINVAL, "null argument");
  if (reduce_out && (!partials_dev || !obs)) return fail(DH_EINVAL, "reduction needs partials and obs");
  return hip_check(dh::launch_add_scalar(in, obs, out, n, a, partials_dev, reduce_out, (hipStream_t)stream),
                   "add_scalar");
}

int dh_box_copy(const float* src, const int64_t* src_shape, float* dst, const int64_t* dst_shape,
                const int64_t* src_start, const int64_t* dst_start, const int64_t* extent, void* stream) {
  if (!src || !dst || !src_shape || !dst_shape || !src_start || !dst_start || !extent)
    return fail(DH_EINVAL, "null argument");
  for (int a = 0; a < 3; ++a) {
    if (extent[a] < 0 || src_start[a] < 0 || dst_start[a] < 0 || src_start[a] + extent[a] > src_shape[a] ||
        dst_start[a] + extent[a] > dst_shape[a])
      return fail(DH_EINVAL, "box outside the volumes");
  }
  return hip_check(dh::launch_box_copy(src, src_shape, dst, dst_shape, src_start, dst_start, extent,
                                       (hipStream_t)stream),
                   "box_copy");
}

int dh_block_sum(const float* fine, float* coarse, int n0, int n1, int n2, int f0, int f1, int f2, void* stream) {
  if (!fine || !coarse) return fail(DH_EINVAL, "null argument");
  if (n0 < 1 || n1 < 1 || n2 < 1 || f0 < 1 || f1 < 1 || f2 < 1 || n0 % f0 || n1 % f1 || n2 % f2)
    return fail(DH_EINVAL, "block factors must divide the fine extents");
  return hip_check(dh::launch_block_sum(fine, coarse, n0, n1, n2, f0, f1, f2, (hipStream_t)stream), "block sum");
}

int dh_block_spread(const float* coarse, float* fine, int n0, int n1, int n2, int f0, int f1, int f2, void* stream) {
  if (!fine || !coarse) return fail(DH_EINVAL, "null argument");
  if (n0 < 1 || n1 < 1 || n2 < 1 || f0 < 1 || f1 < 1 || f2 < 1 || n0 % f0 || n1 % f1 || n2 % f2)
    return fail(DH_EINVAL, "block factors must divide the fine extents");
  return hip_check(dh::launch_block_spread(coarse, fine, n0, n1, n2, f0, f1, f2, (hipStream_t)stream),
                   "block spread");
}

int dh_banded_apply(const float* src, float* dst, int64_t outer, int64_t n_in, int64_t inner, int64_t n_out,
                    const int32_t* idx, const float* wt, int width, void* stream) {
  if (!src || !dst || !idx || !wt) return fail(DH_EINVAL, "null argument");
  if (outer < 1 || n_in < 1 || inner < 1 || n_out < 1 || width < 1) return fail(DH_EINVAL, "bad extents");
  return hip_check(dh::launch_banded(src, dst, outer, n_in, inner, n_out, idx, wt, width, (hipStream_t)stream),
                   "banded_apply");
}

}  // extern "C"
