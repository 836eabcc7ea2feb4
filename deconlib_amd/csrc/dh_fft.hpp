// Shared declarations for the pruned 3-D FFT convolution passes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dh {

constexpr int kMaxPasses = 16;
// Elements of one line a thread keeps in registers per pass (upper bound
// that fixes the per-radix butterfly count, see kMB below).
constexpr int kEPT = 16;

// A Stockham schedule for one axis: radices in pass order and the span
// (product of the previous radices) of each pass.
struct AxisSched {
  int n;
  int np;
  int radix[kMaxPasses];
  int span[kMaxPasses];
};

// Fused prologue (input sample of the forward X pass) and epilogue (cropped
// output sample of the inverse X pass).
enum ProKind { PRO_NONE = 0, PRO_ONES = 1, PRO_SUB_FLOOR = 2, PRO_SPREAD = 3, PRO_PD_Y1 = 4, PRO_DIV_DEV = 100 };
enum EpiKind {
  EPI_NONE = 0,
  EPI_FLOOR = 1,
  EPI_RL_RATIO = 2,
  EPI_RL_UPDATE = 3,
  EPI_ADD = 4,
  EPI_SQDIFF = 5,
  EPI_BLOCKSUM2 = 6  // out = 2x2x2 block sums of the cropped output (k_xi_bs)
};

struct Pro {
  int kind;
  float a, b;
  const double* dptr;  // PRO_DIV_DEV: divide by (float)*dptr
  // PRO_SPREAD: sample (z, y, x) of the input is src[z / f0, y / f1, x / f2] of a
  // coarse volume; rows1 = input rows per plane (the fine grid's y extent).
  int f0 = 1, f1 = 1, f2 = 1, rows1 = 1;
  // PRO_PD_Y1: the sample is PDHG's data-term dual step y1' = prox(y1 + sigma A xbar)
  // of y1 = src and the same sample of A x_k (ax), A x_{k-1} (axp) and the data (obs),
  // with sigma and theta from the float64 scalar block sc and the background in a;
  // y1' is also stored to y1w (which may be src).  loss: 0 Poisson, 1 Gaussian.
  const float* ax = nullptr;
  const float* axp = nullptr;
  const float* obs = nullptr;
  float* y1w = nullptr;
  const double* sc = nullptr;
  int loss = 0;
};

// First sample of input row `row` (plane-major rows of length rowlen): the row
// itself, or under PRO_SPREAD the coarse row it replicates (rows < 2^31).
__host__ __device__ __forceinline__ const float* pro_row(const Pro& p, const float* src, int64_t row, int rowlen) {
  if (p.kind != PRO_SPREAD) return src + row * rowlen;
  const unsigned r = (unsigned)row, n1 = (unsigned)p.rows1;
  const unsigned z = r / n1, y = r - z * n1;
  return src + (int64_t)((z / p.f0) * (n1 / p.f1) + y / p.f1) * (rowlen / p.f2);
}
// Offset within that row of row sample c.
__host__ __device__ __forceinline__ int pro_col(const Pro& p, int c) {
  return p.kind == PRO_SPREAD ? c / p.f2 : c;
}

struct Epi {
  int kind;
  float a, b;
  const float* aux0;
  const float* aux1;
  double* partials;  // per-block partial sums (nullptr: no reduction)
  // Reduction box [z0, z1) x [y0, y1) x [x0, x1) of the output grid: only its
  // samples enter the partial sums (the X inverse passes); z1 <= z0: all.
  int box[6];
};

// Launch helpers (host).
struct ColLaunch {
  int L;        // columns per workgroup
  int T;        // threads per column
};
struct RowLaunch {
  int R;        // rows per workgroup
  int T;        // threads per row
};

// tile = columns (col passes) or rows (row passes) per workgroup; 0 = heuristic.
ColLaunch col_launch_for(int n, int64_t ncols, int tile);
RowLaunch row_launch_for(int m, int tile);
bool col_tile_ok(int n, int L);
bool row_tile_ok(int m, int R);

// Column FFT along a strided axis: for every batch b and column c,
//   line e (canvas index) = src[b*src_b + srow(e)*rstride + c], where rows
//   [0,lo) map to themselves, rows [n-hi,n) map to lo..lo+hi-1, others are 0;
//   stores rows [0, n_out) of the transform (times scale) to dst.
hipError_t launch_col_fft(int dir, const AxisSched& s, const float2* tw, const float2* src,
                          float2* dst, int64_t ncols, int nbatch, int64_t src_b, int64_t dst_b,
                          int64_t src_rstride, int64_t dst_rstride, int lo, int hi, int n_out,
                          float scale, int tile, hipStream_t st);

// Fused Z pass: forward FFT (pruned input rows lo/hi), multiply by otf (or its
// conjugate), inverse FFT, store rows [0, n_out).  In place on buf.
hipError_t launch_z_conv(const AxisSched& f, const AxisSched& iv, const float2* tw, float2* buf,
                         const float2* otf, int64_t ncols, int lo, int n_out, int conj, int tile,
                         hipStream_t st);

// Forward real-to-complex X pass on rows of a real volume (row length
// rowlen = lo + hi source samples mapped into a canvas of 2*m).
hipError_t launch_x_r2c(const AxisSched& s, const float2* twm, const float2* twp, const float* src,
                        float2* dst, int64_t nrows, int rowlen, int lo, int hi, int H,
                        const Pro& pro, int tile, hipStream_t st);

// Inverse complex-to-real X pass; stores the first s2 samples with the epilogue.
// Returns the number of workgroups (= partials written) through nblocks.
hipError_t launch_x_c2r(const AxisSched& s, const float2* twm, const float2* twp, const float2* src,
                        float* dst, int64_t nrows, int s2, int H, const Epi& epi, int tile,
                        int64_t* nblocks, hipStream_t st);

// Deterministic final reduction of n partials into out[0] (sum).
hipError_t launch_sum_partials(const double* partials, int64_t n, double* out, hipStream_t st);

int threads_per_line(int n);

// ---- compile-time specialised passes (dh_fftct*.hpp): per-volume arguments ----
//
// Blocked spectrum layout (the X/Y spectra of specialised plans and their
// OTF): a plane of `rows` rows and H columns is stored as ceil(H / kBlk)
// column blocks one after another.  Block b holds columns [b*kBlk, b*kBlk+w_b),
// w_b = min(kBlk, H - b*kBlk), row-major inside the block, so element (r, c)
// sits at b*kBlk*rows + r*w_b + (c - b*kBlk).  The plane stays dense (rows*H
// elements); a Y pass reads kBlk columns of consecutive rows as one contiguous
// run, and the X passes write/read kBlk samples of several rows contiguously.
constexpr int kBlk = 8;

// LDS elements of the specialised kernels: one column workgroup of L lines of
// length n (min(L, 8) spare slots every 16 line elements, so that narrow tiles of
// long lines, as 2 columns of 4096, keep their padding in proportion and fit two
// workgroups per CU), and one X row of a length-m transform (one spare slot every
// 32 elements, indices 0..m used).
constexpr int64_t col_lds_elems(int n, int L) { return (int64_t)n * L + (L < 8 ? L : 8) * (n >> 4) + 8; }
constexpr int xrow_stride(int m) { return m + (m >> 5) + 1; }
// Lengths whose X workgroups read the twiddle tables from global memory (through
// the caches) instead of staging them in LDS: there the tables would take as much
// LDS as the rows themselves and halve the rows per workgroup.
constexpr bool xrow_tw_global(int m) { return m >= 2048; }
// X workgroup of R rows: the rows, then the twiddle tables (m + m + 1 entries) unless
// they are read from global memory.
constexpr int64_t xrow_lds_elems(int m, int R) {
  return (int64_t)R * xrow_stride(m) + (xrow_tw_global(m) ? 0 : 2 * m + 1);
}

struct ColArgs {  // Y forward/inverse, OTF Z: columns of (batch, n, ncols) planes
  const float2* src;
  float2* dst;
  const float2* tw;
  // blk == 0: element (row r, column c) at r*src_rs + c (dst_rs for dst);
  // blk  > 0: blocked layout, src_rs/dst_rs = rows of the source/destination plane
  int64_t ncols, src_b, dst_b, src_rs, dst_rs, ntiles;
  int lo, hi, n_out;
  float scale;
  int blk;
};
struct ZArgs {  // fused Z: forward, x OTF, inverse, in place
  float2* buf;
  const float2* otf;
  const float2* tw;
  int64_t ncols;
  int lo, n_out, conj;
};
struct XfArgs {  // X real-to-complex rows
  const float* src;  // may be null with PRO_ONES, which never reads it
  float2* dst;
  const float2* twm;
  const float2* twp;
  int64_t nrows;
  int rowlen, lo, hi, H;
  Pro pro;
  int vec2;  // rows start 8-byte aligned and hi == 0: load sample pairs as float2
  int blk;   // 0: spectrum rows at row*H; > 0: blocked layout with prow rows per plane
  int prow;
};
struct XiArgs {  // X complex-to-real rows with the epilogue
  const float2* src;
  float* dst;        // cropped real rows (may be null in a round trip: not stored)
  float2* spec;      // round trip only: forward X spectrum of the epilogue's output (blocked)
  const float2* twm;
  const float2* twp;
  int64_t nrows;
  int s2, H;
  Epi epi;
  int vec2;  // output and epilogue operands 8-byte aligned per row: float2 accesses
  int blk;   // as XfArgs
  int prow;
};

// PSF synthesis (pupil_to_psf): Y pass over the columns of every z-plane, with the
// defocused pupil formed on load; X pass over the rows, storing |.|^2.
struct PsfYArgs {
  const float2* pupil;  // (ny, nx) complex pupil
  const double* kz;     // (ny, nx) axial spatial frequency
  const double* z;      // (nz) axial positions
  float2* work;         // (nz, ny, nx) out: the y-transformed defocused pupils
  const float2* tw;     // pass tables of length ny
  int nx;
  int64_t ntiles;       // column tiles per plane
};
struct PsfXArgs {
  const float2* work;   // (nz*ny, nx) rows
  float* psf;           // (nz*ny, nx) out
  const float2* tw;     // pass tables of length nx
  int64_t nrows;
  float scale;          // channel weight / (ny*nx)^2 (the inverse transform's normalisation, squared)
  int accumulate;       // 0: psf = scale |v|^2; 1: psf += scale |v|^2 (later channels)
  double* partials;     // one per workgroup tile
};

// Up to kPack volumes of one transform length per launch, passed by value.
constexpr int kPack = 16;
template <class A>
struct Pack {
  A a[kPack];
  int64_t prefix[kPack + 1];  // first logical block of each volume
  int nvol;
};

// CT_XR: X round trip = inverse X pass with the epilogue, whose output rows are
// transformed again by the forward X pass (XiArgs.spec) without leaving the CU.
enum CtKind { CT_COL = 0, CT_Z = 1, CT_XF = 2, CT_XI = 3, CT_XR = 4, CT_PSF_Y = 5, CT_PSF_X = 6 };
// True if length n has specialised kernels.
bool ct_supported(int n);
// Threads per line of the specialised kernels for length n.
int ct_threads(int n);
// Threads per row of the specialised X kernels (their row schedule).
int ct_row_threads(int n);
// Launches kind for length n; pk points to the matching Pack<...>.
// tile = columns (CT_COL, CT_Z) or rows (CT_XF, CT_XI) per workgroup, plus
// kTilePersistent for a persistent grid (one workgroup per resident slot, each
// walking tiles with a next-tile prefetch) instead of one workgroup per tile.
constexpr int kTilePersistent = 0x10000;
constexpr int kTileMask = 0xFFFF;
hipError_t ct_launch(int kind, int n, int dir, const void* pk, int tile, int64_t nblocks, hipStream_t st);

// ---- pass ids ----
enum PassId { PASS_X_FWD = 0, PASS_Y_FWD = 1, PASS_Z = 2, PASS_Y_INV = 3, PASS_X_INV = 4, PASS_N = 5 };


}  // namespace dh
