/*
 * deconhip.h -- C ABI of libdeconhip.so, the MI355X (gfx950) implementation of
 * deconlib's FFT-operator deconvolution hot path.
 *
 * Conventions
 *   - Every device buffer is allocated by the caller (PyTorch in the Python
 *     package) and passed in as a plain pointer.  The library never allocates
 *     or frees device memory; plans are host-side structs.
 *   - Every call enqueues work on the caller's HIP stream (passed as void*;
 *     NULL = the null stream) and never synchronises, except the calls that
 *     return a host value (documented below).
 *   - Status: 0 = OK, -1 = invalid argument (the Python layer raises
 *     ValueError, as the reference does for shape errors), -2 = HIP error,
 *     -3 = workspace too small, -4 = unsupported.  dh_last_error() returns a
 *     thread-local message for the last failure.
 *   - Volumes are C-contiguous float32 (z, y, x); 1-D and 2-D inputs are
 *     promoted to 3-D with leading unit axes.  Spectra are interleaved
 *     complex64 (P0, P1, P2/2+1) planes; within a plane the
 *     (P1, H = P2/2+1) samples are stored in the order given by
 *     dh_plan_info.spectrum_block (see there).
 */
#ifndef DECONHIP_H
#define DECONHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DH_ABI_VERSION 11

#define DH_OK 0
#define DH_EINVAL (-1)
#define DH_EHIP (-2)
#define DH_EWORKSPACE (-3)
#define DH_EUNSUPPORTED (-4)

typedef struct dh_conv_plan dh_conv_plan;

typedef struct {
  int32_t ndim;
  int32_t signal_shape[3];     /* promoted to 3-D                          */
  int32_t kernel_shape[3];
  int32_t fft_shape[3];        /* canvas the reference would use            */
  int32_t canvas_shape[3];     /* canvas this plan computes on (x made even) */
  int64_t otf_bytes;           /* complex64 (P0, P1, P2/2+1)                */
  int64_t work_bytes;          /* scratch for the pruned passes + partials  */
  int64_t twiddle_bytes;       /* device twiddle tables                     */
  /* Order of the H = P2/2+1 columns of each OTF plane.  0: row-major (P1, H).
   * B > 0: column blocks of width B, block b holding columns
   * [bB, bB + w_b), w_b = min(B, H - bB), each block row-major (P1, w_b) and
   * the blocks one after another; sample (y, c) sits at
   * bB*P1 + y*w_b + (c - bB), b = c / B.  Planes are dense either way. */
  int32_t spectrum_block;
} dh_plan_info;

/* Fused prologue applied to each input sample of the forward FFT pass
 * (the padded canvas never exists in memory). */
#define DH_PRO_NONE 0        /* v = in[i]                                 */
#define DH_PRO_ONES 1        /* v = 1            (RL sensitivity A^T 1)   */
#define DH_PRO_SUB_FLOOR 2   /* v = max(in[i] - a, b)  (RL max(obs-bg,1e-6)) */
#define DH_PRO_SPREAD 3      /* v[z,y,x] = in[z/f0, y/f1, x/f2], (f0,f1,f2) = spread[], f2 1 or 2: the input is the
                              * coarse volume (signal/f per axis) of FractionalAreaDownsample's
                              * integer-factor adjoint (core_operators.py:205-219, its banded passes
                              * with unit taps), so compose(Downsample, Convolver)'s adjoint is one
                              * application; convolution plans only */
#define DH_PRO_PD_Y1 4       /* v = y1' = prox(in + sigma ((1 + theta) pro_aux[0] - theta pro_aux[1]))
                              * with data pro_aux[2] (PDHG's data-term dual step, pdhg_mlx.py:115-189,
                              * 717-723: the same values as dh_pd_y1), y1' also stored to pro_out
                              * (which may be in); sigma and theta from pro_scalars (dh_pd_* block),
                              * pro_a = background, pro_loss 0 Poisson / 1 Gaussian.  Plans whose
                              * canvas has specialised kernels only (dh_plan_info.spectrum_block > 0) */

/* Fused epilogue applied to each cropped output sample v of the inverse pass. */
#define DH_EPI_NONE 0        /* out = v                                   */
#define DH_EPI_FLOOR 1       /* out = max(v, a)                           */
#define DH_EPI_RL_RATIO 2    /* m = v + a; out = aux0 / max(m, b)         */
                             /*   reduce: sum aux0*log(max(aux0,1e-6)/max(m,1e-6)) - (aux0 - max(m,1e-6)) */
#define DH_EPI_RL_UPDATE 3   /* out = max(aux0 * (v / aux1), b)  (aux1 = +inf off-support) */
#define DH_EPI_ADD 4         /* out = v + a      reduce: same I-divergence as RL_RATIO when aux0 != NULL */
#define DH_EPI_SQDIFF 5      /* out = v          reduce: sum (v - aux0)^2 (PDHG's |A x_new - A x|^2)     */
#define DH_EPI_BLOCKSUM2 6   /* out[z,y,x] = sum of v over the 2x2x2 block at (2z, 2y, 2x): out is the
                              * coarse (S0/2, S1/2, S2/2) volume, FractionalAreaDownsample(2)'s block
                              * sum (core_operators.py:205-219) in its passes' order, so
                              * compose(Downsample(2), Convolver) is one application.  Even signal
                              * extents, dh_conv_apply on plans with specialised kernels only;
                              * no reduction */

typedef struct {
  int32_t prologue;
  float pro_a, pro_b;
  int32_t epilogue;
  float epi_a, epi_b;
  const float* aux0;         /* epilogue operand (observed / x), S-shaped */
  const float* aux1;         /* epilogue operand (sensitivity), S-shaped  */
  double* reduce_out;        /* device; if non-NULL the epilogue's sum is written here */
  int32_t spread[3];         /* DH_PRO_SPREAD factors (z, y, x of the signal promoted to 3-D) */
  int32_t pro_loss;          /* DH_PRO_PD_Y1 loss: 0 Poisson, 1 Gaussian */
  const float* pro_aux[3];   /* DH_PRO_PD_Y1 operands: A x_k, A x_{k-1}, data (signal-shaped) */
  float* pro_out;            /* DH_PRO_PD_Y1: y1' (signal-shaped; may be the input) */
  const double* pro_scalars; /* DH_PRO_PD_Y1: device scalar block of the dh_pd_* calls */
} dh_fuse_args;

int dh_abi_version(void);
const char* dh_last_error(void);

/* Plan for LinearFFTConvolver (linops_mlx.py:435-533) when circular == 0,
 * FFTConvolver (linops_mlx.py:384-432) when circular == 1 (then signal,
 * kernel and fft shapes must all be equal).  fft_shape may be NULL: the
 * reference's fast_padded_shape (linops_mlx.py:975-1032) is used. */
int dh_conv_plan_create(int ndim, const int64_t* signal_shape, const int64_t* kernel_shape,
                        const int64_t* fft_shape, int circular, dh_conv_plan** out);
int dh_conv_plan_info(const dh_conv_plan* plan, dh_plan_info* info);
/* Upload the twiddle tables into twiddles_dev (twiddle_bytes), asynchronously on the stream
 * (dh_conv_plan_destroy waits for the upload). */
int dh_conv_plan_bind(dh_conv_plan* plan, void* twiddles_dev, void* stream);
int dh_conv_plan_destroy(dh_conv_plan* plan);
/* Launch shape of one pass (DH_PASS_*): columns per workgroup for the Y/Z
 * passes, rows per workgroup for the X passes (a power of two), optionally
 * plus DH_TILE_PERSISTENT for a persistent grid (one workgroup per resident
 * slot walking the tiles, prefetching the next tile's operands) instead of one
 * workgroup per tile; 0 restores the built-in heuristic.  Semantics-free: used
 * by the autotuner.  get_tile returns the value in effect. */
#define DH_TILE_PERSISTENT 0x10000
int dh_conv_plan_set_tile(dh_conv_plan* plan, int pass, int tile);
int dh_conv_plan_get_tile(const dh_conv_plan* plan, int pass);

/* OTF = rfftn(pad_corner_origin_kernel(kernel / sum(kernel)))   (linops_mlx.py:408-411,
 * utils/padding.py:142-190), stored pre-scaled by 1/prod(canvas).  Writes
 * max|OTF|^2 (operator_norm_sq, linops_mlx.py:415) to *norm_sq_host:
 * this call synchronises the stream. */
int dh_otf_build(const dh_conv_plan* plan, const float* kernel_dev, int normalize,
                 void* otf_dev, void* work_dev, int64_t work_bytes,
                 double* norm_sq_host, void* stream);

/* dh_otf_build without a host read: max|OTF|^2 goes to the device double *norm_sq_dev
 * (may be NULL) and nothing synchronises, so operators can be built while other work runs
 * (serving admission); the caller reads the value when it needs it. */
int dh_otf_build_dev(const dh_conv_plan* plan, const float* kernel_dev, int normalize,
                     void* otf_dev, void* work_dev, int64_t work_bytes,
                     double* norm_sq_dev, void* stream);

/* One forward (adjoint == 0: convolution, linops_mlx.py:520-524) or adjoint
 * (adjoint == 1: correlation, linops_mlx.py:526-530) application, S -> S.
 * fuse may be NULL.  in and out may alias only for elementwise-safe epilogues. */
int dh_conv_apply(const dh_conv_plan* plan, int adjoint, const float* in, float* out,
                  const void* otf_dev, void* work_dev, int64_t work_bytes,
                  const dh_fuse_args* fuse, void* stream);

/* One Richardson-Lucy iteration (the compiled step _rl_step, rl_mlx.py:146-150):
 *   x <- max(x * A^T(obs / max(A x + background, eps)) / sens, eps)
 * with sens = A^T 1 on the support and +inf off it (dh_rl_support).  The ratio never
 * reaches HBM: the forward convolution's inverse X pass computes it and runs the
 * adjoint's forward X pass on it in the same kernel.  reduce_out (device, may be NULL)
 * receives sum over voxels of the I-divergence of obs against A x + background for
 * the input x (rl_mlx.py:58-80, unnormalised).  spectrum_ready != 0: the workspace still
 * holds the forward X spectrum of x, left by the previous call on this workspace with
 * keep_spectrum != 0 and nothing run on the workspace since.  keep_spectrum != 0: the
 * update pass also leaves the forward X spectrum of the new x there.  Results agree
 * with two dh_conv_apply calls with DH_EPI_RL_RATIO / DH_EPI_RL_UPDATE to float32
 * rounding (the fused kernel may contract multiply-adds differently: relative
 * difference <= 1e-6), not bit for bit.
 * DH_EUNSUPPORTED for canvases without specialised kernels. */
int dh_rl_iteration(const dh_conv_plan* plan, const void* otf_dev, void* work_dev, int64_t work_bytes,
                    float* x, const float* obs, const float* sens, float background, float eps,
                    double* reduce_out, int spectrum_ready, int keep_spectrum, void* stream);

/* dh_conv_apply without a workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The workspace is a file-scope device
 * array of DH_STATIC_SCRATCH_BYTES per GPU, held until the process ends; the call
 * allocates nothing.  Plans whose work_bytes exceed it fail with DH_EWORKSPACE (e.g. a
 * 64x256x256 volume on its 128x512x512 canvas needs 101 MB and fits; 64x512x512 does
 * not).  Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (concurrent streams or threads) race and must be serialised by the caller.  Nothing
 * in the array needs clearing between calls.  Results equal dh_conv_apply's bit for bit. */
#define DH_STATIC_SCRATCH_BYTES (128LL << 20)
int dh_conv_apply_static(const dh_conv_plan* plan, int adjoint, const float* in, float* out,
                         const void* otf_dev, const dh_fuse_args* fuse, void* stream);

/* ---- batched application: many plans (different shapes), few launches ---- */
typedef struct {
  const dh_conv_plan* plan;
  const void* otf;
  void* work;                /* this volume's own workspace (volumes run concurrently) */
  int64_t work_bytes;
  const float* in;
  float* out;
  dh_fuse_args fuse;         /* per-volume prologue/epilogue; reduce_out per volume */
} dh_batch_item;
/* Applies every item's convolution (adjoint == 0) or correlation (adjoint == 1)
 * with its fused prologue/epilogue, with the same results as one dh_conv_apply
 * per item.  Items whose canvas has specialised kernels are grouped per pass by
 * transform length: each group of up to 16 volumes is one launch, descriptors
 * in the kernel arguments.  Serving and batched RL call this once per
 * half-iteration. */
int dh_batch_conv_apply(int nitems, const dh_batch_item* items, int adjoint, void* stream);

/* Batched fused RL iterations: one dh_rl_iteration per item (each item its own plan,
 * OTF, workspace and state; any mix of shapes), run stage by stage with the items
 * grouped per transform length, so a batch iteration is a few launches per stage.
 * Each item's result is bit-identical to dh_rl_iteration on it alone. */
typedef struct {
  const dh_conv_plan* plan;
  const void* otf;
  void* work;
  int64_t work_bytes;
  float* x;
  const float* obs;
  const float* sens;
  float background;
  float eps;
  double* reduce_out;          /* device, may be NULL */
  int32_t spectrum_ready;
  int32_t keep_spectrum;
  /* I-divergence box {z0, z1, y0, y1, x0, x1} of the signal grid: only samples
   * inside it enter reduce_out (a data box embedded in a padded domain, obs zero
   * outside, as a tiled forward model's crop); z1 <= z0: the whole grid.  No
   * counterpart in the reference's RL call: its tiled path crops instead
   * (deconvolution/forward_model.py:106-118). */
  int32_t data_box[6];
} dh_rl_item;
int dh_batch_rl_iteration(int nitems, const dh_rl_item* items, void* stream);

/* ---- PSF synthesis: pupil_to_psf / pupil_to_vectorial_psf (psf/widefield.py:18-73,110-234) ----
 * psf[z, y, x] = sum_c w_c |IFFT2(pupil_c * exp(2 pi i kz z))|^2 per z-plane (numpy's
 * ifft2 normalisation, DC at the corner), normalised to sum 1 when normalize != 0.
 * One channel of weight 1 is pupil_to_psf; the vectorial PSF's channels are the
 * apodised pupil times each dipole-to-field factor, incoherently summed.
 * pupils: (nchan, ny, nx) complex64 interleaved, kz: (ny, nx) float64, z: (nz)
 * float64, psf: (nz, ny, nx) float32 (device); weights: nchan float64 (host).
 * The defocus phase is reduced to [0, 1) turns in float64; the transforms run in
 * float32 and the sums in float64.  total_out (device, may be NULL) receives the
 * sum before normalisation.  ny and nx must be 2^a * {1, 3, 5} <= 4096
 * (DH_EUNSUPPORTED otherwise). */
typedef struct dh_psf_plan dh_psf_plan;
int dh_psf_plan_create(int ny, int nx, dh_psf_plan** out);
int dh_psf_plan_info(const dh_psf_plan* plan, int nz, int nchan, int64_t* twiddle_bytes, int64_t* work_bytes);
int dh_psf_plan_bind(dh_psf_plan* plan, void* twiddles_dev, void* stream);
int dh_psf_synth(const dh_psf_plan* plan, int nchan, const void* pupils, const double* weights, const double* kz,
                 const double* z, int nz, int normalize, float* psf, void* work, int64_t work_bytes,
                 double* total_out, void* stream);
int dh_psf_plan_destroy(dh_psf_plan* plan);

/* ---- reductions and elementwise steps (rl_mlx.py:123-150) ---- */
#define DH_RED_SUM 0
#define DH_RED_MAX 1
#define DH_RED_SUMSQ 2
/* result_dev[0] = reduction of x[0:n] in float64, fixed order (deterministic).
 * partials_dev must hold dh_reduce_partials(n) doubles. */
int64_t dh_reduce_partials(int64_t n);
int dh_reduce(int op, const float* x, int64_t n, double* partials_dev, double* result_dev,
              void* stream);

/* RL support set-up (rl_mlx.py:129-133): floor = max(float(maxval)*1e-6f, 1e-10f);
 * active = sens > floor; sens = active ? sens : +inf; x = active ? x : 0. */
int dh_rl_support(float* sens, float* x, int64_t n, const double* maxval_dev, void* stream);
/* out = max(in, lo) */
int dh_clamp_min(const float* in, float* out, int64_t n, float lo, void* stream);
/* out = max(in - a, b) */
int dh_sub_floor(const float* in, float* out, int64_t n, float a, float b, void* stream);
/* RL ratio on an arbitrary operator's output (rl_mlx.py:147-148):
 * out = obs / max(model + bg, eps); reduce_out (may be NULL) gets the
 * I-divergence sum of obs against model + bg. */
int dh_rl_ratio(const float* model, const float* obs, float* out, int64_t n, float bg,
                float eps, double* partials_dev, double* reduce_out, void* stream);
/* RL update (rl_mlx.py:149-150): x = max(x * (corr / sens), eps). */
int dh_rl_update(float* x, const float* corr, const float* sens, int64_t n, float eps,
                 void* stream);
/* result_dev[0] = sum over i of obs*log(max(obs,eps)/max(model,eps)) - (obs - max(model,eps)),
 * the unnormalised poisson_i_divergence (rl_mlx.py:58-80) for any eps, float64 partials. */
int dh_poisson_idiv(const float* obs, const float* model, int64_t n, float eps, double* partials_dev,
                    double* result_dev, void* stream);
/* out = in + a (forward model plus background); reduce as dh_rl_ratio. */
int dh_add_scalar(const float* in, const float* obs, float* out, int64_t n, float a,
                  double* partials_dev, double* reduce_out, void* stream);

/* ---- PDHG with Hessian regularization (pdhg_mlx.py:507-920) ---- */
/* float64 scalar block (caller-owned device array of DH_PD_NSLOTS doubles). */
#define DH_PD_TAU 0
#define DH_PD_SIGMA 1
#define DH_PD_THETA 2
#define DH_PD_STEP 3      /* 0.99 / sqrt(|A|^2 + |L|^2) */
#define DH_PD_COUNT 4     /* consecutive iterations under tol */
#define DH_PD_DONE 5      /* 1.0 once converged */
#define DH_PD_FINAL 6     /* iteration count at convergence */
#define DH_PD_DX2 7       /* sum (x_new - x)^2 */
#define DH_PD_XX 8        /* sum x^2 */
#define DH_PD_XNXN 9      /* sum x_new^2 */
#define DH_PD_DA2 10      /* sum (A x_new - A x)^2 */
#define DH_PD_DL2 11      /* sum (L x_new - L x)^2 */
#define DH_PD_TMP0 12
#define DH_PD_TMP1 13
#define DH_PD_NSLOTS 16

typedef struct {
  double delta, eta, tol;
  int32_t min_iter, patience;
} dh_pd_params;

/* Partials (doubles) the PDHG reductions need for an (nz, ny, nx) volume whose
 * data domain has at most as many voxels. */
int64_t dh_pd_partials(int nz, int ny, int nx);
/* y1 <- prox_sigma F*(y1 + sigma ((1+theta) ax - theta axp))  (pdhg_mlx.py:115-189, 723-727);
 * loss_type 0 = poisson, 1 = gaussian. */
int dh_pd_y1(float* y1, const float* ax, const float* axp, const float* data, int64_t n, const double* scalars,
             float background, int loss_type, void* stream);
/* y2 (6, nz, ny, nx) <- prox(y2 + sigma H(x + theta (x - xp)))  with clip (l12 = 0) or the
 * per-voxel L2 ball (l12 = 1) of radius alpha  (pdhg_mlx.py:729-734, 197-233).
 * w = Hessian3D weights [r^2, 1, 1, r sqrt2, r sqrt2, sqrt2]. */
int dh_pd_dual(float* y2, const float* x, const float* xp, int nz, int ny, int nx, const double* scalars,
               const float* w, float alpha, int l12, void* stream);
/* x_new = max(0, x - tau (g + H^T y2)); writes DX2, XX, XNXN  (pdhg_mlx.py:741-750, 805-806). */
int dh_pd_primal(float* x_new, const float* x, const float* g, const float* y2, int nz, int ny, int nx,
                 double* scalars, const float* w, double* partials, void* stream);
/* The dual and primal steps in one sweep (pdhg_mlx.py:729-750): y2_new <- prox(y2 + sigma
 * H(x + theta (x - xp))) as dh_pd_dual, then x_new = max(0, x - tau (g + H^T y2_new)) as
 * dh_pd_primal (writes DX2, XX, XNXN), with g = A^T y1_new computed beforehand.  y2_new and
 * x_new must not overlap x, xp, y2 or g (neighbouring tiles re-read those).  Same values as
 * the two calls. */
int dh_pd_dual_primal(float* y2_new, float* x_new, const float* x, const float* xp, const float* y2,
                      const float* g, int nz, int ny, int nx, double* scalars, const float* w, float alpha,
                      int l12, double* partials, void* stream);
/* DL2 = sum |H a - H b|^2; if reg_loss_out, *reg_loss_out = sum |H a| (reg_norm 1, "L1")
 * or sum sqrt(sum_q (H a)_q^2 + 1e-12) (reg_norm 2, "L1_2")  (pdhg_mlx.py:874-882). */
int dh_pd_hess_sums(const float* a, const float* b, int nz, int ny, int nx, const float* w, double* scalars,
                    double* reg_loss_out, int reg_norm, double* partials, void* stream);
/* DA2 = sum (a - b)^2; if data_loss_out, the data loss of a + background against data. */
int dh_pd_data_sums(const float* a, const float* b, const float* data, int64_t n, float background,
                    int loss_type, double* scalars, double* data_loss_out, double* partials, void* stream);
/* Initial scalar block (pdhg_mlx.py:693-701): tau = sigma = step = 0.99 / sqrt(*blur_norm_sq_dev
 * + reg_norm_sq), theta = 1, every other slot 0; the blur norm stays on the device. */
int dh_pd_init_scalars(double* scalars, const double* blur_norm_sq_dev, double reg_norm_sq, void* stream);
/* Malitsky-Pock step update and convergence test for iteration k in float64
 * (pdhg_mlx.py:761-835); hist[3k..3k+2] = tau_new, sigma_new, rel_change. */
int dh_pd_scalars(double* scalars, double* hist, int k, const dh_pd_params* params, void* stream);
/* Hessian3D.forward f -> out (6, nz, ny, nx) and adjoint y -> out (linops_mlx.py:224-259). */
int dh_hessian_forward(const float* f, float* out, int nz, int ny, int nx, const float* w, void* stream);
int dh_hessian_adjoint(const float* y, float* out, int nz, int ny, int nx, const float* w, void* stream);

/* ---- PDHG with any of the reference's regularizers (pdhg_mlx.py:241-499) ----
 * reg selects L: DH_REG_HESSIAN (HessianRegularizer, linops_mlx.py:130-262: 6 components
 * [zz, yy, xx, yz, xz, xy], w = 6 weights; Hessian2D = one plane with [0, 1, 1, 0, 0, sqrt2],
 * Hessian1D = one row with [0, 0, 1, 0, 0, 0]), DH_REG_GRADIENT (GradientRegularizer,
 * linops_mlx.py:46-127: forward differences d1_fwd (linops_core_mlx.py:25-39) along [z, y, x]
 * with w = 3 weights [r, 1, 1]; Gradient2D = [0, 1, 1] on one plane, Gradient1D = [0, 0, 1] on
 * one row) or DH_REG_IDENTITY (IdentityRegularizer, pdhg_mlx.py:241-306: one component, w
 * unused).  y2 holds dh_reg_components(reg) planes of nz*ny*nx.  Same semantics as the
 * Hessian-only calls above, which these extend. */
#define DH_REG_HESSIAN 0
#define DH_REG_GRADIENT 1
#define DH_REG_IDENTITY 2
int dh_reg_components(int reg);
int dh_pd_reg_dual(int reg, float* y2, const float* x, const float* xp, int nz, int ny, int nx,
                   const double* scalars, const float* w, float alpha, int l12, void* stream);
int dh_pd_reg_primal(int reg, float* x_new, const float* x, const float* g, const float* y2, int nz, int ny,
                     int nx, double* scalars, const float* w, double* partials, void* stream);
int dh_pd_reg_sums(int reg, const float* a, const float* b, int nz, int ny, int nx, const float* w,
                   double* scalars, double* reg_loss_out, int reg_norm, double* partials, void* stream);
/* L.forward f -> out (components, nz, ny, nx) and L.adjoint y -> out. */
int dh_reg_forward(int reg, const float* f, float* out, int nz, int ny, int nx, const float* w, void* stream);
int dh_reg_adjoint(int reg, const float* y, float* out, int nz, int ny, int nx, const float* w, void* stream);

/* ---- NLCG: Poisson ML by nonlinear conjugate gradients (nlcg_mlx.py:84-529) ----
 * f = s^2, m = K f + b.  phi = sum m - sum g ln max(m, eps) + beta ||C f||^2.
 * Vectors are float32 device arrays of n elements (domain: s, f, d, r, sd, dd, KTw,
 * CtCf; data: m, g, w, Ksd, Kdd); the convolutions are dh_conv_apply calls made
 * in between.  Scalars live in a caller-owned device block of DH_NL_NSLOTS
 * doubles, zeroed once; partials is a device array of dh_nlcg_partials()
 * doubles.  No call synchronises: the host reads the block when it decides. */
#define DH_NL_LAM 0          /* step length lambda of the current line search   */
#define DH_NL_NEWTON_OK 1    /* 1 while the Newton-Raphson refinement continues  */
#define DH_NL_DPHI1 2        /* last dphi/dlambda                                */
#define DH_NL_DPHI2 3        /* last d2phi/dlambda2                              */
#define DH_NL_PHI0 4         /* phi at the current iterate                       */
#define DH_NL_PHI_TRIAL 5    /* phi at s + lambda d                              */
#define DH_NL_ACCEPT 6       /* 1 if the trial decreases phi (not phi >= phi0)   */
#define DH_NL_TRIAL_DATA 7   /* data part of phi at the trial model              */
#define DH_NL_TRIAL_LOSS 8   /* sum of the Poisson I-divergence of the trial     */
#define DH_NL_MDATA 9        /* data part of phi at the current model            */
#define DH_NL_LOSS 10        /* sum of the Poisson I-divergence of the current m */
#define DH_NL_IDIV 11        /* I-divergence between the last two iterates       */
#define DH_NL_SUMF 12        /* sum f of the current iterate                     */
#define DH_NL_RR 13          /* sum r^2 of the current residual                  */
#define DH_NL_GAMMA 14       /* Fletcher-Reeves gamma of the last direction      */
#define DH_NL_LAM_ACCEPTED 15 /* lambda of the last accepted step                */
#define DH_NL_A0 16          /* regularizer dots: sum Cf.Cf, Cf.Csd, Cf.Cdd,     */
#define DH_NL_NSLOTS 32      /*   Csd.Csd, Csd.Cdd, Cdd.Cdd in slots A0..A0+5     */

/* Partials (doubles) every dh_nlcg_* reduction may use. */
int64_t dh_nlcg_partials(void);
/* s = sqrt(max(f0, floor)), f = s * s; f0 may alias s  (nlcg_mlx.py:407-412, 88-93). */
int dh_nlcg_init(const float* f0, float* s, float* f, int64_t n, float floor_value, void* stream);
/* w = 1 - g / max(m, eps); MDATA = sum m - sum g ln max(m, eps); LOSS = the
 * I-divergence sum of m against g  (nlcg_mlx.py:95-107, rl_mlx.py:58-80). */
int dh_nlcg_model(const float* m, const float* g, float* w, int64_t n, float eps, double* scalars,
                  double* partials, void* stream);
/* Starts a line search along d: sd = s * d, dd = d * d, LAM = 0, NEWTON_OK =
 * (newton_iters > 0)  (nlcg_mlx.py:253-262). */
int dh_nlcg_begin_step(const float* s, const float* d, float* sd, float* dd, int64_t n, double* scalars,
                       int newton_iters, void* stream);
/* A0..A5 = the six dot products of (Cf, C(sd), C(dd)), arrays of n elements
 * each (nlcg_mlx.py:257-260, 273-281, with C f(lambda) expanded in lambda). */
int dh_nlcg_reg_dots(const float* cf, const float* csd, const float* cdd, int64_t n, double* scalars,
                     double* partials, void* stream);
/* One Newton-Raphson refinement of LAM on the exact quadratic-in-lambda
 * model K f(lambda) = (m - b) + 2 lambda Ksd + lambda^2 Kdd; a no-op once
 * NEWTON_OK is 0  (nlcg_mlx.py:263-288). */
int dh_nlcg_newton(const float* m, const float* ksd, const float* kdd, const float* g, int64_t n, float background,
                   double beta, double eps, double* scalars, double* partials, void* stream);
/* Trial model m_new = (m - b) + 2 LAM Ksd + LAM^2 Kdd + b, w = 1 - g / max(m_new, eps),
 * PHI0, PHI_TRIAL, ACCEPT, TRIAL_DATA, TRIAL_LOSS  (nlcg_mlx.py:290, 446-457). */
int dh_nlcg_trial(const float* m, const float* ksd, const float* kdd, const float* g, int64_t n, float background,
                  double beta, double eps, double* scalars, float* m_new, float* w, double* partials, void* stream);
/* Accepts the trial: s += LAM d, f = s * s; IDIV (between f before and after),
 * SUMF; MDATA, LOSS from the trial  (nlcg_mlx.py:450-451, 470-471, 294-314, 498). */
int dh_nlcg_update(float* s, const float* d, float* f, int64_t n, double eps, double* scalars, double* partials,
                   void* stream);
/* r = -(2 s KTw + c4b s CtCf) (ctcf may be NULL), rr_new = sum r^2.  mode 0: RR =
 * rr_new (first gradient); 1: GAMMA = rr_new / RR if RR > eps else 0, RR = rr_new;
 * 2: as 1 with GAMMA = 0 (restart)  (nlcg_mlx.py:104-113, 417-419, 472-479). */
int dh_nlcg_gradient(const float* s, const float* ktw, const float* ctcf, float c4b, float* r, int64_t n, int mode,
                     double eps, double* scalars, double* partials, void* stream);
/* d = r + GAMMA d  (nlcg_mlx.py:480). */
int dh_nlcg_direction(const float* r, float* d, int64_t n, const double* scalars, void* stream);
/* out2[0] = sum (g / max(m*m, eps)) Ksd^2 (skipped when g is NULL), out2[1] = sum v d^2
 * (the Hessian quadratic form's sums, nlcg_mlx.py:155-192). */
int dh_nlcg_quad(const float* g, const float* m, const float* ksd, const float* v, const float* d, int64_t n,
                 double eps, double* out2, double* partials, void* stream);
/* out2[0] = sum fp (ln max(fp, eps) - ln max(fc, eps)) - fp + fc, out2[1] = sum fc
 * (i_divergence_between, nlcg_mlx.py:294-314). */
int dh_nlcg_idiv(const float* fp, const float* fc, int64_t n, double eps, double* out2, double* partials,
                 void* stream);

/* ---- ER-Decon: Hessian-log regularised Gauss-Newton-CG (erdecon_mlx.py:144-692) ----
 * g = s^2, m = K g + b, q = sum_i (H_i g)^2 [+ iw g^2], w = lam / (eps + q) [+ w_floor].
 * Domain vectors have n elements, stacked regulariser responses (nc, n); the
 * convolutions and H^T (dh_hessian_adjoint or the operator's own) are made in between.
 * Scalars live in a caller-owned device block of DH_ER_NSLOTS doubles; partials is a
 * device array of dh_er_partials() doubles.  No call synchronises. */
#define DH_ER_RS 0        /* CG: sum r^2                                       */
#define DH_ER_RS_THR 1    /* CG: tol^2 * sum r0^2                              */
#define DH_ER_DENOM 2     /* CG: sum p (H_GN p)                                */
#define DH_ER_BETA 3      /* CG: rs_new / rs                                   */
#define DH_ER_STOP 4      /* CG: 1 once the solve has ended                    */
#define DH_ER_GG 5        /* sum grad^2                                        */
#define DH_ER_DOT 6       /* a dh_er_dot result slot                           */
#define DH_ER_DATA 7      /* data misfit sum                                   */
#define DH_ER_IDIV 8      /* Poisson I-divergence sum of m against f           */
#define DH_ER_LOGQ 9      /* sum ln(eps + q)                                   */
#define DH_ER_Q 10        /* sum q                                             */
#define DH_ER_NSLOTS 16
/* dh_er_ew operations (out2 only for AXPY_SQUARE, b only where named) */
#define DH_ER_EW_DIV 0           /* out = a / alpha                    */
#define DH_ER_EW_SQUARE_SCALE 1  /* out = (a * a) * alpha              */
#define DH_ER_EW_AXPY_SQUARE 2   /* out = a + alpha b, out2 = out^2    */
#define DH_ER_EW_NEG 3           /* out = -a                           */
#define DH_ER_EW_MUL 4           /* out = a * b                        */
#define DH_ER_EW_SCALE 5         /* out = alpha * a                    */
#define DH_ER_EW_AFFINE 6        /* out = alpha * a + beta             */
#define DH_ER_EW_MAX_DIV 7       /* out = max(a, alpha) / alpha        */
#define DH_ER_EW_AXPY 8          /* out = alpha * a + b                */

int64_t dh_er_partials(void);
/* Hessian3D (weights w6 = [r^2, 1, 1, r sqrt2, r sqrt2, sqrt2]; Hessian2D = one plane with
 * [0, 1, 1, 0, 0, sqrt2]) of g, fused with _weights (erdecon_mlx.py:144-214): LOGQ = sum
 * ln(eps + q), Q = sum q; with w_out (n) also w and u_out (6, n) = w * H g, the operand of
 * H^T in the gradient (erdecon_mlx.py:301-302). */
int dh_er_hess_weights(const float* g, int nz, int ny, int nx, const float* w6, float lam, float eps_reg,
                       float w_floor, float intensity_weight, float* w_out, float* u_out, double* scalars,
                       double* partials, void* stream);
/* The same from any operator's stacked responses hg (nc, n); u_out may alias hg. */
int dh_er_weights_stacked(const float* hg, int nc, const float* g, int64_t n, float lam, float eps_reg,
                          float w_floor, float intensity_weight, float* w_out, float* u_out, double* scalars,
                          double* partials, void* stream);
/* u_out (6, n) = w * H(v) (the regulariser half of erdecon_gn_hvp, erdecon_mlx.py:339-341). */
int dh_er_hess_apply_w(const float* v, const float* w, int nz, int ny, int nx, const float* w6, float* u_out,
                       void* stream);
/* out (nc, n) = w * x for stacked responses x (nc, n). */
int dh_er_scale_stacked(const float* x, const float* w, int nc, int64_t n, float* out, void* stream);
/* DATA = Gaussian sum (m - f)^2 or Poisson sum(mm - f ln mm), mm = max(m, 1e-8); IDIV = the
 * I-divergence sum of m against f; score / hessw (Poisson only) if score is non-NULL
 * (_data_misfit, _data_deriv: erdecon_mlx.py:220-250; rl_mlx.py:58-80). */
int dh_er_data(const float* m, const float* f, int64_t n, int poisson, float* score, float* hessw,
               double* scalars, double* partials, void* stream);
/* grad = 2 s dg + 2 s rh [+ 2 s (iw w g)] (erdecon_mlx.py:300-308); starts CG at x = 0,
 * r = p = -grad, sv = s p: GG = RS = sum grad^2, RS_THR = cg_tol^2 RS, STOP = (RS <= 1e-20)
 * (erdecon_mlx.py:366-372). */
int dh_er_gradient(const float* s, const float* dg, const float* rh, const float* w, const float* g,
                   float intensity_weight, int64_t n, double cg_tol, float* grad, float* x, float* r, float* p,
                   float* sv, double* scalars, double* partials, void* stream);
/* ap = 4 s kt + 4 s htu [+ 4 iw s (w sv)] (erdecon_gn_hvp, erdecon_mlx.py:336-345), DENOM =
 * sum p ap; STOP if DENOM is not finite or <= 1e-20 (erdecon_mlx.py:374-380). */
int dh_er_cg_ap(const float* s, const float* kt, const float* htu, const float* w, const float* sv,
                float intensity_weight, const float* p, float* ap, int64_t n, double* scalars, double* partials,
                void* stream);
/* Unless STOP: x += a p, r -= a ap (a = RS / DENOM); STOP if sum r^2 is not finite, negative
 * or <= RS_THR; else p = r + (rs_new / RS) p, sv = s p, RS = rs_new (erdecon_mlx.py:381-392). */
int dh_er_cg_update(float* x, float* r, float* p, const float* ap, const float* s, float* sv, int64_t n,
                    double* scalars, double* partials, void* stream);
/* scalars[slot] = sum a b (grad^T p, erdecon_mlx.py:611). */
int dh_er_dot(const float* a, const float* b, int64_t n, int slot, double* scalars, double* partials, void* stream);
/* Elementwise steps of the outer loops (DH_ER_EW_*): normalisation by the data scale, g = s^2,
 * the trial point s + alpha p and its square, -grad, d K(sv) (erdecon_mlx.py:544-560, 620-624, 687);
 * s0 K x + b, max(., s0) / s0 and the scaled gradient sum of JetNewton (jetnewton_mlx.py:197-225, 619-624). */
int dh_er_ew(int op, const float* a, const float* b, float alpha, float beta, float* out, float* out2, int64_t n,
             void* stream);

/* ---- JetNewton: active-set projected Newton, exact Hessian (jetnewton_mlx.py:111-785) ----
 * x is the dimensionless iterate (n voxels); r = H x is cached as (6, n) (or the operator's
 * stacked (nc, n)); masks are uint8 (1 = active / frozen).  Scalars live in a caller-owned
 * device block of DH_JN_NSLOTS doubles; partials holds dh_jn_partials() doubles. */
#define DH_JN_EPS 0          /* active-set eps = min(eps_bar, ||x - P+(x - g)||)          */
#define DH_JN_PG 1           /* ||x - P+(x - g)||_inf                                     */
#define DH_JN_KKT2 2         /* ||x - P+(x - g)||^2                                       */
#define DH_JN_ACTIVE 3       /* active-set size                                           */
#define DH_JN_RZ 4           /* CG: r . z (z = r)                                         */
#define DH_JN_TOL_K 5        /* CG: min(0.5, sqrt(r0)) r0                                 */
#define DH_JN_STOP 6         /* CG: 1 once the solve has ended                            */
#define DH_JN_DENOM 7        /* CG: p . A p                                               */
#define DH_JN_BETA 8         /* CG: rz_new / rz                                           */
#define DH_JN_NCG 9          /* CG: steps started                                         */
#define DH_JN_HITNEG 10      /* CG: 1 after non-positive curvature                        */
#define DH_JN_USE_RHS 11     /* CG: 1 if that happened on the first step (d_free = rhs)   */
#define DH_JN_ND 12          /* grad . d                                                  */
#define DH_JN_DIRECTIONAL 13 /* grad . (x_t - x)                                          */
#define DH_JN_LOGU 14        /* sum ln(eta + u)                                           */
#define DH_JN_USUM 15        /* sum u                                                     */
#define DH_JN_R0 16          /* CG: ||r0||                                                */
#define DH_JN_NSLOTS 24

int64_t dh_jn_partials(void);
/* Penalty weights of x (_penalty_weights, jetnewton_mlx.py:111-133) through the Hessian3D
 * stencil with weights w6 (one plane for the 2-D operators): LOGU, USUM; optional outputs
 * r (6, n), w, c = -w^2 / beta, w r (6, n) (the operand of H^T in the gradient), u. */
int dh_jn_penalty(const float* x, int nz, int ny, int nx, const float* w6, float beta, float eta, float* r_out,
                  float* w_out, float* c_out, float* wr_out, float* u_out, double* scalars, double* partials,
                  void* stream);
/* The same from any operator's stacked responses r (nc, n). */
int dh_jn_penalty_stacked(const float* r, int nc, int64_t n, float beta, float eta, float* w_out, float* c_out,
                          float* wr_out, float* u_out, double* scalars, double* partials, void* stream);
/* out (6, n) = w q + 2 (c <r, q>) r with q = H v (jetnewton_hvp, jetnewton_mlx.py:284-293). */
int dh_jn_hvp_operand(const float* v, const float* r, const float* w, const float* c, int nz, int ny, int nx,
                      const float* w6, float* out, void* stream);
/* The same from the operator's stacked q = H v (nc, n); out may alias q. */
int dh_jn_hvp_operand_stacked(const float* q, const float* r, const float* w, const float* c, int nc, int64_t n,
                              float* out, void* stream);
/* KKT2, EPS, PG of (x, g) (identify_active_set, jetnewton_mlx.py:302-315; proj-grad, :654-656). */
int dh_jn_kkt(const float* x, const float* g, int64_t n, double eps_bar, double* scalars, double* partials,
              void* stream);
/* frozen |= (x < tau) & (g > delta) (if frozen is non-NULL); active = ((x <= EPS) & (g > 0)) |
 * frozen; ACTIVE = |active|; with r non-NULL also starts the reduced CG at xcg = 0,
 * r = p = where(free, -g, 0): RZ, R0, TOL_K, STOP = (R0 <= max(TOL_K, eps)), NCG = HITNEG =
 * USE_RHS = 0 (ActiveSetState.update, solve_reduced_newton; jetnewton_mlx.py:334-352, 393-407).
 * x NULL: active is a given mask (read, not written; frozen must be NULL, r non-NULL) and
 * only the CG start is made. */
int dh_jn_active(const float* x, const float* g, int64_t n, float freeze_tau, float freeze_delta, uint8_t* frozen,
                 uint8_t* active, float* r, float* p, float* xcg, double eps, double* scalars, double* partials,
                 void* stream);
/* ap = where(free, scale kt + htu, 0) (htu may be NULL), DENOM = p . ap, NCG += 1; on DENOM not
 * finite or <= eps: STOP, HITNEG and, on the first step, USE_RHS (jetnewton_mlx.py:411-418). */
int dh_jn_cg_ap(const float* kt, float scale, const float* htu, const uint8_t* active, const float* p, float* ap,
                int64_t n, double eps, double* scalars, double* partials, void* stream);
/* Unless STOP: x += a p, r -= a ap (a = RZ / DENOM); STOP if ||r|| is not finite or <= TOL_K;
 * else p = r + (rz_new / RZ) p (jetnewton_mlx.py:419-430). */
int dh_jn_cg_update(float* x, float* r, float* p, const float* ap, int64_t n, double* scalars, double* partials,
                    void* stream);
/* d = d_free + where(active, -g, 0), ND = g . d (jetnewton_mlx.py:627-637). */
int dh_jn_direction(const float* xcg, const float* g, const uint8_t* active, float* d, int64_t n, double* scalars,
                    double* partials, void* stream);
/* x_t = max(x + t d, 0), DIRECTIONAL = g . (x_t - x) (projected_armijo, jetnewton_mlx.py:460-466). */
int dh_jn_trial(const float* x, const float* d, float t, const float* g, float* x_t, int64_t n, double* scalars,
                double* partials, void* stream);

/* ---- measurement ---- */
/* Pass ids for dh_timing_read. */
#define DH_PASS_X_FWD 0
#define DH_PASS_Y_FWD 1
#define DH_PASS_Z 2
#define DH_PASS_Y_INV 3
#define DH_PASS_X_INV 4
#define DH_PASS_COUNT 5
/* When enabled, dh_conv_apply records a HIP event pair around each pass on
 * the caller's stream.  dh_timing_read synchronises on those events and
 * returns per-pass accumulated milliseconds and launch counts
 * (DH_PASS_COUNT entries each); reset != 0 clears the accumulators.
 * dh_timing_select limits the bracketed passes to those whose bit (1 << pass id)
 * is set in mask (default: all), so that a timed run can bracket only the pass it
 * reports.  The events record without a system-scope fence (timing only). */
int dh_timing_enable(int on);
int dh_timing_select(int mask);
int dh_timing_read(double* ms, int64_t* counts, int reset);

/* ---- layout operators (core_operators.py:29-352) ---- */
/* Box copy between C-contiguous 3-D volumes: dst[d] = src[d - dst_start + src_start]
 * inside the box of the given extent, 0 elsewhere in dst.  Pad.forward / Crop.adjoint
 * (zero-pad) and Pad.adjoint / Crop.forward (slice) are both this call. */
int dh_box_copy(const float* src, const int64_t* src_shape, float* dst, const int64_t* dst_shape,
                const int64_t* src_start, const int64_t* dst_start, const int64_t* extent, void* stream);
/* FractionalAreaDownsample with integer factors f that divide the fine extents n (the
 * band tables are then unit block sums): dh_block_sum is its forward, coarse (n/f) =
 * sum over each f0 x f1 x f2 block, accumulated along z, then y, then x in the order the
 * three banded passes below use (bit-identical to them); dh_block_spread its adjoint,
 * fine[z, y, x] = coarse[z / f0, y / f1, x / f2].  One pass each instead of three. */
int dh_block_sum(const float* fine, float* coarse, int n0, int n1, int n2, int f0, int f1, int f2, void* stream);
int dh_block_spread(const float* coarse, float* fine, int n0, int n1, int n2, int f0, int f1, int f2, void* stream);
/* Banded resampling along one axis (core_operators.py:205-219): the volume is
 * viewed as (outer, n_in, inner) -> (outer, n_out, inner) with
 * out[o, i, q] = sum_k wt[i, k] * in[o, idx[i, k], q], k < width.  idx (int32)
 * and wt (float32) are device arrays of n_out * width entries. */
int dh_banded_apply(const float* src, float* dst, int64_t outer, int64_t n_in, int64_t inner, int64_t n_out,
                    const int32_t* idx, const float* wt, int width, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DECONHIP_H */
