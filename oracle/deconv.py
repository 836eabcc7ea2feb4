"""CPU oracle for the FFT-operator deconvolution hot path (TEST INFRASTRUCTURE).

This module is the checker, not the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``deconlib_amd``) never calls into it.

It restates, in NumPy/SciPy, the reference's algorithm for:

* ``fast_padded_shape`` / ``_next_smooth_number``   (deconlib/deconvolution/linops_mlx.py:956-1032)
* ``pad_corner_origin_kernel``                      (deconlib/utils/padding.py:142-190)
* ``Pad`` / ``Crop``                                (deconlib/deconvolution/core_operators.py:29-137)
* ``FractionalAreaDownsample`` and banded weights   (deconlib/deconvolution/core_operators.py:151-352)
* ``Compose`` / ``compose``                         (deconlib/deconvolution/composition.py:52-122)
* ``FFTConvolver`` / ``LinearFFTConvolver``         (deconlib/deconvolution/linops_mlx.py:384-533)
* ``d1_fwd`` / ``d1_fwd_adj``                      (deconlib/deconvolution/linops_core_mlx.py:25-39)
* ``d1_cen`` / ``d1_cen_adj`` / ``d2``              (deconlib/deconvolution/linops_core_mlx.py:42-80)
* ``Gradient1D/2D/3D``                              (deconlib/deconvolution/linops_mlx.py:46-127)
* ``Hessian1D/2D/3D``                               (deconlib/deconvolution/linops_mlx.py:130-262)
* dual proxes                                       (deconlib/deconvolution/pdhg_mlx.py:115-233)
* Identity/Gradient/Hessian regularizers            (deconlib/deconvolution/pdhg_mlx.py:241-456, 485-499)
* ``richardson_lucy_with_operator``                 (deconlib/deconvolution/rl_mlx.py:58-185)
* ``solve_pdhg_mlx`` / ``solve_pdhg_with_operator`` (deconlib/deconvolution/pdhg_mlx.py:459-1242)

Every function takes a ``dtype`` (float32 follows the reference's MLX dtype
flow: fp32 arrays, complex64 spectra; float64 is the FP64 oracle).

Parity pins (see tests/test_oracle_*.py):

* fast_padded_shape      -- the reference docstring's known answers.
* pad_corner_origin      -- brute-force index map on tiny odd/even kernels.
* LinearFFTConvolver     -- brute-force direct linear convolution/correlation in
                            float64; delta PSF => identity; adjoint identity.
* FractionalAreaDownsample -- integer scale == block sum (closed form); flux
                            preservation; adjoint identity incl. scale 1.325.
* Hessian3D              -- quadratic-phantom curvature, adjoint identity,
                            closed-form norm bound via power iteration.
* Gradient / Hessian 1-D/2-D, identity -- linear-ramp and quadratic known
                            answers, brute-force transpose (dense matrix),
                            closed-form norm bounds by power iteration.
* proxes                 -- known values and the Poisson optimality residual.
* RL                     -- fixed point, EM monotone I-divergence decrease,
                            hand-computed first iteration on a 1-D toy.
* PDHG                   -- first iteration in closed form from zero duals;
                            the Malitsky-Pock step rule through its invariants
                            with dx^2 and dK^2 recomputed by four separate
                            operator applications; the patience/min_iter stop
                            on hand-built rel-change sequences; the objective and
                            the I-divergence on hand-computed values; KKT
                            optimality of a converged tiny problem.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Tuple

import numpy as np

try:  # scipy.fft is threaded; numpy.fft is the fallback
    import scipy.fft as _sfft

    def _rfftn(a, s=None, axes=None):
        return _sfft.rfftn(a, s=s, axes=axes, workers=-1)

    def _irfftn(a, s=None, axes=None):
        return _sfft.irfftn(a, s=s, axes=axes, workers=-1)

except ImportError:  # pragma: no cover
    _rfftn = np.fft.rfftn
    _irfftn = np.fft.irfftn

SQRT2 = math.sqrt(2.0)


def _cplx(dtype):
    return np.complex64 if np.dtype(dtype) == np.float32 else np.complex128


# ---------------------------------------------------------------------------
# FFT sizing (linops_mlx.py:956-1032)
# ---------------------------------------------------------------------------


def next_smooth_number(n: int) -> int:
    """Smallest integer >= n with prime factors in {2,3,5} (linops_mlx.py:956-972)."""
    c = int(n)
    while True:
        m = c
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return c
        c += 1


def fast_padded_shape(signal_shape, kernel_shape, min_pad=None):
    """Per-axis next 5-smooth >= max(n + pad, m) (linops_mlx.py:975-1032)."""
    nd = len(signal_shape)
    if min_pad is None:
        pads = (None,) * nd
    elif isinstance(min_pad, int):
        pads = (min_pad,) * nd
    else:
        pads = tuple(min_pad)
    out = []
    for n, m, p in zip(signal_shape, kernel_shape, pads):
        need = (m - 1) if p is None else int(p)
        out.append(next_smooth_number(max(n + need, m)))
    return tuple(out)


def pad_corner_origin_kernel(kernel: np.ndarray, output_shape) -> np.ndarray:
    """Embed a corner-origin kernel in a larger canvas (utils/padding.py:142-190).

    Per axis the first ceil(m/2) samples stay at the low end and the remaining
    floor(m/2) samples move to the high end of the canvas.
    """
    kernel = np.asarray(kernel)
    out_shape = tuple(int(s) for s in output_shape)
    if len(out_shape) != kernel.ndim:
        raise ValueError("output_shape ndim mismatch")
    idx = []
    for m, p in zip(kernel.shape, out_shape):
        if p < m:
            raise ValueError("output size smaller than kernel size")
        npos = (m + 1) // 2
        dst = np.arange(m)
        dst[npos:] = p - (m - npos) + np.arange(m - npos)
        idx.append(dst)
    out = np.zeros(out_shape, dtype=kernel.dtype)
    out[np.ix_(*idx)] = kernel
    return out


# ---------------------------------------------------------------------------
# Pad / Crop (core_operators.py:29-137)
# ---------------------------------------------------------------------------


class Pad:
    def __init__(self, padding):
        self.padding = tuple((int(a), int(b)) for a, b in padding)
        self.operator_norm_sq = 1.0

    def forward(self, x):
        return np.pad(x, self.padding)

    def adjoint(self, y):
        sl = tuple(slice(a, s - b) for (a, b), s in zip(self.padding, y.shape))
        return y[sl]

    __call__ = forward


class Crop:
    def __init__(self, original_shape, target_shape, start=None):
        if len(original_shape) != len(target_shape):
            raise ValueError("ndim mismatch")
        self.original_shape = tuple(int(s) for s in original_shape)
        self.target_shape = tuple(int(s) for s in target_shape)
        self.in_shape, self.out_shape = self.original_shape, self.target_shape
        self.operator_norm_sq = 1.0
        if start is not None and len(start) != len(self.original_shape):
            raise ValueError("start ndim mismatch")
        self._slices, self._padding = (), ()
        for i, (o, t) in enumerate(zip(self.original_shape, self.target_shape)):
            if t > o:
                raise ValueError("target larger than original")
            s = (o - t) // 2 if start is None else int(start[i])
            if s < 0 or s + t > o:
                raise ValueError("crop window outside original")
            self._slices += (slice(s, s + t),)
            self._padding += ((s, o - s - t),)

    def forward(self, x):
        return x[self._slices]

    def adjoint(self, y):
        return np.pad(y, self._padding)

    __call__ = forward


# ---------------------------------------------------------------------------
# Fractional-area downsampling (core_operators.py:151-352)
# ---------------------------------------------------------------------------


def banded_overlap_weights_1d(n_large: int, n_small: int):
    """Interval-overlap weights as banded windows (core_operators.py:151-202)."""
    scale = n_large / n_small
    w_f = min(n_large, int(math.ceil(scale)) + 2)
    i = np.arange(n_small)
    start_f = np.clip(np.floor(i * scale).astype(np.int64), 0, n_large - w_f)
    idx_f = start_f[:, None] + np.arange(w_f)[None, :]
    i0, i1 = (i * scale)[:, None], ((i + 1) * scale)[:, None]
    j0 = idx_f.astype(np.float64)
    wt_f = np.maximum(0.0, np.minimum(i1, j0 + 1) - np.maximum(i0, j0))
    w_a = min(n_small, int(math.ceil(1.0 / scale)) + 3)
    j = np.arange(n_large)
    start_a = np.clip(np.floor(j / scale).astype(np.int64) - 1, 0, n_small - w_a)
    idx_a = start_a[:, None] + np.arange(w_a)[None, :]
    a0 = idx_a.astype(np.float64) * scale
    a1 = (idx_a.astype(np.float64) + 1) * scale
    b0 = j[:, None].astype(np.float64)
    wt_a = np.maximum(0.0, np.minimum(a1, b0 + 1) - np.maximum(a0, b0))
    return (idx_f.astype(np.int32), wt_f.astype(np.float32),
            idx_a.astype(np.int32), wt_a.astype(np.float32))


def apply_banded_1d(x, idx, wt, axis):
    """out[..., i, ...] = sum_k wt[i,k] x[..., idx[i,k], ...] (core_operators.py:205-219)."""
    xm = np.moveaxis(x, axis, -1)
    g = xm[..., idx]                       # (..., out, width)
    out = np.sum(g * wt.astype(x.dtype), axis=-1)
    return np.moveaxis(out, -1, axis)


class FractionalAreaDownsample:
    def __init__(self, scale, axes=None, in_shape=None):
        if isinstance(scale, (int, float)):
            scale = (float(scale),)
        self.scale = tuple(float(s) for s in scale)
        if any(s < 1.0 for s in self.scale):
            raise ValueError("FractionalAreaDownsample requires scale >= 1")
        self.axes = axes
        self._axis_map = {}
        self.in_shape = self.out_shape = None
        if in_shape is not None:
            ax = axes if axes is not None else tuple(range(len(in_shape)))
            sc = self.scale * len(ax) if len(self.scale) == 1 and len(ax) > 1 else self.scale
            for a, s in zip(ax, sc):
                nl = in_shape[a]
                self._axis_map[a] = (nl, max(1, int(round(nl / s))))
            self.in_shape = tuple(int(s) for s in in_shape)
            o = list(self.in_shape)
            for a, (_, ns) in self._axis_map.items():
                o[a % len(o)] = ns
            self.out_shape = tuple(o)
        if self._axis_map:
            n = 1.0
            for nl, ns in self._axis_map.values():
                n *= nl / ns
        else:
            n = math.prod(self.scale)
        self.operator_norm_sq = float(n)

    def _axes_scales(self, ndim):
        ax = self.axes if self.axes is not None else tuple(range(ndim))
        sc = self.scale * len(ax) if len(self.scale) == 1 and len(ax) > 1 else self.scale
        return ax, sc

    def forward(self, x):
        r = x
        for a, s in zip(*self._axes_scales(x.ndim)):
            if a in self._axis_map:
                nl, ns = self._axis_map[a]
            else:
                nl = r.shape[a]
                ns = max(1, int(round(nl / s)))
            if nl == ns:
                continue
            i_f, w_f, _, _ = banded_overlap_weights_1d(nl, ns)
            r = apply_banded_1d(r, i_f, w_f, a)
        return r

    def adjoint(self, y):
        r = y
        for a, s in zip(*self._axes_scales(y.ndim)):
            if a in self._axis_map:
                nl, ns = self._axis_map[a]
            else:
                ns = r.shape[a]
                nl = max(1, int(round(ns * s)))
            if nl == ns:
                continue
            _, _, i_a, w_a = banded_overlap_weights_1d(nl, ns)
            r = apply_banded_1d(r, i_a, w_a, a)
        return r

    __call__ = forward


# ---------------------------------------------------------------------------
# Composition (composition.py:52-122)
# ---------------------------------------------------------------------------


class Compose:
    def __init__(self, outer, inner):
        oi = getattr(outer, "in_shape", None)
        io = getattr(inner, "out_shape", None)
        if oi is not None and io is not None and tuple(oi) != tuple(io):
            raise ValueError(f"Cannot compose: {tuple(oi)} vs {tuple(io)}")
        self.outer, self.inner = outer, inner
        self.in_shape = getattr(inner, "in_shape", None)
        self.out_shape = getattr(outer, "out_shape", None)
        self.operator_norm_sq = float(outer.operator_norm_sq * inner.operator_norm_sq)

    def forward(self, x):
        return self.outer.forward(self.inner.forward(x))

    def adjoint(self, y):
        return self.inner.adjoint(self.outer.adjoint(y))

    __call__ = forward


def compose(*ops):
    if not ops:
        raise ValueError("compose() requires at least one operator")
    r = ops[-1]
    for op in reversed(ops[:-1]):
        r = Compose(op, r)
    return r


# ---------------------------------------------------------------------------
# FFT convolution (linops_mlx.py:384-533)
# ---------------------------------------------------------------------------


class FFTConvolver:
    """Circular convolution with a stored OTF (linops_mlx.py:384-432)."""

    def __init__(self, kernel, normalize=True, dtype=np.float32):
        k = np.asarray(kernel).astype(dtype)
        self.dtype = np.dtype(dtype)
        self.shape = k.shape
        self.in_shape = self.out_shape = tuple(int(s) for s in k.shape)
        if normalize:
            k = k / k.sum(dtype=dtype)
        self.otf = _rfftn(k).astype(_cplx(dtype))
        self.operator_norm_sq = float(np.max(np.abs(self.otf) ** 2))

    def forward(self, x):
        X = _rfftn(x.astype(self.dtype, copy=False))
        return _irfftn(X * self.otf, s=self.shape).astype(self.dtype, copy=False)

    def adjoint(self, y):
        Y = _rfftn(y.astype(self.dtype, copy=False))
        return _irfftn(Y * np.conj(self.otf), s=self.shape).astype(self.dtype, copy=False)

    __call__ = forward


class LinearFFTConvolver:
    """Wrap-free linear convolution through a padded circular FFT (linops_mlx.py:435-533)."""

    def __init__(self, kernel, signal_shape, normalize=True, fft_shape=None,
                 dtype=np.float32):
        k = np.asarray(kernel)
        self.signal_shape = tuple(int(s) for s in signal_shape)
        self.in_shape = self.out_shape = self.signal_shape
        self.kernel_shape = tuple(int(s) for s in k.shape)
        if len(self.signal_shape) != len(self.kernel_shape):
            raise ValueError("signal_shape and kernel must have the same number of dimensions")
        mn = tuple(n + m - 1 for n, m in zip(self.signal_shape, self.kernel_shape))
        if fft_shape is None:
            fft_shape = fast_padded_shape(self.signal_shape, self.kernel_shape)
        self.fft_shape = tuple(int(s) for s in fft_shape)
        if len(self.fft_shape) != len(self.signal_shape):
            raise ValueError("fft_shape must match signal_shape ndim")
        if any(p < q for p, q in zip(self.fft_shape, mn)):
            raise ValueError(f"fft_shape {self.fft_shape} too small; minimum is {mn}")
        # the reference casts the compact kernel to float32 before embedding
        padded = pad_corner_origin_kernel(k.astype(np.float32), self.fft_shape)
        self._domain = Pad(tuple((0, p - n) for n, p in zip(self.signal_shape, self.fft_shape)))
        self._conv = FFTConvolver(padded, normalize=normalize, dtype=dtype)
        self.otf = self._conv.otf
        self.operator_norm_sq = self._conv.operator_norm_sq

    def forward(self, x):
        return self._domain.adjoint(self._conv.forward(self._domain.forward(x)))

    def adjoint(self, y):
        return self._domain.adjoint(self._conv.adjoint(self._domain.forward(y)))

    __call__ = forward


# ---------------------------------------------------------------------------
# Finite differences and Hessian3D (linops_core_mlx.py:42-80, linops_mlx.py:182-262)
# ---------------------------------------------------------------------------


def d1_cen(f, axis):
    f = np.swapaxes(f, axis, 0)
    fp = np.concatenate([f[:1], f, f[-1:]], axis=0)
    return np.swapaxes((fp[2:] - fp[:-2]) / f.dtype.type(2.0), axis, 0)


def d1_cen_adj(g, axis):
    g = np.swapaxes(g, axis, 0)
    fp = np.concatenate([-g[:1], g, -g[-1:]], axis=0)
    return np.swapaxes((fp[:-2] - fp[2:]) / g.dtype.type(2.0), axis, 0)


def d2(f, axis):
    f = np.swapaxes(f, axis, 0)
    fp = np.concatenate([f[:1], f, f[-1:]], axis=0)
    return np.swapaxes(fp[:-2] - f.dtype.type(2.0) * fp[1:-1] + fp[2:], axis, 0)


def d1_fwd(f, axis):
    """Forward difference, Neumann BC: f[i+1] - f[i], last sample 0 (linops_core_mlx.py:25-32)."""
    f = np.swapaxes(f, axis, 0)
    return np.swapaxes(np.concatenate([f[1:], f[-1:]], axis=0) - f, axis, 0)


def d1_fwd_adj(g, axis):
    """Transpose of d1_fwd: [-g0, g[i-1] - g[i] (1 <= i <= n-2), g[n-2]] (linops_core_mlx.py:35-39).

    For n == 1 this is [-g0], as the reference's concatenation gives.
    """
    g = np.swapaxes(g, axis, 0)
    return np.swapaxes(np.concatenate([-g[:1], g[:-2] - g[1:-1], g[-2:-1]], axis=0), axis, 0)


class Gradient1D:
    """d1_fwd along the only axis; ||D||^2 = 4 (linops_mlx.py:46-66)."""

    operator_norm_sq = 4.0

    def forward(self, f):
        return d1_fwd(f, 0)

    def adjoint(self, g):
        return d1_fwd_adj(g, 0)


class Gradient2D:
    """(d/dy, d/dx) forward differences; ||grad||^2 = 8 (linops_mlx.py:69-89)."""

    operator_norm_sq = 8.0

    def forward(self, f):
        return np.stack([d1_fwd(f, 0), d1_fwd(f, 1)])

    def adjoint(self, g):
        return d1_fwd_adj(g[0], 0) + d1_fwd_adj(g[1], 1)


class Gradient3D:
    """(r d/dz, d/dy, d/dx); ||grad||^2 = 4 (r^2 + 2) (linops_mlx.py:92-127)."""

    def __init__(self, r: float = 1.0):
        self.r = r
        self.operator_norm_sq = 4.0 * (r * r + 2.0)

    def forward(self, f):
        return np.stack([self.r * d1_fwd(f, 0), d1_fwd(f, 1), d1_fwd(f, 2)])

    def adjoint(self, g):
        return d1_fwd_adj(self.r * g[0], 0) + d1_fwd_adj(g[1], 1) + d1_fwd_adj(g[2], 2)


class Hessian1D:
    """d2 along the only axis, self-adjoint; ||D2||^2 = 16 (linops_mlx.py:130-150)."""

    operator_norm_sq = 16.0

    def forward(self, f):
        return d2(f, 0)

    def adjoint(self, g):
        return d2(g, 0)


class Hessian2D:
    """[H_yy, H_xx, sqrt2 H_xy]; ||H||^2 bound 48 (linops_mlx.py:153-179)."""

    operator_norm_sq = 48.0

    def forward(self, f):
        s2 = f.dtype.type(SQRT2)
        return np.stack([d2(f, 0), d2(f, 1), s2 * d1_cen(d1_cen(f, 0), 1)])

    def adjoint(self, H):
        s2 = H.dtype.type(SQRT2)
        return d2(H[0], 0) + d2(H[1], 1) + s2 * d1_cen_adj(d1_cen_adj(H[2], 0), 1)


class Hessian3D:
    """Six weighted second differences, Neumann BC (linops_mlx.py:182-262)."""

    def __init__(self, r: float = 1.0):
        self.r = r
        self.operator_norm_sq = 16.0 * r ** 4 + 4.0 * r ** 2 + 34.0

    def weights(self, dtype):
        r = self.r
        return np.array([r * r, 1.0, 1.0, r * SQRT2, r * SQRT2, SQRT2], dtype=dtype)

    def forward(self, f):
        w = self.weights(f.dtype)
        Dz, Dy = d1_cen(f, 0), d1_cen(f, 1)
        H = np.stack([d2(f, 0), d2(f, 1), d2(f, 2), d1_cen(Dz, 1), d1_cen(Dz, 2), d1_cen(Dy, 2)])
        return H * w.reshape(6, 1, 1, 1)

    def adjoint(self, H):
        Hw = H * self.weights(H.dtype).reshape(6, 1, 1, 1)
        return (d2(Hw[0], 0) + d2(Hw[1], 1) + d2(Hw[2], 2)
                + d1_cen_adj(d1_cen_adj(Hw[3], 1), 0)
                + d1_cen_adj(d1_cen_adj(Hw[4], 2), 0)
                + d1_cen_adj(d1_cen_adj(Hw[5], 2), 1))

    __call__ = forward


# ---------------------------------------------------------------------------
# Proximal operators (pdhg_mlx.py:115-233)
# ---------------------------------------------------------------------------


def prox_poisson_dual(y, sigma, data, background):
    """Stable root of z^2 - c z - sigma D = 0, returns 1 - z (pdhg_mlx.py:115-166)."""
    t = y.dtype.type
    c = t(1.0) - y - t(sigma * background)
    ds = np.maximum(data, t(0.0))
    disc = c * c + t(4.0 * sigma) * ds
    den = np.sqrt(disc) - c + t(1e-12)
    z = (t(2.0 * sigma) * ds) / den
    z = np.where(ds < t(1e-6), np.maximum(c, t(1e-6)), z)
    z = np.maximum(z, t(1e-6))
    return t(1.0) - z


def prox_gaussian_dual(y, sigma, data, background):
    return (y - sigma * (data - background)) / (1.0 + sigma)


def prox_nonneg(x):
    return np.maximum(x, x.dtype.type(0.0))


def prox_l1_dual(y, bound):
    return np.clip(y, -bound, bound).astype(y.dtype, copy=False)


def prox_l1_2_dual(y, bound):
    n = np.sqrt(np.sum(y * y, axis=0, keepdims=True) + 1e-12)
    return (y / np.maximum(n / bound, 1.0)).astype(y.dtype, copy=False)


class IdentityRegularizer:
    """L = I with a leading unit component axis; ||L||^2 = 1 (pdhg_mlx.py:241-306)."""

    output_components = 1
    operator_norm_sq = 1.0

    def __init__(self, norm="L1"):
        self.norm = norm

    def forward(self, x):
        return x[None]

    def adjoint(self, y):
        return y[0]


class GradientRegularizer:
    """Gradient1D/2D/3D with a component axis (1-D wrapped to (1, N)) (pdhg_mlx.py:309-377)."""

    def __init__(self, ndim, r=1.0, norm="L1"):
        if ndim not in (1, 2, 3):
            raise ValueError(f"ndim must be 1, 2, or 3, got {ndim}")
        self.ndim, self.r, self.norm = ndim, r, norm
        self._op = Gradient1D() if ndim == 1 else (Gradient2D() if ndim == 2 else Gradient3D(r))
        self.output_components = ndim
        self.operator_norm_sq = self._op.operator_norm_sq

    def forward(self, x):
        return self._op.forward(x)[None] if self.ndim == 1 else self._op.forward(x)

    def adjoint(self, y):
        return self._op.adjoint(y[0]) if self.ndim == 1 else self._op.adjoint(y)


class HessianRegularizer:
    """Hessian1D/2D/3D with a component axis (1, 3 or 6 components) (pdhg_mlx.py:380-456)."""

    def __init__(self, ndim, r=1.0, norm="L1"):
        if ndim not in (1, 2, 3):
            raise ValueError(f"ndim must be 1, 2, or 3, got {ndim}")
        self.ndim, self.r, self.norm = ndim, r, norm
        self._op = Hessian1D() if ndim == 1 else (Hessian2D() if ndim == 2 else Hessian3D(r))
        self.output_components = 1 if ndim == 1 else (3 if ndim == 2 else 6)
        self.operator_norm_sq = self._op.operator_norm_sq

    def forward(self, x):
        return self._op.forward(x)[None] if self.ndim == 1 else self._op.forward(x)

    def adjoint(self, y):
        return self._op.adjoint(y[0]) if self.ndim == 1 else self._op.adjoint(y)


def create_regularizer(regularization, ndim, r, norm):
    """(pdhg_mlx.py:485-499)"""
    if regularization == "identity":
        return IdentityRegularizer(norm)
    if regularization == "gradient":
        return GradientRegularizer(ndim, r, norm)
    if regularization == "hessian":
        return HessianRegularizer(ndim, r, norm)
    raise ValueError(f"Unknown regularization: {regularization}")


# ---------------------------------------------------------------------------
# Richardson-Lucy (rl_mlx.py:35-185)
# ---------------------------------------------------------------------------


@dataclass
class RLResult:
    restored: np.ndarray
    pred: np.ndarray
    iterations: int
    loss_history: list
    background: float = 0.0
    full_shape: Optional[Tuple[int, ...]] = None
    valid_slices: Optional[Tuple[slice, ...]] = None


def poisson_i_divergence(observed, model, eps=1e-6):
    """Mean I-divergence (rl_mlx.py:58-80)."""
    t = observed.dtype.type
    os_ = np.maximum(observed, t(eps))
    ms = np.maximum(model, t(eps))
    return float(np.mean(observed.astype(np.float64) * np.log(os_ / ms) - (observed - ms)))


def richardson_lucy(observed, blur_op, num_iter=50, background=0.0, init=None,
                    callback=None, eval_interval=10, dtype=np.float32):
    """RL multiplicative update (rl_mlx.py:117-185)."""
    t = np.dtype(dtype).type
    obs = np.asarray(observed).astype(dtype)
    eps = t(1e-10)
    dmb = np.maximum(obs - t(background), t(1e-6))
    if init is None:
        x = np.maximum(blur_op.adjoint(dmb), t(1e-6))
    else:
        x = np.maximum(np.asarray(init).astype(dtype), t(1e-6))
    s_raw = blur_op.adjoint(np.ones_like(obs))
    floor = max(float(np.max(s_raw)) * 1e-6, 1e-10)
    active = s_raw > t(floor)
    sens = np.where(active, s_raw, t(1.0))
    x = np.where(active, x, t(0.0))
    losses = []
    bg = t(background)
    k = -1
    for k in range(num_iter):
        model = blur_op.forward(x) + bg
        ratio = obs / np.maximum(model, eps)
        corr = np.where(active, blur_op.adjoint(ratio) / sens, t(0.0))
        x = np.maximum(x * corr, eps)
        if k % eval_interval == 0 or k == num_iter - 1:
            losses.append(poisson_i_divergence(obs, blur_op.forward(x) + bg))
        if callback is not None and callback(k, x):
            break
    if k < 0:
        raise NameError("num_iter must be >= 1")
    pred = blur_op.forward(x) + bg
    return RLResult(restored=x, pred=pred, iterations=k + 1, loss_history=losses,
                    background=float(background), full_shape=tuple(x.shape), valid_slices=None)


# ---------------------------------------------------------------------------
# PDHG (pdhg_mlx.py:459-1242), Hessian/L1/Poisson core
# ---------------------------------------------------------------------------


@dataclass
class PDHGResult:
    restored: np.ndarray
    iterations: int
    loss_history: List[float] = field(default_factory=list)
    converged: bool = False
    tau_history: List[float] = field(default_factory=list)
    sigma_history: List[float] = field(default_factory=list)
    metadata: dict = field(default_factory=dict)


def compute_spacing_ratio(spacing, ndim):
    """(pdhg_mlx.py:459-482)"""
    if spacing is None or ndim in (1, 2):
        return 1.0
    if len(spacing) != ndim:
        raise ValueError("spacing length mismatch")
    dz = spacing[0]
    lat = (spacing[1] + spacing[2]) / 2.0
    return lat / dz if dz > 0 else 1.0


def pdhg_objective(data, fm, Lx, alpha, norm="L1", loss_type="poisson"):
    """Loss recorded by both PDHG entry points (pdhg_mlx.py:860-881, 1171-1190).

    ``fm`` is the forward model A x + b.  Poisson:
    sum D log((D + 1e-6) / max(fm, 1e-6)) - (D - fm); Gaussian: 0.5 sum (D - fm)^2;
    plus alpha sum |Lx| ("L1") or alpha sum_i sqrt(sum_q (Lx)_q,i^2 + 1e-12) ("L1_2").
    Sums in float64.
    """
    t = data.dtype.type
    if loss_type == "poisson":
        fms = np.maximum(fm, t(1e-6))
        dl = np.sum((data * np.log((data + t(1e-6)) / fms) - (data - fm)).astype(np.float64))
    else:
        dl = 0.5 * np.sum(((data - fm) ** 2).astype(np.float64))
    if norm == "L1":
        rl = alpha * np.sum(np.abs(Lx).astype(np.float64))
    else:
        rl = alpha * np.sum(np.sqrt(np.sum(Lx * Lx, axis=0) + t(1e-12)).astype(np.float64))
    return float(dl + rl)


def malitsky_pock_step(tau, sigma, theta, dx2, dK2, step, delta=0.99, eta=0.5):
    """Adaptive step sizes of one iteration (pdhg_mlx.py:761-797), in float64.

    Returns (tau_new, sigma_new, theta_new, backtracked).
    """
    ratio = dx2 / (2.0 * sigma * max(dK2, 1e-12))
    tc = min(math.sqrt(1.0 + theta) * tau, delta * math.sqrt(ratio))
    back = tc * sigma * dK2 > delta * dx2
    tn = eta * tc if back else tc
    tn = max(tn, step * 1e-4)
    thn = tn / tau
    sn = sigma / max(thn, 0.1)
    sn = min(max(sn, step * 1e-4), step * 1e4)
    return tn, sn, thn, back


def convergence_step(k, rel, count, min_iter, tol, patience):
    """Patience/min_iter stopping test (pdhg_mlx.py:805-835).

    Returns (count, stop): from iteration ``min_iter`` on, ``count`` counts
    consecutive iterations with rel < tol and resets otherwise; stop once it
    reaches ``patience``.  Before ``min_iter`` the count is left unchanged.
    """
    if k >= min_iter:
        if rel < tol:
            count += 1
            if count >= patience:
                return count, True
        else:
            count = 0
    return count, False


def pdhg_loop(observed, blur_op, reg_op, x, alpha, num_iter, background, norm="L1",
              loss_type="poisson", delta=0.99, eta=0.5, eval_interval=10, tol=1e-5,
              min_iter=20, patience=5, dtype=np.float32, loss_every_iter=False,
              loss_on_interval=True, callback=None, callback_with_loss=False, trace=None):
    """The Malitsky-Pock loop shared by solve_pdhg_mlx / solve_pdhg_with_operator.

    Scalars (tau, sigma, theta) are Python floats as in the reference
    (pdhg_mlx.py:851-853); array arithmetic is in ``dtype``.  ``trace`` (a list)
    receives one dict per iteration: x, x_new, y1_new, y2_new, tau, sigma, theta
    (before the update), tau_new, sigma_new, theta_new, backtracked, dx2, dK2, rel.
    """
    t = np.dtype(dtype).type
    data = np.asarray(observed).astype(dtype)
    y1 = np.zeros(data.shape, dtype)
    y2 = np.zeros((reg_op.output_components,) + x.shape, dtype)
    xbar = x
    step = 0.99 / (blur_op.operator_norm_sq + reg_op.operator_norm_sq) ** 0.5
    tau = sigma = step
    theta = 1.0
    losses, taus, sigmas = [], [], []
    converged, count, final_iter = False, 0, num_iter
    bg = float(background)
    # A x and L x of the current iterate: the previous iteration's A x' and L x' (the same
    # arrays the reference recomputes, so the values are identical; this only saves time)
    Ax_cur = Lx_cur = None
    for k in range(num_iter):
        Axb = blur_op.forward(xbar)
        if loss_type == "poisson":
            y1n = prox_poisson_dual(y1 + t(sigma) * Axb, sigma, data, bg)
        else:
            y1n = prox_gaussian_dual(y1 + t(sigma) * Axb, sigma, data, bg).astype(dtype)
        y2a = y2 + t(sigma) * reg_op.forward(xbar)
        y2n = prox_l1_dual(y2a, alpha) if norm == "L1" else prox_l1_2_dual(y2a, alpha)
        grad = blur_op.adjoint(y1n) + reg_op.adjoint(y2n)
        xn = prox_nonneg(x - t(tau) * grad)
        dx = xn - x
        dx2 = float(np.sum(dx.astype(np.float64) ** 2))
        Axn, Lxn = blur_op.forward(xn), reg_op.forward(xn)
        dA = Axn - (blur_op.forward(x) if Ax_cur is None else Ax_cur)
        dL = Lxn - (reg_op.forward(x) if Lx_cur is None else Lx_cur)
        dK2 = float(np.sum(dA.astype(np.float64) ** 2) + np.sum(dL.astype(np.float64) ** 2))
        tn, sn, thn, back = malitsky_pock_step(tau, sigma, theta, dx2, dK2, step, delta, eta)
        xx = float(np.sum(x.astype(np.float64) ** 2))
        xnxn = float(np.sum(xn.astype(np.float64) ** 2))
        rel = math.sqrt(dx2 / (max(xx, xnxn) + 1e-12))
        if trace is not None:
            trace.append(dict(x=x, x_new=xn, y1_new=y1n, y2_new=y2n, tau=tau, sigma=sigma, theta=theta,
                              tau_new=tn, sigma_new=sn, theta_new=thn, backtracked=back, dx2=dx2, dK2=dK2,
                              rel=rel))
        count, stop = convergence_step(k, rel, count, min_iter, tol, patience)
        if stop:
            converged = True
            final_iter = k + 1
            x, y1, y2 = xn, y1n, y2n
            taus.append(tn)
            sigmas.append(sn)
            break
        xbar = xn + t(thn) * (xn - x)
        x, y1, y2 = xn, y1n, y2n
        Ax_cur, Lx_cur = Axn, Lxn
        tau, sigma, theta = tn, sn, thn
        taus.append(tau)
        sigmas.append(sigma)
        loss = None
        if loss_every_iter or (loss_on_interval and (k + 1) % eval_interval == 0):
            loss = pdhg_objective(data, Ax_cur + t(bg), Lx_cur, alpha, norm, loss_type)
            losses.append(loss)
        if callback is not None:
            if callback_with_loss:
                callback(k, x, loss)
            elif callback(k, x):
                final_iter = k + 1
                break
    return PDHGResult(restored=x, iterations=final_iter, loss_history=losses,
                      converged=converged, tau_history=taus, sigma_history=sigmas,
                      metadata=dict(y1=y1, y2=y2))


def solve_pdhg(observed, psf, alpha=0.01, num_iter=200, background=0.0, spacing=None,
               bin_factors=None, sampling_factors=None, init=None, norm="L1", loss_type="poisson",
               regularization="hessian", delta=0.99, eta=0.5, eval_interval=10, tol=1e-5, min_iter=20,
               patience=5, callback=None, dtype=np.float32, trace=None):
    """``solve_pdhg_mlx`` (pdhg_mlx.py:507-920)."""
    obs = np.asarray(observed).astype(dtype)
    psf = np.asarray(psf)
    ndim = obs.ndim
    if bin_factors is not None and sampling_factors is not None:
        raise ValueError("bin_factors and sampling_factors are mutually exclusive")
    if sampling_factors is not None:
        f = ((float(sampling_factors),) * ndim if isinstance(sampling_factors, (int, float))
             else tuple(float(v) for v in sampling_factors))
        hi = tuple(max(1, int(round(n * v))) for n, v in zip(obs.shape, f))
        if tuple(psf.shape) != hi:
            raise ValueError("psf.shape does not match the sampling_factors high-resolution shape")
        A = compose(FractionalAreaDownsample(scale=f), LinearFFTConvolver(psf, hi, normalize=True, dtype=dtype))
    elif bin_factors is not None:
        f = (bin_factors,) * ndim if isinstance(bin_factors, int) else tuple(int(b) for b in bin_factors)
        hi = tuple(psf.shape)
        for i, (n, b) in enumerate(zip(hi, f)):
            if n % b:
                raise ValueError(f"PSF dim {i} size {n} is not divisible by bin factor {b}")
        lo = tuple(n // b for n, b in zip(hi, f))
        if tuple(obs.shape) != lo:
            raise ValueError("observed shape does not match psf.shape / bin_factors")
        A = compose(FractionalAreaDownsample(scale=tuple(float(b) for b in f)),
                    LinearFFTConvolver(psf, hi, normalize=True, dtype=dtype))
    else:
        A = LinearFFTConvolver(psf, obs.shape, normalize=True, dtype=dtype)
    r = compute_spacing_ratio(spacing, ndim)
    L = create_regularizer(regularization, ndim, r, norm)
    if init is not None:
        x = np.asarray(init).astype(dtype)
    else:
        x = A.adjoint(obs) if (bin_factors is not None or sampling_factors is not None) else obs.copy()
        x = np.maximum(x, np.dtype(dtype).type(0.0))
    return pdhg_loop(obs, A, L, x, alpha, num_iter, background, norm=norm,
                     loss_type=loss_type, delta=delta, eta=eta, eval_interval=eval_interval,
                     tol=tol, min_iter=min_iter, patience=patience, dtype=dtype, callback=callback,
                     trace=trace)


def solve_pdhg_with_operator(observed, blur_op, alpha=0.01, regularization="identity", norm="L1",
                             loss_type="poisson", num_iter=200, background=0.0, init=None, verbose=False,
                             callback=None, delta=0.99, eta=0.5, eval_interval=10, tol=1e-5, min_iter=20,
                             patience=5, dtype=np.float32, trace=None):
    """``solve_pdhg_with_operator`` (pdhg_mlx.py:928-1242).

    Solution shape (pdhg_mlx.py:1011-1018): (n,) for an operator with a 2-D
    ``shape`` (m, n); else the operator's ``in_shape`` when it declares one (the
    reference assumes ``observed.shape``, equal for the operators it documents);
    else ``observed.shape``.  r = 1; x0 = max(init or 0.01, 0); the loss is
    recorded (and ``callback(k, x, loss)`` called, its return value ignored) every
    iteration when ``verbose`` or ``callback`` is given, never otherwise.
    """
    obs = np.asarray(observed).astype(dtype)
    shp = getattr(blur_op, "shape", None)
    if shp is not None and len(shp) == 2:
        shape, ndim = (int(shp[1]),), 1
    else:
        shape = tuple(getattr(blur_op, "in_shape", None) or obs.shape)
        ndim = len(shape)
    L = create_regularizer(regularization, ndim, 1.0, norm)
    if init is not None:
        x = np.asarray(init).astype(dtype)
    else:
        x = np.full(shape, 0.01, dtype)
    x = np.maximum(x, np.dtype(dtype).type(0.0))
    every = bool(verbose or callback is not None)
    return pdhg_loop(obs, blur_op, L, x, alpha, num_iter, background, norm=norm, loss_type=loss_type,
                     delta=delta, eta=eta, eval_interval=eval_interval, tol=tol, min_iter=min_iter,
                     patience=patience, dtype=dtype, loss_every_iter=every, loss_on_interval=False,
                     callback=callback, callback_with_loss=True, trace=trace)
