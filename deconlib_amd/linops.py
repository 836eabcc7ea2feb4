"""FFT convolution operators on MI355X (drop-in for deconlib's linops_mlx).

``FFTConvolver`` (linops_mlx.py:384-432) and ``LinearFFTConvolver``
(linops_mlx.py:435-533) keep the reference's constructor signatures,
attributes (``signal_shape``, ``in_shape``, ``out_shape``, ``kernel_shape``,
``fft_shape``, ``operator_norm_sq``) and ``ValueError`` behaviour.  Every
application runs the pruned five-pass HIP pipeline of libdeconhip; inputs are
device ``torch.Tensor`` (numpy is accepted and uploaded), outputs are new
device tensors, matching the reference's "every operator returns a new array".
"""

from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np
import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32


def _next_smooth_number(n: int) -> int:
    """Smallest integer >= n whose prime factors are 2, 3, 5 (linops_mlx.py:956-972)."""
    c = int(n)
    while True:
        m = c
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return c
        c += 1


def fast_padded_shape(signal_shape, kernel_shape, min_pad=None) -> Tuple[int, ...]:
    """Smallest 5-smooth wrap-free canvas (linops_mlx.py:975-1032)."""
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
        out.append(_next_smooth_number(max(n + need, m)))
    return tuple(out)


SQRT2 = float(np.float32(np.sqrt(np.float32(2.0))))


class Hessian3D:
    """Six weighted second differences with Neumann boundaries (linops_mlx.py:182-262).

    Components [zz, yy, xx, yz, xz, xy] weighted [r^2, 1, 1, r sqrt2, r sqrt2, sqrt2];
    ``operator_norm_sq = 16 r^4 + 4 r^2 + 34``.  Forward and adjoint are the
    libdeconhip stencil kernels.
    """

    _kind = N.REG_HESSIAN

    def __init__(self, r: float = 1.0):
        self.r = r
        self.operator_norm_sq = 16.0 * (r ** 4) + 4.0 * (r ** 2) + 34.0
        self._w = np.array([r * r, 1.0, 1.0, r * SQRT2, r * SQRT2, SQRT2], dtype=np.float32)

    @classmethod
    def from_spacing(cls, spacing):
        dz, dy, dx = spacing
        if abs(dy - dx) > 1e-2 * max(dy, dx):
            raise ValueError(
                f"Hessian3D.from_spacing assumes isotropic lateral sampling "
                f"(dy == dx); got dy={dy}, dx={dx}. Construct Hessian3D(r=...) "
                f"directly for anisotropic lateral spacing."
            )
        return cls(r=dy / dz)

    def _wp(self):
        return self._w.ctypes.data_as(C.POINTER(C.c_float))

    def forward(self, f) -> torch.Tensor:
        f = as_f32(f)
        if f.dim() != 3:
            raise ValueError("Hessian3D expects a 3-D volume")
        out = empty_f32((6,) + tuple(f.shape))
        N.call("dh_hessian_forward", f.data_ptr(), out.data_ptr(), *f.shape, self._wp(), N.stream_ptr())
        return out

    def adjoint(self, H) -> torch.Tensor:
        H = as_f32(H)
        if H.dim() != 4 or H.shape[0] != 6:
            raise ValueError("Hessian3D.adjoint expects a (6, Z, Y, X) array")
        out = empty_f32(tuple(H.shape[1:]))
        N.call("dh_hessian_adjoint", H.data_ptr(), out.data_ptr(), *H.shape[1:], self._wp(), N.stream_ptr())
        return out

    def __call__(self, f):
        return self.forward(f)


class Hessian2D:
    """[H_yy, H_xx, sqrt2 H_xy] with Neumann boundaries (linops_mlx.py:153-179).

    Runs the Hessian3D stencil kernels on a single plane, where the z terms
    vanish (replicate padding of one sample), with weights [0, 1, 1, 0, 0, sqrt2].
    """

    operator_norm_sq = 48.0
    _kind = N.REG_HESSIAN

    def __init__(self):
        self._w = np.array([0.0, 1.0, 1.0, 0.0, 0.0, SQRT2], dtype=np.float32)

    def _wp(self):
        return self._w.ctypes.data_as(C.POINTER(C.c_float))

    def forward(self, f) -> torch.Tensor:
        f = as_f32(f)
        if f.dim() != 2:
            raise ValueError("Hessian2D expects a 2-D image")
        ny, nx = f.shape
        h6 = empty_f32((6, 1, ny, nx))
        N.call("dh_hessian_forward", f.data_ptr(), h6.data_ptr(), 1, ny, nx, self._wp(), N.stream_ptr())
        return h6[[1, 2, 5], 0].contiguous()

    def adjoint(self, H) -> torch.Tensor:
        H = as_f32(H)
        if H.dim() != 3 or H.shape[0] != 3:
            raise ValueError("Hessian2D.adjoint expects a (3, H, W) array")
        ny, nx = H.shape[1:]
        h6 = torch.zeros((6, 1, ny, nx), dtype=torch.float32, device=H.device)
        h6[[1, 2, 5], 0] = H
        out = empty_f32((ny, nx))
        N.call("dh_hessian_adjoint", h6.data_ptr(), out.data_ptr(), 1, ny, nx, self._wp(), N.stream_ptr())
        return out

    def __call__(self, f):
        return self.forward(f)


class _StencilOp:
    """A regularizer operator run by the library's 3-D stencil kernels on the input
    promoted to (nz, ny, nx) (``dh_reg_forward`` / ``dh_reg_adjoint``).

    ``_kind`` is the DH_REG_* operator family, ``_w`` its weights, ``_ndim`` the
    input rank and ``_comps`` the internal components (of 6 Hessian or 3 gradient
    planes) that form the reference's output, in order; unset components are zero.
    """

    _kind = N.REG_HESSIAN
    _ndim = 3
    _comps = (0, 1, 2, 3, 4, 5)
    _squeeze = False  # a single-component output without the component axis (Gradient1D, Hessian1D)

    def _wp(self):
        return self._w.ctypes.data_as(C.POINTER(C.c_float))

    def _dims(self, shape):
        return (1,) * (3 - len(shape)) + tuple(int(v) for v in shape)

    def forward(self, f) -> torch.Tensor:
        f = as_f32(f)
        if f.dim() != self._ndim:
            raise ValueError(f"{type(self).__name__} expects a {self._ndim}-D array")
        nc = N.fns["dh_reg_components"](self._kind)
        d = self._dims(f.shape)
        full = empty_f32((nc,) + d)
        N.call("dh_reg_forward", self._kind, f.data_ptr(), full.data_ptr(), *d, self._wp(), N.stream_ptr())
        if len(self._comps) == nc:
            out = full.reshape((nc,) + tuple(f.shape))
        else:
            out = full[list(self._comps)].reshape((len(self._comps),) + tuple(f.shape)).contiguous()
        return out[0] if self._squeeze else out

    def adjoint(self, y) -> torch.Tensor:
        y = as_f32(y)
        if self._squeeze:
            y = y.unsqueeze(0)
        if y.dim() != self._ndim + 1 or y.shape[0] != len(self._comps):
            raise ValueError(f"{type(self).__name__}.adjoint expects a ({len(self._comps)}, *spatial) array")
        shape = tuple(y.shape[1:])
        nc = N.fns["dh_reg_components"](self._kind)
        d = self._dims(shape)
        if len(self._comps) == nc:
            full = y.reshape((nc,) + d)
        else:
            full = torch.zeros((nc,) + d, dtype=torch.float32, device=y.device)
            full[list(self._comps)] = y.reshape((len(self._comps),) + d)
        out = empty_f32(shape)
        N.call("dh_reg_adjoint", self._kind, full.data_ptr(), out.data_ptr(), *d, self._wp(), N.stream_ptr())
        return out

    def __call__(self, f):
        return self.forward(f)


class Gradient1D(_StencilOp):
    """Forward difference with Neumann boundary, ``||D||^2 = 4`` (linops_mlx.py:46-66)."""

    operator_norm_sq = 4.0
    _kind, _ndim, _comps, _squeeze = N.REG_GRADIENT, 1, (2,), True

    def __init__(self):
        self._w = np.array([0.0, 0.0, 1.0], dtype=np.float32)


class Gradient2D(_StencilOp):
    """(d/dy, d/dx) forward differences, ``||grad||^2 = 8`` (linops_mlx.py:69-89)."""

    operator_norm_sq = 8.0
    _kind, _ndim, _comps = N.REG_GRADIENT, 2, (1, 2)

    def __init__(self):
        self._w = np.array([0.0, 1.0, 1.0], dtype=np.float32)


class Gradient3D(_StencilOp):
    """(r d/dz, d/dy, d/dx), ``||grad||^2 = 4 (r^2 + 2)`` (linops_mlx.py:92-127)."""

    _kind, _ndim, _comps = N.REG_GRADIENT, 3, (0, 1, 2)

    def __init__(self, r: float = 1.0):
        self.r = r
        self.operator_norm_sq = 4.0 * (r * r + 2.0)
        self._w = np.array([r, 1.0, 1.0], dtype=np.float32)


class Hessian1D(_StencilOp):
    """Second difference with Neumann boundary, ``||D2||^2 = 16`` (linops_mlx.py:130-150)."""

    operator_norm_sq = 16.0
    _kind, _ndim, _comps, _squeeze = N.REG_HESSIAN, 1, (2,), True

    def __init__(self):
        self._w = np.array([0.0, 0.0, 1.0, 0.0, 0.0, 0.0], dtype=np.float32)


def _f32(v) -> np.float32:
    return np.float32(v)


class AnisotropicHessian3D(Hessian3D):
    """Hessian with per-axis coefficients kappa = (kz, ky, kx) (linops_mlx.py:314-381).

    Components [zz, yy, xx, yz, xz, xy] weighted [kz^2, ky^2, kx^2, sqrt2 ky kz,
    sqrt2 kx kz, sqrt2 kx ky] on the Hessian3D stencil kernels;
    ``Hessian3D(r)`` is ``AnisotropicHessian3D((r, 1, 1))``.
    """

    def __init__(self, kappa):
        self.kappa = tuple(float(k) for k in kappa)
        kz, ky, kx = self.kappa
        self.r = kz
        self.operator_norm_sq = 16.0 * (kz ** 4 + ky ** 4 + kx ** 4) + 2.0 * (
            (ky ** 2) * (kz ** 2) + (kx ** 2) * (kz ** 2) + (kx ** 2) * (ky ** 2))
        s2 = _f32(SQRT2)
        self._w = np.array([_f32(kz ** 2), _f32(ky ** 2), _f32(kx ** 2), _f32(ky * kz) * s2, _f32(kx * kz) * s2,
                            _f32(kx * ky) * s2], dtype=np.float32)

    @classmethod
    def from_lengths(cls, ell, spacing):
        """kappa_a = ell_a / h_a (z, y, x order)."""
        return cls(kappa=tuple(e / h for e, h in zip(ell, spacing)))


class AnisotropicHessian2D(Hessian2D):
    """[ky^2 H_yy, kx^2 H_xx, sqrt2 ky kx H_xy] (linops_mlx.py:265-311), on one plane of
    the Hessian3D stencil kernels; ``Hessian2D()`` is ``AnisotropicHessian2D((1, 1))``."""

    def __init__(self, kappa):
        self.kappa = tuple(float(k) for k in kappa)
        ky, kx = self.kappa
        self.operator_norm_sq = 16.0 * (ky ** 4 + kx ** 4) + 2.0 * (ky ** 2) * (kx ** 2)
        self._w = np.array([0.0, _f32(ky ** 2), _f32(kx ** 2), 0.0, 0.0, _f32(ky * kx) * _f32(SQRT2)],
                           dtype=np.float32)

    @classmethod
    def from_lengths(cls, ell, spacing):
        """kappa_a = ell_a / h_a (y, x order)."""
        return cls(kappa=tuple(e / h for e, h in zip(ell, spacing)))


class _ConvPlan:
    """Owns one native plan and its device buffers (twiddles, OTF, workspace)."""

    def __init__(self, kernel: torch.Tensor, signal_shape, fft_shape, circular: bool, normalize: bool):
        nd = len(signal_shape)
        sig = (C.c_int64 * nd)(*signal_shape)
        ker = (C.c_int64 * nd)(*kernel.shape)
        fft = (C.c_int64 * nd)(*fft_shape) if fft_shape is not None else None
        h = C.c_void_p()
        N.call("dh_conv_plan_create", nd, sig, ker, fft, int(circular), C.byref(h))
        self._h = h
        info = N.PlanInfo()
        N.call("dh_conv_plan_info", h, C.byref(info))
        self.info = info
        dev = device()
        self.twiddles = torch.empty(int(info.twiddle_bytes), dtype=torch.uint8, device=dev)
        self.otf_buf = torch.empty(int(info.otf_bytes), dtype=torch.uint8, device=dev)
        self.work = torch.empty(int(info.work_bytes), dtype=torch.uint8, device=dev)
        N.call("dh_conv_plan_bind", h, self.twiddles.data_ptr(), N.stream_ptr())
        # max|OTF|^2 stays on the device until someone asks for it (no synchronisation here)
        self.norm_dev = torch.empty(1, dtype=torch.float64, device=dev)
        N.call("dh_otf_build_dev", h, kernel.data_ptr(), int(normalize), self.otf_buf.data_ptr(),
               self.work.data_ptr(), int(info.work_bytes), self.norm_dev.data_ptr(), N.stream_ptr())
        self._norm_sq = None
        self.canvas_shape = tuple(int(v) for v in info.canvas_shape)[3 - nd:]
        # bumped by everything that writes the workspace, so a caller that left a
        # spectrum there (dh_rl_iteration keep_spectrum) can tell whether it survives
        self.gen = 0

    @property
    def norm_sq(self) -> float:
        """max|OTF|^2 (read from the device on first use)."""
        if self._norm_sq is None:
            self._norm_sq = float(self.norm_dev.item())
        return self._norm_sq

    def apply(self, x: Optional[torch.Tensor], adjoint: bool, out: torch.Tensor,
              fuse: Optional[N.FuseArgs] = None) -> torch.Tensor:
        self.gen += 1
        N.call("dh_conv_apply", self._h, int(adjoint), N.ptr(x), out.data_ptr(),
               self.otf_buf.data_ptr(), self.work.data_ptr(), int(self.info.work_bytes),
               C.byref(fuse) if fuse is not None else None, N.stream_ptr())
        return out

    def otf(self) -> torch.Tensor:
        """Unscaled complex64 OTF (P0, P1, P2//2+1), as the reference stores it."""
        p = self.info.canvas_shape
        h = p[2] // 2 + 1
        flat = self.otf_buf.view(torch.complex64).reshape(p[0], p[1] * h)
        b = int(self.info.spectrum_block)
        if b <= 0 or b >= h:
            z = flat.reshape(p[0], p[1], h)
        else:
            # column blocks of width b (the last one narrower), each (P1, width) row-major
            nfull = (h - 1) // b
            full = flat[:, :nfull * b * p[1]].reshape(p[0], nfull, p[1], b).permute(0, 2, 1, 3)
            last = flat[:, nfull * b * p[1]:].reshape(p[0], p[1], h - nfull * b)
            z = torch.cat([full.reshape(p[0], p[1], nfull * b), last], dim=2)
        return z * float(p[0] * p[1] * p[2])

    def __del__(self):
        h = getattr(self, "_h", None)
        fns = getattr(N, "fns", None)  # None while the interpreter tears modules down
        if h is not None and h.value and fns is not None:
            fns["dh_conv_plan_destroy"](h)
            self._h = None


class FFTConvolver:
    """Circular convolution with a stored OTF (linops_mlx.py:384-432)."""

    def __init__(self, kernel, normalize: bool = True):
        k = as_f32(kernel)
        self.shape = tuple(int(s) for s in k.shape)
        self.in_shape = self.shape
        self.out_shape = self.in_shape
        self.axes = tuple(range(-len(self.shape), 0))
        self._plan = _ConvPlan(k, self.shape, self.shape, True, normalize)

    @property
    def operator_norm_sq(self) -> float:
        """max|OTF|^2 (linops_mlx.py:415), read from the device on first use."""
        return self._plan.norm_sq

    @property
    def otf(self) -> torch.Tensor:
        return self._plan.otf()

    @property
    def otf_conj(self) -> torch.Tensor:
        return torch.conj(self.otf)

    def forward(self, x) -> torch.Tensor:
        x = as_f32(x)
        return self._plan.apply(x, False, empty_f32(self.shape))

    def adjoint(self, y) -> torch.Tensor:
        y = as_f32(y)
        return self._plan.apply(y, True, empty_f32(self.shape))

    def adjoint_spread(self, y, factors) -> torch.Tensor:
        """adjoint(spread(y)) in one correlation (see LinearFFTConvolver.adjoint_spread)."""
        return _adjoint_spread(self._plan, self.shape, y, factors)

    def __call__(self, x) -> torch.Tensor:
        return self.forward(x)


def _adjoint_spread(plan: _ConvPlan, shape, y, factors) -> torch.Tensor:
    """Correlation of the integer-factor block spread of ``y`` (each coarse sample
    repeated f0 x f1 x f2 times, the adjoint of FractionalAreaDownsample's block sum,
    core_operators.py:205-219) with the plan's kernel: DH_PRO_SPREAD."""
    y = as_f32(y)
    f = tuple(int(v) for v in factors)
    full = (1,) * (3 - len(shape)) + tuple(int(v) for v in shape)
    got = (1,) * (3 - y.dim()) + tuple(int(v) for v in y.shape)
    if len(f) != 3 or y.dim() != len(shape) or any(n != g * q for n, g, q in zip(full, got, f)):
        raise ValueError(f"spread of input {tuple(y.shape)} by {f} does not give the signal shape {tuple(shape)}")
    fz = N.FuseArgs(prologue=N.PRO_SPREAD)
    fz.spread[:] = f
    return plan.apply(y, True, empty_f32(tuple(shape)), fz)


class LinearFFTConvolver:
    """Wrap-free linear convolution through a pruned padded FFT (linops_mlx.py:435-533).

    Domain and range are both ``signal_shape``; the ``N + M - 1`` padding
    happens inside (never materialised: the HIP passes skip the zero rows and
    crop on store).
    """

    def __init__(self, kernel, signal_shape: Tuple[int, ...], normalize: bool = True,
                 fft_shape: Optional[Tuple[int, ...]] = None):
        if isinstance(kernel, torch.Tensor):
            kshape = tuple(int(s) for s in kernel.shape)
        else:
            kshape = tuple(int(s) for s in np.shape(kernel))
        self.signal_shape = tuple(int(s) for s in signal_shape)
        self.in_shape = self.signal_shape
        self.out_shape = self.signal_shape
        self.kernel_shape = kshape
        if len(self.signal_shape) != len(self.kernel_shape):
            raise ValueError("signal_shape and kernel must have the same number of dimensions")
        minimum = tuple(n + m - 1 for n, m in zip(self.signal_shape, self.kernel_shape))
        if fft_shape is None:
            fft_shape = fast_padded_shape(self.signal_shape, self.kernel_shape)
        self.fft_shape = tuple(int(s) for s in fft_shape)
        if len(self.fft_shape) != len(self.signal_shape):
            raise ValueError("fft_shape must match signal_shape ndim")
        if any(p < q for p, q in zip(self.fft_shape, minimum)):
            raise ValueError(
                f"fft_shape {self.fft_shape} is too small for linear "
                f"convolution; minimum is {minimum}"
            )
        k = as_f32(kernel)
        self._plan = _ConvPlan(k, self.signal_shape, self.fft_shape, False, normalize)
        self.canvas_shape = self._plan.canvas_shape

    @property
    def operator_norm_sq(self) -> float:
        """max|OTF|^2 (linops_mlx.py:415-518), read from the device on first use."""
        return self._plan.norm_sq

    @property
    def otf(self) -> torch.Tensor:
        return self._plan.otf()

    def forward(self, x) -> torch.Tensor:
        """Zero-boundary linear convolution."""
        x = as_f32(x)
        self._check(x)
        return self._plan.apply(x, False, empty_f32(self.signal_shape))

    def adjoint(self, y) -> torch.Tensor:
        """Adjoint (zero-boundary correlation)."""
        y = as_f32(y)
        self._check(y)
        return self._plan.apply(y, True, empty_f32(self.signal_shape))

    def forward_block_sum(self, x) -> Optional[torch.Tensor]:
        """FractionalAreaDownsample(2)(forward(x)) of a 3-D signal with even extents in one
        application: the inverse X pass sums each 2x2x2 block (DH_EPI_BLOCKSUM2) and the
        fine result never exists.  None where the plan cannot (odd extents, a canvas
        without specialised kernels or with rows beyond 4096 samples)."""
        if not self.block_sum_fusable():
            return None
        x = as_f32(x)
        self._check(x)
        return self._plan.apply(x, False, empty_f32(tuple(v // 2 for v in self.signal_shape)),
                                N.FuseArgs(epilogue=N.EPI_BLOCKSUM2))

    def block_sum_fusable(self) -> bool:
        """Whether the 2x2x2 block-sum epilogue applies: a 3-D signal with even extents on a
        canvas with specialised kernels and rows of at most 4096 samples."""
        s = self.signal_shape
        return (len(s) == 3 and not any(v % 2 for v in s) and int(self._plan.info.spectrum_block) > 0
                and self.canvas_shape[2] <= 4096)

    def adjoint_spread(self, y, factors) -> torch.Tensor:
        """adjoint(spread(y)) for a coarse ``y`` whose integer-factor block spread
        (FractionalAreaDownsample's adjoint) is the signal: one correlation whose first
        pass reads ``y`` through the spread prologue (the fine volume never exists)."""
        return _adjoint_spread(self._plan, self.signal_shape, y, factors)

    def _check(self, x: torch.Tensor) -> None:
        if tuple(x.shape) != self.signal_shape:
            raise ValueError(f"expected input of shape {self.signal_shape}, got {tuple(x.shape)}")

    def __call__(self, x) -> torch.Tensor:
        return self.forward(x)
