"""Malitsky–Pock adaptive PDHG on MI355X (drop-in for pdhg_mlx.solve_pdhg_mlx).

Per iteration (pdhg_mlx.py:717-900) the HIP kernels run, all enqueued on one
stream with the step sizes kept in a float64 block on the device:

1. ``dh_pd_y1``      y1 <- prox(y1 + sigma A xbar), with A xbar = (1+theta) A x_k - theta A x_{k-1}
2. ``dh_pd_dual``    y2 <- clip(y2 + sigma H(xbar), -alpha, alpha), xbar formed in the stencil
3. adjoint          g = A^T y1 (pruned FFT correlation)
4. ``dh_pd_primal``  x_new = max(0, x - tau (g + H^T y2)) and the sums of dx^2, x^2, x_new^2
5. forward          A x_new (pruned FFT convolution); A x_{k+1} is reused next iteration
6. ``dh_pd_data_sums`` / ``dh_pd_hess_sums``  |A dx|^2, |H x_new - H x|^2 (+ loss terms)
7. ``dh_pd_scalars`` FP64 tau / sigma / theta update and the convergence test

so one iteration costs two FFT convolutions instead of the reference's four
A-applications.  The host reads the device scalars only where the reference
can stop (after ``min_iter``) and at the end.
"""

from __future__ import annotations

import math
from typing import Callable, List, Optional, Tuple, Union

import numpy as np
import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32
from .base import MLXDeconvolutionResult
from .composition import compose
from .core_operators import FractionalAreaDownsample
from .linops import (FFTConvolver, Gradient1D, Gradient2D, Gradient3D, Hessian1D, Hessian2D, Hessian3D,
                     LinearFFTConvolver)

__all__ = [
    "solve_pdhg_mlx",
    "solve_pdhg_with_operator",
    "IdentityRegularizer",
    "GradientRegularizer",
    "HessianRegularizer",
]

_TAU, _SIGMA, _THETA, _STEP, _COUNT, _DONE, _FINAL = 0, 1, 2, 3, 4, 5, 6
_DA2 = 10  # DH_PD_DA2


class IdentityRegularizer:
    """L = I for sparsity on x itself (pdhg_mlx.py:241-306); one component, ``||L||^2 = 1``."""

    output_components = 1
    operator_norm_sq = 1.0
    _kind = N.REG_IDENTITY

    def __init__(self, norm: str = "L1"):
        self.norm = norm
        self._w = np.ones(3, dtype=np.float32)

    def forward(self, x):
        return as_f32(x).unsqueeze(0).clone()

    def adjoint(self, y):
        return as_f32(y)[0].clone()

    def __call__(self, x):
        return self.forward(x)


class _StencilRegularizer:
    """Gradient/Hessian regularizer around a stencil operator; 1-D outputs get a leading
    unit component axis (pdhg_mlx.py:309-456)."""

    def __init__(self, ndim: int, r: float, norm: str, ops):
        if ndim not in (1, 2, 3):
            raise ValueError(f"ndim must be 1, 2, or 3, got {ndim}")
        self.ndim, self.r, self.norm = ndim, r, norm
        self._op = ops[ndim - 1]() if ndim < 3 else ops[2](r=r)
        self._kind = self._op._kind
        self._w = self._op._w

    @property
    def operator_norm_sq(self) -> float:
        return self._op.operator_norm_sq

    def forward(self, x):
        out = self._op.forward(x)
        return out.unsqueeze(0) if self.ndim == 1 else out

    def adjoint(self, y):
        return self._op.adjoint(as_f32(y)[0] if self.ndim == 1 else y)

    def __call__(self, x):
        return self.forward(x)


class GradientRegularizer(_StencilRegularizer):
    """Gradient1D/2D/3D (total variation) regularizer (pdhg_mlx.py:309-377)."""

    def __init__(self, ndim: int, r: float = 1.0, norm: str = "L1"):
        super().__init__(ndim, r, norm, (Gradient1D, Gradient2D, Gradient3D))

    @property
    def output_components(self) -> int:
        return self.ndim


class HessianRegularizer(_StencilRegularizer):
    """Hessian1D/2D/3D second-order regularizer (pdhg_mlx.py:380-456)."""

    def __init__(self, ndim: int, r: float = 1.0, norm: str = "L1"):
        super().__init__(ndim, r, norm, (Hessian1D, Hessian2D, Hessian3D))

    @property
    def output_components(self) -> int:
        return 1 if self.ndim == 1 else (3 if self.ndim == 2 else 6)


def _create_regularizer(regularization: str, ndim: int, r: float, norm: str):
    """(pdhg_mlx.py:485-499)"""
    if regularization == "identity":
        return IdentityRegularizer(norm=norm)
    if regularization == "gradient":
        return GradientRegularizer(ndim=ndim, r=r, norm=norm)
    if regularization == "hessian":
        return HessianRegularizer(ndim=ndim, r=r, norm=norm)
    raise ValueError(f"Unknown regularization: {regularization}")


def _compute_spacing_ratio(spacing, ndim: int) -> float:
    """r = mean(dy, dx) / dz for 3-D, else 1 (pdhg_mlx.py:459-482)."""
    if spacing is None or ndim in (1, 2):
        return 1.0
    if len(spacing) != ndim:
        raise ValueError(f"spacing must have {ndim} elements, got {len(spacing)}")
    dz = spacing[0]
    lat = (spacing[1] + spacing[2]) / 2.0
    return lat / dz if dz > 0 else 1.0


def _normalize_bin_factors(factors, ndim):
    if isinstance(factors, int):
        factors = (factors,) * ndim
    elif len(factors) != ndim:
        raise ValueError(f"bin_factors has length {len(factors)}, expected {ndim}")
    else:
        factors = tuple(int(f) for f in factors)
    if any(f <= 0 for f in factors):
        raise ValueError("bin_factors must all be positive")
    return factors


def _normalize_sampling_factors(factors, ndim):
    if isinstance(factors, (int, float)):
        factors = (float(factors),) * ndim
    elif len(factors) != ndim:
        raise ValueError(f"sampling_factors has length {len(factors)}, expected {ndim}")
    else:
        factors = tuple(float(f) for f in factors)
    if any(f <= 0 for f in factors):
        raise ValueError("sampling_factors must all be positive")
    return factors


def _nonneg(x: torch.Tensor) -> torch.Tensor:
    """max(x, 0) into a new buffer (prox_nonneg, pdhg_mlx.py:192-194)."""
    out = empty_f32(x.shape)
    N.call("dh_clamp_min", x.data_ptr(), out.data_ptr(), x.numel(), 0.0, N.stream_ptr())
    return out


class _PDState:
    """Device-resident PDHG state for one solve (also driven by the batch/serving layers).

    The regularizer runs on the volume promoted to 3-D (``dh_pd_reg_*``); its dual
    holds the kernel family's internal components (6 Hessian, 3 gradient, 1
    identity planes), of which the 1-D and 2-D operators leave the unused ones zero.
    """

    def __init__(self, observed, blur_op, x0, alpha, num_iter, background=0.0, reg=None, norm="L1",
                 loss_type="poisson", delta=0.99, eta=0.5, tol=1e-5, min_iter=20, patience=5, r=1.0):
        if norm not in ("L1", "L1_2"):
            raise ValueError(f"Unknown norm: {norm}")
        if loss_type not in ("poisson", "gaussian"):
            raise ValueError(f"Unknown loss_type: {loss_type}")
        self.op = blur_op
        self.data = as_f32(observed)
        self.x = as_f32(x0).clone()
        if self.x.dim() not in (1, 2, 3):
            raise ValueError("PDHG solves 1-D, 2-D and 3-D arrays")
        self.S = tuple(self.x.shape)
        self.D3 = (1,) * (3 - len(self.S)) + self.S
        if reg is None:
            reg = HessianRegularizer(len(self.S), r, norm)
        self.reg = reg
        self.kind = int(reg._kind)
        self.alpha = float(alpha)
        self.bg = float(background)
        self.l12 = 1 if norm == "L1_2" else 0
        self.reg_norm = 2 if norm == "L1_2" else 1
        self.lt = 0 if loss_type == "poisson" else 1
        self.num_iter = int(num_iter)
        self.w = np.ascontiguousarray(reg._w, dtype=np.float32)
        self.wp = self.w.ctypes.data_as(N.C.POINTER(N.C.c_float))
        self.params = N.PdParams(delta=float(delta), eta=float(eta), tol=float(tol),
                                 min_iter=int(min_iter), patience=int(patience))
        self._plan = blur_op._plan if isinstance(blur_op, (LinearFFTConvolver, FFTConvolver)) else None
        dev = device()
        self.xp = self.x.clone()
        self.Ax = self._fwd(self.x, empty_f32(self.data.shape))
        self.Axp = self.Ax.clone()
        self.y1 = torch.zeros(self.data.shape, dtype=torch.float32, device=dev)
        nc = int(N.fns["dh_reg_components"](self.kind))
        self.y2 = torch.zeros((nc,) + self.D3, dtype=torch.float32, device=dev)
        self.g = empty_f32(self.S)
        # Hessian: the dual and primal steps run as one sweep (dh_pd_dual_primal), which
        # reads x, xp and y2 around each tile while it writes: the new iterate and dual
        # go to buffers of their own (xn, y2n), rotated after the step
        self.fused = self.kind == N.REG_HESSIAN and self.D3[1] * self.D3[2] < 2 ** 31
        if self.fused:
            self.xn = empty_f32(self.S)
            self.y2n = torch.empty_like(self.y2)
        # the y1 dual step runs in the adjoint's first pass (DH_PRO_PD_Y1) on canvases with
        # specialised kernels
        self.y1_pro = self._plan is not None and int(self._plan.info.spectrum_block) > 0
        if self._plan is not None:
            # the step sizes start from the OTF norm on the device: no host round trip
            self.sc = torch.empty(N.PD_NSLOTS, dtype=torch.float64, device=dev)
            N.call("dh_pd_init_scalars", self.sc.data_ptr(), self._plan.norm_dev.data_ptr(),
                   float(reg.operator_norm_sq), N.stream_ptr())
        else:
            step = 0.99 / math.sqrt(blur_op.operator_norm_sq + reg.operator_norm_sq)
            sc = np.zeros(N.PD_NSLOTS, np.float64)
            sc[_TAU] = sc[_SIGMA] = sc[_STEP] = step
            sc[_THETA] = 1.0
            self.sc = torch.from_numpy(sc).to(dev)
        self.hist = torch.zeros(3 * max(self.num_iter, 1), dtype=torch.float64, device=dev)
        self.loss = torch.zeros(2 * (self.num_iter + 1), dtype=torch.float64, device=dev)
        if self.data.numel() > int(np.prod(self.S)):
            raise ValueError("the data domain must not be larger than the reconstruction domain")
        self.part = torch.empty(int(N.fns["dh_pd_partials"](*self.D3)), dtype=torch.float64, device=dev)

    @property
    def step_scale(self) -> float:
        """0.99 / sqrt(|A|^2 + |L|^2) (pdhg_mlx.py:693-701)."""
        return float(self.sc[_STEP].item())

    def _fwd(self, x, out):
        """A x into ``out`` for a plan-backed operator; any other operator's own output
        is returned as is (callers keep the returned tensor, no copy)."""
        if self._plan is not None:
            return self._plan.apply(x, False, out)
        return self._own(as_f32(self.op.forward(x)), x)

    def _adj(self, y, out):
        if self._plan is not None:
            return self._plan.apply(y, True, out)
        return self._own(as_f32(self.op.adjoint(y)), y)

    def _own(self, r, arg):
        """``r`` unless it shares memory with the argument or with a buffer the state
        keeps (an operator that returns a persistent buffer): then a copy."""
        held = [arg] + [getattr(self, n) for n in ("x", "xp", "xn", "Ax", "Axp", "g", "y1") if hasattr(self, n)]
        if any(r.data_ptr() == h.data_ptr() for h in held):
            return r.clone()
        return r

    @property
    def x_next(self) -> torch.Tensor:
        """The buffer the primal step writes x_{k+1} into."""
        return self.xn if self.fused else self.xp

    # one iteration in five phases, so that a batch of solves can group its convolutions
    def _pre(self, k: int) -> None:
        """y1 dual step (unless the adjoint's first pass runs it), and the y2 dual step
        unless it is fused with the primal one (pdhg_mlx.py:717-734)."""
        st = N.stream_ptr()
        nz, ny, nx = self.D3
        sc = self.sc.data_ptr()
        if not self.y1_pro:
            N.call("dh_pd_y1", self.y1.data_ptr(), self.Ax.data_ptr(), self.Axp.data_ptr(), self.data.data_ptr(),
                   self.data.numel(), sc, self.bg, self.lt, st)
        if not self.fused:
            N.call("dh_pd_reg_dual", self.kind, self.y2.data_ptr(), self.x.data_ptr(), self.xp.data_ptr(), nz, ny,
                   nx, sc, self.wp, self.alpha, self.l12, st)

    def _mid(self, k: int) -> None:
        """Primal step with the regularizer's adjoint (pdhg_mlx.py:741-750); for the
        Hessian the y2 dual step runs in the same sweep (y2 does not enter A^T y1)."""
        nz, ny, nx = self.D3
        if self.fused:
            N.call("dh_pd_dual_primal", self.y2n.data_ptr(), self.xn.data_ptr(), self.x.data_ptr(),
                   self.xp.data_ptr(), self.y2.data_ptr(), self.g.data_ptr(), nz, ny, nx, self.sc.data_ptr(),
                   self.wp, self.alpha, self.l12, self.part.data_ptr(), N.stream_ptr())
            self.y2, self.y2n = self.y2n, self.y2
        else:
            N.call("dh_pd_reg_primal", self.kind, self.xp.data_ptr(), self.x.data_ptr(), self.g.data_ptr(),
                   self.y2.data_ptr(), nz, ny, nx, self.sc.data_ptr(), self.wp, self.part.data_ptr(),
                   N.stream_ptr())

    def _adjoint(self) -> None:
        """g = A^T y1_new, the y1 step in the first pass where the operator allows."""
        afz = self._adj_fuse()
        if afz is not None:
            self.g = self._plan.apply(self.y1, True, self.g, afz)
        else:
            self.g = self._adj(self.y1, self.g)

    def _forward(self, want_loss: bool) -> None:
        """A x_new into Axp, |A x_new - A x|^2 in its last pass where the operator allows."""
        fuse = self._fwd_fuse(want_loss)
        if fuse is not None:
            self.Axp = self._plan.apply(self.x_next, False, self.Axp, fuse)
        else:
            self.Axp = self._fwd(self.x_next, self.Axp)

    def _adj_fuse(self) -> Optional[N.FuseArgs]:
        """Prologue of the adjoint A^T y1 on a plan with specialised kernels: the y1
        dual step (dh_pd_y1's values) on each sample as the first pass reads it, y1
        updated in place."""
        if not self.y1_pro:
            return None
        f = N.FuseArgs(prologue=N.PRO_PD_Y1, pro_a=self.bg, pro_loss=self.lt, pro_out=self.y1.data_ptr(),
                       pro_scalars=self.sc.data_ptr())
        f.pro_aux[:] = (self.Ax.data_ptr(), self.Axp.data_ptr(), self.data.data_ptr())
        return f

    def _fwd_fuse(self, want_loss: bool) -> Optional[N.FuseArgs]:
        """Epilogue of the forward convolution A x_new on a plan: |A x_new - A x|^2 summed
        into the DA2 slot in the inverse X pass (the loss iterations keep the separate
        data-sums kernel, which also evaluates the data term)."""
        if self._plan is None or want_loss:
            return None
        return N.FuseArgs(epilogue=N.EPI_SQDIFF, aux0=self.Ax.data_ptr(),
                          reduce_out=self.sc.data_ptr() + 8 * _DA2)

    def _post(self, k: int, want_loss: bool) -> None:
        """Step-size sums, the Malitsky-Pock update and the rotation of the iterates
        (pdhg_mlx.py:752-860)."""
        st = N.stream_ptr()
        nz, ny, nx = self.D3
        sc = self.sc.data_ptr()
        dl = self.loss[2 * k:].data_ptr() if want_loss else None
        rl = self.loss[2 * k + 1:].data_ptr() if want_loss else None
        xnew = self.x_next
        if self._fwd_fuse(want_loss) is None:
            N.call("dh_pd_data_sums", self.Axp.data_ptr(), self.Ax.data_ptr(), self.data.data_ptr(),
                   self.data.numel(), self.bg, self.lt, sc, dl, self.part.data_ptr(), st)
        N.call("dh_pd_reg_sums", self.kind, xnew.data_ptr(), self.x.data_ptr(), nz, ny, nx, self.wp, sc, rl,
               self.reg_norm, self.part.data_ptr(), st)
        N.call("dh_pd_scalars", sc, self.hist.data_ptr(), k, N.C.byref(self.params), st)
        if self.fused:
            self.x, self.xp, self.xn = self.xn, self.x, self.xp
        else:
            self.x, self.xp = self.xp, self.x
        self.Ax, self.Axp = self.Axp, self.Ax

    def _conv_item(self, adjoint: bool, want_loss: bool = False) -> N.BatchItem:
        p = self._plan
        src, dst = (self.y1, self.g) if adjoint else (self.x_next, self.Axp)
        fuse = self._adj_fuse() if adjoint else self._fwd_fuse(want_loss)
        return N.BatchItem(plan=p._h, otf=p.otf_buf.data_ptr(), work=p.work.data_ptr(),
                           work_bytes=int(p.info.work_bytes), input=src.data_ptr(), out=dst.data_ptr(),
                           fuse=fuse if fuse is not None else N.FuseArgs())

    def step(self, k: int, want_loss: bool) -> None:
        self._pre(k)
        self._adjoint()
        self._mid(k)
        self._forward(want_loss)
        self._post(k, want_loss)

    def done(self) -> bool:
        return bool(self.sc[_DONE].item() != 0.0)

    def loss_value(self, k: int) -> float:
        v = self.loss[2 * k: 2 * k + 2].cpu().numpy()
        return float(v[0] + self.alpha * v[1])


def pd_batch_step(states: List[_PDState], ks: List[int], want_loss=None, max_items: int = 256) -> None:
    """One iteration of several PDHG solves (any shapes): the same kernels as
    ``_PDState.step`` on each, with the adjoint and forward convolutions of the
    plan-backed solves grouped per transform length (``dh_batch_conv_apply``), so
    each solve's result equals its solo run bit for bit."""
    want_loss = want_loss or [False] * len(states)
    planned = [s for s in states if s._plan is not None]
    wl_of = {id(s): w for s, w in zip(states, want_loss)}
    for s, k in zip(states, ks):
        s._pre(k)
    for adjoint in (True, False):
        for j in range(0, len(planned), max_items):
            chunk = planned[j:j + max_items]
            wl = [wl_of[id(s)] for s in chunk]
            arr = (N.BatchItem * len(chunk))(*[s._conv_item(adjoint, w) for s, w in zip(chunk, wl)])
            N.call("dh_batch_conv_apply", len(chunk), arr, int(adjoint), N.stream_ptr())
            for s in chunk:
                s._plan.gen += 1
        for s in states:
            if s._plan is None:
                if adjoint:
                    s._adjoint()
                else:
                    s._forward(wl_of[id(s)])
        if adjoint:
            for s, k in zip(states, ks):
                s._mid(k)
    for s, k, w in zip(states, ks, want_loss):
        s._post(k, w)


def _run(state: _PDState, num_iter, min_iter, patience, tol, eval_interval, callback, verbose,
         loss_every_iter, callback_with_loss, loss_on_interval=True):
    """Iteration driver shared by both entry points; returns (iterations, converged, losses)."""
    can_stop_from = min_iter + patience - 1 if tol > 0 else num_iter
    losses_at: List[int] = []
    final_iter = num_iter
    converged = False
    k_done = -1
    for k in range(num_iter):
        want = loss_every_iter or (loss_on_interval and (k + 1) % eval_interval == 0)
        state.step(k, want)
        k_done = k
        if k >= can_stop_from and state.done():
            converged = True
            final_iter = k + 1
            break
        if want:
            losses_at.append(k)
        if callback is not None:
            torch.cuda.current_stream().synchronize()
            # a copy: the solver reuses its buffers (the reference hands over an immutable array)
            if callback_with_loss:
                callback(k, state.x.clone(), state.loss_value(k) if want else None)
            elif callback(k, state.x.clone()):
                final_iter = k + 1
                break
        if verbose and want and not loss_every_iter:
            print(f"Iter {k + 1:4d}: loss={state.loss_value(k):.4e}")
    hist = state.hist[: 3 * (k_done + 1)].cpu().numpy().reshape(-1, 3)
    losses = [state.loss_value(k) for k in losses_at]
    return final_iter, converged, losses, hist


def solve_pdhg_mlx(
    observed,
    psf,
    alpha: float = 0.01,
    regularization: str = "hessian",
    norm: str = "L1",
    loss_type: str = "poisson",
    num_iter: int = 200,
    background: float = 0.0,
    spacing: Optional[Tuple[float, ...]] = None,
    bin_factors=None,
    sampling_factors=None,
    icf_sigmas=None,
    icf_spacings=None,
    init=None,
    verbose: bool = False,
    callback: Optional[Callable] = None,
    delta: float = 0.99,
    eta: float = 0.5,
    eval_interval: int = 10,
    tol: float = 1e-5,
    min_iter: int = 20,
    patience: int = 5,
) -> MLXDeconvolutionResult:
    """Adaptive PDHG deconvolution (pdhg_mlx.py:507-920), same signature and result."""
    obs = as_f32(observed)
    ndim = obs.dim()
    if bin_factors is not None and sampling_factors is not None:
        raise ValueError("bin_factors and sampling_factors are mutually exclusive")
    if icf_sigmas is not None:
        raise NotImplementedError(
            "ICF (icf_sigmas) is not yet supported with atomic operators. "
            "Use the composed operator API directly or remove icf_sigmas parameter."
        )
    if regularization not in ("identity", "gradient", "hessian"):
        raise ValueError(f"Unknown regularization: {regularization}")
    psf_shape = tuple(int(s) for s in (psf.shape if hasattr(psf, "shape") else np.shape(psf)))
    if sampling_factors is not None:
        f = _normalize_sampling_factors(sampling_factors, ndim)
        hi = tuple(max(1, int(round(n * s))) for n, s in zip(obs.shape, f))
        if psf_shape != hi:
            raise ValueError(
                "For sampling_factors, psf.shape must match the reconstructed "
                f"high-resolution shape {hi}; got {psf_shape}"
            )
        A = compose(FractionalAreaDownsample(scale=f), LinearFFTConvolver(psf, hi, normalize=True))
    elif bin_factors is not None:
        f = _normalize_bin_factors(bin_factors, ndim)
        hi = psf_shape
        for i, (n, b) in enumerate(zip(hi, f)):
            if n % b != 0:
                raise ValueError(f"PSF dim {i} size {n} is not divisible by bin factor {b}")
        lo = tuple(n // b for n, b in zip(hi, f))
        if tuple(obs.shape) != lo:
            raise ValueError(
                "For bin_factors, observed.shape must match psf.shape divided "
                f"by bin_factors; expected {lo}, got {tuple(obs.shape)}"
            )
        A = compose(FractionalAreaDownsample(scale=f), LinearFFTConvolver(psf, hi, normalize=True))
    else:
        A = LinearFFTConvolver(psf, tuple(obs.shape), normalize=True)
    r = _compute_spacing_ratio(spacing, ndim)
    if init is not None:
        x0 = as_f32(init)
    else:
        if sampling_factors is not None or bin_factors is not None:
            x0 = as_f32(A.adjoint(obs))
        else:
            x0 = obs
        x0 = _nonneg(x0)
    reg = _create_regularizer(regularization, ndim, r, norm)
    state = _PDState(obs, A, x0, alpha, num_iter, background, reg, norm, loss_type, delta, eta, tol,
                     min_iter, patience)
    iters, conv, losses, hist = _run(state, num_iter, min_iter, patience, tol, eval_interval, callback,
                                     verbose, False, False)
    return MLXDeconvolutionResult(
        restored=state.x,
        iterations=iters,
        loss_history=losses,
        converged=conv,
        tau_history=[float(v) for v in hist[:, 0]],
        sigma_history=[float(v) for v in hist[:, 1]],
        metadata={
            "algorithm": "malitsky_pock_pdhg",
            "regularization": regularization,
            "norm": norm,
            "loss_type": loss_type,
            "alpha": alpha,
            "background": background,
            "tol": tol,
            "min_iter": min_iter,
            "patience": patience,
        },
    )


def solve_pdhg_with_operator(
    observed,
    blur_op,
    alpha: float = 0.01,
    regularization: str = "identity",
    norm: str = "L1",
    loss_type: str = "poisson",
    num_iter: int = 200,
    background: float = 0.0,
    init=None,
    verbose: bool = False,
    callback: Optional[Callable] = None,
    delta: float = 0.99,
    eta: float = 0.5,
    eval_interval: int = 10,
    tol: float = 1e-5,
    min_iter: int = 20,
    patience: int = 5,
) -> MLXDeconvolutionResult:
    """PDHG on a pre-built operator (pdhg_mlx.py:928-1242).

    Deviation (documented): the solution shape is ``blur_op.in_shape`` when the
    operator declares it (the reference assumes ``observed.shape``, which breaks
    operators whose domain and range differ).
    """
    obs = as_f32(observed)
    shp = getattr(blur_op, "shape", None)
    if shp is not None and len(shp) == 2:
        # a matrix operator (m, n): the solution is a length-n vector (pdhg_mlx.py:1011-1014)
        shape = (int(shp[1]),)
    else:
        shape = tuple(getattr(blur_op, "in_shape", None) or obs.shape)
    reg = _create_regularizer(regularization, len(shape), 1.0, norm)
    x0 = as_f32(init) if init is not None else torch.full(shape, 0.01, dtype=torch.float32, device=device())
    x0 = _nonneg(x0)
    state = _PDState(obs, blur_op, x0, alpha, num_iter, background, reg, norm, loss_type, delta, eta, tol,
                     min_iter, patience)
    every = bool(verbose or callback is not None)
    iters, conv, losses, hist = _run(state, num_iter, min_iter, patience, tol, eval_interval, callback,
                                     verbose, every, True, loss_on_interval=False)
    return MLXDeconvolutionResult(
        restored=state.x,
        iterations=iters,
        loss_history=losses,
        converged=conv,
        tau_history=[float(v) for v in hist[:, 0]],
        sigma_history=[float(v) for v in hist[:, 1]],
        metadata={
            "algorithm": "malitsky_pock_pdhg",
            "regularization": regularization,
            "norm": norm,
            "loss_type": loss_type,
            "alpha": alpha,
            "background": background,
            "tol": tol,
            "min_iter": min_iter,
            "patience": patience,
            "operator_type": "generic",
        },
    )
