"""Edge-preserving Hessian-log deconvolution by Gauss-Newton-CG on MI355X
(drop-in for deconlib's erdecon_mlx: erdecon_with_operator, erdecon_solver,
ERDeconResult, erdecon_objective, erdecon_gradient, erdecon_gn_hvp).

Semantics follow erdecon_mlx.py:396-692 step for step: amplitude
normalisation, g = s^2, the Gaussian or Poisson data term, the IRLS weights
w = lam / (eps + q) of the curvature q = sum_i (H_i g)^2 (with the optional
floor and intensity term), inner CG on the Gauss-Newton Hessian from zero with
its curvature guard and relative-residual stop, the steepest-descent fallback,
Armijo backtracking, and the Newton-decrement / data-misfit stopping tests.

Per outer iteration: the gradient (one adjoint convolution; the model of the
current iterate is the accepted trial's, not recomputed), each CG step (two
convolutions, the weighted Hessian w * H(s v) formed by one stencil kernel and
its adjoint), and each line-search trial (one convolution plus one stencil
pass that only sums).  For ``Hessian3D`` / ``Hessian2D`` the curvature and
weights come from the stencil on the fly; any other regulariser goes through
its own forward/adjoint and the library's stacked-channel kernels.  The CG
scalars (step, beta, the stopping tests) live on the device; the host reads
them once per CG step, and once per line-search trial, as the reference does.
Sums are float64 (the reference's are float32).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Any, Callable, Optional, Tuple

import numpy as np
import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32
from .linops import Hessian2D, Hessian3D
from .nlcg import _conv
from .rl import _plan_of


@dataclass
class ERDeconResult:
    """Result of ER-Decon (erdecon_mlx.py:91-129); arrays are device tensors."""

    restored: torch.Tensor
    pred: torch.Tensor
    iterations: int
    loss_history: list
    data_misfit_history: list = field(default_factory=list)
    converged: bool = False
    background: float = 0.0
    data_scale: float = 1.0
    full_shape: Optional[Tuple[int, ...]] = None
    valid_slices: Optional[Tuple[slice, ...]] = None


def _default_hessian(ndim: int):
    """Hessian2D / Hessian3D by rank (erdecon_mlx.py:132-141)."""
    if ndim == 2:
        return Hessian2D()
    if ndim == 3:
        return Hessian3D()
    raise ValueError(
        f"ER-Decon supports 2D or 3D reconstructions; got ndim={ndim}. "
        "Pass an explicit `hessian` operator for other ranks."
    )


def _check_data_term(data_term: str) -> None:
    if data_term not in ("gaussian", "poisson"):
        raise ValueError(f"data_term must be 'gaussian' or 'poisson'; got {data_term!r}")


class _Scalars:
    def __init__(self):
        dev = device()
        self.sc = torch.zeros(N.ER_NSLOTS, dtype=torch.float64, device=dev)
        self.part = torch.empty(int(N.fns["dh_er_partials"]()), dtype=torch.float64, device=dev)
        self.host = torch.empty(N.ER_NSLOTS, dtype=torch.float64, pin_memory=True)

    def read(self):
        self.host.copy_(self.sc, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        return self.host.tolist()


def _ew(op, a, out, alpha=0.0, b=None, out2=None, beta=0.0):
    N.call("dh_er_ew", op, a.data_ptr(), N.ptr(b), float(alpha), float(beta), out.data_ptr(), N.ptr(out2), a.numel(),
           N.stream_ptr())
    return out


class _Reg:
    """The regulariser H: fused stencil kernels for Hessian3D / Hessian2D, otherwise
    the operator's own forward/adjoint with the stacked-channel kernels."""

    def __init__(self, hessian, shape):
        self.h = hessian
        self.shape = tuple(shape)
        self.n = int(np.prod(self.shape))
        if isinstance(hessian, Hessian3D) and len(self.shape) == 3:
            self.dims = self.shape
        elif isinstance(hessian, Hessian2D) and len(self.shape) == 2:
            self.dims = (1,) + self.shape
        else:
            self.dims = None
        self.u = empty_f32((6, self.n)) if self.dims is not None else None

    def _w6(self):
        return self.h._w.ctypes.data_as(C.POINTER(C.c_float))

    def weights(self, g, p, w_out, S) -> Optional[torch.Tensor]:
        """LOGQ and Q of g; with w_out, also w and the adjoint's operand w * H g (returned)."""
        lam, eps, w_floor, iw = p
        st = N.stream_ptr()
        if self.dims is not None:
            N.call("dh_er_hess_weights", g.data_ptr(), *self.dims, self._w6(), lam, eps, w_floor, iw, N.ptr(w_out),
                   N.ptr(self.u) if w_out is not None else None, S.sc.data_ptr(), S.part.data_ptr(), st)
            return self.u if w_out is not None else None
        hg = as_f32(self.h.forward(g))
        if hg.numel() % self.n:
            raise ValueError("the regulariser's forward must return stacked channels of the domain's shape")
        nc = hg.numel() // self.n
        N.call("dh_er_weights_stacked", hg.data_ptr(), nc, g.data_ptr(), self.n, lam, eps, w_floor, iw,
               N.ptr(w_out), hg.data_ptr() if w_out is not None else None, S.sc.data_ptr(), S.part.data_ptr(), st)
        return hg

    def apply_w(self, v, w) -> torch.Tensor:
        """w * H(v)."""
        if self.dims is not None:
            N.call("dh_er_hess_apply_w", v.data_ptr(), w.data_ptr(), *self.dims, self._w6(), self.u.data_ptr(),
                   N.stream_ptr())
            return self.u
        hv = as_f32(self.h.forward(v))
        N.call("dh_er_scale_stacked", hv.data_ptr(), w.data_ptr(), hv.numel() // self.n, self.n, hv.data_ptr(),
               N.stream_ptr())
        return hv

    def adjoint(self, u, out=None) -> torch.Tensor:
        if self.dims is not None:
            out = out if out is not None else empty_f32(self.shape)
            N.call("dh_hessian_adjoint", u.data_ptr(), out.data_ptr(), *self.dims, self._w6(), N.stream_ptr())
            return out
        nc = u.numel() // self.n
        return as_f32(self.h.adjoint(u.reshape((nc,) + self.shape)))


class _ERState:
    """Device-resident state of one ER-Decon solve (normalised units)."""

    def __init__(self, f, op, hessian, b, lam, eps_reg, data_term, floor_frac, iw, init):
        self.op, self.f, self.b = op, f, float(b)
        self.lam, self.eps = float(lam), float(eps_reg)
        self.poisson = 1 if data_term == "poisson" else 0
        self.floor_frac, self.iw = float(floor_frac), float(iw)
        self.wp = (self.lam, self.eps, float(floor_frac * lam / eps_reg) if floor_frac > 0.0 else 0.0, self.iw)
        self.obj_wp = (self.lam, self.eps, 0.0, self.iw)
        self.S = _Scalars()
        st = N.stream_ptr()
        if init is None:
            if _has_plan(op):
                g0 = _conv(op, f, True, fuse=N.FuseArgs(prologue=N.PRO_SUB_FLOOR, pro_a=self.b, pro_b=1e-6))
            else:
                dmb = empty_f32(f.shape)
                N.call("dh_sub_floor", f.data_ptr(), dmb.data_ptr(), f.numel(), self.b, 1e-6, st)
                g0 = _conv(op, dmb, True)
        else:
            g0 = init
        shape = tuple(g0.shape)
        self.s, self.g = empty_f32(shape), empty_f32(shape)
        N.call("dh_nlcg_init", g0.data_ptr(), self.s.data_ptr(), self.g.data_ptr(), g0.numel(), 1e-6, st)
        self._alloc(hessian, shape)
        self.s_try, self.g_try = empty_f32(shape), empty_f32(shape)
        self.m = self.model(self.g)

    def _alloc(self, hessian, shape) -> None:
        """Working vectors: domain (n) and data (nd) sized, one per operand."""
        self.n, self.nd = int(np.prod(shape)), self.f.numel()
        self.reg = _Reg(hessian, shape)
        self.w, self.grad = empty_f32(shape), empty_f32(shape)
        self.x, self.r, self.p = empty_f32(shape), empty_f32(shape), empty_f32(shape)
        self.sv, self.ap = empty_f32(shape), empty_f32(shape)
        self.dg, self.rh, self.kt, self.htu = empty_f32(shape), empty_f32(shape), empty_f32(shape), empty_f32(shape)
        self.score = empty_f32(self.f.shape)
        self.hessw = empty_f32(self.f.shape) if self.poisson else None
        self.ksv, self.tvec = empty_f32(self.f.shape), empty_f32(self.f.shape)

    def model(self, g, out=None) -> torch.Tensor:
        """K g + b."""
        if _has_plan(self.op):
            out = out if out is not None else empty_f32(self.f.shape)
            return _conv(self.op, g, False, out, N.FuseArgs(epilogue=N.EPI_ADD, epi_a=self.b))
        y = _conv(self.op, g, False)
        out = empty_f32(y.shape)
        N.call("dh_add_scalar", y.data_ptr(), None, out.data_ptr(), y.numel(), self.b, None, None, N.stream_ptr())
        return out

    def objective_sums(self, g, m) -> float:
        """phi from the model m of g (erdecon_mlx.py:253-278): one data pass, one stencil pass."""
        S = self.S
        N.call("dh_er_data", m.data_ptr(), self.f.data_ptr(), self.nd, self.poisson, None, None, S.sc.data_ptr(),
               S.part.data_ptr(), N.stream_ptr())
        self.reg.weights(g, self.obj_wp, None, S)
        v = S.read()
        phi = v[N.ER_DATA] + 0.5 * self.lam * v[N.ER_LOGQ]
        if self.floor_frac > 0.0:
            phi += 0.5 * self.floor_frac * (self.lam / self.eps) * v[N.ER_Q]
        return phi, v

    def gradient(self, cg_tol: float) -> None:
        """grad at s from the current model, and the CG start (erdecon_mlx.py:281-308, 366-372)."""
        S, st = self.S, N.stream_ptr()
        N.call("dh_er_data", self.m.data_ptr(), self.f.data_ptr(), self.nd, self.poisson, self.score.data_ptr(),
               N.ptr(self.hessw), S.sc.data_ptr(), S.part.data_ptr(), st)
        dg = _conv(self.op, self.score, True, self.dg if _has_plan(self.op) else None)
        u = self.reg.weights(self.g, self.wp, self.w, S)
        rh = self.reg.adjoint(u, self.rh)
        N.call("dh_er_gradient", self.s.data_ptr(), dg.data_ptr(), rh.data_ptr(), self.w.data_ptr(),
               self.g.data_ptr(), self.iw, self.n, float(cg_tol), self.grad.data_ptr(), self.x.data_ptr(),
               self.r.data_ptr(), self.p.data_ptr(), self.sv.data_ptr(), S.sc.data_ptr(), S.part.data_ptr(), st)

    def hvp_parts(self):
        """(K^T(d K(sv)), H^T(w H(sv))) for the current sv."""
        plan = _has_plan(self.op)
        ksv = _conv(self.op, self.sv, False, self.ksv if plan else None)
        if self.poisson:
            t = _ew(N.ER_EW_MUL, ksv, self.tvec, b=self.hessw)
        else:
            t = _ew(N.ER_EW_SCALE, ksv, self.tvec, alpha=2.0)
        kt = _conv(self.op, t, True, self.kt if plan else None)
        htu = self.reg.adjoint(self.reg.apply_w(self.sv, self.w), self.htu)
        return kt, htu

    def cg(self, max_steps: int, stopped: bool) -> int:
        """Inner CG for H_GN x = -grad (erdecon_mlx.py:348-393); returns the HVPs made."""
        S, st = self.S, N.stream_ptr()
        if stopped:
            return 0
        steps = 0
        for _ in range(max_steps):
            kt, htu = self.hvp_parts()
            N.call("dh_er_cg_ap", self.s.data_ptr(), kt.data_ptr(), htu.data_ptr(), self.w.data_ptr(),
                   self.sv.data_ptr(), self.iw, self.p.data_ptr(), self.ap.data_ptr(), self.n, S.sc.data_ptr(),
                   S.part.data_ptr(), st)
            N.call("dh_er_cg_update", self.x.data_ptr(), self.r.data_ptr(), self.p.data_ptr(), self.ap.data_ptr(),
                   self.s.data_ptr(), self.sv.data_ptr(), self.n, S.sc.data_ptr(), S.part.data_ptr(), st)
            steps += 1
            if S.read()[N.ER_STOP] != 0.0:
                break
        return steps


def _has_plan(op) -> bool:
    return _plan_of(op)[0] is not None


def erdecon_with_operator(
    observed,
    blur_op,
    hessian=None,
    reg_weight: float = 0.05,
    eps_reg: float = 1e-2,
    data_term: str = "gaussian",
    floor_frac: float = 0.0,
    intensity_weight: float = 0.0,
    num_iter: int = 50,
    background: float = 0.0,
    normalize: bool = True,
    data_scale: Optional[float] = None,
    init=None,
    cg_max_steps: int = 25,
    cg_tol: float = 0.1,
    callback: Optional[Callable[[int, torch.Tensor], bool]] = None,
    eval_interval: int = 5,
    newton_tol: float = 1e-3,
    tol: float = 0.0,
    min_iter: int = 5,
    ls_max_backtracks: int = 30,
    ls_c1: float = 1e-4,
    verbose: bool = False,
) -> ERDeconResult:
    """Hessian-log deconvolution by Gauss-Newton-CG with a positive operator (erdecon_mlx.py:396-692)."""
    f = as_f32(observed)
    st_ptr = N.stream_ptr()
    if data_scale is not None:
        c = float(data_scale)
    elif normalize:
        part = torch.empty(N.fns["dh_reduce_partials"](f.numel()), dtype=torch.float64, device=device())
        mx = torch.empty(1, dtype=torch.float64, device=device())
        N.call("dh_reduce", N.RED_MAX, f.data_ptr(), f.numel(), part.data_ptr(), mx.data_ptr(), st_ptr)
        c = float(np.float32(mx.item()))
    else:
        c = 1.0
    if not math.isfinite(c) or c <= 0.0:
        c = 1.0
    ini = as_f32(init) if init is not None else None
    if c != 1.0:
        f = _ew(N.ER_EW_DIV, f, empty_f32(f.shape), alpha=c)
        if ini is not None:
            ini = _ew(N.ER_EW_DIV, ini, empty_f32(ini.shape), alpha=c)
    _check_data_term(data_term)
    b = float(background) / c
    lam = float(reg_weight)
    if hessian is None:
        hessian = _default_hessian(f.dim())
    st = _ERState(f, blur_op, hessian, b, lam, eps_reg, data_term, floor_frac, intensity_weight, ini)
    nd = st.nd

    loss_history: list = []
    misfit_history: list = []
    prev_misfit = None
    nd0 = None
    converged = False
    phi, v = st.objective_sums(st.g, st.m)
    k = 0
    for k in range(num_iter):
        st.gradient(cg_tol)
        v = st.S.read()
        misfit_cur = c * v[N.ER_IDIV] / nd
        gg = v[N.ER_GG]
        st.cg(cg_max_steps, v[N.ER_STOP] != 0.0)
        N.call("dh_er_dot", st.grad.data_ptr(), st.x.data_ptr(), st.n, N.ER_DOT, st.S.sc.data_ptr(),
               st.S.part.data_ptr(), N.stream_ptr())
        gtp = st.S.read()[N.ER_DOT]
        if not math.isfinite(gtp) or gtp >= 0.0:
            _ew(N.ER_EW_NEG, st.grad, st.x)  # steepest descent
            gtp = -gg
        ndec = abs(gtp)
        if nd0 is None:
            nd0 = ndec
        alpha = 1.0
        accepted = False
        m_try = None
        for _ in range(ls_max_backtracks):
            _ew(N.ER_EW_AXPY_SQUARE, st.s, st.s_try, alpha=alpha, b=st.x, out2=st.g_try)
            m_try = st.model(st.g_try)
            phi_try, vt = st.objective_sums(st.g_try, m_try)
            if math.isfinite(phi_try) and phi_try <= phi + ls_c1 * alpha * gtp:
                accepted = True
                break
            alpha *= 0.5
        if not accepted:
            converged = True
            loss_history.append(phi)
            misfit_history.append(misfit_cur)
            if verbose:
                print(f"  Stopped at iter {k}: line search found no decrease")
            break
        st.s, st.s_try = st.s_try, st.s
        st.g, st.g_try = st.g_try, st.g
        st.m = m_try
        phi = phi_try
        logged = k % eval_interval == 0 or k == num_iter - 1
        misfit = c * vt[N.ER_IDIV] / nd
        if logged:
            loss_history.append(phi)
            misfit_history.append(misfit)
            if verbose:
                print(f"  Iter {k:4d}: I-div = {misfit:.6g}, phi = {phi:.6g}, alpha = {alpha:.3g}, "
                      f"|grad^T p| = {abs(gtp):.3g}")
        stop = False
        if callback is not None:
            stop = bool(callback(k, _ew(N.ER_EW_SQUARE_SCALE, st.s, empty_f32(st.s.shape), alpha=c)))
        if (not stop and newton_tol > 0.0 and k + 1 >= min_iter and nd0 is not None and nd0 > 0.0
                and ndec < newton_tol * nd0):
            converged = stop = True
            if verbose:
                print(f"  Stopped at iter {k}: Newton decrement {ndec:.3g} < {newton_tol:.1g} * {nd0:.3g}")
        if not stop and k + 1 >= min_iter and tol > 0.0 and prev_misfit is not None:
            rel = abs(prev_misfit - misfit) / max(abs(prev_misfit), 1e-30)
            if rel < tol:
                converged = stop = True
                if verbose:
                    print(f"  Stopped at iter {k}: relative I-div change {rel:.3g}")
        prev_misfit = misfit
        if stop:
            if not logged:
                loss_history.append(phi)
                misfit_history.append(misfit)
            break

    g_final = _ew(N.ER_EW_SQUARE_SCALE, st.s, empty_f32(st.s.shape), alpha=c)
    if _has_plan(blur_op):
        pred = _conv(blur_op, g_final, False, empty_f32(f.shape), N.FuseArgs(epilogue=N.EPI_ADD, epi_a=b * c))
    else:
        y = _conv(blur_op, g_final, False)
        pred = empty_f32(y.shape)
        N.call("dh_add_scalar", y.data_ptr(), None, pred.data_ptr(), y.numel(), b * c, None, None, N.stream_ptr())
    return ERDeconResult(
        restored=g_final,
        pred=pred,
        iterations=k + 1,
        loss_history=loss_history,
        data_misfit_history=misfit_history,
        converged=converged,
        background=float(background),
        data_scale=c,
        full_shape=tuple(g_final.shape),
        valid_slices=None,
    )


def erdecon_solver(reg_weight: float = 0.05, eps_reg: float = 1e-2, num_iter: int = 50, background: float = 0.0,
                   hessian=None, init_value: Optional[float] = None, newton_tol: float = 1e-3,
                   **erdecon_kwargs: Any) -> Callable:
    """ER-Decon adapted to the ``solve(data, model) -> visible array`` contract (erdecon_mlx.py:695-733)."""

    def solve(data, model):
        init = None
        if init_value is not None:
            init = torch.full(model.padded_shape, float(init_value), dtype=torch.float32, device=device())
        result = erdecon_with_operator(observed=data, blur_op=model.op, hessian=hessian, reg_weight=reg_weight,
                                       eps_reg=eps_reg, num_iter=num_iter, background=background, init=init,
                                       newton_tol=newton_tol, **erdecon_kwargs)
        return result.restored[model.valid_slices]

    solve.accepts_device_tiles = True
    return solve


# ---------------------------------------------------------------------------
# The functional, its gradient and the Gauss-Newton product (erdecon_mlx.py:253-345)
# ---------------------------------------------------------------------------


def _helper_state(s, blur_op, observed, hessian, background, data_term):
    _check_data_term(data_term)
    s = as_f32(s)
    st = _ERState.__new__(_ERState)
    st.op, st.f, st.b = blur_op, as_f32(observed), float(background)
    st.poisson = 1 if data_term == "poisson" else 0
    st.S = _Scalars()
    st.s = s
    st.g = _ew(N.ER_EW_SQUARE_SCALE, s, empty_f32(s.shape), alpha=1.0)
    st._alloc(hessian, tuple(s.shape))
    st.m = st.model(st.g)
    return st


def erdecon_objective(s, blur_op, observed, hessian, background: float = 0.0, reg_weight: float = 0.05,
                      eps_reg: float = 1e-2, data_term: str = "gaussian", floor_frac: float = 0.0,
                      intensity_weight: float = 0.0) -> float:
    """phi(s) = D(K g, f) + (lam/2) sum log(eps + q) [+ (w_floor/2) sum q] (erdecon_mlx.py:253-278)."""
    st = _helper_state(s, blur_op, observed, hessian, background, data_term)
    st.lam, st.eps, st.floor_frac = float(reg_weight), float(eps_reg), float(floor_frac)
    st.obj_wp = (st.lam, st.eps, 0.0, float(intensity_weight))
    return st.objective_sums(st.g, st.m)[0]


def erdecon_gradient(s, blur_op, observed, hessian, background: float = 0.0, reg_weight: float = 0.05,
                     eps_reg: float = 1e-2, data_term: str = "gaussian", floor_frac: float = 0.0,
                     intensity_weight: float = 0.0):
    """(grad, (w, data_hess_w)) at s (erdecon_mlx.py:281-308); w has shape (1, *spatial)."""
    st = _helper_state(s, blur_op, observed, hessian, background, data_term)
    st.iw = float(intensity_weight)
    st.wp = (float(reg_weight), float(eps_reg),
             float(floor_frac * reg_weight / eps_reg) if floor_frac > 0.0 else 0.0, st.iw)
    shape = tuple(st.s.shape)
    st.gradient(0.1)
    return st.grad, (st.w.reshape((1,) + shape), st.hessw if st.poisson else 2.0)


def erdecon_gn_hvp(s, v, blur_op, hessian, w, data_hess_w=2.0, intensity_weight: float = 0.0) -> torch.Tensor:
    """H_GN v = 4 s K^T(d K(s v)) + 4 s H^T(w H(s v)) [+ 4 iw s (w s v)] (erdecon_mlx.py:311-345)."""
    s, v = as_f32(s), as_f32(v)
    shape = tuple(s.shape)
    w = as_f32(w).reshape(shape)
    sv = _ew(N.ER_EW_MUL, s, empty_f32(shape), b=v)
    ksv = _conv(blur_op, sv, False)
    if isinstance(data_hess_w, (int, float)):
        t = _ew(N.ER_EW_SCALE, ksv, empty_f32(ksv.shape), alpha=float(data_hess_w))
    else:
        t = _ew(N.ER_EW_MUL, ksv, empty_f32(ksv.shape), b=as_f32(data_hess_w))
    kt = _conv(blur_op, t, True)
    reg = _Reg(hessian, shape)
    htu = reg.adjoint(reg.apply_w(sv, w))
    out = empty_f32(shape)
    S = _Scalars()
    N.call("dh_er_cg_ap", s.data_ptr(), kt.data_ptr(), htu.data_ptr(), w.data_ptr(), sv.data_ptr(),
           float(intensity_weight), v.data_ptr(), out.data_ptr(), s.numel(), S.sc.data_ptr(), S.part.data_ptr(),
           N.stream_ptr())
    return out
