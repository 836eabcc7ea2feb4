"""Accelerated Poisson ML deconvolution by nonlinear conjugate gradients on MI355X
(drop-in for deconlib's nlcg_mlx: nlcg_with_operator, nlcg_solver, NLCGResult
and the helper functions).

Semantics follow nlcg_mlx.py:317-541 step for step (Schaefer et al. 2001 with
the exact Newton-Raphson step length of COSM): f = s^2, m = K f + b, the
gradient 2 s K^T(1 - g/m) + 4 beta s C^T C f, Fletcher-Reeves directions with
a steepest-descent fallback, and the discrepancy-principle / Eq. 17 stops.

An iteration is three convolutions (K(s d), K(d d), K^T w) and the library's
NLCG kernels; the Newton refinements of lambda, the accept test and gamma
run on the device (float64 scalar block), so an iteration enqueues without a
host round trip until the host reads the accept flag.  The stopping tests of
iteration k are read together with the accept flag of iteration k+1 (whose
line search, which changes no iterate, is dropped when k stops), so an
iteration costs one host synchronisation.  With a callback every iteration
is settled before the next begins.

Sums are float64 (the reference sums in float32), so decisions that hinge on
a float32 rounding of a sum may differ.  With a regulariser, the trial
objective's ||C f(lambda)||^2 and the Newton terms come from six dot products
of (Cf, C(s d), C(d d)): C f(lambda) is exactly Cf + 2 lambda C(s d) +
lambda^2 C(d d), so no fourth regulariser application is needed.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Callable, Optional, Tuple

import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32
from .rl import _plan_of

_EPS = 1e-10


@dataclass
class NLCGResult:
    """Result of NLCG (nlcg_mlx.py:61-85); arrays are device tensors."""

    restored: torch.Tensor
    pred: torch.Tensor
    iterations: int
    loss_history: list
    converged: bool = False
    background: float = 0.0
    full_shape: Optional[Tuple[int, ...]] = None
    valid_slices: Optional[Tuple[slice, ...]] = None


def _conv(op, x, adjoint: bool, out=None, fuse=None) -> torch.Tensor:
    """One application of the blur operator (plan-backed: into ``out``, fused)."""
    plan, shape = _plan_of(op)
    if plan is not None:
        return plan.apply(x, adjoint, out if out is not None else empty_f32(shape), fuse)
    if fuse is not None:
        raise ValueError("fused epilogues need a plan-backed operator")
    return as_f32(op.adjoint(x) if adjoint else op.forward(x))


def _reg_apply(reg, x, adjoint=False) -> torch.Tensor:
    return as_f32(reg.adjoint(x) if adjoint else reg.forward(x))


class _Scalars:
    """The float64 scalar block, its partials, and a pinned host mirror."""

    def __init__(self):
        dev = device()
        self.sc = torch.zeros(N.NL_NSLOTS, dtype=torch.float64, device=dev)
        self.part = torch.empty(int(N.fns["dh_nlcg_partials"]()), dtype=torch.float64, device=dev)
        self.host = torch.empty(N.NL_NSLOTS, dtype=torch.float64, pin_memory=True)

    def read(self):
        self.host.copy_(self.sc, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        return self.host.tolist()


class _NLCGState:
    """Device-resident state of one NLCG solve."""

    def __init__(self, observed, blur_op, background, regularizer, reg_weight, init, newton_iters):
        self.op = blur_op
        self.g = as_f32(observed)
        self.bg = float(background)
        self.beta = float(reg_weight)
        self.reg = regularizer if (regularizer is not None and self.beta != 0.0) else None
        self.newton_iters = int(newton_iters)
        self.plan, _ = _plan_of(blur_op)
        self.S = _Scalars()
        self.n_data = self.g.numel()
        st = N.stream_ptr()
        # f0 = max(K^T max(g - b, 1e-6), 1e-6) or max(init, 1e-6); s = sqrt(f0), f = s^2
        if init is None:
            if self.plan is not None:
                fz = N.FuseArgs(prologue=N.PRO_SUB_FLOOR, pro_a=self.bg, pro_b=1e-6)
                f0 = _conv(blur_op, self.g, True, fuse=fz)
            else:
                dmb = empty_f32(self.g.shape)
                N.call("dh_sub_floor", self.g.data_ptr(), dmb.data_ptr(), self.n_data, self.bg, 1e-6, st)
                f0 = _conv(blur_op, dmb, True)
        else:
            f0 = as_f32(init)
        self.n = f0.numel()
        self.s = empty_f32(f0.shape)
        self.f = empty_f32(f0.shape)
        N.call("dh_nlcg_init", f0.data_ptr(), self.s.data_ptr(), self.f.data_ptr(), self.n, 1e-6, st)
        dshape, xshape = self.g.shape, self.s.shape
        self.m = self._forward_plus_b(self.f)
        self.m_new = empty_f32(dshape)
        self.w = empty_f32(dshape)
        N.call("dh_nlcg_model", self.m.data_ptr(), self.g.data_ptr(), self.w.data_ptr(), self.n_data, _EPS,
               self.S.sc.data_ptr(), self.S.part.data_ptr(), st)
        self.sd, self.dd = empty_f32(xshape), empty_f32(xshape)
        self.r, self.d = empty_f32(xshape), empty_f32(xshape)
        if self.plan is not None:
            self.ksd, self.kdd, self.ktw = empty_f32(dshape), empty_f32(dshape), empty_f32(xshape)
        else:
            self.ksd = self.kdd = self.ktw = None
        self.cf = self.ctcf = None
        self._gradient(mode=0)
        self.d.copy_(self.r)

    def _forward_plus_b(self, x) -> torch.Tensor:
        if self.plan is not None:
            return _conv(self.op, x, False, empty_f32(self.g.shape), N.FuseArgs(epilogue=N.EPI_ADD, epi_a=self.bg))
        y = _conv(self.op, x, False)
        out = empty_f32(y.shape)
        N.call("dh_add_scalar", y.data_ptr(), None, out.data_ptr(), y.numel(), self.bg, None, None, N.stream_ptr())
        return out

    def _gradient(self, mode: int) -> None:
        """r = -grad at the current (s, f, w), sum r^2 and gamma (nlcg_mlx.py:104-113, 472-479)."""
        st = N.stream_ptr()
        if self.reg is not None:
            self.cf = _reg_apply(self.reg, self.f)
            self.ctcf = _reg_apply(self.reg, self.cf, adjoint=True)
        ktw = _conv(self.op, self.w, True, self.ktw)
        N.call("dh_nlcg_gradient", self.s.data_ptr(), ktw.data_ptr(), N.ptr(self.ctcf), 4.0 * self.beta,
               self.r.data_ptr(), self.n, mode, _EPS, self.S.sc.data_ptr(), self.S.part.data_ptr(), st)

    def line_search(self) -> None:
        """Exact step length along d and the trial model (nlcg_mlx.py:245-291, 428-454)."""
        st = N.stream_ptr()
        sc, part = self.S.sc.data_ptr(), self.S.part.data_ptr()
        N.call("dh_nlcg_begin_step", self.s.data_ptr(), self.d.data_ptr(), self.sd.data_ptr(), self.dd.data_ptr(),
               self.n, sc, self.newton_iters, st)
        ksd = _conv(self.op, self.sd, False, self.ksd)
        kdd = _conv(self.op, self.dd, False, self.kdd)
        if self.reg is not None:
            csd = _reg_apply(self.reg, self.sd)
            cdd = _reg_apply(self.reg, self.dd)
            N.call("dh_nlcg_reg_dots", self.cf.data_ptr(), csd.data_ptr(), cdd.data_ptr(), csd.numel(), sc, part, st)
        args = (self.m.data_ptr(), ksd.data_ptr(), kdd.data_ptr(), self.g.data_ptr(), self.n_data, self.bg,
                self.beta if self.reg is not None else 0.0, _EPS, sc)
        for _ in range(self.newton_iters):
            N.call("dh_nlcg_newton", *args, part, st)
        N.call("dh_nlcg_trial", *args, self.m_new.data_ptr(), self.w.data_ptr(), part, st)

    def accept(self, restart: bool) -> None:
        """s += lambda d, m = trial model, new gradient and direction (nlcg_mlx.py:470-481)."""
        st = N.stream_ptr()
        sc, part = self.S.sc.data_ptr(), self.S.part.data_ptr()
        N.call("dh_nlcg_update", self.s.data_ptr(), self.d.data_ptr(), self.f.data_ptr(), self.n, _EPS, sc, part, st)
        self.m, self.m_new = self.m_new, self.m
        self._gradient(mode=2 if restart else 1)
        N.call("dh_nlcg_direction", self.r.data_ptr(), self.d.data_ptr(), self.n, sc, st)

    def fall_back_to_steepest_descent(self) -> None:
        self.d.copy_(self.r)

    def pred(self) -> torch.Tensor:
        return self._forward_plus_b(self.f)


def nlcg_with_operator(
    observed,
    blur_op,
    num_iter: int = 50,
    background: float = 0.0,
    regularizer=None,
    reg_weight: float = 0.0,
    init=None,
    callback: Optional[Callable[[int, torch.Tensor], bool]] = None,
    eval_interval: int = 10,
    slack: float = 1.0,
    tol: float = 1e-4,
    min_iter: int = 10,
    restart_interval: Optional[int] = None,
    newton_iters: int = 3,
    verbose: bool = False,
) -> NLCGResult:
    """Poisson-ML NLCG with a pre-built positive operator (nlcg_mlx.py:317-541)."""
    st = _NLCGState(observed, blur_op, background, regularizer, reg_weight, init, newton_iters)
    use_reg = st.reg is not None
    n_data = st.n_data
    loss_history: list = []
    window: list = []
    converged = False
    d_is_r = True  # the reference's `d is r`
    defer = callback is None
    pending = None

    def settle(kp: int, v) -> bool:
        """End-of-iteration bookkeeping of iteration kp; True stops (nlcg_mlx.py:485-523)."""
        nonlocal converged
        loss = v[N.NL_LOSS] / n_data
        logged = kp % eval_interval == 0 or kp == num_iter - 1
        if logged:
            loss_history.append(loss)
            if verbose:
                print(f"  Iter {kp:4d}: mean I-div = {loss:.6g}, lambda = {v[N.NL_LAM_ACCEPTED]:.4g}, "
                      f"gamma = {v[N.NL_GAMMA]:.4g}")
        window.append(v[N.NL_IDIV] / max(v[N.NL_SUMF], _EPS))
        if len(window) > 5:
            window.pop(0)
        rel_avg = sum(window) / len(window)
        # a copy: the next iteration updates st.f in place
        stop = bool(callback(kp, st.f.clone())) if callback is not None else False
        reason = None
        ready = not stop and kp + 1 >= min_iter
        if ready and not use_reg and slack > 0.0 and loss <= 0.5 * slack:
            converged = stop = True
            reason = "discrepancy principle"
        elif ready and tol > 0.0 and len(window) == 5 and rel_avg < tol:
            converged = stop = True
            reason = "relative divergence change (Eq. 17)"
        if stop:
            if not logged:
                loss_history.append(loss)
            if verbose and reason is not None:
                print(f"  Stopped at iter {kp}: {reason}")
        return stop

    k = 0
    stopped = False
    for k in range(num_iter):
        st.line_search()
        v = st.S.read()
        if pending is not None:
            kp, pending = pending, None
            if settle(kp, v):
                k, stopped = kp, True  # the line search just enqueued is dropped
                break
        if v[N.NL_ACCEPT] == 0.0 and not d_is_r:
            st.fall_back_to_steepest_descent()
            d_is_r = True
            st.line_search()
            v = st.S.read()
        if v[N.NL_ACCEPT] == 0.0:
            converged = stopped = True
            loss_history.append(v[N.NL_LOSS] / n_data)
            break
        st.accept(restart=restart_interval is not None and (k + 1) % restart_interval == 0)
        d_is_r = False
        if defer:
            pending = k
        elif settle(k, st.S.read()):
            stopped = True
            break
    if not stopped and pending is not None:
        settle(pending, st.S.read())

    pred = st.pred()
    return NLCGResult(
        restored=st.f,
        pred=pred,
        iterations=k + 1,
        loss_history=loss_history,
        converged=converged,
        background=float(background),
        full_shape=tuple(st.f.shape),
        valid_slices=None,
    )


def nlcg_solver(num_iter: int = 50, background: float = 0.0, regularizer=None, reg_weight: float = 0.0,
                init_value: Optional[float] = None, **nlcg_kwargs: Any) -> Callable:
    """NLCG adapted to the ``solve(data, model) -> visible array`` contract (nlcg_mlx.py:544-588)."""

    def solve(data, model):
        init = None
        if init_value is not None:
            init = torch.full(model.padded_shape, float(init_value), dtype=torch.float32, device=device())
        result = nlcg_with_operator(observed=data, blur_op=model.op, num_iter=num_iter, background=background,
                                    regularizer=regularizer, reg_weight=reg_weight, init=init, **nlcg_kwargs)
        return result.restored[model.valid_slices]

    solve.accepts_device_tiles = True
    return solve


# ---------------------------------------------------------------------------
# Helper functions of nlcg_mlx (objective, gradient, quadratic form, step length)
# ---------------------------------------------------------------------------


def _square(s: torch.Tensor, d: Optional[torch.Tensor] = None):
    """(s * d, d * d) (d = s: (s^2, s^2)) through the line-search set-up kernel."""
    S = _Scalars()
    d = s if d is None else d
    sd, dd = empty_f32(s.shape), empty_f32(s.shape)
    N.call("dh_nlcg_begin_step", s.data_ptr(), d.data_ptr(), sd.data_ptr(), dd.data_ptr(), s.numel(),
           S.sc.data_ptr(), 0, N.stream_ptr())
    return sd, dd


def _model_sums(m, g, eps, w=None):
    S = _Scalars()
    N.call("dh_nlcg_model", m.data_ptr(), g.data_ptr(), N.ptr(w), m.numel(), eps, S.sc.data_ptr(),
           S.part.data_ptr(), N.stream_ptr())
    return S.read()


def _sumsq(x: torch.Tensor) -> float:
    S = _Scalars()
    N.call("dh_nlcg_reg_dots", x.data_ptr(), x.data_ptr(), x.data_ptr(), x.numel(), S.sc.data_ptr(),
           S.part.data_ptr(), N.stream_ptr())
    return S.read()[N.NL_A0]


def _quad_sums(g, m, ksd, v, d, eps):
    S = _Scalars()
    out = S.sc[:2]
    N.call("dh_nlcg_quad", N.ptr(g), N.ptr(m), N.ptr(ksd), v.data_ptr(), d.data_ptr(), d.numel(), eps,
           out.data_ptr(), S.part.data_ptr(), N.stream_ptr())
    vals = S.read()
    return vals[0], vals[1]


def _forward_model(s, blur_op, background):
    f, _ = _square(s)
    y = _conv(blur_op, f, False)
    m = empty_f32(y.shape)
    N.call("dh_add_scalar", y.data_ptr(), None, m.data_ptr(), y.numel(), float(background), None, None,
           N.stream_ptr())
    return f, m


def nlcg_objective(s, blur_op, observed, background: float = 0.0, regularizer=None, reg_weight: float = 0.0,
                   eps: float = 1e-10) -> float:
    """phi(s) = sum Kf - sum g ln(Kf + b) + beta ||C f||^2 (nlcg_mlx.py:116-131)."""
    s, g = as_f32(s), as_f32(observed)
    f, m = _forward_model(s, blur_op, background)
    val = _model_sums(m, g, eps)[N.NL_MDATA]
    if regularizer is not None and reg_weight != 0.0:
        val += float(reg_weight) * _sumsq(_reg_apply(regularizer, f))
    return val


def nlcg_gradient(s, blur_op, observed, background: float = 0.0, regularizer=None, reg_weight: float = 0.0,
                  eps: float = 1e-10):
    """(grad, (m, K^T w, C^T C f)) at s (nlcg_mlx.py:134-152)."""
    s, g = as_f32(s), as_f32(observed)
    f, m = _forward_model(s, blur_op, background)
    w = empty_f32(m.shape)
    _model_sums(m, g, eps, w)
    ktw = _conv(blur_op, w, True)
    ctcf = None
    use_reg = regularizer is not None and reg_weight != 0.0
    if use_reg:
        ctcf = _reg_apply(regularizer, _reg_apply(regularizer, f), adjoint=True)
    S = _Scalars()
    r = empty_f32(s.shape)
    N.call("dh_nlcg_gradient", s.data_ptr(), ktw.data_ptr(), N.ptr(ctcf), 4.0 * float(reg_weight), r.data_ptr(),
           s.numel(), 0, eps, S.sc.data_ptr(), S.part.data_ptr(), N.stream_ptr())
    grad = empty_f32(s.shape)
    N.call("dh_er_ew", N.ER_EW_NEG, r.data_ptr(), None, 0.0, 0.0, grad.data_ptr(), None, r.numel(), N.stream_ptr())
    return grad, (m, ktw, ctcf)


def nlcg_hessian_quadform(s, d, blur_op, observed, background: float = 0.0, regularizer=None,
                          reg_weight: float = 0.0, aux=None, eps: float = 1e-10) -> float:
    """<d, A(s) d> with one extra forward convolution (nlcg_mlx.py:155-192)."""
    s, d, g = as_f32(s), as_f32(d), as_f32(observed)
    if aux is None:
        _, aux = nlcg_gradient(s, blur_op, g, background, regularizer, reg_weight, eps)
    m, ktw, ctcf = aux
    sd, _ = _square(s, d)
    ksd = _conv(blur_op, sd, False)
    q0, q1 = _quad_sums(g, m, ksd, ktw, d, eps)
    quad = 4.0 * q0 + 2.0 * q1
    if regularizer is not None and reg_weight != 0.0:
        csd2 = _sumsq(_reg_apply(regularizer, sd))
        _, c1 = _quad_sums(None, None, None, ctcf, d, eps)
        quad += 8.0 * float(reg_weight) * csd2 + 4.0 * float(reg_weight) * c1
    return quad


def nlcg_step_length(s, d, f, m, blur_op, observed, background: float = 0.0, regularizer=None,
                     reg_weight: float = 0.0, newton_iters: int = 3, eps: float = 1e-10):
    """(lambda, K f(lambda) + b) by Newton-Raphson on the exact quadratic model (nlcg_mlx.py:195-291)."""
    s, d, f, m, g = as_f32(s), as_f32(d), as_f32(f), as_f32(m), as_f32(observed)
    S = _Scalars()
    st = N.stream_ptr()
    sc, part = S.sc.data_ptr(), S.part.data_ptr()
    sd, dd = empty_f32(s.shape), empty_f32(s.shape)
    N.call("dh_nlcg_begin_step", s.data_ptr(), d.data_ptr(), sd.data_ptr(), dd.data_ptr(), s.numel(), sc,
           int(newton_iters), st)
    ksd = _conv(blur_op, sd, False)
    kdd = _conv(blur_op, dd, False)
    use_reg = regularizer is not None and reg_weight != 0.0
    if use_reg:
        cf, csd, cdd = (_reg_apply(regularizer, a) for a in (f, sd, dd))
        N.call("dh_nlcg_reg_dots", cf.data_ptr(), csd.data_ptr(), cdd.data_ptr(), csd.numel(), sc, part, st)
    n = m.numel()
    args = (m.data_ptr(), ksd.data_ptr(), kdd.data_ptr(), g.data_ptr(), n, float(background),
            float(reg_weight) if use_reg else 0.0, eps, sc)
    for _ in range(int(newton_iters)):
        N.call("dh_nlcg_newton", *args, part, st)
    m_new, w = empty_f32(m.shape), empty_f32(m.shape)
    N.call("dh_nlcg_trial", *args, m_new.data_ptr(), w.data_ptr(), part, st)
    lam = S.read()[N.NL_LAM]
    if not math.isfinite(lam):
        m_new.fill_(float("nan"))
    return lam, m_new


def i_divergence_between(f_prev, f_curr, eps: float = 1e-10) -> float:
    """sum f_k ln(f_k / f_{k+1}) - f_k + f_{k+1} (Eq. 17, nlcg_mlx.py:294-314)."""
    a, b = as_f32(f_prev), as_f32(f_curr)
    S = _Scalars()
    out = S.sc[:2]
    N.call("dh_nlcg_idiv", a.data_ptr(), b.data_ptr(), a.numel(), eps, out.data_ptr(), S.part.data_ptr(),
           N.stream_ptr())
    return S.read()[0]
