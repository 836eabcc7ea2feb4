"""Non-dimensional log-penalty deconvolution by active-set projected Newton on
MI355X (drop-in for deconlib's jetnewton_mlx).

Semantics follow jetnewton_mlx.py:547-785 step for step: x = s0 x_tilde, the
data term in physical units with its chain-rule scale factors, the penalty
(beta/2) sum log(eta + u) of u = |H~ x_tilde|^2, the exact (indefinite)
Hessian product, the Bertsekas active set with persistent freezing, the masked
reduced CG with the absolute forcing tolerance min(0.5, sqrt(r0)) r0 and its
first-step negative-curvature return, the two-metric direction, the projected
Armijo search, and the Newton-decrement / projected-gradient stops.

Per outer iteration: the gradient (one adjoint convolution; the model of the
current iterate is the accepted trial's), each CG step (two convolutions, the
product's regulariser operand w q + 2 c <r, q> r formed from the stencil on the
fly and its adjoint), each line-search trial (one convolution and one stencil
pass that only sums).  ``AnisotropicHessian3D/2D``, ``Hessian3D/2D`` go
through the fused stencil kernels; any other regulariser through its own
forward/adjoint and the stacked-channel kernels.  The CG scalars live on the
device; the host reads them once per CG step and line-search trial, as the
reference does.  Sums are float64 (the reference's are float32).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Tuple

import numpy as np
import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32
from .erdecon import _check_data_term, _ew
from .linops import Hessian2D, Hessian3D
from .nlcg import _conv
from .rl import _plan_of

_CG_EPS = 1e-20


def _data_scale_factors(data_term: str, s0: float) -> Tuple[float, float, float]:
    """(objective, gradient, Hessian) scale of the data term (jetnewton_mlx.py:83-108)."""
    if data_term == "poisson":
        return 1.0, s0, s0 * s0
    return 1.0 / (2.0 * s0 * s0), 1.0 / (2.0 * s0), 0.5


class _Blocks:
    """A JetNewton scalar block, an ER-layout block for the data sums, partials, host mirrors."""

    def __init__(self):
        dev = device()
        self.jn = torch.zeros(N.JN_NSLOTS, dtype=torch.float64, device=dev)
        self.er = torch.zeros(N.ER_NSLOTS, dtype=torch.float64, device=dev)
        n = max(int(N.fns["dh_jn_partials"]()), int(N.fns["dh_er_partials"]()))
        self.part = torch.empty(n, dtype=torch.float64, device=dev)
        self.host = torch.empty(N.JN_NSLOTS + N.ER_NSLOTS, dtype=torch.float64, pin_memory=True)

    def read(self):
        """(jn slots, er slots) after the stream drains."""
        self.host[:N.JN_NSLOTS].copy_(self.jn, non_blocking=True)
        self.host[N.JN_NSLOTS:].copy_(self.er, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        v = self.host.tolist()
        return v[:N.JN_NSLOTS], v[N.JN_NSLOTS:]


class _Penalty:
    """r = H~ x and its weights; fused stencil kernels for the Hessian family."""

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
        self.nc = 6 if self.dims is not None else None
        self.r = self.wr = self.w = self.c = None
        self._q = None

    def _w6(self):
        return self.h._w.ctypes.data_as(C.POINTER(C.c_float))

    def _ensure(self, nc):
        if self.r is None or self.nc != nc:
            self.nc = nc
            self.r = empty_f32((nc, self.n))
            self.wr = empty_f32((nc, self.n))
            self.w, self.c = empty_f32(self.shape), empty_f32(self.shape)

    def weights(self, x, beta, eta, B, full: bool, u_out=None) -> None:
        """LOGU, USUM of x; with ``full`` also r, w, c and w r."""
        st = N.stream_ptr()
        if self.dims is not None:
            if full:
                self._ensure(6)
            N.call("dh_jn_penalty", x.data_ptr(), *self.dims, self._w6(), float(beta), float(eta),
                   N.ptr(self.r) if full else None, N.ptr(self.w) if full else None,
                   N.ptr(self.c) if full else None, N.ptr(self.wr) if full else None, N.ptr(u_out),
                   B.jn.data_ptr(), B.part.data_ptr(), st)
            return
        r = as_f32(self.h.forward(x))
        if r.numel() % self.n:
            raise ValueError("the regulariser's forward must return stacked channels of the domain's shape")
        nc = r.numel() // self.n
        if full:
            self._ensure(nc)
            self.r.copy_(r.reshape(nc, self.n))
            r = self.r
        N.call("dh_jn_penalty_stacked", r.data_ptr(), nc, self.n, float(beta), float(eta),
               N.ptr(self.w) if full else None, N.ptr(self.c) if full else None, N.ptr(self.wr) if full else None,
               N.ptr(u_out), B.jn.data_ptr(), B.part.data_ptr(), st)

    def adjoint(self, u, out=None) -> torch.Tensor:
        if self.dims is not None:
            out = out if out is not None else empty_f32(self.shape)
            N.call("dh_hessian_adjoint", u.data_ptr(), out.data_ptr(), *self.dims, self._w6(), N.stream_ptr())
            return out
        return as_f32(self.h.adjoint(u.reshape((self.nc,) + self.shape)))

    def hvp_operand(self, v) -> torch.Tensor:
        """w q + 2 c <r, q> r with q = H~ v."""
        if self.dims is not None:
            if self._q is None:
                self._q = empty_f32((6, self.n))
            N.call("dh_jn_hvp_operand", v.data_ptr(), self.r.data_ptr(), self.w.data_ptr(), self.c.data_ptr(),
                   *self.dims, self._w6(), self._q.data_ptr(), N.stream_ptr())
            return self._q
        q = as_f32(self.h.forward(v)).reshape(self.nc, self.n)
        N.call("dh_jn_hvp_operand_stacked", q.data_ptr(), self.r.data_ptr(), self.w.data_ptr(), self.c.data_ptr(),
               self.nc, self.n, q.data_ptr(), N.stream_ptr())
        return q


def _has_plan(op) -> bool:
    return _plan_of(op)[0] is not None


class _JNState:
    """Device-resident state of one JetNewton solve (the dimensionless iterate x)."""

    def __init__(self, f, op, hessian, s0, b, beta, eta, data_term, init):
        self.op, self.f = op, f
        self.s0, self.b, self.beta, self.eta = float(s0), float(b), float(beta), float(eta)
        self.poisson = 1 if data_term == "poisson" else 0
        self.obj_scale, self.grad_scale, self.hvp_scale = _data_scale_factors(data_term, self.s0)
        self.B = _Blocks()
        st = N.stream_ptr()
        if init is None:
            if _has_plan(op):
                g0 = _conv(op, f, True, fuse=N.FuseArgs(prologue=N.PRO_SUB_FLOOR, pro_a=self.b, pro_b=self.s0))
            else:
                dmb = empty_f32(f.shape)
                N.call("dh_sub_floor", f.data_ptr(), dmb.data_ptr(), f.numel(), self.b, self.s0, st)
                g0 = _conv(op, dmb, True)
        else:
            g0 = as_f32(init)
        shape = tuple(g0.shape)
        self.x = _ew(N.ER_EW_MAX_DIV, g0, empty_f32(shape), alpha=self.s0)
        self.n, self.nd = self.x.numel(), f.numel()
        self.pen = _Penalty(hessian, shape)
        dev = device()
        self.frozen = torch.zeros(shape, dtype=torch.bool, device=dev)
        self.active = torch.zeros(shape, dtype=torch.bool, device=dev)
        self.grad, self.d, self.x_t = empty_f32(shape), empty_f32(shape), empty_f32(shape)
        self.xcg, self.rcg, self.pcg, self.ap = empty_f32(shape), empty_f32(shape), empty_f32(shape), empty_f32(shape)
        self.kt, self.rh, self.htu = empty_f32(shape), empty_f32(shape), empty_f32(shape)
        self.score = empty_f32(f.shape)
        self.hessw = empty_f32(f.shape) if self.poisson else None
        self.av, self.tvec = empty_f32(f.shape), empty_f32(f.shape)
        self.mu = self.model(self.x)

    def model(self, x, out=None) -> torch.Tensor:
        """mu = s0 K x + b."""
        y = _conv(self.op, x, False, empty_f32(self.f.shape) if _has_plan(self.op) else None)
        return _ew(N.ER_EW_AFFINE, y, y, alpha=self.s0, beta=self.b)

    def data_sums(self, mu, score=None) -> None:
        N.call("dh_er_data", mu.data_ptr(), self.f.data_ptr(), self.nd, self.poisson, N.ptr(score),
               N.ptr(self.hessw) if score is not None else None, self.B.er.data_ptr(), self.B.part.data_ptr(),
               N.stream_ptr())

    def objective(self, jn, er) -> float:
        return self.obj_scale * er[N.ER_DATA] + 0.5 * self.beta * jn[N.JN_LOGU]

    def gradient(self) -> None:
        """grad at x from mu (jetnewton_mlx.py:215-225), then the active set and CG start."""
        self.data_sums(self.mu, self.score)
        kt = _conv(self.op, self.score, True, self.kt if _has_plan(self.op) else None)
        self.pen.weights(self.x, self.beta, self.eta, self.B, full=True)
        rh = self.pen.adjoint(self.pen.wr, self.rh)
        _ew(N.ER_EW_AXPY, kt, self.grad, alpha=self.grad_scale, b=rh)

    def active_set(self, eps_bar, tau, delta) -> None:
        B, st = self.B, N.stream_ptr()
        N.call("dh_jn_kkt", self.x.data_ptr(), self.grad.data_ptr(), self.n, float(eps_bar), B.jn.data_ptr(),
               B.part.data_ptr(), st)
        N.call("dh_jn_active", self.x.data_ptr(), self.grad.data_ptr(), self.n, float(tau), float(delta),
               self.frozen.data_ptr(), self.active.data_ptr(), self.rcg.data_ptr(), self.pcg.data_ptr(),
               self.xcg.data_ptr(), _CG_EPS, B.jn.data_ptr(), B.part.data_ptr(), st)

    def hvp_apply(self) -> None:
        """ap = where(free, hvp_scale K^T(d K p) + H~^T(w q + 2 c <r, q> r), 0), DENOM."""
        plan = _has_plan(self.op)
        av = _conv(self.op, self.pcg, False, self.av if plan else None)
        if self.poisson:
            t = _ew(N.ER_EW_MUL, av, self.tvec, b=self.hessw)
        else:
            t = _ew(N.ER_EW_SCALE, av, self.tvec, alpha=2.0)
        kt = _conv(self.op, t, True, self.kt if plan else None)
        htu = self.pen.adjoint(self.pen.hvp_operand(self.pcg), self.htu)
        N.call("dh_jn_cg_ap", kt.data_ptr(), self.hvp_scale, htu.data_ptr(), self.active.data_ptr(),
               self.pcg.data_ptr(), self.ap.data_ptr(), self.n, _CG_EPS, self.B.jn.data_ptr(),
               self.B.part.data_ptr(), N.stream_ptr())


def _cg_steps(jn) -> int:
    """The reduced solve's reported step count (jetnewton_mlx.py:411-430)."""
    ncg = int(jn[N.JN_NCG])
    if jn[N.JN_HITNEG] != 0.0:
        return 1 if ncg == 1 else ncg - 1
    return ncg


def _reduced_cg(apply_hvp: Callable[[], None], xcg, rcg, pcg, ap, B, max_steps: int, stopped: bool):
    """Run the masked CG on the device; apply_hvp() forms ap (and DENOM) from pcg."""
    jn = None
    if not stopped:
        for _ in range(max_steps):
            apply_hvp()
            N.call("dh_jn_cg_update", xcg.data_ptr(), rcg.data_ptr(), pcg.data_ptr(), ap.data_ptr(), xcg.numel(),
                   B.jn.data_ptr(), B.part.data_ptr(), N.stream_ptr())
            jn, _ = B.read()
            if jn[N.JN_STOP] != 0.0:
                break
    if jn is None:
        jn, _ = B.read()
    return jn


@dataclass
class JetNewtonResult:
    """Result of jetnewton_with_operator (jetnewton_mlx.py:471-544); arrays are device tensors."""

    restored: torch.Tensor
    pred: torch.Tensor
    iterations: int
    active_set_size_history: list = field(default_factory=list)
    cg_iterations_history: list = field(default_factory=list)
    loss_history: list = field(default_factory=list)
    data_misfit_history: list = field(default_factory=list)
    idiv_history: list = field(default_factory=list)
    curvature_term_history: list = field(default_factory=list)
    proj_grad_history: list = field(default_factory=list)
    newton_decrement_history: list = field(default_factory=list)
    converged: bool = False
    s0: float = 1.0


def jetnewton_with_operator(
    observed,
    blur_op,
    hessian,
    s0: float,
    background: float = 0.0,
    beta: float = 1.0,
    eta: float = 1.0,
    data_term: str = "poisson",
    num_iter: int = 100,
    init=None,
    cg_max_steps: int = 50,
    eps_bar: float = 1e-2,
    freeze_tau: float = 1e-3,
    freeze_delta: float = 1e-6,
    newton_tol: float = 1e-3,
    tol: float = 0.0,
    min_iter: int = 3,
    ls_sigma: float = 1e-4,
    ls_max_backtracks: int = 30,
    callback: Optional[Callable[[int, torch.Tensor], Optional[bool]]] = None,
    eval_interval: int = 5,
    verbose: bool = False,
) -> JetNewtonResult:
    """Single-run active-set projected Newton solve (jetnewton_mlx.py:547-785)."""
    _check_data_term(data_term)
    f = as_f32(observed)
    st = _JNState(f, blur_op, hessian, s0, background, beta, eta, data_term, init)
    res = JetNewtonResult(restored=None, pred=None, iterations=0, s0=float(s0))
    pg0 = nd0 = None
    converged = False
    k = 0
    for k in range(num_iter):
        st.gradient()
        st.active_set(eps_bar, freeze_tau, freeze_delta)
        jn0, er0 = st.B.read()
        F0 = st.objective(jn0, er0)
        jn = _reduced_cg(st.hvp_apply, st.xcg, st.rcg, st.pcg, st.ap, st.B, cg_max_steps, jn0[N.JN_STOP] != 0.0)
        n_cg = _cg_steps(jn)
        res.cg_iterations_history.append(n_cg)
        N.call("dh_jn_direction", st.xcg.data_ptr(), st.grad.data_ptr(), st.active.data_ptr(), st.d.data_ptr(),
               st.n, st.B.jn.data_ptr(), st.B.part.data_ptr(), N.stream_ptr())
        jn, _ = st.B.read()
        nd = abs(jn[N.JN_ND])
        if nd0 is None:
            nd0 = nd
        # projected Armijo along x(t) = max(x + t d, 0) (jetnewton_mlx.py:438-468)
        t = 1.0
        accepted = False
        mu_t = None
        for _ in range(ls_max_backtracks):
            N.call("dh_jn_trial", st.x.data_ptr(), st.d.data_ptr(), t, st.grad.data_ptr(), st.x_t.data_ptr(), st.n,
                   st.B.jn.data_ptr(), st.B.part.data_ptr(), N.stream_ptr())
            mu_t = st.model(st.x_t)
            st.data_sums(mu_t)
            st.pen.weights(st.x_t, st.beta, st.eta, st.B, full=False)
            jt, et = st.B.read()
            Ft = st.objective(jt, et)
            if math.isfinite(Ft) and Ft <= F0 + ls_sigma * jt[N.JN_DIRECTIONAL]:
                accepted = True
                break
            t *= 0.5
        if not accepted:
            t = 0.0
        pg = jn0[N.JN_PG]
        if pg0 is None:
            pg0 = pg
        logged = k % eval_interval == 0 or k == num_iter - 1 or not accepted
        if logged:
            res.loss_history.append(F0)
            res.data_misfit_history.append(er0[N.ER_DATA])
            res.idiv_history.append(er0[N.ER_IDIV] / st.nd)
            res.curvature_term_history.append(jn0[N.JN_USUM] / st.n)
            res.active_set_size_history.append(int(jn0[N.JN_ACTIVE]))
            res.proj_grad_history.append(pg)
            res.newton_decrement_history.append(nd)
            if verbose:
                print(f"  Iter {k:4d}: F = {F0:.6g}, I-div = {res.idiv_history[-1]:.4g}, nd = {nd:.3g}, "
                      f"|proj-grad| = {pg:.3g}, cg_iter = {n_cg}, |I| = {res.active_set_size_history[-1]}, "
                      f"t = {t:.3g}, u = {res.curvature_term_history[-1]:.3g}")
        if not accepted:
            converged = True
            if verbose:
                print(f"  Stopped at iter {k}: line search found no decrease")
            break
        st.x, st.x_t = st.x_t, st.x
        st.mu = mu_t
        stop = False
        if callback is not None:
            stop = bool(callback(k, _ew(N.ER_EW_SCALE, st.x, empty_f32(st.x.shape), alpha=st.s0)))
        if not stop and newton_tol > 0.0 and k + 1 >= min_iter and nd0 is not None and nd0 > 0.0 \
                and nd < newton_tol * nd0:
            converged = stop = True
            if verbose:
                print(f"  Stopped at iter {k}: Newton decrement {nd:.3g} < {newton_tol:.1g} * {nd0:.3g}")
        if not stop and tol > 0.0 and k + 1 >= min_iter and pg0 is not None and pg0 > 0.0 and pg < tol * pg0:
            converged = stop = True
            if verbose:
                print(f"  Stopped at iter {k}: |proj-grad| {pg:.3g} < {tol:.1g} * {pg0:.3g}")
        if stop:
            break
    res.restored = _ew(N.ER_EW_SCALE, st.x, empty_f32(st.x.shape), alpha=st.s0)
    res.pred = st.model(st.x)
    res.iterations = k + 1
    res.converged = converged
    return res


# ---------------------------------------------------------------------------
# The reference's building blocks (jetnewton_mlx.py:111-468)
# ---------------------------------------------------------------------------


def _helper(x_tilde, blur_op, observed, hessian, s0, background, beta, eta, data_term):
    _check_data_term(data_term)
    x = as_f32(x_tilde)
    st = _JNState.__new__(_JNState)
    st.op, st.f = blur_op, as_f32(observed)
    st.s0, st.b, st.beta, st.eta = float(s0), float(background), float(beta), float(eta)
    st.poisson = 1 if data_term == "poisson" else 0
    st.obj_scale, st.grad_scale, st.hvp_scale = _data_scale_factors(data_term, st.s0)
    st.B = _Blocks()
    st.x = x
    st.n, st.nd = x.numel(), st.f.numel()
    st.pen = _Penalty(hessian, tuple(x.shape))
    shape = tuple(x.shape)
    st.grad, st.kt, st.rh, st.htu = empty_f32(shape), empty_f32(shape), empty_f32(shape), empty_f32(shape)
    st.score = empty_f32(st.f.shape)
    st.hessw = empty_f32(st.f.shape) if st.poisson else None
    st.mu = st.model(x)
    return st


def jetnewton_objective(x_tilde, blur_op, observed, hessian, s0: float, background: float = 0.0, beta: float = 1.0,
                        eta: float = 1e-2, data_term: str = "poisson") -> float:
    """F = D(mu) + (beta/2) sum log(eta + u) (jetnewton_mlx.py:186-202)."""
    st = _helper(x_tilde, blur_op, observed, hessian, s0, background, beta, eta, data_term)
    st.data_sums(st.mu)
    st.pen.weights(st.x, st.beta, st.eta, st.B, full=False)
    jn, er = st.B.read()
    return st.objective(jn, er)


def jetnewton_gradient(x_tilde, blur_op, observed, hessian, s0: float, background: float = 0.0, beta: float = 1.0,
                       eta: float = 1e-2, data_term: str = "poisson"):
    """(grad, (r, w, c, data_hess_w)) (jetnewton_mlx.py:205-225); r is (nc, *shape)."""
    st = _helper(x_tilde, blur_op, observed, hessian, s0, background, beta, eta, data_term)
    st.gradient()
    shape = tuple(st.x.shape)
    r = st.pen.r.reshape((st.pen.nc,) + shape)
    if st.pen.dims is not None and len(shape) == 2:
        r = r[[1, 2, 5]]  # the 2-D operators' (yy, xx, xy) channels
    return st.grad, (r, st.pen.w, st.pen.c, st.hessw if st.poisson else 2.0)


def jetnewton_hvp(v, blur_op, hessian, r, w, c, data_hess_w, s0: float, data_term: str = "poisson") -> torch.Tensor:
    """Exact Hessian product (jetnewton_mlx.py:228-293): hvp_scale A^T(d A v) + H~^T(w q + 2 c <r, q> r)."""
    _check_data_term(data_term)
    v = as_f32(v)
    shape = tuple(v.shape)
    pen = _Penalty(hessian, shape)
    r = as_f32(r).reshape(-1, pen.n)
    if pen.dims is not None and r.shape[0] == 3:  # (yy, xx, xy) of a 2-D operator into the 6-channel layout
        r6 = torch.zeros((6, pen.n), dtype=torch.float32, device=r.device)
        r6[[1, 2, 5]] = r
        r = r6
    pen.nc = r.shape[0]
    pen.r, pen.w, pen.c = r.contiguous(), as_f32(w).reshape(shape), as_f32(c).reshape(shape)
    _, _, hvp_scale = _data_scale_factors(data_term, float(s0))
    av = _conv(blur_op, v, False)
    if isinstance(data_hess_w, (int, float)):
        t = _ew(N.ER_EW_SCALE, av, empty_f32(av.shape), alpha=float(data_hess_w))
    else:
        t = _ew(N.ER_EW_MUL, av, empty_f32(av.shape), b=as_f32(data_hess_w))
    kt = _conv(blur_op, t, True)
    htu = pen.adjoint(pen.hvp_operand(v))
    return _ew(N.ER_EW_AXPY, kt, empty_f32(shape), alpha=hvp_scale, b=htu)


def identify_active_set(x, grad_F, eps_bar: float = 1e-2):
    """(active mask, eps) by the Bertsekas rule (jetnewton_mlx.py:302-315)."""
    x, g = as_f32(x), as_f32(grad_F)
    B = _Blocks()
    active = torch.empty(x.shape, dtype=torch.bool, device=device())
    N.call("dh_jn_kkt", x.data_ptr(), g.data_ptr(), x.numel(), float(eps_bar), B.jn.data_ptr(), B.part.data_ptr(),
           N.stream_ptr())
    N.call("dh_jn_active", x.data_ptr(), g.data_ptr(), x.numel(), 0.0, 0.0, None, active.data_ptr(), None, None,
           None, _CG_EPS, B.jn.data_ptr(), B.part.data_ptr(), N.stream_ptr())
    return active, B.read()[0][N.JN_EPS]


@dataclass
class ActiveSetState:
    """Persistent freeze mask (jetnewton_mlx.py:318-352)."""

    frozen: Optional[torch.Tensor] = None

    def reset(self, shape) -> None:
        self.frozen = torch.zeros(tuple(shape), dtype=torch.bool, device=device())

    def update(self, x, grad_F, eps_bar: float = 1e-2, freeze_tau: float = 1e-3, freeze_delta: float = 1e-6):
        x, g = as_f32(x), as_f32(grad_F)
        if self.frozen is None:
            self.reset(x.shape)
        B = _Blocks()
        active = torch.empty(x.shape, dtype=torch.bool, device=device())
        N.call("dh_jn_kkt", x.data_ptr(), g.data_ptr(), x.numel(), float(eps_bar), B.jn.data_ptr(),
               B.part.data_ptr(), N.stream_ptr())
        N.call("dh_jn_active", x.data_ptr(), g.data_ptr(), x.numel(), float(freeze_tau), float(freeze_delta),
               self.frozen.data_ptr(), active.data_ptr(), None, None, None, _CG_EPS, B.jn.data_ptr(),
               B.part.data_ptr(), N.stream_ptr())
        return active, B.read()[0][N.JN_EPS]


def solve_reduced_newton(hvp_full, grad_F, active_mask, precond_apply=None, max_steps: int = 50,
                         eps: float = 1e-20):
    """(d, n_iterations, hit_negative_curvature) of the masked reduced CG (jetnewton_mlx.py:355-430)."""
    if precond_apply is not None:
        raise NotImplementedError("solve_reduced_newton runs unpreconditioned CG (the solver's own use)")
    g = as_f32(grad_F)
    shape = tuple(g.shape)
    active = active_mask if isinstance(active_mask, torch.Tensor) else torch.from_numpy(np.asarray(active_mask))
    active = active.to(device=device(), dtype=torch.bool).contiguous()
    B = _Blocks()
    xcg, rcg, pcg, ap = empty_f32(shape), empty_f32(shape), empty_f32(shape), empty_f32(shape)
    N.call("dh_jn_active", None, g.data_ptr(), g.numel(), 0.0, 0.0, None, active.data_ptr(), rcg.data_ptr(),
           pcg.data_ptr(), xcg.data_ptr(), float(eps), B.jn.data_ptr(), B.part.data_ptr(), N.stream_ptr())
    rhs = rcg.clone()

    def apply_hvp():
        hv = as_f32(hvp_full(pcg))
        N.call("dh_jn_cg_ap", hv.data_ptr(), 1.0, None, active.data_ptr(), pcg.data_ptr(), ap.data_ptr(), g.numel(),
               float(eps), B.jn.data_ptr(), B.part.data_ptr(), N.stream_ptr())

    jn = _reduced_cg(apply_hvp, xcg, rcg, pcg, ap, B, max_steps, B.read()[0][N.JN_STOP] != 0.0)
    n_cg = _cg_steps(jn)
    d = rhs if jn[N.JN_USE_RHS] != 0.0 else xcg
    return d, n_cg, jn[N.JN_HITNEG] != 0.0


def projected_armijo(objective_fn, grad_F, x, d, sigma: float = 1e-4, max_backtracks: int = 30):
    """(x_new, t, accepted) along x(t) = max(x + t d, 0) (jetnewton_mlx.py:438-468)."""
    x, d, g = as_f32(x), as_f32(d), as_f32(grad_F)
    B = _Blocks()
    F0 = objective_fn(x)
    t = 1.0
    for _ in range(max_backtracks):
        xt = empty_f32(x.shape)
        N.call("dh_jn_trial", x.data_ptr(), d.data_ptr(), t, g.data_ptr(), xt.data_ptr(), x.numel(),
               B.jn.data_ptr(), B.part.data_ptr(), N.stream_ptr())
        directional = B.read()[0][N.JN_DIRECTIONAL]
        Ft = objective_fn(xt)
        if math.isfinite(Ft) and Ft <= F0 + sigma * directional:
            return xt, t, True
        t *= 0.5
    return x, 0.0, False


def estimate_penalty_noise_floor(hessian, shape, n_trials: int = 8, seed: int = 0) -> dict:
    """u statistics of unit white noise through the regulariser (jetnewton_mlx.py:136-183)."""
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    pen = _Penalty(hessian, shape)
    B = _Blocks()
    samples = []
    for _ in range(n_trials):
        noise = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(device())
        u = empty_f32(shape)
        if pen.dims is not None:
            pen.weights(noise, 1.0, 1.0, B, full=False, u_out=u)
        else:
            r = as_f32(hessian.forward(noise))
            N.call("dh_jn_penalty_stacked", r.data_ptr(), r.numel() // pen.n, pen.n, 1.0, 1.0, None, None, None,
                   u.data_ptr(), B.jn.data_ptr(), B.part.data_ptr(), N.stream_ptr())
        samples.append(u.cpu().numpy().ravel())
    arr = np.concatenate(samples)
    return {"curvature": {"mean": float(arr.mean()), "median": float(np.median(arr)),
                          "p1": float(np.percentile(arr, 1)), "p99": float(np.percentile(arr, 99))}}
