"""CPU oracle for JetNewton (TEST INFRASTRUCTURE, not the product).

Only ``tests/`` may import this module.  It restates in NumPy the reference's
non-dimensional log-penalty deconvolution by two-metric active-set projected
Newton (deconlib/deconvolution/jetnewton_mlx.py) and the per-axis Hessians it
regularises with (deconlib/deconvolution/linops_mlx.py:265-381):

* ``data_scale_factors``   (jetnewton_mlx.py:83-108);
* ``penalty_weights``      (:111-133): r = H x, u = sum r^2, w = beta / (eta + u), c = -w^2 / beta;
* ``objective`` / ``gradient`` / ``hvp`` (:186-293), the exact (indefinite) Hessian;
* ``identify_active_set`` / ``ActiveSetState`` (:302-352);
* ``solve_reduced_newton`` (:355-430), masked CG with the absolute forcing tolerance;
* ``projected_armijo``     (:438-468);
* ``jetnewton_with_operator`` (:547-785);
* ``AnisotropicHessian2D/3D``.

Arithmetic follows ``dtype`` (float64 by default); sums are float64.

Parity pins (tests/test_jetnewton.py): the gradient against central differences
of the objective and the Hessian product against central differences of the
gradient (both data terms); symmetry of the product; the regulariser's dense
Hessian indefinite and symmetric; CG against a dense solve on a tiny SPD case;
the active-set rules on hand-made inputs; the first-step negative-curvature
return; the per-axis curvature against its continuum value on refined grids;
and the reference's end-to-end assertions run on the oracle.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .deconv import SQRT2, d1_cen, d1_cen_adj, d2
from .erdecon import data_deriv, data_misfit
from .deconv import poisson_i_divergence


def _sum(a) -> float:
    return float(np.sum(a, dtype=np.float64))


class AnisotropicHessian3D:
    """[kz^2 zz, ky^2 yy, kx^2 xx, sqrt2 ky kz yz, sqrt2 kx kz xz, sqrt2 kx ky xy] (linops_mlx.py:314-381)."""

    def __init__(self, kappa):
        self.kappa = tuple(float(k) for k in kappa)

    def _w(self, dtype):
        kz, ky, kx = self.kappa
        return np.array([kz * kz, ky * ky, kx * kx, ky * kz * SQRT2, kx * kz * SQRT2, kx * ky * SQRT2], dtype=dtype)

    def forward(self, f):
        w = self._w(f.dtype)
        dz, dy = d1_cen(f, 0), d1_cen(f, 1)
        H = np.stack([d2(f, 0), d2(f, 1), d2(f, 2), d1_cen(dz, 1), d1_cen(dz, 2), d1_cen(dy, 2)])
        return H * w.reshape(6, 1, 1, 1)

    def adjoint(self, H):
        Hw = H * self._w(H.dtype).reshape(6, 1, 1, 1)
        return (d2(Hw[0], 0) + d2(Hw[1], 1) + d2(Hw[2], 2) + d1_cen_adj(d1_cen_adj(Hw[3], 1), 0)
                + d1_cen_adj(d1_cen_adj(Hw[4], 2), 0) + d1_cen_adj(d1_cen_adj(Hw[5], 2), 1))


class AnisotropicHessian2D:
    """[ky^2 yy, kx^2 xx, sqrt2 ky kx xy] (linops_mlx.py:265-311)."""

    def __init__(self, kappa):
        self.kappa = tuple(float(k) for k in kappa)

    @classmethod
    def from_lengths(cls, ell, spacing):
        return cls(tuple(e / h for e, h in zip(ell, spacing)))

    def forward(self, f):
        ky, kx = self.kappa
        t = f.dtype.type
        return np.stack([t(ky * ky) * d2(f, 0), t(kx * kx) * d2(f, 1), t(ky * kx * SQRT2) * d1_cen(d1_cen(f, 0), 1)])

    def adjoint(self, H):
        ky, kx = self.kappa
        t = H.dtype.type
        return (t(ky * ky) * d2(H[0], 0) + t(kx * kx) * d2(H[1], 1)
                + t(ky * kx * SQRT2) * d1_cen_adj(d1_cen_adj(H[2], 0), 1))


def data_scale_factors(data_term, s0):
    if data_term == "poisson":
        return 1.0, s0, s0 * s0
    return 1.0 / (2.0 * s0 * s0), 1.0 / (2.0 * s0), 0.5


def penalty_weights(x, hessian, beta, eta):
    t = x.dtype.type
    r = hessian.forward(x)
    u = np.sum(r * r, axis=0)
    w = t(beta) / (t(eta) + u)
    c = -(w * w) / t(beta)
    return r, u, w, c


def _mu(x, op, s0, b):
    t = x.dtype.type
    return t(s0) * op.forward(x) + t(b)


def objective(x, op, f, hessian, s0, background=0.0, beta=1.0, eta=1e-2, data_term="poisson"):
    obj_scale, _, _ = data_scale_factors(data_term, s0)
    mu = _mu(x, op, s0, background)
    _, u, _, _ = penalty_weights(x, hessian, beta, eta)
    return obj_scale * data_misfit(mu, f, data_term) + 0.5 * beta * _sum(np.log(x.dtype.type(eta) + u))


def gradient(x, op, f, hessian, s0, background=0.0, beta=1.0, eta=1e-2, data_term="poisson"):
    t = x.dtype.type
    _, gs, _ = data_scale_factors(data_term, s0)
    mu = _mu(x, op, s0, background)
    score, dhw = data_deriv(mu, f, data_term)
    r, _, w, c = penalty_weights(x, hessian, beta, eta)
    return t(gs) * op.adjoint(score) + hessian.adjoint(w[None] * r), (r, w, c, dhw)


def hvp(v, op, hessian, r, w, c, dhw, s0, data_term="poisson"):
    t = v.dtype.type
    _, _, hs = data_scale_factors(data_term, s0)
    d = t(dhw) if np.isscalar(dhw) else dhw
    data = t(hs) * op.adjoint(d * op.forward(v))
    q = hessian.forward(v)
    m = c * np.sum(r * q, axis=0)
    return data + hessian.adjoint(w[None] * q + t(2.0) * m[None] * r)


def identify_active_set(x, grad, eps_bar=1e-2):
    t = x.dtype.type
    diff = x - np.maximum(x - grad, t(0.0))
    eps = min(eps_bar, float(np.sqrt(_sum(diff * diff))))
    return (x <= eps) & (grad > 0.0), eps


@dataclass
class ActiveSetState:
    frozen: Optional[np.ndarray] = None

    def reset(self, shape):
        self.frozen = np.zeros(shape, dtype=bool)

    def update(self, x, grad, eps_bar=1e-2, freeze_tau=1e-3, freeze_delta=1e-6):
        if self.frozen is None:
            self.reset(x.shape)
        active, eps = identify_active_set(x, grad, eps_bar)
        self.frozen = self.frozen | ((x < freeze_tau) & (grad > freeze_delta))
        return active | self.frozen, eps


def solve_reduced_newton(hvp_full, grad, active, max_steps=50, eps=1e-20):
    """(d, n, hit_negative_curvature) (jetnewton_mlx.py:355-430), unpreconditioned."""
    t = grad.dtype.type
    free = ~active
    rhs = np.where(free, -grad, t(0.0))

    def masked(v):
        return np.where(free, hvp_full(np.where(free, v, t(0.0))), t(0.0))

    x = np.zeros_like(rhs)
    r = rhs
    r0 = float(np.sqrt(_sum(r * r)))
    tol_k = min(0.5, r0 ** 0.5) * r0
    if r0 <= max(tol_k, eps):
        return x, 0, False
    p = r
    rz = _sum(r * r)
    n = 0
    hit = False
    for n in range(1, max_steps + 1):
        ap = masked(p)
        denom = _sum(p * ap)
        if not np.isfinite(denom) or denom <= eps:
            hit = True
            if n == 1:
                return rhs, 1, True
            return x, n - 1, True
        a = rz / denom
        x = x + t(a) * p
        r = r - t(a) * ap
        rn = float(np.sqrt(_sum(r * r)))
        if not np.isfinite(rn):
            break
        if rn <= tol_k:
            break
        rz_new = _sum(r * r)
        p = r + t(rz_new / rz) * p
        rz = rz_new
    return x, n, hit


def projected_armijo(objective_fn, grad, x, d, sigma=1e-4, max_backtracks=30, f0=None):
    t = x.dtype.type
    F0 = objective_fn(x) if f0 is None else f0
    step = 1.0
    for _ in range(max_backtracks):
        xt = np.maximum(x + t(step) * d, t(0.0))
        directional = _sum(grad * (xt - x))
        Ft = objective_fn(xt)
        if np.isfinite(Ft) and Ft <= F0 + sigma * directional:
            return xt, step, True
        step *= 0.5
    return x, 0.0, False


@dataclass
class JetNewtonResult:
    restored: np.ndarray
    pred: np.ndarray
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


def jetnewton_with_operator(observed, blur_op, hessian, s0, background=0.0, beta=1.0, eta=1.0, data_term="poisson",
                            num_iter=100, init=None, cg_max_steps=50, eps_bar=1e-2, freeze_tau=1e-3,
                            freeze_delta=1e-6, newton_tol=1e-3, tol=0.0, min_iter=3, ls_sigma=1e-4,
                            ls_max_backtracks=30, callback=None, eval_interval=5, dtype=np.float64):
    """The reference's outer loop step for step (jetnewton_mlx.py:597-785)."""
    t = np.dtype(dtype).type
    f = np.asarray(observed).astype(dtype)
    if init is None:
        g0 = np.maximum(blur_op.adjoint(np.maximum(f - t(background), t(s0))), t(s0))
        x = g0 / t(s0)
    else:
        x = np.maximum(np.asarray(init).astype(dtype), t(s0)) / t(s0)
    state = ActiveSetState()
    state.reset(x.shape)

    def F(xt):
        return objective(xt, blur_op, f, hessian, s0, background, beta, eta, data_term)

    res = JetNewtonResult(restored=None, pred=None, iterations=0, s0=float(s0))
    pg0 = nd0 = None
    converged = False
    k = 0
    for k in range(num_iter):
        grad, (r, w, c, dhw) = gradient(x, blur_op, f, hessian, s0, background, beta, eta, data_term)
        active, _ = state.update(x, grad, eps_bar, freeze_tau, freeze_delta)
        d_free, n_cg, _ = solve_reduced_newton(lambda v: hvp(v, blur_op, hessian, r, w, c, dhw, s0, data_term),
                                               grad, active, max_steps=cg_max_steps)
        res.cg_iterations_history.append(n_cg)
        d = d_free + np.where(active, -grad, t(0.0))
        nd = abs(_sum(grad * d))
        if nd0 is None:
            nd0 = nd
        x_try, step, accepted = projected_armijo(F, grad, x, d, ls_sigma, ls_max_backtracks)
        pg = float(np.max(np.abs(x - np.maximum(x - grad, t(0.0)))))
        if pg0 is None:
            pg0 = pg
        logged = k % eval_interval == 0 or k == num_iter - 1 or not accepted
        if logged:
            mu = _mu(x, blur_op, s0, background)
            res.loss_history.append(F(x))
            res.data_misfit_history.append(data_misfit(mu, f, data_term))
            res.idiv_history.append(poisson_i_divergence(f, mu))
            res.curvature_term_history.append(float(np.mean(np.sum(r * r, axis=0))))
            res.active_set_size_history.append(int(np.sum(active)))
            res.proj_grad_history.append(pg)
            res.newton_decrement_history.append(nd)
        if not accepted:
            converged = True
            break
        x = x_try
        stop = bool(callback(k, x * t(s0))) if callback is not None else False
        if (not stop and newton_tol > 0.0 and k + 1 >= min_iter and nd0 is not None and nd0 > 0.0
                and nd < newton_tol * nd0):
            converged = stop = True
        if not stop and tol > 0.0 and k + 1 >= min_iter and pg0 is not None and pg0 > 0.0 and pg < tol * pg0:
            converged = stop = True
        if stop:
            break
    res.restored = x * t(s0)
    res.pred = _mu(x, blur_op, s0, background)
    res.iterations = k + 1
    res.converged = converged
    return res
