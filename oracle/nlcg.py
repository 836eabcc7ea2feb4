"""CPU oracle for NLCG deconvolution (TEST INFRASTRUCTURE, not the product).

Only ``tests/`` may import this module.  It restates in NumPy the reference's
accelerated Poisson maximum-likelihood solver by nonlinear conjugate
gradients (deconlib/deconvolution/nlcg_mlx.py):

* ``objective`` / ``gradient``              (nlcg_mlx.py:88-152): f = s^2, m = K f + b,
  phi = sum m - sum g ln max(m, eps) + beta ||C f||^2,
  grad = 2 s K^T(1 - g / max(m, eps)) + 4 beta s C^T C f;
* ``hessian_quadform``                      (nlcg_mlx.py:155-192);
* ``step_length``                           (nlcg_mlx.py:195-291): Newton-Raphson on the
  exact quadratic-in-lambda model K f(lambda) = Kss + 2 lambda Ksd + lambda^2 Kdd;
* ``i_divergence_between``                  (nlcg_mlx.py:294-314);
* ``nlcg_with_operator`` / ``nlcg_solver``  (nlcg_mlx.py:317-588): Fletcher-Reeves with
  steepest-descent fallback, the discrepancy-principle and Eq. 17 stopping rules;
* ``Hessian2D``                             (linops_mlx.py:153-179).

All arithmetic follows ``dtype`` (float64 by default, the FP64 oracle; float32
follows the reference's array dtype flow).  Sums are float64 either way.

Parity pins (tests/test_nlcg.py): the gradient against central differences of
the objective, the quadratic form against central differences of the
gradient, the step length as the line minimiser (dphi collapses) with its
algebraic model equal to a direct forward evaluation, descent of the
analytical step, the separable closed form (delta PSF: f -> g - b), and the
reference's own end-to-end assertions (loss decrease, the two stopping rules,
bounded high-count runs, the regulariser's smoothing) run on the oracle.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from .deconv import SQRT2, d1_cen, d1_cen_adj, d2, poisson_i_divergence


class Hessian2D:
    """[H_yy, H_xx, sqrt2 H_xy] with Neumann boundaries (linops_mlx.py:153-179)."""

    operator_norm_sq = 48.0

    def forward(self, f):
        t = f.dtype.type
        return np.stack([d2(f, 0), d2(f, 1), t(SQRT2) * d1_cen(d1_cen(f, 0), 1)])

    def adjoint(self, H):
        t = H.dtype.type
        return d2(H[0], 0) + d2(H[1], 1) + t(SQRT2) * d1_cen_adj(d1_cen_adj(H[2], 0), 1)

    __call__ = forward


def _sum(a) -> float:
    return float(np.sum(a, dtype=np.float64))


def _model(s, op, b):
    f = s * s
    return f, op.forward(f) + f.dtype.type(b)


def _objective_from_m(m, g, f, reg, beta, eps):
    t = m.dtype.type
    val = _sum(m) - _sum(g * np.log(np.maximum(m, t(eps))))
    if reg is not None and beta != 0.0:
        cf = reg.forward(f)
        val += beta * _sum(cf * cf)
    return val


def _gradient_from_m(s, f, m, g, op, reg, beta, eps):
    t = s.dtype.type
    w = t(1.0) - g / np.maximum(m, t(eps))
    ktw = op.adjoint(w)
    grad = t(2.0) * s * ktw
    ctcf = None
    if reg is not None and beta != 0.0:
        ctcf = reg.adjoint(reg.forward(f))
        grad = grad + t(4.0 * beta) * s * ctcf
    return grad, (m, ktw, ctcf)


def objective(s, op, g, background=0.0, reg=None, beta=0.0, eps=1e-10):
    f, m = _model(s, op, background)
    return _objective_from_m(m, g, f, reg, beta, eps)


def gradient(s, op, g, background=0.0, reg=None, beta=0.0, eps=1e-10):
    f, m = _model(s, op, background)
    return _gradient_from_m(s, f, m, g, op, reg, beta, eps)


def hessian_quadform(s, d, op, g, background=0.0, reg=None, beta=0.0, aux=None, eps=1e-10):
    t = s.dtype.type
    if aux is None:
        _, aux = gradient(s, op, g, background, reg, beta, eps)
    m, ktw, ctcf = aux
    sd = s * d
    ksd = op.forward(sd)
    quad = 4.0 * _sum((g / np.maximum(m * m, t(eps))) * (ksd * ksd)) + 2.0 * _sum(ktw * (d * d))
    if reg is not None and beta != 0.0:
        csd = reg.forward(sd)
        quad += 8.0 * beta * _sum(csd * csd) + 4.0 * beta * _sum(ctcf * (d * d))
    return quad


def step_length(s, d, f, m, op, g, background=0.0, reg=None, beta=0.0, newton_iters=3, eps=1e-10):
    """(lam, m_lam) (nlcg_mlx.py:195-291)."""
    t = s.dtype.type
    b = t(background)
    use_reg = reg is not None and beta != 0.0
    kss = m - b
    ksd = op.forward(s * d)
    kdd = op.forward(d * d)
    if use_reg:
        cf0, csd, cdd = reg.forward(f), reg.forward(s * d), reg.forward(d * d)
    lam = 0.0
    for _ in range(newton_iters):
        ml = np.maximum(kss + t(2.0 * lam) * ksd + t(lam * lam) * kdd + b, t(eps))
        ksdl = ksd + t(lam) * kdd
        wl = t(1.0) - g / ml
        dphi1 = 2.0 * _sum(ksdl * wl)
        dphi2 = 2.0 * _sum(kdd * wl) + 4.0 * _sum(g * (ksdl * ksdl) / (ml * ml))
        if use_reg:
            u = cf0 + t(2.0 * lam) * csd + t(lam * lam) * cdd
            csdl = csd + t(lam) * cdd
            dphi1 += 4.0 * beta * _sum(u * csdl)
            dphi2 += 8.0 * beta * _sum(csdl * csdl) + 4.0 * beta * _sum(u * cdd)
        if not np.isfinite(dphi2) or abs(dphi2) < eps:
            break
        lam = lam - dphi1 / dphi2
        if not np.isfinite(lam):
            break
    m_lam = kss + t(2.0 * lam) * ksd + t(lam * lam) * kdd + b
    return lam, m_lam


def i_divergence_between(f_prev, f_curr, eps=1e-10) -> float:
    t = f_prev.dtype.type
    fp = np.maximum(f_prev, t(eps))
    fc = np.maximum(f_curr, t(eps))
    return _sum(f_prev * (np.log(fp) - np.log(fc)) - f_prev + f_curr)


@dataclass
class NLCGResult:
    restored: np.ndarray
    pred: np.ndarray
    iterations: int
    loss_history: list
    converged: bool = False
    background: float = 0.0
    full_shape: Optional[Tuple[int, ...]] = None
    valid_slices: Optional[Tuple[slice, ...]] = None


def nlcg_with_operator(observed, blur_op, num_iter=50, background=0.0, regularizer=None, reg_weight=0.0,
                       init=None, callback=None, eval_interval=10, slack=1.0, tol=1e-4, min_iter=10,
                       restart_interval=None, newton_iters=3, dtype=np.float64):
    """The reference's control flow step for step (nlcg_mlx.py:317-541)."""
    t = np.dtype(dtype).type
    g = np.asarray(observed).astype(dtype)
    eps = 1e-10
    b = float(background)
    beta = float(reg_weight)
    use_reg = regularizer is not None and beta != 0.0
    dmb = np.maximum(g - t(b), t(1e-6))
    if init is None:
        f0 = np.maximum(blur_op.adjoint(dmb), t(1e-6))
    else:
        f0 = np.maximum(np.asarray(init).astype(dtype), t(1e-6))
    s = np.sqrt(f0)
    reg = regularizer if use_reg else None
    f, m = _model(s, blur_op, b)
    grad, aux = _gradient_from_m(s, f, m, g, blur_op, reg, beta, eps)
    r = -grad
    d = r
    rr = _sum(r * r)
    loss_history, window = [], []
    converged = False
    k = 0

    def _step(direction, phi0):
        lam, m_lam = step_length(s, direction, f, aux[0], blur_op, g, b, reg, beta, newton_iters, eps)
        if not np.isfinite(lam):
            return None
        s_new = s + t(lam) * direction
        f_new = s_new * s_new
        if _objective_from_m(m_lam, g, f_new, reg, beta, eps) >= phi0:
            return None
        return s_new, f_new, m_lam, lam

    for k in range(num_iter):
        phi0 = _objective_from_m(aux[0], g, f, reg, beta, eps)
        step = _step(d, phi0)
        if step is None and d is not r:
            d = r
            step = _step(d, phi0)
        if step is None:
            converged = True
            loss_history.append(poisson_i_divergence(g, aux[0]))
            break
        f_prev = f
        s, f, m, lam = step
        grad, aux = _gradient_from_m(s, f, m, g, blur_op, reg, beta, eps)
        r_new = -grad
        rr_new = _sum(r_new * r_new)
        gamma = rr_new / rr if rr > eps else 0.0
        if restart_interval is not None and (k + 1) % restart_interval == 0:
            gamma = 0.0
        d = r_new + t(gamma) * d
        r = r_new
        rr = rr_new
        loss = poisson_i_divergence(g, m)
        logged = k % eval_interval == 0 or k == num_iter - 1
        if logged:
            loss_history.append(loss)
        rel = i_divergence_between(f_prev, f) / max(_sum(f), eps)
        window.append(rel)
        if len(window) > 5:
            window.pop(0)
        rel_avg = sum(window) / len(window)
        stop = bool(callback(k, f)) if callback is not None else False
        ready = not stop and k + 1 >= min_iter
        if ready and not use_reg and slack > 0.0 and loss <= 0.5 * slack:
            converged = stop = True
        elif ready and tol > 0.0 and len(window) == 5 and rel_avg < tol:
            converged = stop = True
        if stop:
            if not logged:
                loss_history.append(loss)
            break
    x = s * s
    pred = blur_op.forward(x) + t(b)
    return NLCGResult(restored=x, pred=pred, iterations=k + 1, loss_history=loss_history, converged=converged,
                      background=b, full_shape=tuple(x.shape), valid_slices=None)


def nlcg_solver(num_iter=50, background=0.0, regularizer=None, reg_weight=0.0, init_value=None, dtype=np.float64,
                **kw):
    """solve(data, model) -> restored[model.valid_slices] (nlcg_mlx.py:544-588)."""

    def solve(data, model):
        init = None if init_value is None else np.full(model.padded_shape, float(init_value), dtype=dtype)
        res = nlcg_with_operator(data, model.op, num_iter=num_iter, background=background,
                                 regularizer=regularizer, reg_weight=reg_weight, init=init, dtype=dtype, **kw)
        return res.restored[model.valid_slices]

    return solve
