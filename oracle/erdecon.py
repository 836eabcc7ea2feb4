"""CPU oracle for ER-Decon (TEST INFRASTRUCTURE, not the product).

Only ``tests/`` may import this module.  It restates in NumPy the reference's
edge-preserving Hessian-log deconvolution by Gauss-Newton-CG
(deconlib/deconvolution/erdecon_mlx.py):

* ``weights``        (erdecon_mlx.py:144-214): q = sum_i (H_i g)^2 [+ iw g^2],
  w = lam / (eps + q) [+ floor_frac lam / eps];
* ``data_misfit`` / ``data_deriv`` (:220-250): Gaussian ||m - f||^2 or the Poisson
  sum(m - f ln m) with m floored at 1e-8;
* ``objective`` / ``gradient`` / ``gn_hvp`` (:253-345);
* ``cg_solve``       (:348-393): matrix-free CG from zero with the non-positive
  curvature guard and the relative-residual stop;
* ``erdecon_with_operator`` / ``erdecon_solver`` (:396-733): amplitude
  normalisation, Gauss-Newton-CG outer loop, Armijo backtracking, the Newton
  decrement and data-misfit stopping tests.

Arithmetic follows ``dtype`` (float64 by default); sums are float64.

Parity pins (tests/test_erdecon.py): the gradient against central differences
of the objective for both data terms, the floor and the intensity term; the
Gauss-Newton quadratic form against its explicit sum of weighted squares,
symmetry and positive semidefiniteness; CG against a dense solve of the same
SPD system on a tiny problem; the floor's weight identity; descent of a small
gradient step; and the reference's end-to-end assertions run on the oracle.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

from .deconv import Hessian3D, poisson_i_divergence
from .nlcg import Hessian2D

POISSON_M_FLOOR = 1e-8


def _sum(a) -> float:
    return float(np.sum(a, dtype=np.float64))


def default_hessian(ndim):
    if ndim == 2:
        return Hessian2D()
    if ndim == 3:
        return Hessian3D()
    raise ValueError(f"ER-Decon supports 2D or 3D reconstructions; got ndim={ndim}. "
                     "Pass an explicit `hessian` operator for other ranks.")


def weights(g, hessian, lam, eps_reg, floor_frac=0.0, intensity_weight=0.0):
    """(Hg, q, w) with q, w of shape (1, *spatial)."""
    t = g.dtype.type
    hg = hessian.forward(g)
    q = np.sum(hg * hg, axis=0, keepdims=True)
    if intensity_weight > 0.0:
        q = q + t(intensity_weight) * (g * g)[None]
    w = t(lam) / (t(eps_reg) + q)
    if floor_frac > 0.0:
        w = w + t(floor_frac * lam / eps_reg)
    return hg, q, w


def data_misfit(m, f, data_term):
    t = m.dtype.type
    if data_term == "poisson":
        mm = np.maximum(m, t(POISSON_M_FLOOR))
        return _sum(mm - f * np.log(mm))
    r = m - f
    return _sum(r * r)


def data_deriv(m, f, data_term):
    t = m.dtype.type
    if data_term == "poisson":
        mm = np.maximum(m, t(POISSON_M_FLOOR))
        return t(1.0) - f / mm, f / (mm * mm)
    return t(2.0) * (m - f), 2.0


def _model(s, op, b):
    g = s * s
    return g, op.forward(g) + s.dtype.type(b)


def objective(s, op, f, hessian, background=0.0, lam=0.05, eps_reg=1e-2, data_term="gaussian",
              floor_frac=0.0, intensity_weight=0.0):
    g, m = _model(s, op, background)
    _, q, _ = weights(g, hessian, lam, eps_reg, intensity_weight=intensity_weight)
    val = data_misfit(m, f, data_term) + 0.5 * lam * _sum(np.log(s.dtype.type(eps_reg) + q))
    if floor_frac > 0.0:
        val += 0.5 * floor_frac * (lam / eps_reg) * _sum(q)
    return val


def gradient(s, op, f, hessian, background=0.0, lam=0.05, eps_reg=1e-2, data_term="gaussian",
             floor_frac=0.0, intensity_weight=0.0):
    t = s.dtype.type
    g, m = _model(s, op, background)
    hg, _, w = weights(g, hessian, lam, eps_reg, floor_frac, intensity_weight)
    score, dhw = data_deriv(m, f, data_term)
    grad = t(2.0) * s * op.adjoint(score) + t(2.0) * s * hessian.adjoint(w * hg)
    if intensity_weight > 0.0:
        grad = grad + t(2.0) * s * (t(intensity_weight) * w[0] * g)
    return grad, (w, dhw)


def gn_hvp(s, v, op, hessian, w, data_hess_w=2.0, intensity_weight=0.0):
    t = s.dtype.type
    sv = s * v
    dhw = t(data_hess_w) if np.isscalar(data_hess_w) else data_hess_w
    data = t(4.0) * s * op.adjoint(dhw * op.forward(sv))
    reg = t(4.0) * s * hessian.adjoint(w * hessian.forward(sv))
    if intensity_weight > 0.0:
        reg = reg + t(4.0 * intensity_weight) * s * (w[0] * sv)
    return data + reg


def cg_solve(hvp, b, max_steps, tol, eps=1e-20):
    """(x, steps) for H x = b from x = 0 (erdecon_mlx.py:348-393)."""
    t = b.dtype.type
    x = np.zeros_like(b)
    r = b
    p = r
    rs = _sum(r * r)
    if rs <= eps:
        return x, 0
    thr = (tol ** 2) * rs
    steps = 0
    for _ in range(max_steps):
        ap = hvp(p)
        denom = _sum(p * ap)
        steps += 1
        if not np.isfinite(denom) or denom <= eps:
            break
        step = rs / denom
        x = x + t(step) * p
        r = r - t(step) * ap
        rs_new = _sum(r * r)
        if not np.isfinite(rs_new) or rs_new < 0.0:
            break
        if rs_new <= thr:
            break
        p = r + t(rs_new / rs) * p
        rs = rs_new
    return x, steps


@dataclass
class ERDeconResult:
    restored: np.ndarray
    pred: np.ndarray
    iterations: int
    loss_history: list
    data_misfit_history: list = field(default_factory=list)
    converged: bool = False
    background: float = 0.0
    data_scale: float = 1.0
    full_shape: Optional[Tuple[int, ...]] = None
    valid_slices: Optional[Tuple[slice, ...]] = None


def erdecon_with_operator(observed, blur_op, hessian=None, reg_weight=0.05, eps_reg=1e-2, data_term="gaussian",
                          floor_frac=0.0, intensity_weight=0.0, num_iter=50, background=0.0, normalize=True,
                          data_scale=None, init=None, cg_max_steps=25, cg_tol=0.1, callback=None, eval_interval=5,
                          newton_tol=1e-3, tol=0.0, min_iter=5, ls_max_backtracks=30, ls_c1=1e-4,
                          dtype=np.float64):
    """The reference's outer loop step for step (erdecon_mlx.py:564-692)."""
    t = np.dtype(dtype).type
    f = np.asarray(observed).astype(dtype)
    if init is not None:
        init = np.asarray(init).astype(dtype)
    if data_scale is not None:
        c = float(data_scale)
    elif normalize:
        c = float(np.max(f))
    else:
        c = 1.0
    if not np.isfinite(c) or c <= 0.0:
        c = 1.0
    if c != 1.0:
        f = f / t(c)
        if init is not None:
            init = init / t(c)
    if data_term not in ("gaussian", "poisson"):
        raise ValueError(f"data_term must be 'gaussian' or 'poisson'; got {data_term!r}")
    b = float(background) / c
    lam = float(reg_weight)
    if hessian is None:
        hessian = default_hessian(f.ndim)
    if init is None:
        g0 = np.maximum(blur_op.adjoint(np.maximum(f - t(b), t(1e-6))), t(1e-6))
    else:
        g0 = np.maximum(init, t(1e-6))
    s = np.sqrt(g0)
    kw = dict(background=b, lam=lam, eps_reg=eps_reg, data_term=data_term, floor_frac=floor_frac,
              intensity_weight=intensity_weight)

    def misfit(s_):
        _, m = _model(s_, blur_op, b)
        return c * poisson_i_divergence(f, m)

    loss_history, misfit_history = [], []
    prev_misfit = None
    nd0 = None
    converged = False
    phi = objective(s, blur_op, f, hessian, **kw)
    k = 0
    for k in range(num_iter):
        grad, (w, dhw) = gradient(s, blur_op, f, hessian, **kw)
        p, _ = cg_solve(lambda v: gn_hvp(s, v, blur_op, hessian, w, dhw, intensity_weight), -grad,
                        cg_max_steps, cg_tol)
        gtp = _sum(grad * p)
        if not np.isfinite(gtp) or gtp >= 0.0:
            p = -grad
            gtp = -_sum(grad * grad)
        nd = abs(gtp)
        if nd0 is None:
            nd0 = nd
        alpha = 1.0
        accepted = False
        for _ in range(ls_max_backtracks):
            s_try = s + t(alpha) * p
            phi_try = objective(s_try, blur_op, f, hessian, **kw)
            if np.isfinite(phi_try) and phi_try <= phi + ls_c1 * alpha * gtp:
                accepted = True
                break
            alpha *= 0.5
        if not accepted:
            converged = True
            loss_history.append(phi)
            misfit_history.append(misfit(s))
            break
        s = s_try
        phi = phi_try
        logged = k % eval_interval == 0 or k == num_iter - 1
        mf = misfit(s)
        if logged:
            loss_history.append(phi)
            misfit_history.append(mf)
        stop = bool(callback(k, (s * s) * t(c))) if callback is not None else False
        if (not stop and newton_tol > 0.0 and k + 1 >= min_iter and nd0 is not None and nd0 > 0.0
                and nd < newton_tol * nd0):
            converged = stop = True
        if not stop and k + 1 >= min_iter and tol > 0.0 and prev_misfit is not None:
            rel = abs(prev_misfit - mf) / max(abs(prev_misfit), 1e-30)
            if rel < tol:
                converged = stop = True
        prev_misfit = mf
        if stop:
            if not logged:
                loss_history.append(phi)
                misfit_history.append(mf)
            break
    g_final = (s * s) * t(c)
    pred = blur_op.forward(g_final) + t(b * c)
    return ERDeconResult(restored=g_final, pred=pred, iterations=k + 1, loss_history=loss_history,
                         data_misfit_history=misfit_history, converged=converged, background=float(background),
                         data_scale=c, full_shape=tuple(g_final.shape), valid_slices=None)


def erdecon_solver(reg_weight=0.05, eps_reg=1e-2, num_iter=50, background=0.0, hessian=None, init_value=None,
                   newton_tol=1e-3, dtype=np.float64, **kw):
    """solve(data, model) -> restored[model.valid_slices] (erdecon_mlx.py:695-733)."""

    def solve(data, model):
        init = None if init_value is None else np.full(model.padded_shape, float(init_value), dtype=dtype)
        res = erdecon_with_operator(data, model.op, hessian=hessian, reg_weight=reg_weight, eps_reg=eps_reg,
                                    num_iter=num_iter, background=background, init=init, newton_tol=newton_tol,
                                    dtype=dtype, **kw)
        return res.restored[model.valid_slices]

    return solve
