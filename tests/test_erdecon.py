"""ER-Decon oracle pinned on the CPU (the GPU solver is tested in test_gpu_erdecon.py).

Pins of oracle/erdecon.py against what the mathematics fixes: the gradient
against central differences of the objective (Gaussian and Poisson data terms,
the weight floor, the intensity term); the Gauss-Newton quadratic form against
its explicit sum of weighted squares, its symmetry and positive
semidefiniteness; CG against a dense solve of the same SPD system; the floor's
weight identity; descent of a small gradient step; and the reference's own
end-to-end assertions (tests/test_erdecon.py of the reference) on the oracle.
"""

import numpy as np
import pytest

from oracle import deconv as O
from oracle import erdecon as ER
from oracle import tiles as OT
from oracle.nlcg import Hessian2D


def gaussian_psf(shape, sigma):
    grids = np.meshgrid(*[np.arange(n) - (n - 1) / 2.0 for n in shape], indexing="ij")
    r2 = sum(gc ** 2 for gc in grids)
    psf = np.exp(-r2 / (2.0 * sigma ** 2)).astype(np.float32)
    return psf / psf.sum()


def _setup(shape=(24, 24)):
    rng = np.random.default_rng(0)
    blur = O.LinearFFTConvolver(gaussian_psf((9, 9), 1.5), shape, dtype=np.float64)
    return blur, rng.random(shape), rng.random(shape) + 0.5, Hessian2D()


def _fd(fun, s, e, h=1e-5):
    return (fun(s + h * e) - fun(s - h * e)) / (2 * h)


@pytest.mark.parametrize("data_term,floor,iw", [("gaussian", 0.0, 0.0), ("poisson", 0.0, 0.0),
                                                ("gaussian", 0.05, 0.0), ("gaussian", 0.0, 0.7),
                                                ("poisson", 0.05, 0.7)])
def test_gradient_matches_finite_difference(data_term, floor, iw):
    blur, f, s, H = _setup()
    kw = dict(background=0.05, lam=0.1, eps_reg=1e-2, data_term=data_term, floor_frac=floor, intensity_weight=iw)
    grad, _ = ER.gradient(s, blur, f, H, **kw)
    e = np.random.default_rng(1).standard_normal(s.shape)
    fd = _fd(lambda x: ER.objective(x, blur, f, H, **kw), s, e)
    analytic = float(np.sum(grad * e))
    assert abs(fd - analytic) <= 1e-6 * (abs(analytic) + 1.0)


@pytest.mark.parametrize("iw", [0.0, 0.7])
def test_gn_quadratic_form_symmetry_psd(iw):
    blur, f, s, H = _setup()
    _, _, w = ER.weights(s * s, H, 0.1, 1e-2, intensity_weight=iw)
    rng = np.random.default_rng(2)
    u, v = rng.standard_normal(s.shape), rng.standard_normal(s.shape)
    quad = float(np.sum(v * ER.gn_hvp(s, v, blur, H, w, 2.0, iw)))
    sv = s * v
    ksv, hsv = blur.forward(sv), H.forward(sv)
    explicit = 8.0 * np.sum(ksv * ksv) + 4.0 * np.sum(w * hsv * hsv) + 4.0 * iw * np.sum(w[0] * sv * sv)
    assert abs(quad - explicit) <= 1e-10 * explicit and quad > 0
    uhv = float(np.sum(u * ER.gn_hvp(s, v, blur, H, w, 2.0, iw)))
    vhu = float(np.sum(v * ER.gn_hvp(s, u, blur, H, w, 2.0, iw)))
    assert abs(uhv - vhu) <= 1e-10 * (abs(uhv) + 1.0)
    # the Poisson data diagonal f/m^2 keeps the form an explicit weighted sum
    _, m = ER._model(s, blur, 0.05)
    _, dhw = ER.data_deriv(m, f, "poisson")
    quad_p = float(np.sum(v * ER.gn_hvp(s, v, blur, H, w, dhw, iw)))
    explicit_p = 4.0 * np.sum(dhw * ksv * ksv) + 4.0 * np.sum(w * hsv * hsv) + 4.0 * iw * np.sum(w[0] * sv * sv)
    assert abs(quad_p - explicit_p) <= 1e-10 * explicit_p


def test_cg_matches_dense_solve():
    """On a 6x7 problem CG with enough steps reaches the dense solution of the SPD system."""
    blur, f, s, H = _setup((6, 7))
    _, _, w = ER.weights(s * s, H, 0.1, 1e-2)
    n = s.size
    cols = [ER.gn_hvp(s, np.eye(n)[i].reshape(s.shape), blur, H, w, 2.0).ravel() for i in range(n)]
    A = np.stack(cols, axis=1)
    np.testing.assert_allclose(A, A.T, rtol=1e-10, atol=1e-10)
    b = np.random.default_rng(3).standard_normal(n)
    x, steps = ER.cg_solve(lambda v: ER.gn_hvp(s, v, blur, H, w, 2.0), b.reshape(s.shape), 200, 1e-12)
    np.testing.assert_allclose(x.ravel(), np.linalg.solve(A, b), rtol=1e-6, atol=1e-8)
    # the first CG step alone is a positive multiple of b
    x1, _ = ER.cg_solve(lambda v: ER.gn_hvp(s, v, blur, H, w, 2.0), b.reshape(s.shape), 1, 0.0)
    ratio = x1.ravel() / b
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-10)
    assert ratio[0] > 0


def test_floor_identity_and_descent():
    blur, f, s, H = _setup()
    _, q, w0 = ER.weights(s * s, H, 0.1, 1e-2)
    _, _, w1 = ER.weights(s * s, H, 0.1, 1e-2, floor_frac=0.05)
    np.testing.assert_allclose(w1, w0 + 0.05 * 0.1 / 1e-2, rtol=1e-12)
    np.testing.assert_allclose(w0, 0.1 / (1e-2 + q), rtol=1e-12)
    kw = dict(background=0.05, lam=0.1, eps_reg=1e-2)
    grad, _ = ER.gradient(s, blur, f, H, **kw)
    step = 1e-3 / (1.0 + float(np.sum(grad * grad)))
    assert ER.objective(s - step * grad, blur, f, H, **kw) < ER.objective(s, blur, f, H, **kw)


def _make_problem(seed=42):
    rng = np.random.default_rng(seed)
    shape = (48, 48)
    truth = np.zeros(shape)
    for _ in range(6):
        y, x = rng.integers(8, 40, size=2)
        truth[y, x] += rng.uniform(0.5, 1.0)
    truth += 0.02
    fm = OT.make_forward_model(gaussian_psf((15, 15), 2.0), shape, 1.0, dtype=np.float64)
    tp = np.zeros(fm.padded_shape)
    tp[fm.valid_slices] = truth
    clean = fm.op.forward(tp)
    observed = np.maximum(clean + 0.01 * rng.standard_normal(clean.shape), 0.0)
    return fm, truth, observed


def test_reconstruction_reference_assertions():
    fm, truth, observed = _make_problem()
    res = ER.erdecon_with_operator(observed, fm.op, num_iter=30, eval_interval=1, tol=0.0)
    assert res.restored.shape == fm.padded_shape and res.pred.shape == fm.data_shape
    assert res.restored.min() >= 0.0 and res.loss_history[-1] < res.loss_history[0]
    for term in ("gaussian", "poisson"):
        r = ER.erdecon_with_operator(observed, fm.op, num_iter=50, reg_weight=0.02, data_term=term)
        assert np.linalg.norm(r.restored[fm.valid_slices] - truth) < np.linalg.norm(observed - truth)

    def rough(x):
        a = x[fm.valid_slices]
        return float(np.abs(np.diff(a, axis=0)).sum() + np.abs(np.diff(a, axis=1)).sum())

    weak = ER.erdecon_with_operator(observed, fm.op, num_iter=40, reg_weight=0.01, tol=0.0)
    strong = ER.erdecon_with_operator(observed, fm.op, num_iter=40, reg_weight=0.5, tol=0.0)
    assert rough(strong.restored) < rough(weak.restored)
    conv = ER.erdecon_with_operator(observed, fm.op, num_iter=500, min_iter=5, reg_weight=0.05)
    assert conv.converged and conv.iterations < 500


def test_normalisation_is_scale_equivariant():
    """Scaling the data by c scales restored and pred by c (the solve runs on data / max)."""
    fm, _, observed = _make_problem()
    a = ER.erdecon_with_operator(observed, fm.op, num_iter=5, tol=0.0, newton_tol=0.0)
    b = ER.erdecon_with_operator(observed * 37.0, fm.op, num_iter=5, tol=0.0, newton_tol=0.0)
    np.testing.assert_allclose(b.restored, 37.0 * a.restored, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(b.data_misfit_history, 37.0 * np.asarray(a.data_misfit_history), rtol=1e-9)
    assert b.loss_history == pytest.approx(a.loss_history, rel=1e-9)
