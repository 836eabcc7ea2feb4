"""NLCG oracle pinned on the CPU (the GPU solver is tested in test_gpu_nlcg.py).

Pins of oracle/nlcg.py that a plausible slip (a dropped factor 2, a wrong sign,
a missing regulariser term, a mis-stepped lambda) would fail: the gradient
against central differences of the objective, the Hessian quadratic form
against central differences of the gradient, the Newton step as the line
minimiser with its algebraic model equal to a direct forward evaluation,
descent of the analytical step, the separable closed form of a delta PSF
(f -> g - b), and the reference's own end-to-end assertions run on the oracle
(tests/test_nlcg.py of the reference: loss decrease, discrepancy and Eq. 17
stopping, improvement over the data, bounded high-count runs, smoothing).
"""

import numpy as np
import pytest

from oracle import deconv as O
from oracle import nlcg as NL
from oracle import tiles as OT


def gaussian_psf(shape, sigma):
    grids = np.meshgrid(*[np.arange(n) - (n - 1) / 2.0 for n in shape], indexing="ij")
    r2 = sum(gc ** 2 for gc in grids)
    psf = np.exp(-r2 / (2.0 * sigma ** 2)).astype(np.float32)
    return psf / psf.sum()


def _math_setup(reg):
    rng = np.random.default_rng(0)
    shape = (24, 24)
    blur = O.LinearFFTConvolver(gaussian_psf((9, 9), 1.5), shape, dtype=np.float64)
    g = rng.random(shape) * 5.0 + 0.5
    s = rng.random(shape) + 0.5
    return blur, g, s, (NL.Hessian2D() if reg else None), (0.3 if reg else 0.0)


@pytest.mark.parametrize("reg", [False, True])
def test_gradient_matches_finite_difference(reg):
    blur, g, s, C, beta = _math_setup(reg)
    grad, _ = NL.gradient(s, blur, g, 0.5, C, beta)
    e = np.random.default_rng(1).standard_normal(s.shape)
    h = 1e-5
    fd = (NL.objective(s + h * e, blur, g, 0.5, C, beta) - NL.objective(s - h * e, blur, g, 0.5, C, beta)) / (2 * h)
    analytic = float(np.sum(grad * e))
    assert abs(fd - analytic) <= 1e-6 * (abs(analytic) + 1.0)


@pytest.mark.parametrize("reg", [False, True])
def test_hessian_quadform_matches_finite_difference(reg):
    blur, g, s, C, beta = _math_setup(reg)
    d = np.random.default_rng(2).standard_normal(s.shape)
    quad = NL.hessian_quadform(s, d, blur, g, 0.5, C, beta)
    h = 1e-5
    gp, _ = NL.gradient(s + h * d, blur, g, 0.5, C, beta)
    gm, _ = NL.gradient(s - h * d, blur, g, 0.5, C, beta)
    fd = (float(np.sum(gp * d)) - float(np.sum(gm * d))) / (2 * h)
    assert abs(fd - quad) <= 1e-6 * (abs(quad) + 1.0)


@pytest.mark.parametrize("reg", [False, True])
def test_step_length_is_line_minimizer(reg):
    blur, g, s, C, beta = _math_setup(reg)
    grad, aux = NL.gradient(s, blur, g, 0.5, C, beta)
    d = -grad
    lam, m_lam = NL.step_length(s, d, s * s, aux[0], blur, g, 0.5, C, beta, newton_iters=8)
    direct = blur.forward((s + lam * d) ** 2) + 0.5
    np.testing.assert_allclose(m_lam, direct, rtol=1e-10, atol=1e-10)
    h = 1e-6

    def dphi(l0):
        return (NL.objective(s + (l0 + h) * d, blur, g, 0.5, C, beta)
                - NL.objective(s + (l0 - h) * d, blur, g, 0.5, C, beta)) / (2 * h)

    assert abs(dphi(lam)) <= 1e-6 * abs(dphi(0.0))
    # three refinements (the default) already collapse dphi, as the reference asserts
    lam3, _ = NL.step_length(s, d, s * s, aux[0], blur, g, 0.5, C, beta)
    assert abs(dphi(lam3)) <= 0.05 * abs(dphi(0.0))


def test_analytical_step_is_descent():
    blur, g, s, C, beta = _math_setup(False)
    grad, aux = NL.gradient(s, blur, g, 0.5, C, beta)
    d = -grad
    quad = NL.hessian_quadform(s, d, blur, g, 0.5, C, beta, aux=aux)
    assert quad > 0
    lam = float(np.sum(grad * grad)) / quad
    assert NL.objective(s + lam * d, blur, g, 0.5) < NL.objective(s, blur, g, 0.5)


def test_hessian2d_adjoint_and_3d_restriction():
    rng = np.random.default_rng(3)
    f = rng.standard_normal((9, 11))
    H = rng.standard_normal((3, 9, 11))
    C = NL.Hessian2D()
    assert abs(np.sum(C.forward(f) * H) - np.sum(f * C.adjoint(H))) <= 1e-10 * np.sum(np.abs(f))
    # on a single plane the 3-D operator's yy, xx, xy components are the 2-D ones
    H3 = O.Hessian3D(1.0).forward(f[None])
    np.testing.assert_allclose(C.forward(f), H3[[1, 2, 5], 0], rtol=1e-12, atol=1e-12)


def test_delta_psf_converges_to_data_minus_background():
    """K = I makes the problem separable with minimiser f = g - b wherever g > b."""
    rng = np.random.default_rng(4)
    psf = np.zeros((1, 1, 1), np.float32)
    psf[0, 0, 0] = 1.0
    shape = (3, 6, 5)
    blur = O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    g = rng.uniform(20.0, 200.0, shape)
    res = NL.nlcg_with_operator(g, blur, num_iter=200, background=5.0, tol=0.0, slack=0.0, min_iter=1)
    np.testing.assert_allclose(res.restored, g - 5.0, rtol=1e-6)
    np.testing.assert_allclose(res.pred, g, rtol=1e-6)


def _make_problem(seed=42):
    rng = np.random.default_rng(seed)
    shape = (48, 48)
    truth = np.zeros(shape)
    for _ in range(6):
        y, x = rng.integers(8, 40, size=2)
        truth[y, x] += rng.uniform(50, 150)
    truth += 2.0
    fm = OT.make_forward_model(gaussian_psf((15, 15), 2.0), shape, 1.0, dtype=np.float64)
    tp = np.zeros(fm.padded_shape)
    tp[fm.valid_slices] = truth
    observed = rng.poisson(np.maximum(fm.op.forward(tp), 0)).astype(np.float64)
    return fm, truth, observed


def test_reconstruction_loss_decreases_and_improves():
    fm, truth, observed = _make_problem()
    res = NL.nlcg_with_operator(observed, fm.op, num_iter=40, eval_interval=1, tol=0.0, slack=0.0)
    assert res.restored.shape == fm.padded_shape and res.pred.shape == fm.data_shape
    assert res.restored.min() >= 0.0
    assert len(res.loss_history) == 40 and res.loss_history[-1] < res.loss_history[0]
    res = NL.nlcg_with_operator(observed, fm.op, num_iter=60)
    assert np.linalg.norm(res.restored[fm.valid_slices] - truth) < np.linalg.norm(observed - truth)


def test_stopping_rules():
    fm, truth, observed = _make_problem()
    stopped = NL.nlcg_with_operator(observed, fm.op, num_iter=500, min_iter=10, tol=0.0)
    assert stopped.converged and stopped.iterations < 500 and stopped.loss_history[-1] <= 0.6
    overfit = NL.nlcg_with_operator(observed, fm.op, num_iter=500, tol=0.0, slack=0.0)
    assert overfit.loss_history[-1] < stopped.loss_history[-1]
    eq17 = NL.nlcg_with_operator(observed, fm.op, num_iter=500, min_iter=10, slack=0.0)
    assert eq17.converged and eq17.iterations < 500
    assert eq17.loss_history[-1] <= 1.2 * overfit.loss_history[-1]
    reg = NL.nlcg_with_operator(observed, fm.op, num_iter=300, min_iter=10, regularizer=NL.Hessian2D(),
                                reg_weight=0.05)
    assert reg.converged and reg.iterations < 300


def test_high_counts_stay_bounded_and_regularizer_smooths():
    rng = np.random.default_rng(3)
    shape = (40, 40)
    truth = np.full(shape, 20.0)
    truth[12:28, 12:28] = 1500.0
    fm = OT.make_forward_model(gaussian_psf((13, 13), 2.0), shape, 1.0, dtype=np.float64)
    tp = np.zeros(fm.padded_shape)
    tp[fm.valid_slices] = truth
    observed = rng.poisson(np.maximum(fm.op.forward(tp) + 100.0, 0)).astype(np.float64)
    res = NL.nlcg_with_operator(observed, fm.op, num_iter=60, background=100.0,
                                init=np.full(fm.padded_shape, 100.0), eval_interval=1, tol=0.0)
    assert np.isfinite(res.restored.max()) and res.restored.max() < 50.0 * observed.max()
    assert res.loss_history[-1] < res.loss_history[0]

    fm, _, observed = _make_problem()
    base = NL.nlcg_with_operator(observed, fm.op, num_iter=60, tol=0.0)
    smooth = NL.nlcg_with_operator(observed, fm.op, num_iter=60, tol=0.0, regularizer=NL.Hessian2D(),
                                   reg_weight=1.0)

    def rough(x):
        a = x[fm.valid_slices]
        return float(np.abs(np.diff(a, axis=0)).sum() + np.abs(np.diff(a, axis=1)).sum())

    assert rough(smooth.restored) < rough(base.restored)


def test_float32_flow_tracks_float64():
    fm, _, observed = _make_problem()
    a = NL.nlcg_with_operator(observed, fm.op, num_iter=15, tol=0.0, slack=0.0)
    fm32 = OT.make_forward_model(gaussian_psf((15, 15), 2.0), (48, 48), 1.0, dtype=np.float32)
    b = NL.nlcg_with_operator(observed.astype(np.float32), fm32.op, num_iter=15, tol=0.0, slack=0.0,
                              dtype=np.float32)
    err = np.linalg.norm(b.restored - a.restored) / np.linalg.norm(a.restored)
    assert err <= 1e-4, err
