"""JetNewton oracle pinned on the CPU (the GPU solver is tested in test_gpu_jetnewton.py).

Pins of oracle/jetnewton.py: the gradient against central differences of the
objective and the exact Hessian product against central differences of the
gradient (Gaussian and Poisson data terms); the product's symmetry; the
regulariser's dense Hessian symmetric and indefinite; the reduced CG against a
dense solve on a positive-definite case and its first-step negative-curvature
return; the active-set rules on hand-made inputs; the per-axis curvature
against its continuum value on refined grids; and the reference's own
end-to-end assertions (tests/test_jetnewton.py of the reference) on the oracle.
"""

import numpy as np
import pytest

from oracle import deconv as O
from oracle import jetnewton as JN


def gaussian_psf(shape, sigma):
    grids = np.meshgrid(*[np.arange(n) - (n - 1) / 2.0 for n in shape], indexing="ij")
    r2 = sum(gc ** 2 for gc in grids)
    psf = np.exp(-r2 / (2.0 * sigma ** 2)).astype(np.float32)
    return psf / psf.sum()


def _setup(shape=(24, 24), kappa=(1.3, 0.7)):
    rng = np.random.default_rng(0)
    blur = O.LinearFFTConvolver(gaussian_psf((9, 9), 1.5), shape, dtype=np.float64)
    return blur, rng.random(shape) * 5.0 + 0.5, rng.random(shape) + 0.5, JN.AnisotropicHessian2D(kappa)


@pytest.mark.parametrize("data_term", ["gaussian", "poisson"])
def test_gradient_and_hvp_match_finite_differences(data_term):
    blur, f, x, H = _setup()
    args = (0.8, 0.1, 0.3, 1e-2, data_term)
    grad, (r, w, c, dhw) = JN.gradient(x, blur, f, H, *args)
    v = np.random.default_rng(1).standard_normal(x.shape)
    h = 1e-5
    fd = (JN.objective(x + h * v, blur, f, H, *args) - JN.objective(x - h * v, blur, f, H, *args)) / (2 * h)
    analytic = float(np.sum(grad * v))
    assert abs(fd - analytic) <= 1e-6 * (abs(analytic) + 1.0)
    hv = JN.hvp(v, blur, H, r, w, c, dhw, 0.8, data_term)
    gp, _ = JN.gradient(x + h * v, blur, f, H, *args)
    gm, _ = JN.gradient(x - h * v, blur, f, H, *args)
    np.testing.assert_allclose(hv, (gp - gm) / (2 * h), rtol=1e-5, atol=1e-6 * np.abs(hv).max())
    u = np.random.default_rng(3).standard_normal(x.shape)
    hu = JN.hvp(u, blur, H, r, w, c, dhw, 0.8, data_term)
    assert abs(np.sum(v * hu) - np.sum(u * hv)) <= 1e-10 * (abs(np.sum(v * hu)) + 1.0)
    # the s0 chain-rule factor: the data gradient scales with s0
    g2, _ = JN.gradient(x, blur, f, H, 2.0, *args[1:])
    assert np.max(np.abs(g2 - grad)) > 1e-3


def test_regularizer_hessian_indefinite_and_symmetric():
    rng = np.random.default_rng(0)
    x = rng.random((6, 6)) * 2.0
    x[3, 3] += 3.0
    H = JN.AnisotropicHessian2D((1.0, 1.0))
    r, _, w, c = JN.penalty_weights(x, H, 0.1, 0.05)
    n = x.size
    D = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        q = H.forward(e.reshape(x.shape))
        m = c * np.sum(r * q, axis=0)
        D[:, i] = H.adjoint(w[None] * q + 2.0 * m[None] * r).ravel()
    assert np.max(np.abs(D - D.T)) < 1e-12
    assert np.linalg.eigvalsh((D + D.T) / 2).min() < -1e-3


def test_reduced_cg_dense_solve_and_negative_curvature():
    shape = (6, 7)
    _, f, x, H = _setup(shape)
    blur = O.LinearFFTConvolver(gaussian_psf((3, 3), 0.5), shape, dtype=np.float64)  # well conditioned
    grad, (r, w, c, dhw) = JN.gradient(x, blur, f, H, 0.5, 0.0, 1e-4, 1.0, "gaussian")
    n = x.size
    A = np.stack([JN.hvp(np.eye(n)[i].reshape(shape), blur, H, r, w, c, dhw, 0.5, "gaussian").ravel()
                  for i in range(n)], axis=1)
    assert np.linalg.eigvalsh((A + A.T) / 2).min() > 0  # data-dominated: positive definite
    active = np.zeros(shape, bool)
    active[0, :3] = True
    free = ~active.ravel()
    Af = A[np.ix_(free, free)]

    def solve(gr):
        return JN.solve_reduced_newton(lambda v: JN.hvp(v, blur, H, r, w, c, dhw, 0.5, "gaussian"), gr, active,
                                       max_steps=500, eps=1e-300)

    # the absolute forcing tolerance min(0.5, sqrt(r0)) r0 bounds the residual ...
    d, steps, hit = solve(grad)
    r0 = np.linalg.norm(grad.ravel()[free])
    res = np.linalg.norm(Af @ d.ravel()[free] + grad.ravel()[free])
    assert not hit and steps >= 1 and res <= min(0.5, r0 ** 0.5) * r0 * (1 + 1e-9)
    assert np.all(d[active] == 0) and float(np.sum(grad * d)) < 0
    # ... and is relative sqrt(r0) for a small right-hand side: CG then reaches the dense solution
    gs = grad * (1e-10 / np.linalg.norm(grad))
    d, steps, hit = solve(gs)
    exact = np.linalg.solve(Af, -gs.ravel()[free])
    assert not hit and np.linalg.norm(d.ravel()[free] - exact) <= 1e-3 * np.linalg.norm(exact)
    g = np.random.default_rng(1).standard_normal((6, 6))
    d, steps, hit = JN.solve_reduced_newton(lambda v: -v, g, np.zeros((6, 6), bool))
    assert hit and steps == 1
    np.testing.assert_array_equal(d, -g)


def test_active_set_rules():
    x = np.array([[0.0, 0.0, 0.5], [1e-4, 2.0, 0.0]])
    g = np.array([[1.0, -1.0, 1.0], [1e-5, 1.0, 0.0]])
    active, eps = JN.identify_active_set(x, g, eps_bar=1e-2)
    diff = x - np.maximum(x - g, 0.0)
    assert eps == pytest.approx(min(1e-2, np.linalg.norm(diff)))
    np.testing.assert_array_equal(active, (x <= eps) & (g > 0))
    st = JN.ActiveSetState()
    st.reset(x.shape)
    act, _ = st.update(x, g, freeze_tau=1e-3, freeze_delta=1e-6)
    np.testing.assert_array_equal(st.frozen, (x < 1e-3) & (g > 1e-6))
    x2 = x + 1.0  # nothing would be active now, the frozen voxels stay
    act2, _ = st.update(x2, g)
    np.testing.assert_array_equal(act2, st.frozen)
    st.reset(x.shape)
    assert not st.frozen.any()


def test_per_axis_curvature_matches_continuum():
    fov, ell, sig, amp = 20.0, (1.5, 0.8), (2.5, 1.8), 5.0
    analytic = (ell[0] ** 2 * amp / sig[0] ** 2) ** 2 + (ell[1] ** 2 * amp / sig[1] ** 2) ** 2
    errs = []
    for hy, hx in [(0.5, 0.5), (0.5, 0.25), (0.25, 0.5), (0.2, 0.1)]:
        ny, nx = int(round(fov / hy)), int(round(fov / hx))
        Y, X = np.meshgrid((np.arange(ny) - ny / 2.0) * hy, (np.arange(nx) - nx / 2.0) * hx, indexing="ij")
        f = amp * np.exp(-(Y ** 2 / (2 * sig[0] ** 2) + X ** 2 / (2 * sig[1] ** 2)))
        Hx = JN.AnisotropicHessian2D.from_lengths(ell, (hy, hx)).forward(f)
        errs.append(abs(np.sum(Hx * Hx, axis=0)[ny // 2, nx // 2] - analytic) / analytic)
    assert all(e < 0.05 for e in errs) and errs[-1] < errs[0]


def _problem(seed=42, sigma=0.02, gain=1.0, shape=(48, 48), rng_seed=None):
    r = np.random.default_rng(seed)
    truth = np.zeros(shape)
    for _ in range(6):
        y, x = r.integers(8, shape[0] - 8, size=2)
        truth[y, x] += r.uniform(0.5, 1.0)
    truth = (truth + 0.02) * gain
    blur = O.LinearFFTConvolver(gaussian_psf((15, 15), 2.0), shape, dtype=np.float64)
    clean = blur.forward(truth)
    noise = np.random.default_rng(seed if rng_seed is None else rng_seed).standard_normal(clean.shape)
    return blur, truth, np.maximum(clean + sigma * gain * noise, 0.0)


def test_solver_reference_assertions():
    blur, truth, observed = _problem()
    H = JN.AnisotropicHessian2D((1.0, 1.0))
    res = JN.jetnewton_with_operator(observed, blur, H, 0.02, beta=20.0, eta=1.0, data_term="gaussian", num_iter=60,
                                     tol=1e-3)
    assert res.converged and res.iterations < 60 and res.restored.min() >= 0.0
    assert np.linalg.norm(res.restored - truth) < np.linalg.norm(observed - truth)
    res = JN.jetnewton_with_operator(observed, blur, H, 0.02, beta=20.0, eta=1.0, data_term="gaussian", num_iter=20,
                                     tol=1e-3, eval_interval=1)
    n = len(res.loss_history)
    assert n > 0 and len(res.idiv_history) == n and len(res.curvature_term_history) == n
    assert all(v > 0 for v in res.curvature_term_history)
    # gain invariance of the dimensionless reconstruction
    xt = {}
    for gain in (1.0, 5.0):
        b, _, obs = _problem(seed=1, gain=gain, rng_seed=123)
        r5 = JN.jetnewton_with_operator(obs, b, H, 0.02 * gain, beta=20.0, eta=1.0, data_term="gaussian",
                                        num_iter=80, tol=1e-3)
        assert r5.converged
        xt[gain] = r5.restored / (0.02 * gain)
    assert np.abs(xt[1.0] - xt[5.0]).mean() / (np.abs(xt[1.0]).mean() + 1e-8) < 0.02


def test_poisson_no_stall():
    rng = np.random.default_rng(7)
    shape = (32, 32)
    truth = np.zeros(shape)
    for _ in range(4):
        y, x = rng.integers(6, 26, size=2)
        truth[y, x] += rng.uniform(20.0, 50.0)
    truth += 2.0
    blur = O.LinearFFTConvolver(gaussian_psf((15, 15), 2.0), shape, dtype=np.float64)
    observed = rng.poisson(blur.forward(truth)).astype(np.float64)
    res = JN.jetnewton_with_operator(observed, blur, JN.AnisotropicHessian2D((1.0, 1.0)), 1.5, beta=0.1, eta=1.0,
                                     data_term="poisson", num_iter=15, tol=1e-4, eval_interval=1, cg_max_steps=150)
    losses = np.array(res.loss_history)
    assert losses[-1] < losses[0] and np.sum(np.diff(losses) >= -1e-6) < len(losses) - 2
