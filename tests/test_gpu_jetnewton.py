"""JetNewton on the GPU against the oracle (oracle/jetnewton.py) on the same inputs.

Bars: the objective, gradient, penalty weights and exact Hessian product within
1e-5 relative of the float64 oracle (both data terms; the 2-D and 3-D per-axis
Hessians through the fused stencil kernels and a regulariser through its own
forward/adjoint); the active-set rules and the reduced CG's negative-curvature
return exactly; the solver's restored image within 1e-4 relative L2 of the
oracle with the same iteration count and CG/active-set histories for a 2-D
Gaussian and a 3-D Poisson problem with active voxels; the reference's
end-to-end assertions on the GPU solver; the noise-floor probe's kappa^4 scaling.
"""

import numpy as np
import pytest
import torch

from oracle import deconv as O
from oracle import jetnewton as JN

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _host(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def gaussian_psf(shape, sigma):
    grids = np.meshgrid(*[np.arange(n) - (n - 1) / 2.0 for n in shape], indexing="ij")
    r2 = sum(gc ** 2 for gc in grids)
    psf = np.exp(-r2 / (2.0 * sigma ** 2)).astype(np.float32)
    return psf / psf.sum()


class _OwnForwardAdjoint:
    def __init__(self, inner):
        self.inner = inner

    def forward(self, x):
        return self.inner.forward(x)

    def adjoint(self, y):
        return self.inner.adjoint(y)


def _setup(kind):
    from deconlib_amd import AnisotropicHessian2D, AnisotropicHessian3D, LinearFFTConvolver

    rng = np.random.default_rng(0)
    if kind == "2d":
        shape, psf = (24, 24), gaussian_psf((9, 9), 1.5)
        H, OH = AnisotropicHessian2D((1.3, 0.7)), JN.AnisotropicHessian2D((1.3, 0.7))
    elif kind == "3d":
        shape, psf = (6, 20, 24), gaussian_psf((3, 7, 7), 1.5)
        H, OH = AnisotropicHessian3D((0.6, 1.3, 0.7)), JN.AnisotropicHessian3D((0.6, 1.3, 0.7))
    else:
        shape, psf = (6, 20, 24), gaussian_psf((3, 7, 7), 1.5)
        H, OH = _OwnForwardAdjoint(AnisotropicHessian3D((0.6, 1.3, 0.7))), JN.AnisotropicHessian3D((0.6, 1.3, 0.7))
    blur, oblur = LinearFFTConvolver(psf, shape), O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    f = (rng.random(shape) * 5.0 + 0.5).astype(np.float32)
    x = (rng.random(shape) + 0.5).astype(np.float32)
    return blur, oblur, f, x, H, OH


@pytest.mark.parametrize("kind", ["2d", "3d", "stacked"])
@pytest.mark.parametrize("data_term", ["gaussian", "poisson"])
def test_helpers_match_oracle(gpu, kind, data_term):
    from deconlib_amd.jetnewton import jetnewton_gradient, jetnewton_hvp, jetnewton_objective

    blur, oblur, f, x, H, OH = _setup(kind)
    f64, x64 = f.astype(np.float64), x.astype(np.float64)
    args = (0.8, 0.1, 0.3, 1e-2, data_term)
    F = jetnewton_objective(x, blur, f, H, *args)
    oF = JN.objective(x64, oblur, f64, OH, *args)
    assert abs(F - oF) <= 1e-5 * abs(oF), (F, oF)
    grad, (r, w, c, dhw) = jetnewton_gradient(x, blur, f, H, *args)
    ograd, (orr, ow, oc, odhw) = JN.gradient(x64, oblur, f64, OH, *args)
    assert rel_l2(_host(grad), ograd) <= 1e-5
    assert rel_l2(_host(r), orr) <= 1e-5 and rel_l2(_host(w), ow) <= 1e-5 and rel_l2(_host(c), oc) <= 1e-5
    v = np.random.default_rng(2).standard_normal(x.shape).astype(np.float32)
    hv = jetnewton_hvp(v, blur, H, r, w, c, dhw, 0.8, data_term)
    ohv = JN.hvp(v.astype(np.float64), oblur, OH, orr, ow, oc, odhw, 0.8, data_term)
    assert rel_l2(_host(hv), ohv) <= 1e-5, rel_l2(_host(hv), ohv)


def test_active_set_and_reduced_cg_rules(gpu):
    from deconlib_amd.jetnewton import ActiveSetState, identify_active_set, solve_reduced_newton

    x = np.array([[0.0, 0.0, 0.5], [1e-4, 2.0, 0.0]], np.float32)
    g = np.array([[1.0, -1.0, 1.0], [1e-5, 1.0, 0.0]], np.float32)
    active, eps = identify_active_set(x, g, eps_bar=1e-2)
    oactive, oeps = JN.identify_active_set(x.astype(np.float64), g.astype(np.float64), 1e-2)
    assert eps == pytest.approx(oeps, rel=1e-6)
    np.testing.assert_array_equal(_host(active), oactive)
    st, ost = ActiveSetState(), JN.ActiveSetState()
    st.reset(x.shape)
    ost.reset(x.shape)
    for xx in (x, x + 1.0):
        a, _ = st.update(xx, g)
        oa, _ = ost.update(xx.astype(np.float64), g.astype(np.float64))
        np.testing.assert_array_equal(_host(a), oa)
        np.testing.assert_array_equal(_host(st.frozen), ost.frozen)
    gg = torch.from_numpy(np.random.default_rng(1).standard_normal((6, 6)).astype(np.float32)).cuda()
    d, n, hit = solve_reduced_newton(lambda v: -v, gg, torch.zeros((6, 6), dtype=torch.bool, device="cuda"))
    assert hit and n == 1
    np.testing.assert_array_equal(_host(d), -_host(gg))


def _problem_2d(seed=42, sigma=0.02):
    from deconlib_amd import LinearFFTConvolver

    rng = np.random.default_rng(seed)
    shape = (48, 48)
    truth = np.zeros(shape)
    for _ in range(6):
        y, x = rng.integers(8, 40, size=2)
        truth[y, x] += rng.uniform(0.5, 1.0)
    truth += 0.02
    psf = gaussian_psf((15, 15), 2.0)
    oblur = O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    clean = oblur.forward(truth)
    obs = np.maximum(clean + sigma * np.random.default_rng(seed).standard_normal(clean.shape), 0.0).astype(np.float32)
    return LinearFFTConvolver(psf, shape), oblur, truth, obs


def _problem_3d_poisson(seed=9):
    from deconlib_amd import LinearFFTConvolver

    rng = np.random.default_rng(seed)
    shape = (8, 32, 32)
    truth = np.zeros(shape)
    for _ in range(6):
        truth[rng.integers(2, 6), rng.integers(6, 26), rng.integers(6, 26)] += rng.uniform(20.0, 50.0)
    truth += 0.5
    psf = gaussian_psf((5, 9, 9), 1.5)
    oblur = O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    obs = rng.poisson(oblur.forward(truth) + 2.0).astype(np.float32)
    return LinearFFTConvolver(psf, shape), oblur, obs


@pytest.mark.parametrize("kind", ["gaussian2d", "poisson3d"])
def test_solver_matches_oracle(gpu, kind):
    from deconlib_amd import AnisotropicHessian2D, AnisotropicHessian3D, JetNewtonResult, jetnewton_with_operator

    if kind == "gaussian2d":
        op, oop, _, obs = _problem_2d()
        H, OH = AnisotropicHessian2D((1.0, 1.0)), JN.AnisotropicHessian2D((1.0, 1.0))
        kw = dict(s0=0.02, beta=20.0, eta=1.0, data_term="gaussian", tol=1e-3)
    else:
        op, oop, obs = _problem_3d_poisson()
        H, OH = AnisotropicHessian3D((0.5, 1.0, 1.0)), JN.AnisotropicHessian3D((0.5, 1.0, 1.0))
        kw = dict(s0=1.5, beta=0.1, eta=1.0, data_term="poisson", background=2.0, cg_max_steps=60)
    for it, tol in ((1, 1e-5), (4, 1e-4)):
        got = jetnewton_with_operator(obs, op, H, num_iter=it, eval_interval=1, **kw)
        ref = JN.jetnewton_with_operator(obs.astype(np.float64), oop, OH, num_iter=it, eval_interval=1, **kw)
        assert isinstance(got, JetNewtonResult)
        err = rel_l2(_host(got.restored), ref.restored)
        print(f"{kind} it={it}: restored rel-L2 {err:.3g}, cg {got.cg_iterations_history} vs "
              f"{ref.cg_iterations_history}, |I| {got.active_set_size_history} vs {ref.active_set_size_history}")
        assert got.iterations == ref.iterations and got.converged == ref.converged
        assert got.cg_iterations_history == ref.cg_iterations_history
        assert got.active_set_size_history == ref.active_set_size_history
        assert err <= tol, (kind, it, err)
        assert rel_l2(_host(got.pred), ref.pred) <= tol
        np.testing.assert_allclose(got.loss_history, ref.loss_history, rtol=1e-5)
        np.testing.assert_allclose(got.idiv_history, ref.idiv_history, rtol=1e-4)
        np.testing.assert_allclose(got.curvature_term_history, ref.curvature_term_history, rtol=1e-4)


def test_reference_behaviour_on_gpu(gpu):
    """The reference's TestSolver assertions, on the GPU solver."""
    from deconlib_amd import AnisotropicHessian2D, LinearFFTConvolver, jetnewton_with_operator

    _, _, truth, observed = _problem_2d()
    blur = LinearFFTConvolver(gaussian_psf((15, 15), 2.0), (48, 48))
    H = AnisotropicHessian2D((1.0, 1.0))
    res = jetnewton_with_operator(observed, blur, H, 0.02, beta=20.0, eta=1.0, data_term="gaussian", num_iter=60,
                                  tol=1e-3)
    restored = _host(res.restored)
    assert res.converged and res.iterations < 60 and restored.min() >= 0.0
    assert np.linalg.norm(restored - truth) < np.linalg.norm(observed - truth)
    res = jetnewton_with_operator(observed, blur, H, 0.02, beta=20.0, eta=1.0, data_term="gaussian", num_iter=20,
                                  tol=1e-3, eval_interval=1)
    n = len(res.loss_history)
    assert n > 0 and len(res.idiv_history) == n and all(v > 0.0 for v in res.curvature_term_history)
    rng = np.random.default_rng(7)
    truth = np.zeros((32, 32))
    for _ in range(4):
        y, x = rng.integers(6, 26, size=2)
        truth[y, x] += rng.uniform(20.0, 50.0)
    truth += 2.0
    blur = LinearFFTConvolver(gaussian_psf((15, 15), 2.0), (32, 32))
    clean = _host(blur.forward(truth.astype(np.float32)))
    obs = np.maximum(rng.poisson(np.maximum(clean, 0)).astype(np.float32), 0)
    res = jetnewton_with_operator(obs, blur, H, 1.5, beta=0.1, eta=1.0, data_term="poisson", num_iter=15, tol=1e-4,
                                  eval_interval=1, cg_max_steps=150)
    losses = np.array(res.loss_history)
    assert losses[-1] < losses[0] and np.sum(np.diff(losses) >= -1e-6) < len(losses) - 2


def test_noise_floor_scales_with_kappa(gpu):
    from deconlib_amd import AnisotropicHessian3D
    from deconlib_amd.jetnewton import estimate_penalty_noise_floor

    small = estimate_penalty_noise_floor(AnisotropicHessian3D((1.0, 1.0, 1.0)), (16, 20, 20), n_trials=4, seed=0)
    large = estimate_penalty_noise_floor(AnisotropicHessian3D((10.0, 10.0, 10.0)), (16, 20, 20), n_trials=4, seed=0)
    ratio = large["curvature"]["median"] / small["curvature"]["median"]
    assert 5e3 < ratio < 2e4
