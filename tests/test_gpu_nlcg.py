"""NLCG on the GPU against the oracle (oracle/nlcg.py) on the same inputs.

Bars: the helpers (objective, gradient, quadratic form, step length, model at
lambda, I-divergence between iterates) within 1e-5 relative of the float64
oracle; the solver's restored image within 1e-5 relative L2 after one
iteration and 1e-4 after 20 (plan-backed LinearFFTConvolver in 3-D and the
composed ForwardModel operator in 2-D), with the same iteration count,
convergence flag and loss history (1e-4 relative); then the reference's own
end-to-end assertions (tests/test_nlcg.py of the reference) on the GPU solver.
"""

import numpy as np
import pytest
import torch

from oracle import deconv as O
from oracle import nlcg as NL
from oracle import tiles as OT

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


def _math_setup(reg, three_d=False):
    from deconlib_amd import Hessian2D, Hessian3D, LinearFFTConvolver

    rng = np.random.default_rng(0)
    shape = (6, 20, 24) if three_d else (24, 24)
    psf = gaussian_psf((3, 7, 7) if three_d else (9, 9), 1.5)
    blur, oblur = LinearFFTConvolver(psf, shape), O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    g = (rng.random(shape) * 5.0 + 0.5).astype(np.float32)
    s = (rng.random(shape) + 0.5).astype(np.float32)
    if not reg:
        return blur, oblur, g, s, None, None, 0.0
    C, OC = (Hessian3D(), O.Hessian3D()) if three_d else (Hessian2D(), NL.Hessian2D())
    return blur, oblur, g, s, C, OC, 0.3


def test_hessian2d_matches_oracle(gpu):
    from deconlib_amd import Hessian2D

    rng = np.random.default_rng(5)
    f = rng.standard_normal((17, 23)).astype(np.float32)
    H = rng.standard_normal((3, 17, 23)).astype(np.float32)
    assert rel_l2(_host(Hessian2D().forward(f)), NL.Hessian2D().forward(f.astype(np.float64))) < 1e-6
    assert rel_l2(_host(Hessian2D().adjoint(H)), NL.Hessian2D().adjoint(H.astype(np.float64))) < 1e-6


@pytest.mark.parametrize("reg,three_d", [(False, False), (True, False), (True, True)])
def test_helpers_match_oracle(gpu, reg, three_d):
    from deconlib_amd.nlcg import (i_divergence_between, nlcg_gradient, nlcg_hessian_quadform, nlcg_objective,
                                   nlcg_step_length)

    blur, oblur, g, s, C, OC, beta = _math_setup(reg, three_d)
    g64, s64 = g.astype(np.float64), s.astype(np.float64)
    phi = nlcg_objective(s, blur, g, 0.5, C, beta)
    assert abs(phi - NL.objective(s64, oblur, g64, 0.5, OC, beta)) <= 1e-5 * abs(phi)
    grad, aux = nlcg_gradient(s, blur, g, 0.5, C, beta)
    ograd, oaux = NL.gradient(s64, oblur, g64, 0.5, OC, beta)
    assert rel_l2(_host(grad), ograd) <= 1e-5
    d = -_host(grad)
    quad = nlcg_hessian_quadform(s, d, blur, g, 0.5, C, beta, aux=aux)
    oquad = NL.hessian_quadform(s64, d.astype(np.float64), oblur, g64, 0.5, OC, beta)
    assert abs(quad - oquad) <= 1e-5 * abs(oquad)
    lam, m_lam = nlcg_step_length(s, d, s * s, aux[0], blur, g, 0.5, C, beta)
    olam, om = NL.step_length(s64, d.astype(np.float64), s64 * s64, oaux[0], oblur, g64, 0.5, OC, beta)
    assert abs(lam - olam) <= 1e-5 * abs(olam)
    assert rel_l2(_host(m_lam), om) <= 1e-5
    f1 = (s + np.float32(lam) * d) ** 2
    assert abs(i_divergence_between(s * s, f1) - NL.i_divergence_between(s64 * s64, f1.astype(np.float64))) <= \
        1e-4 * abs(NL.i_divergence_between(s64 * s64, f1.astype(np.float64)))


def _problem_3d(seed=11):
    from deconlib_amd import LinearFFTConvolver

    rng = np.random.default_rng(seed)
    shape = (8, 32, 40)
    psf = gaussian_psf((5, 9, 9), 1.6)
    truth = np.full(shape, 2.0)
    for _ in range(12):
        z, y, x = rng.integers(1, 7), rng.integers(4, 28), rng.integers(4, 36)
        truth[z, y, x] += rng.uniform(50, 150)
    oblur = O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    observed = rng.poisson(np.maximum(oblur.forward(truth), 0)).astype(np.float32)
    return LinearFFTConvolver(psf, shape), oblur, observed


def _problem_2d(seed=42):
    from deconlib_amd import make_forward_model

    rng = np.random.default_rng(seed)
    shape = (48, 48)
    truth = np.zeros(shape)
    for _ in range(6):
        y, x = rng.integers(8, 40, size=2)
        truth[y, x] += rng.uniform(50, 150)
    truth += 2.0
    psf = gaussian_psf((15, 15), 2.0)
    ofm = OT.make_forward_model(psf, shape, 1.0, dtype=np.float64)
    tp = np.zeros(ofm.padded_shape)
    tp[ofm.valid_slices] = truth
    observed = rng.poisson(np.maximum(ofm.op.forward(tp), 0)).astype(np.float32)
    return make_forward_model(psf, shape, zoom=1.0), ofm, truth, observed


@pytest.mark.parametrize("kind", ["plan3d", "composed2d"])
def test_solver_matches_oracle(gpu, kind):
    from deconlib_amd import NLCGResult, nlcg_with_operator

    if kind == "plan3d":
        op, oop, obs = _problem_3d()
    else:
        fm, ofm, _, obs = _problem_2d()
        op, oop = fm.op, ofm.op
    for it, tol in ((1, 1e-5), (20, 1e-4)):
        kw = dict(num_iter=it, background=0.5, eval_interval=3, tol=0.0, slack=0.0)
        got = nlcg_with_operator(obs, op, **kw)
        ref = NL.nlcg_with_operator(obs.astype(np.float64), oop, **kw)
        assert isinstance(got, NLCGResult)
        assert got.iterations == ref.iterations and got.converged == ref.converged
        assert rel_l2(_host(got.restored), ref.restored) <= tol, (kind, it, rel_l2(_host(got.restored), ref.restored))
        assert rel_l2(_host(got.pred), ref.pred) <= tol
        np.testing.assert_allclose(got.loss_history, ref.loss_history, rtol=1e-4)
        assert got.full_shape == ref.full_shape


@pytest.mark.parametrize("reg_kind", ["hessian3d", "restart"])
def test_solver_options_match_oracle(gpu, reg_kind):
    from deconlib_amd import Hessian3D, nlcg_with_operator

    op, oop, obs = _problem_3d(seed=12)
    kw = dict(num_iter=12, eval_interval=1, tol=0.0, slack=0.0)
    if reg_kind == "hessian3d":
        got = nlcg_with_operator(obs, op, regularizer=Hessian3D(), reg_weight=0.05, **kw)
        ref = NL.nlcg_with_operator(obs.astype(np.float64), oop, regularizer=O.Hessian3D(), reg_weight=0.05, **kw)
    else:
        init = np.full(obs.shape, 7.0, np.float32)
        got = nlcg_with_operator(obs, op, restart_interval=4, init=init, newton_iters=5, **kw)
        ref = NL.nlcg_with_operator(obs.astype(np.float64), oop, restart_interval=4, init=init.astype(np.float64),
                                    newton_iters=5, **kw)
    assert got.iterations == ref.iterations
    assert rel_l2(_host(got.restored), ref.restored) <= 1e-4
    np.testing.assert_allclose(got.loss_history, ref.loss_history, rtol=1e-4)


def test_callback_and_stopping_rules_match_oracle(gpu):
    from deconlib_amd import nlcg_with_operator

    fm, ofm, _, obs = _problem_2d()
    seen = []
    got = nlcg_with_operator(obs, fm.op, num_iter=50, callback=lambda k, f: seen.append(k) or k == 6)
    ref = NL.nlcg_with_operator(obs.astype(np.float64), ofm.op, num_iter=50, callback=lambda k, f: k == 6)
    assert seen == list(range(7)) and got.iterations == ref.iterations == 7
    assert got.loss_history == pytest.approx(ref.loss_history, rel=1e-4)
    for kw in (dict(num_iter=500, min_iter=10, tol=0.0), dict(num_iter=500, min_iter=10, slack=0.0)):
        got = nlcg_with_operator(obs, fm.op, **kw)
        ref = NL.nlcg_with_operator(obs.astype(np.float64), ofm.op, **kw)
        assert got.converged and ref.converged
        # the stop may move by an iteration where a float32 sum decides it
        assert abs(got.iterations - ref.iterations) <= 1, (kw, got.iterations, ref.iterations)
        assert got.loss_history[-1] == pytest.approx(ref.loss_history[-1], rel=2e-2)


def test_reference_behaviour_on_gpu(gpu):
    """The reference's TestReconstruction assertions, on the GPU solver."""
    from deconlib_amd import Hessian2D, make_forward_model, nlcg_with_operator

    fm, _, truth, observed = _problem_2d()
    res = nlcg_with_operator(observed, fm.op, num_iter=40, eval_interval=1, tol=0.0, slack=0.0)
    assert tuple(res.restored.shape) == fm.padded_shape and tuple(res.pred.shape) == fm.data_shape
    assert float(res.restored.min()) >= 0.0 and res.loss_history[-1] < res.loss_history[0]
    res = nlcg_with_operator(observed, fm.op, num_iter=60)
    restored = _host(res.restored[fm.valid_slices])
    assert np.linalg.norm(restored - truth) < np.linalg.norm(observed - truth)
    stopped = nlcg_with_operator(observed, fm.op, num_iter=500, min_iter=10, tol=0.0)
    overfit = nlcg_with_operator(observed, fm.op, num_iter=500, tol=0.0, slack=0.0)
    assert stopped.converged and stopped.iterations < 500 and stopped.loss_history[-1] <= 0.6
    assert overfit.loss_history[-1] < stopped.loss_history[-1]
    regd = nlcg_with_operator(observed, fm.op, num_iter=300, min_iter=10, regularizer=Hessian2D(), reg_weight=0.05)
    assert regd.converged and regd.iterations < 300
    base = nlcg_with_operator(observed, fm.op, num_iter=60, tol=0.0)
    smooth = nlcg_with_operator(observed, fm.op, num_iter=60, tol=0.0, regularizer=Hessian2D(), reg_weight=1.0)

    def rough(x):
        a = _host(x[fm.valid_slices])
        return float(np.abs(np.diff(a, axis=0)).sum() + np.abs(np.diff(a, axis=1)).sum())

    assert rough(smooth.restored) < rough(base.restored)
    # high counts, background 100, flat init: bounded and decreasing
    rng = np.random.default_rng(3)
    tr = np.full((40, 40), 20.0)
    tr[12:28, 12:28] = 1500.0
    fm2 = make_forward_model(gaussian_psf((13, 13), 2.0), (40, 40), zoom=1.0)
    tp = torch.zeros(fm2.padded_shape, device="cuda")
    tp[fm2.valid_slices] = torch.from_numpy(tr.astype(np.float32)).cuda()
    obs2 = rng.poisson(np.maximum(_host(fm2.op.forward(tp)) + 100.0, 0)).astype(np.float32)
    r2 = nlcg_with_operator(obs2, fm2.op, num_iter=60, background=100.0,
                            init=torch.full(fm2.padded_shape, 100.0, device="cuda"), eval_interval=1, tol=0.0)
    assert np.isfinite(float(r2.restored.max())) and float(r2.restored.max()) < 50.0 * float(obs2.max())
    assert r2.loss_history[-1] < r2.loss_history[0]


def test_nlcg_solver_under_process_tiles(gpu):
    from deconlib_amd import nlcg_solver, process_tiles

    rng = np.random.default_rng(7)
    shape = (16, 40, 40)
    psf = gaussian_psf((5, 7, 7), 1.5)
    observed = rng.poisson(20.0 + 10.0 * rng.random(shape)).astype(np.float32)
    out = process_tiles(observed, psf, zoom=1.0, solve=nlcg_solver(num_iter=10), tile_size=24)
    ref = OT.process_tiles(observed, psf, 1.0, NL.nlcg_solver(num_iter=10), tile_size=24, dtype=np.float64)
    assert out.shape == shape and np.all(out >= 0)
    assert rel_l2(out, ref) <= 1e-4


def test_resident_tiles_equal_host_tiles(gpu):
    """The device-resident tile path (volume uploaded once, tiles cut and cores
    stitched on the GPU) gives exactly the host loop's result (same kernels, same tiles);
    a per-tile callback selects the host loop."""
    from deconlib_amd import erdecon_solver, nlcg_solver, process_tiles

    rng = np.random.default_rng(8)
    shape = (12, 40, 44)
    psf = gaussian_psf((5, 7, 7), 1.5)
    observed = rng.poisson(20.0 + 10.0 * rng.random(shape)).astype(np.float32)
    for solve in (nlcg_solver(num_iter=6), erdecon_solver(num_iter=2)):
        assert solve.accepts_device_tiles
        resident = process_tiles(observed, psf, zoom=1.0, solve=solve, tile_size=24)
        seen = []
        host = process_tiles(observed, psf, zoom=1.0, solve=solve, tile_size=24,
                             on_tile_done=lambda spec, out: seen.append(spec.index) and False)
        assert len(seen) > 1
        np.testing.assert_array_equal(resident, host)

