"""ER-Decon on the GPU against the oracle (oracle/erdecon.py) on the same inputs.

Bars: the functional, gradient and Gauss-Newton product within 1e-5 relative
of the float64 oracle (both data terms, the weight floor, the intensity term;
Hessian2D and Hessian3D through the fused stencil kernels and a regulariser
through its own forward/adjoint); the solver's restored image within 1e-5
relative L2 after one outer iteration and 1e-4 after five, with the same
iteration count and objective / misfit histories (1e-5 relative), for the
composed 2-D ForwardModel operator and a plan-backed 3-D LinearFFTConvolver
with the Poisson data term; then the reference's end-to-end assertions on the
GPU solver and erdecon_solver under process_tiles.
"""

import numpy as np
import pytest

from oracle import deconv as O
from oracle import erdecon as ER
from oracle import tiles as OT
from oracle.nlcg import Hessian2D as OHessian2D

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
    """A regulariser the solver only knows through forward/adjoint (the stacked path)."""

    def __init__(self, inner):
        self.inner = inner

    def forward(self, x):
        return self.inner.forward(x)

    def adjoint(self, y):
        return self.inner.adjoint(y)


def _setup(kind):
    from deconlib_amd import Hessian2D, Hessian3D, LinearFFTConvolver

    rng = np.random.default_rng(0)
    if kind == "2d":
        shape, psf, H, OH = (24, 24), gaussian_psf((9, 9), 1.5), Hessian2D(), OHessian2D()
    elif kind == "3d":
        shape, psf, H, OH = (6, 20, 24), gaussian_psf((3, 7, 7), 1.5), Hessian3D(0.7), O.Hessian3D(0.7)
    else:
        shape, psf, H, OH = (6, 20, 24), gaussian_psf((3, 7, 7), 1.5), _OwnForwardAdjoint(Hessian3D()), O.Hessian3D()
    blur, oblur = LinearFFTConvolver(psf, shape), O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    f = rng.random(shape).astype(np.float32)
    s = (rng.random(shape) + 0.5).astype(np.float32)
    return blur, oblur, f, s, H, OH


@pytest.mark.parametrize("kind", ["2d", "3d", "stacked"])
@pytest.mark.parametrize("data_term,floor,iw", [("gaussian", 0.0, 0.0), ("poisson", 0.05, 0.0),
                                                ("gaussian", 0.0, 0.7)])
def test_helpers_match_oracle(gpu, kind, data_term, floor, iw):
    from deconlib_amd.erdecon import erdecon_gn_hvp, erdecon_gradient, erdecon_objective

    blur, oblur, f, s, H, OH = _setup(kind)
    f64, s64 = f.astype(np.float64), s.astype(np.float64)
    args = (0.05, 0.1, 1e-2, data_term, floor, iw)
    phi = erdecon_objective(s, blur, f, H, *args)
    ophi = ER.objective(s64, oblur, f64, OH, 0.05, 0.1, 1e-2, data_term, floor, iw)
    assert abs(phi - ophi) <= 1e-5 * abs(ophi), (phi, ophi)
    grad, (w, dhw) = erdecon_gradient(s, blur, f, H, *args)
    ograd, (ow, odhw) = ER.gradient(s64, oblur, f64, OH, 0.05, 0.1, 1e-2, data_term, floor, iw)
    assert rel_l2(_host(grad), ograd) <= 1e-5, rel_l2(_host(grad), ograd)
    assert rel_l2(_host(w), ow) <= 1e-5
    v = np.random.default_rng(2).standard_normal(s.shape).astype(np.float32)
    hv = erdecon_gn_hvp(s, v, blur, H, w, dhw, iw)
    ohv = ER.gn_hvp(s64, v.astype(np.float64), oblur, OH, ow, odhw, iw)
    assert rel_l2(_host(hv), ohv) <= 1e-5, rel_l2(_host(hv), ohv)


def _problem_2d(seed=42):
    from deconlib_amd import make_forward_model

    rng = np.random.default_rng(seed)
    shape = (48, 48)
    truth = np.zeros(shape)
    for _ in range(6):
        y, x = rng.integers(8, 40, size=2)
        truth[y, x] += rng.uniform(0.5, 1.0)
    truth += 0.02
    psf = gaussian_psf((15, 15), 2.0)
    ofm = OT.make_forward_model(psf, shape, 1.0, dtype=np.float64)
    tp = np.zeros(ofm.padded_shape)
    tp[ofm.valid_slices] = truth
    clean = ofm.op.forward(tp)
    observed = np.maximum(clean + 0.01 * rng.standard_normal(clean.shape), 0.0).astype(np.float32)
    return make_forward_model(psf, shape, zoom=1.0), ofm, truth, observed


def _problem_3d(seed=13):
    from deconlib_amd import LinearFFTConvolver

    rng = np.random.default_rng(seed)
    shape = (8, 32, 40)
    psf = gaussian_psf((5, 9, 9), 1.6)
    truth = np.full(shape, 2.0)
    for _ in range(12):
        z, y, x = rng.integers(1, 7), rng.integers(4, 28), rng.integers(4, 36)
        truth[z, y, x] += rng.uniform(50, 150)
    oblur = O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    observed = rng.poisson(np.maximum(oblur.forward(truth) + 5.0, 0)).astype(np.float32)
    return LinearFFTConvolver(psf, shape), oblur, observed


@pytest.mark.parametrize("kind", ["composed2d", "plan3d_poisson"])
def test_solver_matches_oracle(gpu, kind):
    from deconlib_amd import ERDeconResult, erdecon_with_operator

    if kind == "composed2d":
        fm, ofm, _, obs = _problem_2d()
        op, oop, kw = fm.op, ofm.op, dict(reg_weight=0.05)
    else:
        op, oop, obs = _problem_3d()
        kw = dict(reg_weight=0.02, data_term="poisson", background=5.0)
    for it, tol in ((1, 1e-5), (5, 1e-4)):
        kw.update(num_iter=it, eval_interval=1, newton_tol=0.0)
        got = erdecon_with_operator(obs, op, **kw)
        ref = ER.erdecon_with_operator(obs.astype(np.float64), oop, **kw)
        assert isinstance(got, ERDeconResult)
        err = rel_l2(_host(got.restored), ref.restored)
        print(f"{kind} it={it}: restored rel-L2 {err:.3g}")
        assert got.iterations == ref.iterations and got.converged == ref.converged
        assert err <= tol, (kind, it, err)
        assert rel_l2(_host(got.pred), ref.pred) <= tol
        np.testing.assert_allclose(got.loss_history, ref.loss_history, rtol=1e-5)
        np.testing.assert_allclose(got.data_misfit_history, ref.data_misfit_history, rtol=1e-4)
        assert got.data_scale == pytest.approx(ref.data_scale, rel=1e-7)


def test_reference_behaviour_on_gpu(gpu):
    """The reference's TestReconstruction assertions, on the GPU solver."""
    from deconlib_amd import erdecon_with_operator

    fm, _, truth, observed = _problem_2d()
    res = erdecon_with_operator(observed, fm.op, num_iter=30, eval_interval=1, tol=0.0)
    assert tuple(res.restored.shape) == fm.padded_shape and tuple(res.pred.shape) == fm.data_shape
    assert float(res.restored.min()) >= 0.0 and res.loss_history[-1] < res.loss_history[0]
    for term in ("gaussian", "poisson"):
        r = erdecon_with_operator(observed, fm.op, num_iter=50, reg_weight=0.02, data_term=term)
        restored = _host(r.restored[fm.valid_slices])
        assert restored.min() >= 0.0
        assert np.linalg.norm(restored - truth) < np.linalg.norm(observed - truth)

    def rough(x):
        a = _host(x[fm.valid_slices])
        return float(np.abs(np.diff(a, axis=0)).sum() + np.abs(np.diff(a, axis=1)).sum())

    weak = erdecon_with_operator(observed, fm.op, num_iter=40, reg_weight=0.01, tol=0.0)
    strong = erdecon_with_operator(observed, fm.op, num_iter=40, reg_weight=0.5, tol=0.0)
    assert rough(strong.restored) < rough(weak.restored)
    conv = erdecon_with_operator(observed, fm.op, num_iter=500, min_iter=5, reg_weight=0.05)
    assert conv.converged and conv.iterations < 500


def test_erdecon_solver_under_process_tiles(gpu):
    from deconlib_amd import erdecon_solver, process_tiles

    rng = np.random.default_rng(7)
    shape = (16, 40, 40)
    psf = gaussian_psf((5, 7, 7), 1.5)
    observed = (rng.random(shape) * 0.5 + 0.25).astype(np.float32)
    out = process_tiles(observed, psf, zoom=1.0, solve=erdecon_solver(num_iter=3), tile_size=24)
    ref = OT.process_tiles(observed, psf, 1.0, ER.erdecon_solver(num_iter=3), tile_size=24, dtype=np.float64)
    assert out.shape == shape and np.all(out >= 0)
    assert rel_l2(out, ref) <= 1e-4, rel_l2(out, ref)
