"""GPU parity of Hessian3D and PDHG-Hessian against the fp32 oracle.

Bars: Hessian forward/adjoint relative L2 <= 1e-6; PDHG relative L2 <= 1e-5
after one iteration and <= 1e-4 after 100 iterations; tau/sigma histories
within 1e-4 relative (the GPU and the oracle both keep them in float64).
"""

import numpy as np
import pytest
import torch

from oracle import deconv as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape,r", [((7, 9, 11), 1.0), ((1, 5, 6), 0.5), ((12, 2, 3), 2.0), ((16, 40, 48), 0.43)])
def test_hessian_matches_oracle(gpu, shape, r):
    from deconlib_amd import Hessian3D

    rng = np.random.default_rng(31)
    f = rng.standard_normal(shape).astype(np.float32)
    y = rng.standard_normal((6,) + shape).astype(np.float32)
    H, OH = Hessian3D(r), O.Hessian3D(r)
    assert rel_l2(H.forward(f).cpu().numpy(), OH.forward(f.astype(np.float64))) < 1e-6
    assert rel_l2(H.adjoint(y).cpu().numpy(), OH.adjoint(y.astype(np.float64))) < 1e-6


def _workload(shape=(12, 40, 40), seed=32, n_blobs=20):
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS, blob_volume

    rng = np.random.default_rng(seed)
    psf = widefield_psf(shape, 0.085, 0.1, WIDEFIELD_OPTICS)
    obj = blob_volume(shape, n_blobs, rng)
    A = O.LinearFFTConvolver(psf, shape, dtype=np.float64)
    obs = rng.poisson(np.maximum(A.forward(obj), 0)).astype(np.float32)
    return psf, obs


@pytest.mark.parametrize("num_iter,bar", [(1, 1e-5), (100, 1e-4)])
def test_pdhg_matches_oracle(gpu, num_iter, bar):
    from deconlib_amd import solve_pdhg_mlx

    print(f"pdhg {num_iter} iterations", flush=True)
    psf, obs = _workload()
    ref = O.solve_pdhg(obs, psf, alpha=1e-3, num_iter=num_iter, background=1.0, tol=0.0)
    res = solve_pdhg_mlx(torch.from_numpy(obs).cuda(), psf, alpha=1e-3, num_iter=num_iter, background=1.0,
                         tol=0.0)
    assert res.iterations == ref.iterations == num_iter
    assert rel_l2(res.restored.cpu().numpy(), ref.restored) <= bar
    np.testing.assert_allclose(res.tau_history, ref.tau_history, rtol=bar)
    np.testing.assert_allclose(res.sigma_history, ref.sigma_history, rtol=bar)
    assert len(res.loss_history) == len(ref.loss_history)
    np.testing.assert_allclose(res.loss_history, ref.loss_history, rtol=1e-4)


def test_pdhg_binned_operator(gpu):
    """bin_factors=2: A = FractionalAreaDownsample(2) o LinearFFTConvolver (the binned recipe)."""
    from deconlib_amd import solve_pdhg_mlx

    rng = np.random.default_rng(33)
    psf = (rng.random((8, 24, 24)) + 0.01).astype(np.float32)
    obs = rng.poisson(30 * rng.random((4, 12, 12)) + 5).astype(np.float32)
    ref = O.solve_pdhg(obs, psf, alpha=1e-3, num_iter=30, bin_factors=2, tol=0.0)
    res = solve_pdhg_mlx(torch.from_numpy(obs).cuda(), psf, alpha=1e-3, num_iter=30, bin_factors=2, tol=0.0)
    assert tuple(res.restored.shape) == (8, 24, 24)
    assert rel_l2(res.restored.cpu().numpy(), ref.restored) <= 1e-4


def test_pdhg_convergence_stop_matches_oracle(gpu):
    """A loose tolerance makes both stop early at the same iteration."""
    from deconlib_amd import solve_pdhg_mlx

    psf, obs = _workload((8, 32, 32), seed=34, n_blobs=6)
    kw = dict(alpha=1e-2, num_iter=200, tol=2e-3, min_iter=5, patience=3)
    ref = O.solve_pdhg(obs, psf, **kw)
    res = solve_pdhg_mlx(torch.from_numpy(obs).cuda(), psf, **kw)
    assert ref.converged and res.converged
    assert res.iterations == ref.iterations
    assert len(res.tau_history) == len(ref.tau_history)
    assert rel_l2(res.restored.cpu().numpy(), ref.restored) <= 1e-4


def test_pdhg_errors_like_reference(gpu):
    from deconlib_amd import solve_pdhg_mlx

    obs = torch.ones((4, 8, 8), device="cuda")
    psf = np.ones((4, 8, 8), np.float32)
    with pytest.raises(ValueError):
        solve_pdhg_mlx(obs, psf, bin_factors=2, sampling_factors=2.0)
    with pytest.raises(NotImplementedError):
        solve_pdhg_mlx(obs, psf, icf_sigmas=(1.0, 1.0, 1.0))
    with pytest.raises(ValueError):
        solve_pdhg_mlx(obs, psf, bin_factors=3)


@pytest.mark.parametrize("shape", [(37, 21, 70), (1, 13, 130), (1, 1, 77), (16, 8, 64), (18, 9, 65), (65, 9, 66),
                                   (130, 17, 40), (128, 8, 64)])
@pytest.mark.parametrize("l12", [0, 1])
def test_fused_dual_primal_equals_separate_kernels(gpu, shape, l12):
    """dh_pd_dual_primal equals dh_pd_dual followed by dh_pd_primal: y2_new and x_new bit
    for bit, the DX2 / XX / XNXN sums to float64 rounding (one to three 64-plane z chunks,
    a last chunk of one or two planes, partial tiles in x and y, 1-D and 2-D volumes,
    clipped and L1_2-projected duals)."""
    import ctypes as C

    from deconlib_amd import _native as N

    print(f"fused dual/primal {shape} l12={l12}", flush=True)
    rng = np.random.default_rng(40 + l12)
    nz, ny, nx = shape
    dev = "cuda"
    x = torch.from_numpy(rng.random(shape).astype(np.float32)).to(dev)
    xp = torch.from_numpy(rng.random(shape).astype(np.float32)).to(dev)
    y2 = torch.from_numpy(0.02 * rng.standard_normal((6,) + shape).astype(np.float32)).to(dev)
    g = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(dev)
    w = (C.c_float * 6)(1.7, 1.0, 1.0, 1.3, 1.3, 2 ** 0.5)
    sc0 = np.zeros(16, np.float64)
    sc0[0], sc0[1], sc0[2] = 0.3, 0.7, 0.9  # tau, sigma, theta
    part = torch.empty(int(N.fns["dh_pd_partials"](nz, ny, nx)), dtype=torch.float64, device=dev)
    st = N.stream_ptr()
    alpha = 0.01
    # separate kernels
    sc_a = torch.from_numpy(sc0.copy()).to(dev)
    y2_a = y2.clone()
    xn_a = torch.empty_like(x)
    N.call("dh_pd_dual", y2_a.data_ptr(), x.data_ptr(), xp.data_ptr(), nz, ny, nx, sc_a.data_ptr(), w, alpha,
           l12, st)
    N.call("dh_pd_primal", xn_a.data_ptr(), x.data_ptr(), g.data_ptr(), y2_a.data_ptr(), nz, ny, nx,
           sc_a.data_ptr(), w, part.data_ptr(), st)
    # fused
    sc_b = torch.from_numpy(sc0.copy()).to(dev)
    y2_b = torch.full_like(y2, float("nan"))
    xn_b = torch.full_like(x, float("nan"))
    N.call("dh_pd_dual_primal", y2_b.data_ptr(), xn_b.data_ptr(), x.data_ptr(), xp.data_ptr(), y2.data_ptr(),
           g.data_ptr(), nz, ny, nx, sc_b.data_ptr(), w, alpha, l12, part.data_ptr(), st)
    torch.cuda.synchronize()
    assert torch.equal(y2_a, y2_b)
    assert torch.equal(xn_a, xn_b)
    a, b = sc_a.cpu().numpy(), sc_b.cpu().numpy()
    # the same float32 terms summed in float64 over longer z chunks: equal to rounding
    np.testing.assert_allclose(b[7:10], a[7:10], rtol=1e-12)
    # the clip (or the projection onto the ball) was active somewhere
    mag = y2_b.norm(dim=0) if l12 else y2_b.abs()
    assert float(mag.max()) <= alpha * (1 + 1e-5)
    assert float((mag >= alpha * 0.999).float().mean()) > 0


@pytest.mark.parametrize("shape,off", [((16, 24, 32), 0), ((5, 7, 33), 0), ((6, 10, 40), 1), ((64, 128, 128), 0)])
@pytest.mark.parametrize("loss", [0, 1])
def test_y1_prologue_equals_separate_kernel(gpu, shape, off, loss):
    """DH_PRO_PD_Y1: the adjoint's first pass runs the y1 dual step.  y1' and A^T y1'
    equal dh_pd_y1 followed by the plain adjoint bit for bit, through dh_conv_apply and
    dh_batch_conv_apply (row pairs as 8-byte loads; odd rows and operands 4 bytes off an
    8-byte boundary take the scalar loads; data with zeros takes the D ~ 0 branch)."""
    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd import _native as N

    print(f"y1 prologue {shape} off={off} loss={loss}", flush=True)
    rng = np.random.default_rng(90 + loss)
    n = int(np.prod(shape))
    dev = "cuda"

    def arr(v):
        t = torch.empty(n + off, dtype=torch.float32, device=dev)[off:].view(shape)
        t.copy_(torch.from_numpy(v.astype(np.float32)))
        return t

    k = (rng.random((5, 5, 7)) + 0.05).astype(np.float32)
    op = LinearFFTConvolver(k, shape)
    p = op._plan
    assert int(p.info.spectrum_block) > 0
    y1 = arr(rng.standard_normal(shape) * 0.3)
    ax = arr(rng.random(shape) * 5)
    axp = arr(rng.random(shape) * 5)
    d = rng.poisson(3.0, shape).astype(np.float32)
    d[rng.random(shape) < 0.2] = 0.0
    data = arr(d)
    sc = torch.zeros(16, dtype=torch.float64, device=dev)
    sc[0], sc[1], sc[2] = 0.4, 0.6, 0.85
    bg = 0.5
    st = N.stream_ptr()
    # separate
    y1_a = y1.clone()
    N.call("dh_pd_y1", y1_a.data_ptr(), ax.data_ptr(), axp.data_ptr(), data.data_ptr(), n, sc.data_ptr(), bg, loss, st)
    g_a = op.adjoint(y1_a)
    # fused, in place and through the batch entry point into another buffer
    y1_b = arr(y1.cpu().numpy())
    g_b = torch.full(shape, float("nan"), device=dev)

    def fz(out):
        f = N.FuseArgs(prologue=N.PRO_PD_Y1, pro_a=bg, pro_loss=loss, pro_out=out.data_ptr(),
                       pro_scalars=sc.data_ptr())
        f.pro_aux[:] = (ax.data_ptr(), axp.data_ptr(), data.data_ptr())
        return f

    p.apply(y1_b, True, g_b, fz(y1_b))
    y1_c = arr(np.full(shape, np.nan))
    g_c = torch.full(shape, float("nan"), device=dev)
    item = N.BatchItem(plan=p._h, otf=p.otf_buf.data_ptr(), work=p.work.data_ptr(),
                       work_bytes=int(p.info.work_bytes), input=y1.data_ptr(), out=g_c.data_ptr(), fuse=fz(y1_c))
    N.call("dh_batch_conv_apply", 1, (N.BatchItem * 1)(item), 1, st)
    torch.cuda.synchronize()
    assert torch.equal(y1_a, y1_b) and torch.equal(y1_a, y1_c)
    assert torch.equal(g_a, g_b) and torch.equal(g_a, g_c)
    # the prox moved y1 (and the Poisson branch's floor held)
    assert not torch.equal(y1_a, y1)
    if loss == 0:
        assert float(y1_a.max()) <= 1.0 - 1e-6


def test_y1_prologue_rejects_missing_operands(gpu):
    import ctypes as C

    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd import _native as N

    op = LinearFFTConvolver(np.ones((3, 3, 3), np.float32), (8, 12, 16))
    p = op._plan
    y = torch.zeros((8, 12, 16), device="cuda")
    out = torch.empty_like(y)
    sc = torch.zeros(16, dtype=torch.float64, device="cuda")
    good = [y.data_ptr()] * 3
    for aux, po, ps, lt in [(good, 0, sc.data_ptr(), 0), ([y.data_ptr(), 0, y.data_ptr()], y.data_ptr(),
                                                         sc.data_ptr(), 0),
                            (good, y.data_ptr(), 0, 0), (good, y.data_ptr(), sc.data_ptr(), 2)]:
        f = N.FuseArgs(prologue=N.PRO_PD_Y1, pro_loss=lt, pro_out=po or None, pro_scalars=ps or None)
        f.pro_aux[:] = [a or None for a in aux]
        assert N.fns["dh_conv_apply"](p._h, 1, y.data_ptr(), out.data_ptr(), p.otf_buf.data_ptr(), p.work.data_ptr(),
                                      int(p.info.work_bytes), C.byref(f), N.stream_ptr()) == N.DH_EINVAL
    # a canvas without specialised kernels (10000 samples) keeps dh_pd_y1: the prologue is refused
    rl = LinearFFTConvolver(np.ones(3001, np.float32), (7000,))
    assert int(rl._plan.info.spectrum_block) == 0
    q = rl._plan
    y7 = torch.zeros(7000, device="cuda")
    o7 = torch.empty_like(y7)
    f = N.FuseArgs(prologue=N.PRO_PD_Y1, pro_out=y7.data_ptr(), pro_scalars=sc.data_ptr())
    f.pro_aux[:] = [y7.data_ptr()] * 3
    assert N.fns["dh_conv_apply"](q._h, 1, y7.data_ptr(), o7.data_ptr(), q.otf_buf.data_ptr(), q.work.data_ptr(),
                                  int(q.info.work_bytes), C.byref(f), N.stream_ptr()) == N.DH_EINVAL
