"""GPU parity of the pruned HIP convolution against the oracle (same inputs).

Tolerances: relative L2 <= 2e-6 against the float64 oracle (fp32 FFT
rounding), adjoint dot-product identity <= 1e-5.
"""

import numpy as np
import pytest
import torch

from oracle import deconv as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    """Relative L2 error; complex inputs (OTFs) keep their imaginary parts."""
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


SHAPES = [
    ((16, 32, 32), (16, 32, 32)),      # canvas 32x64x64 (radix 16/4/2)
    ((8, 12, 20), (5, 7, 9)),          # canvas 12x18x28: radix 3, odd kernels
    ((10, 15, 9), (4, 6, 8)),          # even kernels, 5-smooth mixed (13->15, 20, 16)
    ((7, 11, 30), (7, 11, 30)),        # canvas 13->15, 21->24, 59->60
    ((1, 50, 64), (1, 9, 16)),         # single plane
    ((33, 40), (9, 10)),               # 2-D
    ((250,), (31,)),                   # 1-D
    ((3, 5, 6), (3, 5, 6)),            # tiny
    ((40, 96, 96), (40, 96, 96)),      # canvas 80x192x192 (radix 3 and 5 on every axis)
    ((7000,), (3001,)),                # canvas 10000: no specialised length, runtime-length kernels
    ((5000, 6), (1201, 3)),            # y canvas 6250: runtime-length column pass
    ((2, 3, 2048), (2, 3, 2049)),      # x canvas 4096: 2048-point rows, twiddles read from global memory
    ((1, 2, 4096), (1, 1, 4097)),      # x canvas 8192: 4096-point rows
    ((2, 2048, 4), (1, 2049, 3)),      # y canvas 4096: blocks split over 2-column tiles (cached, XCD-major)
    ((2, 2048, 40), (1, 2049, 9)),     # y canvas 4096, x canvas 48 (H = 25): 2-column tiles over three full
                                       # 8-column blocks and a partial one
]


@pytest.mark.parametrize("sshape,kshape", SHAPES)
def test_linear_conv_matches_oracle(gpu, sshape, kshape):
    from deconlib_amd import LinearFFTConvolver

    print(f"linear conv {sshape} * {kshape}", flush=True)
    rng = np.random.default_rng(11)
    k = (rng.random(kshape) + 0.05).astype(np.float32)
    x = rng.standard_normal(sshape).astype(np.float32)
    y = rng.standard_normal(sshape).astype(np.float32)
    ref = O.LinearFFTConvolver(k, sshape, dtype=np.float64)
    op = LinearFFTConvolver(k, sshape)
    assert op.fft_shape == ref.fft_shape
    fx = op.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ay = op.adjoint(torch.from_numpy(y).cuda()).cpu().numpy()
    assert fx.shape == sshape and ay.shape == sshape
    assert rel_l2(fx, ref.forward(x.astype(np.float64))) < 2e-6
    assert rel_l2(ay, ref.adjoint(y.astype(np.float64))) < 2e-6
    lhs = float(np.vdot(fx.astype(np.float64), y))
    rhs = float(np.vdot(x.astype(np.float64), ay))
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1e-12) + 1e-6
    assert abs(op.operator_norm_sq - ref.operator_norm_sq) < 1e-5


def test_linear_conv_explicit_fft_shape_and_odd_canvas(gpu):
    from deconlib_amd import LinearFFTConvolver

    rng = np.random.default_rng(12)
    k = (rng.random((5, 8, 9)) + 0.1).astype(np.float32)
    x = rng.random((12, 20, 21)).astype(np.float32)
    # odd x canvas (29 -> computed on 30) and a user fft_shape larger than needed
    for fft in [(16, 27, 29), (20, 30, 45)]:
        ref = O.LinearFFTConvolver(k, x.shape, fft_shape=fft, dtype=np.float64)
        op = LinearFFTConvolver(k, x.shape, fft_shape=fft)
        assert op.fft_shape == fft
        out = op.forward(x).cpu().numpy()
        assert rel_l2(out, ref.forward(x.astype(np.float64))) < 2e-6


def test_delta_psf_is_identity(gpu):
    from deconlib_amd import LinearFFTConvolver

    k = np.zeros((9, 16, 16), np.float32)
    k[0, 0, 0] = 1.0
    x = np.random.default_rng(13).random((20, 40, 48)).astype(np.float32)
    op = LinearFFTConvolver(k, x.shape)
    np.testing.assert_allclose(op.forward(x).cpu().numpy(), x, atol=2e-6)
    np.testing.assert_allclose(op.adjoint(x).cpu().numpy(), x, atol=2e-6)


def test_circular_convolver(gpu):
    from deconlib_amd import FFTConvolver

    rng = np.random.default_rng(14)
    k = (rng.random((8, 12, 10)) + 0.1).astype(np.float32)
    x = rng.standard_normal((8, 12, 10)).astype(np.float32)
    ref = O.FFTConvolver(k, dtype=np.float64)
    op = FFTConvolver(k)
    assert rel_l2(op.forward(x).cpu().numpy(), ref.forward(x.astype(np.float64))) < 2e-6
    assert rel_l2(op.adjoint(x).cpu().numpy(), ref.adjoint(x.astype(np.float64))) < 2e-6
    otf = op.otf.cpu().numpy()
    assert rel_l2(otf, ref.otf) < 2e-6


def test_config_shape_conv(gpu):
    """64x256x256 signal with a full-size PSF (canvas 128x512x512)."""
    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd.psf import widefield_psf

    print("config-shape conv 64x256x256", flush=True)
    psf = widefield_psf((64, 256, 256), 0.085, 0.1)
    rng = np.random.default_rng(15)
    x = rng.random((64, 256, 256)).astype(np.float32)
    ref = O.LinearFFTConvolver(psf, x.shape, dtype=np.float64)
    op = LinearFFTConvolver(psf, x.shape)
    assert rel_l2(op.forward(x).cpu().numpy(), ref.forward(x.astype(np.float64))) < 2e-6
    assert rel_l2(op.adjoint(x).cpu().numpy(), ref.adjoint(x.astype(np.float64))) < 2e-6


def test_layout_operators(gpu):
    from deconlib_amd import Crop, FractionalAreaDownsample, Pad

    rng = np.random.default_rng(16)
    x = rng.random((12, 15, 18)).astype(np.float32)
    c = Crop((12, 15, 18), (6, 9, 10), start=(1, 2, 3))
    oc = O.Crop((12, 15, 18), (6, 9, 10), start=(1, 2, 3))
    np.testing.assert_array_equal(c.forward(x).cpu().numpy(), oc.forward(x))
    y = rng.random((6, 9, 10)).astype(np.float32)
    np.testing.assert_array_equal(c.adjoint(y).cpu().numpy(), oc.adjoint(y))
    p = Pad(((1, 2), (0, 3), (4, 0)))
    op = O.Pad(((1, 2), (0, 3), (4, 0)))
    np.testing.assert_array_equal(p.forward(x).cpu().numpy(), op.forward(x))
    np.testing.assert_array_equal(p.adjoint(op.forward(x)).cpu().numpy(), x)
    for scale in (2.0, 1.325, (1.5, 2.0, 3.0)):
        d = FractionalAreaDownsample(scale, in_shape=x.shape)
        od = O.FractionalAreaDownsample(scale, in_shape=x.shape)
        fx = d.forward(x).cpu().numpy()
        np.testing.assert_allclose(fx, od.forward(x), rtol=2e-6, atol=1e-6)
        z = rng.random(fx.shape).astype(np.float32)
        np.testing.assert_allclose(d.adjoint(z).cpu().numpy(), od.adjoint(z), rtol=2e-6, atol=1e-6)


@pytest.mark.parametrize("sshape,kshape", [((16, 32, 32), (16, 32, 32)), ((10, 15, 9), (4, 6, 8)),
                                           ((12, 40, 64), (7, 21, 33))])
def test_launch_shapes_are_bit_identical(gpu, sshape, kshape):
    """Every tile size and the persistent grid compute the same values in the
    same order: results are bit-identical to the default launch shapes."""
    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd import _native as N

    print(f"launch shapes {sshape} * {kshape}", flush=True)
    rng = np.random.default_rng(12)
    k = (rng.random(kshape) + 0.05).astype(np.float32)
    x = torch.from_numpy(rng.standard_normal(sshape).astype(np.float32)).cuda()
    op = LinearFFTConvolver(k, sshape)
    ref_f = op.forward(x).cpu().numpy()
    ref_a = op.adjoint(x).cpu().numpy()
    h = op._plan._h
    tried = 0
    for p in range(5):
        for t in (1, 2, 4, 8, 16, 32):
            for v in (t, t | N.TILE_PERSISTENT):
                if N.fns["dh_conv_plan_set_tile"](h, p, v) != 0:
                    continue
                tried += 1
                assert np.array_equal(op.forward(x).cpu().numpy(), ref_f), (p, v)
                assert np.array_equal(op.adjoint(x).cpu().numpy(), ref_a), (p, v)
        N.call("dh_conv_plan_set_tile", h, p, 0)
    assert tried >= 10


@pytest.mark.parametrize("sshape,kshape", [((16, 32, 32), (16, 32, 32)), ((9, 15, 11), (4, 6, 8)),
                                           ((5, 7, 1000), (3, 3, 21))])
def test_sqdiff_epilogue(gpu, sshape, kshape):
    """DH_EPI_SQDIFF leaves the output unchanged and sums (out - aux0)^2 in float64
    (each square formed in float32, as PDHG's data sums form it); the batched entry
    point gives the same output and sum.  Shapes: specialised even canvas, odd rows,
    and a runtime-length canvas."""
    import ctypes as C

    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd import _native as N

    print(f"sqdiff epilogue {sshape} * {kshape}", flush=True)
    rng = np.random.default_rng(13)
    k = (rng.random(kshape) + 0.05).astype(np.float32)
    x = torch.from_numpy(rng.standard_normal(sshape).astype(np.float32)).cuda()
    prev = torch.from_numpy(rng.standard_normal(sshape).astype(np.float32)).cuda()
    op = LinearFFTConvolver(k, sshape)
    ref = op.forward(x)
    red = torch.zeros(2, dtype=torch.float64, device="cuda")
    out = torch.empty_like(ref)
    fuse = N.FuseArgs(epilogue=N.EPI_SQDIFF, aux0=prev.data_ptr(), reduce_out=red.data_ptr())
    op._plan.apply(x, False, out, fuse)
    p = op._plan
    out_b = torch.empty_like(ref)
    fuse_b = N.FuseArgs(epilogue=N.EPI_SQDIFF, aux0=prev.data_ptr(), reduce_out=red.data_ptr() + 8)
    item = N.BatchItem(plan=p._h, otf=p.otf_buf.data_ptr(), work=p.work.data_ptr(),
                       work_bytes=int(p.info.work_bytes), input=x.data_ptr(), out=out_b.data_ptr(), fuse=fuse_b)
    N.call("dh_batch_conv_apply", 1, (N.BatchItem * 1)(item), 0, N.stream_ptr())
    torch.cuda.synchronize()
    assert torch.equal(out, ref) and torch.equal(out_b, ref)
    d = (ref - prev).double().cpu().numpy().astype(np.float32)
    want = float(np.sum((d * d).astype(np.float64)))
    got = red.cpu().numpy()
    np.testing.assert_allclose(got, [want, want], rtol=1e-12)
    assert got[0] == got[1]
    # aux0 is required
    bad = N.FuseArgs(epilogue=N.EPI_SQDIFF, reduce_out=red.data_ptr())
    assert N.fns["dh_conv_apply"](p._h, 0, x.data_ptr(), out.data_ptr(), p.otf_buf.data_ptr(), p.work.data_ptr(),
                                  int(p.info.work_bytes), C.byref(bad), N.stream_ptr()) == N.DH_EINVAL


def test_autotune_restricted_passes(gpu):
    """autotune(passes=...) times and sets only the named passes; the others keep their
    launch shapes, and the tuned operator computes the same values."""
    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd import _native as N
    from deconlib_amd.autotune import autotune

    rng = np.random.default_rng(21)
    shape = (16, 64, 96)
    op = LinearFFTConvolver((rng.random((5, 7, 9)) + 0.05).astype(np.float32), shape)
    x = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda()
    ref = op.forward(x)
    h = op._plan._h
    before = [N.fns["dh_conv_plan_get_tile"](h, p) for p in range(5)]
    best = autotune(op, reps=1, candidates=(1, 2), passes=(1, 3))
    assert set(best) <= {1, 3}
    after = [N.fns["dh_conv_plan_get_tile"](h, p) for p in range(5)]
    assert [after[p] for p in (0, 2, 4)] == [before[p] for p in (0, 2, 4)]
    for p, t in best.items():
        assert after[p] == t
    assert torch.equal(op.forward(x), ref)
