"""FractionalAreaDownsample with integer factors: the one-pass block sum / spread
(dh_block_sum / dh_block_spread) against the three banded passes (bit for bit: the
same accumulation order, core_operators.py:205-219) and against the float64 oracle;
non-integer scales keep the banded passes."""

import numpy as np
import pytest
import torch

from oracle import deconv as O

pytestmark = pytest.mark.gpu


def _banded_chain(op, x, forward):
    from deconlib_amd.core_operators import _apply_banded

    r = x
    for a, nl, ns in op.axis_pairs(r.dim(), r.shape, forward):
        i_f, w_f, i_a, w_a = op._get_band(nl, ns)
        r = _apply_banded(r, i_f if forward else i_a, w_f if forward else w_a, a)
    return r


@pytest.mark.parametrize("shape,scale", [
    ((16, 24, 32), 2.0), ((16, 24, 32), (1.0, 2.0, 2.0)), ((9, 12, 10), (3.0, 1.0, 2.0)), ((20, 36), (2.0, 3.0)),
    ((128, 512, 512), 2.0), ((8, 10, 18), 2.0),
])
def test_block_sum_equals_banded_passes(gpu, shape, scale):
    from deconlib_amd import FractionalAreaDownsample

    rng = np.random.default_rng(70)
    op = FractionalAreaDownsample(scale)
    x = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda()
    pairs = op.axis_pairs(x.dim(), x.shape, True)
    assert op._block_factors(pairs, x.dim()) is not None
    got = op.forward(x)
    assert torch.equal(got, _banded_chain(op, x, True))
    y = torch.from_numpy(rng.standard_normal(tuple(got.shape)).astype(np.float32)).cuda()
    back = op.adjoint(y)
    assert tuple(back.shape) == shape
    assert torch.equal(back, _banded_chain(op, y, False))
    if np.prod(shape) < 1e6:
        ref = O.FractionalAreaDownsample(scale)
        np.testing.assert_allclose(got.cpu().numpy(), ref.forward(x.cpu().numpy().astype(np.float64)), rtol=1e-5,
                                   atol=1e-5)
        np.testing.assert_allclose(back.cpu().numpy(), ref.adjoint(y.cpu().numpy().astype(np.float64)), rtol=1e-6)


def test_block_sum_and_spread_on_unaligned_views(gpu):
    """Inputs and outputs 4 bytes off a 16-byte boundary take the scalar paths."""
    from deconlib_amd import FractionalAreaDownsample
    from deconlib_amd import _native as N

    rng = np.random.default_rng(72)
    op = FractionalAreaDownsample(2.0)
    base = torch.from_numpy(rng.standard_normal(16 * 24 * 32 + 1).astype(np.float32)).cuda()
    x = base[1:].view(16, 24, 32)
    assert x.data_ptr() % 16 == 4
    assert torch.equal(op.forward(x), _banded_chain(op, x, True))
    cb = torch.from_numpy(rng.standard_normal(8 * 12 * 16 + 1).astype(np.float32)).cuda()
    y = cb[1:].view(8, 12, 16)
    fine = torch.empty(16 * 24 * 32 + 1, device="cuda")[1:].view(16, 24, 32)
    N.call("dh_block_spread", y.data_ptr(), fine.data_ptr(), 16, 24, 32, 2, 2, 2, N.stream_ptr())
    assert torch.equal(fine, _banded_chain(op, y, False))


def test_non_integer_scale_keeps_banded_passes(gpu):
    from deconlib_amd import FractionalAreaDownsample

    op = FractionalAreaDownsample(1.325)
    x = torch.rand((10, 16, 20), device="cuda")
    assert op._block_factors(op.axis_pairs(3, x.shape, True), 3) is None
    ref = O.FractionalAreaDownsample(1.325)
    np.testing.assert_allclose(op.forward(x).cpu().numpy(), ref.forward(x.cpu().numpy().astype(np.float64)),
                               rtol=1e-5)


@pytest.mark.parametrize("shape,kshape,scale", [
    ((16, 24, 32), (5, 7, 9), 2.0),
    ((16, 24, 32), (3, 3, 3), (1.0, 2.0, 2.0)),
    ((9, 12, 30), (4, 5, 6), (3.0, 1.0, 1.0)),   # x factor 1
    ((21, 36), (5, 7), (3.0, 2.0)),
    ((7000,), (3001,), 2.0),                      # canvas 10000: runtime-length kernels
    ((128, 512, 512), (33, 65, 65), 2.0),         # the binned configuration's object grid
])
def test_composed_adjoint_spreads_in_the_first_pass(gpu, shape, kshape, scale):
    """compose(FractionalAreaDownsample, convolver).adjoint reads the coarse input
    through DH_PRO_SPREAD: bit for bit the block spread followed by the correlation,
    through dh_conv_apply and dh_batch_conv_apply; the float64 oracle on small cases."""
    from deconlib_amd import FFTConvolver, FractionalAreaDownsample, LinearFFTConvolver, compose
    from deconlib_amd import _native as N

    print(f"spread prologue {shape} * {kshape} scale {scale}", flush=True)
    rng = np.random.default_rng(71)
    k = (rng.random(kshape) + 0.05).astype(np.float32)
    conv = LinearFFTConvolver(k, shape)
    ds = FractionalAreaDownsample(scale)
    A = compose(ds, conv)
    coarse = tuple(int(v) for v in ds.forward(torch.zeros(shape, device="cuda")).shape)
    y = torch.from_numpy(rng.standard_normal(coarse).astype(np.float32)).cuda()
    f = ds.spread_factors(y)
    assert f is not None
    two = conv.adjoint(ds.adjoint(y))
    got = A.adjoint(y)
    assert torch.equal(got, two)
    p = conv._plan
    out_b = torch.empty_like(two)
    fz = N.FuseArgs(prologue=N.PRO_SPREAD)
    fz.spread[:] = f
    item = N.BatchItem(plan=p._h, otf=p.otf_buf.data_ptr(), work=p.work.data_ptr(),
                       work_bytes=int(p.info.work_bytes), input=y.data_ptr(), out=out_b.data_ptr(), fuse=fz)
    N.call("dh_batch_conv_apply", 1, (N.BatchItem * 1)(item), 1, N.stream_ptr())
    assert torch.equal(out_b, two)
    if np.prod(shape) < 1e5:
        conv64 = O.LinearFFTConvolver(k.astype(np.float64), shape, dtype=np.float64)
        ref = O.compose(O.FractionalAreaDownsample(scale), conv64)
        want = ref.adjoint(y.cpu().numpy().astype(np.float64))
        g = got.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(g - want) <= 1e-5 * np.linalg.norm(want)
    if len(shape) == 3 and np.prod(shape) < 1e5:
        circ = FFTConvolver(np.pad(k, [(0, s - q) for s, q in zip(shape, kshape)]))
        assert torch.equal(compose(ds, circ).adjoint(y), circ.adjoint(ds.adjoint(y)))


def test_spread_prologue_rejects_bad_factors(gpu):
    import ctypes as C

    from deconlib_amd import FractionalAreaDownsample, LinearFFTConvolver
    from deconlib_amd import _native as N

    conv = LinearFFTConvolver(np.ones((3, 3, 3), np.float32), (8, 12, 16))
    p = conv._plan
    y = torch.zeros((4, 6, 8), device="cuda")
    out = torch.empty((8, 12, 16), device="cuda")
    for f in [(2, 2, 3), (0, 2, 2), (3, 1, 1), (1, 1, 4)]:
        fz = N.FuseArgs(prologue=N.PRO_SPREAD)
        fz.spread[:] = f
        assert N.fns["dh_conv_apply"](p._h, 1, y.data_ptr(), out.data_ptr(), p.otf_buf.data_ptr(), p.work.data_ptr(),
                                      int(p.info.work_bytes), C.byref(fz), N.stream_ptr()) == N.DH_EINVAL
    with pytest.raises(ValueError):
        conv.adjoint_spread(torch.zeros((4, 6, 9), device="cuda"), (2, 2, 2))
    # an x factor above 2 keeps the two steps
    assert FractionalAreaDownsample((1.0, 1.0, 4.0)).spread_factors(torch.zeros((8, 12, 4), device="cuda")) is None


@pytest.mark.parametrize("shape,kshape", [
    ((16, 24, 32), (5, 7, 9)),
    ((8, 12, 18), (3, 4, 5)),        # coarse rows of 9 samples
    ((6, 10, 64), (5, 5, 5)),
    ((4, 6, 1600), (3, 3, 3)),       # 1024-point rows: the general load walk
    ((128, 512, 512), (33, 65, 65)),  # the binned configuration's object grid
])
def test_composed_forward_block_sums_in_the_last_pass(gpu, shape, kshape):
    """compose(FractionalAreaDownsample(2), convolver).forward sums the 2x2x2 blocks in
    the inverse X pass (DH_EPI_BLOCKSUM2): bit for bit the convolution followed by the
    block sum; the float64 oracle on small cases."""
    from deconlib_amd import FractionalAreaDownsample, LinearFFTConvolver, compose

    print(f"block-sum epilogue {shape} * {kshape}", flush=True)
    rng = np.random.default_rng(73)
    k = (rng.random(kshape) + 0.05).astype(np.float32)
    conv = LinearFFTConvolver(k, shape)
    ds = FractionalAreaDownsample(2.0)
    assert ds.sum_factors(shape) == (2, 2, 2)
    x = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda()
    got = conv.forward_block_sum(x)
    assert got is not None and tuple(got.shape) == tuple(v // 2 for v in shape)
    two = ds.forward(conv.forward(x))
    assert torch.equal(got, two)
    assert torch.equal(compose(ds, conv).forward(x), two)
    if np.prod(shape) < 1e5:
        conv64 = O.LinearFFTConvolver(k.astype(np.float64), shape, dtype=np.float64)
        want = O.compose(O.FractionalAreaDownsample(2.0), conv64).forward(x.cpu().numpy().astype(np.float64))
        g = got.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(g - want) <= 1e-5 * np.linalg.norm(want)


def test_block_sum_epilogue_limits(gpu):
    """Odd extents and 2-D signals keep the two steps; the C-ABI rejects the epilogue there
    and in the batched entry point."""
    import ctypes as C

    from deconlib_amd import FractionalAreaDownsample, LinearFFTConvolver, compose
    from deconlib_amd import _native as N

    k = np.ones((3, 3, 3), np.float32)
    odd = LinearFFTConvolver(k, (6, 9, 8))
    assert odd.forward_block_sum(torch.zeros((6, 9, 8), device="cuda")) is None
    p = odd._plan
    x = torch.zeros((6, 9, 8), device="cuda")
    out = torch.empty((3, 4, 4), device="cuda")
    f = N.FuseArgs(epilogue=N.EPI_BLOCKSUM2)
    assert N.fns["dh_conv_apply"](p._h, 0, x.data_ptr(), out.data_ptr(), p.otf_buf.data_ptr(), p.work.data_ptr(),
                                  int(p.info.work_bytes), C.byref(f), N.stream_ptr()) == N.DH_EINVAL
    flat = LinearFFTConvolver(np.ones((3, 3), np.float32), (8, 8))
    y = torch.rand((8, 8), device="cuda")
    ds = FractionalAreaDownsample(2.0)
    assert torch.equal(compose(ds, flat).forward(y), ds.forward(flat.forward(y)))
    ev = LinearFFTConvolver(k, (6, 8, 8))
    pe = ev._plan
    xe = torch.rand((6, 8, 8), device="cuda")
    oe = torch.empty((3, 4, 4), device="cuda")
    item = N.BatchItem(plan=pe._h, otf=pe.otf_buf.data_ptr(), work=pe.work.data_ptr(),
                       work_bytes=int(pe.info.work_bytes), input=xe.data_ptr(), out=oe.data_ptr(), fuse=f)
    assert N.fns["dh_batch_conv_apply"](1, (N.BatchItem * 1)(item), 0, N.stream_ptr()) == N.DH_EINVAL
