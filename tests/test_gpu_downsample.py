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
    ((128, 512, 512), 2.0),
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


def test_non_integer_scale_keeps_banded_passes(gpu):
    from deconlib_amd import FractionalAreaDownsample

    op = FractionalAreaDownsample(1.325)
    x = torch.rand((10, 16, 20), device="cuda")
    assert op._block_factors(op.axis_pairs(3, x.shape, True), 3) is None
    ref = O.FractionalAreaDownsample(1.325)
    np.testing.assert_allclose(op.forward(x).cpu().numpy(), ref.forward(x.cpu().numpy().astype(np.float64)),
                               rtol=1e-5)
