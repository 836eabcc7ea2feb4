"""Pins of the CPU oracle: known answers, brute force and closed forms.

These tests run on the CPU only.  Each one is chosen so that a plausible slip
in the oracle (a dropped term, a wrong sign, a transposed operand, an
off-by-one index) fails it.
"""

import math

import numpy as np
import pytest

from oracle import deconv as O


# -- fast_padded_shape: the reference docstring's known answers (linops_mlx.py:1009-1017)
def test_fast_padded_shape_known_answers():
    assert O.fast_padded_shape((101, 589, 549), (101, 256, 256)) == (216, 864, 810)
    assert O.fast_padded_shape((101, 589, 549), (101, 256, 256), min_pad=(0, None, None)) == (108, 864, 810)
    assert O.fast_padded_shape((34, 311, 311), (101, 160, 160), min_pad=(0, None, None)) == (108, 480, 480)
    assert O.fast_padded_shape((64, 256, 256), (64, 256, 256)) == (128, 512, 512)
    # 5-smoothness and minimality (tests/test_tile_processing.py:78-86 style)
    for n in range(1, 400):
        s = O.next_smooth_number(n)
        assert s >= n
        m = s
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        assert m == 1
        for c in range(n, s):
            mm = c
            for p in (2, 3, 5):
                while mm % p == 0:
                    mm //= p
            assert mm != 1


# -- pad_corner_origin_kernel: brute-force index map on tiny odd/even kernels
@pytest.mark.parametrize("kshape,oshape", [((3,), (8,)), ((4,), (9,)), ((3, 4), (7, 10)), ((2, 5, 3), (6, 9, 8))])
def test_pad_corner_origin_bruteforce(kshape, oshape):
    rng = np.random.default_rng(0)
    k = rng.standard_normal(kshape)
    out = O.pad_corner_origin_kernel(k, oshape)
    ref = np.zeros(oshape)
    for idx in np.ndindex(*kshape):
        # signed offset of each compact sample, then wrap into the canvas
        off = [i if i < (m + 1) // 2 else i - m for i, m in zip(idx, kshape)]
        ref[tuple(o % p for o, p in zip(off, oshape))] = k[idx]
    np.testing.assert_array_equal(out, ref)


def _direct_linear_conv(x, k):
    """O(N*M) float64 direct linear convolution with a corner-origin kernel, cropped to x."""
    out = np.zeros(x.shape)
    for kidx in np.ndindex(*k.shape):
        off = [i if i < (m + 1) // 2 else i - m for i, m in zip(kidx, k.shape)]
        src = []
        dst = []
        for o, n in zip(off, x.shape):
            if o >= 0:
                dst.append(slice(o, n))
                src.append(slice(0, n - o))
            else:
                dst.append(slice(0, n + o))
                src.append(slice(-o, n))
        out[tuple(dst)] += k[kidx] * x[tuple(src)]
    return out


@pytest.mark.parametrize("sshape,kshape", [((11,), (5,)), ((12,), (4,)), ((9, 10), (5, 4)),
                                          ((6, 7, 8), (3, 4, 5)), ((5, 6, 4), (5, 6, 4))])
def test_linear_conv_bruteforce(sshape, kshape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(sshape)
    k = rng.random(kshape) + 0.1
    op = O.LinearFFTConvolver(k, sshape, dtype=np.float64)
    kn = k.astype(np.float32).astype(np.float64)
    kn = kn / kn.sum()
    np.testing.assert_allclose(op.forward(x), _direct_linear_conv(x, kn), atol=1e-12)
    # adjoint = correlation = convolution with the flipped (negated-offset) kernel
    kf = np.zeros_like(kn)
    for idx in np.ndindex(*kn.shape):
        kf[tuple((-i) % m if m > 1 else 0 for i, m in zip(idx, kn.shape))] = kn[idx]
    # flipping a corner-origin kernel maps offset o -> -o only when the compact
    # extent is odd; check the adjoint through the dot-product identity instead
    y = rng.standard_normal(sshape)
    lhs = np.vdot(op.forward(x), y)
    rhs = np.vdot(x, op.adjoint(y))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    if all(m % 2 == 1 for m in kshape):
        np.testing.assert_allclose(op.adjoint(x), _direct_linear_conv(x, kf), atol=1e-12)


def test_linear_conv_delta_identity_and_norm():
    # delta PSF => identity (tests/test_forward_model_registration.py:18-54)
    k = np.zeros((5, 6, 7), np.float32)
    k[0, 0, 0] = 1.0
    x = np.random.default_rng(2).random((8, 9, 10)).astype(np.float32)
    op = O.LinearFFTConvolver(k, x.shape)
    np.testing.assert_allclose(op.forward(x), x, atol=1e-6)
    # normalized non-negative PSF => max|OTF|^2 = 1 (DC term)
    k2 = np.random.default_rng(3).random((5, 6, 7))
    op2 = O.LinearFFTConvolver(k2, (8, 9, 10), dtype=np.float64)
    assert abs(op2.operator_norm_sq - 1.0) < 1e-12


def test_linear_conv_shape_errors():
    with pytest.raises(ValueError):
        O.LinearFFTConvolver(np.ones((3, 3)), (8, 8, 8))
    with pytest.raises(ValueError):
        O.LinearFFTConvolver(np.ones((3, 3)), (8, 8), fft_shape=(9, 9))


# -- Pad / Crop (tests/test_core_operators.py:53-143)
def test_crop_pad_adjoint():
    rng = np.random.default_rng(4)
    c = O.Crop((10, 12), (6, 7))
    x = rng.standard_normal((10, 12))
    y = rng.standard_normal((6, 7))
    assert abs(np.vdot(c.forward(x), y) - np.vdot(x, c.adjoint(y))) < 1e-12
    np.testing.assert_array_equal(c.forward(x), x[2:8, 2:9])
    c2 = O.Crop((10, 12), (6, 7), start=(1, 5))
    np.testing.assert_array_equal(c2.forward(x), x[1:7, 5:12])
    with pytest.raises(ValueError):
        O.Crop((10, 12), (6, 7), start=(5, 6))


# -- FractionalAreaDownsample (core_operators.py:222-352)
def test_downsample_integer_is_block_sum():
    rng = np.random.default_rng(5)
    x = rng.random((8, 12, 6))
    d = O.FractionalAreaDownsample(2)
    d3 = O.FractionalAreaDownsample((2.0, 2.0, 2.0))
    ref = x.reshape(4, 2, 6, 2, 3, 2).sum(axis=(1, 3, 5))
    np.testing.assert_allclose(d3.forward(x), ref, rtol=1e-6)
    # scalar scale with axes=None resamples every axis
    np.testing.assert_allclose(d.forward(x), ref, rtol=1e-6)
    assert d3.operator_norm_sq == 8.0


@pytest.mark.parametrize("scale", [2.0, 1.325, (1.5, 2.0, 3.0)])
def test_downsample_adjoint_and_flux(scale):
    rng = np.random.default_rng(6)
    shape = (12, 15, 18)
    d = O.FractionalAreaDownsample(scale, in_shape=shape)
    x = rng.random(shape)
    y = rng.random(d.out_shape)
    fx = d.forward(x)
    assert fx.shape == d.out_shape
    lhs, rhs = np.vdot(fx, y), np.vdot(x, d.adjoint(y))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    # area-weight rows sum to the scale, columns to 1: total flux is preserved
    assert abs(fx.sum() - x.sum()) <= 1e-5 * x.sum()


# -- Hessian3D (tests/test_operators_mlx_norms.py:109-170, tests/test_pdhg_mlx.py:318-358)
def test_hessian_quadratic_phantom():
    r = 0.065 / 0.15
    H = O.Hessian3D(r)
    dz, dy, dx = 0.15, 0.065, 0.065
    a, b, c = 2.0, 3.0, 5.0
    shape = (12, 12, 12)
    iz, iy, ix = np.meshgrid(*[np.arange(n) - n // 2 for n in shape], indexing="ij")
    f = 0.5 * (a * (iz * dz) ** 2 + b * (iy * dy) ** 2 + c * (ix * dx) ** 2)
    Hg = H.forward(f)
    i = (6, 6, 6)
    dy2 = dy * dy
    assert abs(Hg[0][i] - a * dy2) < 1e-10
    assert abs(Hg[1][i] - b * dy2) < 1e-10
    assert abs(Hg[2][i] - c * dy2) < 1e-10
    for q in (3, 4, 5):
        assert abs(Hg[q][i]) < 1e-10
    # a mixed quadratic exercises the cross terms: f = x*y*z gradient-free diag
    g = (ix * iy + 2 * iy * iz + 3 * ix * iz).astype(np.float64)
    Hh = O.Hessian3D(1.0).forward(g)
    assert abs(Hh[5][i] - math.sqrt(2) * 1.0) < 1e-12   # d2/dxdy
    assert abs(Hh[3][i] - math.sqrt(2) * 2.0) < 1e-12   # d2/dydz
    assert abs(Hh[4][i] - math.sqrt(2) * 3.0) < 1e-12   # d2/dxdz


@pytest.mark.parametrize("r", [1.0, 2.0, 0.45])
def test_hessian_adjoint_and_norm(r):
    rng = np.random.default_rng(7)
    H = O.Hessian3D(r)
    x = rng.standard_normal((7, 8, 9))
    y = rng.standard_normal((6, 7, 8, 9))
    lhs, rhs = np.vdot(H.forward(x), y), np.vdot(x, H.adjoint(y))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    assert H.operator_norm_sq == 16 * r ** 4 + 4 * r ** 2 + 34
    v = rng.standard_normal((8, 10, 10))
    for _ in range(60):
        v = H.adjoint(H.forward(v))
        nrm = np.linalg.norm(v)
        v /= nrm
    assert nrm <= H.operator_norm_sq * 1.0001


# -- proxes (tests/test_pdhg_mlx.py:58-229)
def test_prox_known_values():
    y = np.array([-2.0, -0.5, 0.0, 0.5, 2.0], np.float32)
    np.testing.assert_array_equal(O.prox_l1_dual(y, 1.0), [-1.0, -0.5, 0.0, 0.5, 1.0])
    np.testing.assert_array_equal(O.prox_nonneg(y), [0, 0, 0, 0.5, 2.0])
    yy = np.array([[3.0], [4.0]])
    out = O.prox_l1_2_dual(yy, 1.0)
    assert abs(np.linalg.norm(out) - 1.0) < 1e-6
    g = O.prox_gaussian_dual(np.array([1.0]), 0.5, np.array([3.0]), 1.0)
    assert abs(g[0] - (1.0 - 0.5 * 2.0) / 1.5) < 1e-12


def test_prox_poisson_optimality():
    # u = prox_{sigma F*}(v) satisfies v - u in sigma dF*(u): for the Poisson dual
    # F*(u) = -sum D log(1-u) - b u  =>  (v - u)/sigma = -D/(1-u) ... residual < 1e-4
    rng = np.random.default_rng(8)
    for sigma in (0.01, 0.1, 1.0):
        v = rng.standard_normal(1000)
        D = rng.random(1000) * 10 + 0.1
        b = 0.3
        u = O.prox_poisson_dual(v, sigma, D, b)
        assert np.all(u < 1.0)
        res = (v - u) / sigma - (-D / (1.0 - u) + b) * -1.0
        # closed form check of the quadratic: z = 1-u solves z^2 - c z - sigma D = 0
        z = 1.0 - u
        c = 1.0 - v - sigma * b
        assert np.max(np.abs(z * z - c * z - sigma * D) / (1 + np.abs(sigma * D))) < 1e-8
        del res
    # D = 0: z = max(c, 1e-6)
    u0 = O.prox_poisson_dual(np.array([0.5, 2.0]), 0.1, np.zeros(2), 0.0)
    np.testing.assert_allclose(u0, [1.0 - 0.5, 1.0 - 1e-6])


# -- RL (rl_mlx.py:117-185)
def test_rl_first_iteration_1d_by_hand():
    k = np.array([0.5, 0.25, 0.25])          # corner origin: offsets 0, +1, -1
    obs = np.array([1.0, 2.0, 3.0, 4.0])
    A = O.LinearFFTConvolver(k, (4,), dtype=np.float64)
    Am = np.array([[0.5, 0.25, 0.0, 0.0], [0.25, 0.5, 0.25, 0.0], [0.0, 0.25, 0.5, 0.25], [0.0, 0.0, 0.25, 0.5]])
    cols = np.stack([A.forward(e) for e in np.eye(4)], axis=1)
    np.testing.assert_allclose(cols, Am, atol=1e-14)
    x0 = np.maximum(Am.T @ np.maximum(obs, 1e-6), 1e-6)
    s = Am.T @ np.ones(4)
    x1 = np.maximum(x0 * (Am.T @ (obs / np.maximum(Am @ x0, 1e-10))) / s, 1e-10)
    res = O.richardson_lucy(obs, A, num_iter=1, dtype=np.float64)
    np.testing.assert_allclose(res.restored, x1, rtol=1e-12)
    assert res.iterations == 1 and len(res.loss_history) == 1


def test_rl_fixed_point_and_monotone():
    rng = np.random.default_rng(9)
    k = rng.random((3, 5, 5)) + 0.05
    A = O.LinearFFTConvolver(k, (6, 10, 10), dtype=np.float64)
    xs = rng.random((6, 10, 10)) + 0.5
    obs = A.forward(xs)
    res = O.richardson_lucy(obs, A, num_iter=3, init=xs, dtype=np.float64)
    np.testing.assert_allclose(res.restored, xs, rtol=1e-9)
    obs2 = rng.poisson(obs * 50).astype(np.float64)
    res2 = O.richardson_lucy(obs2, A, num_iter=20, eval_interval=1, dtype=np.float64)
    L = np.array(res2.loss_history)
    assert np.all(np.diff(L) <= 1e-12 * np.abs(L[:-1]))


# -- PDHG first iteration in closed form from zero duals (pdhg_mlx.py:717-760)
def test_pdhg_first_iteration_closed_form():
    rng = np.random.default_rng(10)
    psf = rng.random((3, 5, 5)) + 0.05
    obs = (rng.random((6, 8, 8)) * 20).astype(np.float64)
    A = O.LinearFFTConvolver(psf, obs.shape, dtype=np.float64)
    L = O.Hessian3D(1.0)
    alpha = 0.01
    x0 = np.maximum(obs, 0.0)
    s = 0.99 / math.sqrt(A.operator_norm_sq + L.operator_norm_sq)
    c = 1.0 - s * A.forward(x0)
    z = 2 * s * obs / (np.sqrt(c * c + 4 * s * obs) - c + 1e-12)
    z = np.where(obs < 1e-6, np.maximum(c, 1e-6), z)
    y1 = 1.0 - np.maximum(z, 1e-6)
    y2 = np.clip(s * L.forward(x0), -alpha, alpha)
    x1 = np.maximum(x0 - s * (A.adjoint(y1) + L.adjoint(y2)), 0.0)
    res = O.solve_pdhg(obs, psf, alpha=alpha, num_iter=1, dtype=np.float64)
    np.testing.assert_allclose(res.restored, x1, rtol=1e-10, atol=1e-12)
    assert len(res.tau_history) == 1
    # step-size invariants after a few iterations
    res5 = O.solve_pdhg(obs, psf, alpha=alpha, num_iter=8, dtype=np.float64)
    for t, sg in zip(res5.tau_history, res5.sigma_history):
        assert t >= 1e-4 * s * (1 - 1e-12)
        assert 1e-4 * s * (1 - 1e-12) <= sg <= 1e4 * s * (1 + 1e-12)
