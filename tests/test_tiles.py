"""Tiling arithmetic and the tiled oracle, on the CPU.

The product's host logic (deconlib_amd.tile_processing / shapes) and the
oracle's restatement (oracle.tiles) both face the reference tests' known
answers and invariants (tests/test_tile_processing.py:76-150 of the
reference): balanced cores, 5-smooth windows, uniform read shape, writes
partitioning the output, reads in bounds.  The oracle's forward model and
tiled RL are pinned by the passthrough stitch, the adjoint identity, the
delta-PSF registration phase (tests/test_forward_model_registration.py) and
tiled-versus-full agreement in the interior.
"""

import numpy as np
import pytest

from deconlib_amd import shapes as S
from deconlib_amd import tile_processing as TP
from oracle import tiles as OT


def _is_5_smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


@pytest.mark.parametrize("ots", [TP.optimal_tile_size, OT.optimal_tile_size], ids=["product", "oracle"])
def test_optimal_tile_size_known_answers(ots):
    assert ots(256, guard=32, tile_size=256) == 256
    assert ots(200, guard=32, tile_size=256) == 200
    assert ots(512, guard=0, tile_size=256) == 256
    for N in (61, 67, 100, 337, 1000, 4095):
        T = ots(N, guard=8, tile_size=64)
        assert _is_5_smooth(T) or T == N
    for N in range(50, 80):
        assert ots(N, guard=4, tile_size=32) <= N
    with pytest.raises(ValueError):
        ots(64, guard=20, tile_size=32)


def _plans(shape, zoom, guard, tile_size, min_z=48):
    """(tile_shape, visible_shape, [(read, write, crop)]) from both implementations."""
    p = TP.plan_tiles(shape, zoom, guard=guard, tile_size=tile_size, min_z_slices=min_z)
    prod = (p.tile_shape, p.visible_shape, [(t.read, t.write, t.crop) for t in p.tiles])
    return prod, OT.plan_tiles(shape, zoom, guard, tile_size, min_z)


@pytest.mark.parametrize("shape,zoom", [((4, 61, 67), (1.0, 1.3, 1.3)), ((4, 128, 128), (1.0, 1.3, 1.3)),
                                        ((4, 100, 45), (1.0, 1.3, 1.3)), ((4, 61, 67), (1.0, 1.25, 1.25)),
                                        ((60, 61, 67), 1.5), ((4, 61, 67), (1.0, 1.5, 1.5))])
def test_plan_invariants(shape, zoom):
    prod, orac = _plans(shape, zoom, guard=4, tile_size=32)
    assert prod == orac
    tile_shape, visible, tiles = prod
    counts = np.zeros(visible, dtype=int)
    for read, write, crop in tiles:
        assert tuple(s.stop - s.start for s in read) == tile_shape  # uniform reads
        for sl, n in zip(read, shape):
            assert 0 <= sl.start and sl.stop <= n  # in bounds
        assert tuple(s.stop - s.start for s in crop) == tuple(s.stop - s.start for s in write)
        counts[write] += 1
    assert (counts == 1).all()  # the writes partition the output


def test_plan_shapes_and_counts():
    p = TP.plan_tiles((8, 64, 64), 1.0, guard=4, tile_size=32)
    assert isinstance(p, TP.TilePlan) and len(p.tiles) > 1
    assert all(isinstance(t, TP.TileSpec) for t in p.tiles)
    p = TP.plan_tiles((4, 32, 32), 1.0, guard=4, tile_size=64)
    assert len(p.tiles) == 1 and p.tile_shape == (4, 32, 32)
    p = TP.plan_tiles((10, 64, 64), 1.0, guard=4, tile_size=32, min_z_slices=48)
    assert all(t.read[0] == slice(0, 10) for t in p.tiles)
    p = TP.plan_tiles((1000, 1000), 1.0, guard=32, tile_size=448)
    cores = [t.write[0].stop - t.write[0].start for t in p.tiles]
    assert min(cores) / max(cores) > 0.9  # balanced, no trailing sliver
    with pytest.raises(ValueError):
        TP.plan_tiles((4, 64, 64), 1.0, guard=20, tile_size=32)
    with pytest.raises(ValueError):
        TP.plan_tiles((64,), 1.0, guard=2)


def test_shapes_known_answers():
    assert S.compute_padded_shape((100, 100), (31, 31)) == ((130, 130), ((15, 15), (15, 15)))
    assert S.compute_padded_shape((10, 10), (4, 5)) == ((13, 14), ((1, 2), (2, 2)))
    assert S.compute_padded_shape((10, 10), (4, 5), min_pad=(None, 0)) == ((13, 10), ((1, 2), (0, 0)))
    assert S.get_valid_slices((120, 120), (100, 100)) == (slice(10, 110), slice(10, 110))
    assert S.get_valid_slices((115,), (100,), padding=((10, 5),)) == (slice(10, 110),)
    assert S.compute_visible_shape((100, 100), bin_factor=0.85) == (118, 118)
    assert S.compute_visible_shape((100, 100), bin_factor=1.2) == (83, 83)
    assert S.compute_visible_shape((100, 100, 50), bin_factor=(1.0, 1.0, 2.0)) == (100, 100, 25)
    with pytest.raises(ValueError):
        S.compute_padded_shape((10, 10), (3,))


def test_make_forward_model_rejects_bad_zoom():
    from deconlib_amd import make_forward_model

    psf = np.ones((3, 5, 5), np.float32) / 75
    with pytest.raises(ValueError):
        make_forward_model(psf, (4, 16, 16), zoom=0.8)
    with pytest.raises(ValueError):
        make_forward_model(psf, (4, 16, 16), zoom=(1.0, 1.0))
    with pytest.raises(ValueError):
        make_forward_model(np.ones((5, 5), np.float32), (4, 16, 16))


# ---- the oracle's forward model and tiled solve ----

def _flat_psf(shape):
    p = np.ones(shape, np.float64)
    return p / p.sum()


def test_oracle_passthrough_stitches_input():
    data = np.random.default_rng(0).random((3, 61, 67)).astype(np.float32)
    out = OT.process_tiles(data, _flat_psf((3, 5, 5)), 1.0, lambda t, m: t, guard=4, tile_size=32)
    np.testing.assert_array_equal(out, data)


@pytest.mark.parametrize("zoom", [1.0, (1.0, 1.25, 1.25)])
def test_oracle_model_adjoint_identity(zoom):
    rng = np.random.default_rng(1)
    m = OT.make_forward_model(_flat_psf((3, 5, 5)), (4, 16, 16), zoom, dtype=np.float64)
    x = rng.random(m.padded_shape)
    y = rng.random(m.data_shape)
    fx = m.op.forward(x)
    assert fx.shape == m.data_shape
    lhs, rhs = float(np.vdot(fx, y)), float(np.vdot(x, m.op.adjoint(y)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert tuple(s.stop - s.start for s in m.valid_slices) == m.visible_shape


def _delta_peak(psf_size, data=48, zoom=1.25):
    psf = np.zeros((psf_size, psf_size))
    psf[0, 0] = 1.0  # corner-origin delta: identity blur
    m = OT.make_forward_model(psf, (data, data), zoom, dtype=np.float64)
    vs = m.valid_slices
    x = np.zeros(m.padded_shape)
    x[(vs[0].start + vs[0].stop) // 2, (vs[1].start + vs[1].stop) // 2] = 1.0
    y = m.op.forward(x)
    py, px = np.unravel_index(np.argmax(y), y.shape)
    win = y[py - 1:py + 2, px]
    return (py, px), win / win.sum()


def test_oracle_registration_independent_of_psf_parity():
    ref_peak, ref_split = _delta_peak(25)
    for n in range(20, 31):
        peak, split = _delta_peak(n)
        assert peak == ref_peak
        np.testing.assert_allclose(split, ref_split, atol=1e-5)
        assert split.max() > 0.999


def test_oracle_tiled_rl_matches_full_interior():
    truth = np.zeros((4, 40, 40), np.float32)
    truth[2, 10:30:6, 10:30:6] = 200.0
    psf = _flat_psf((3, 5, 5))
    full_model = OT.make_forward_model(psf, truth.shape, 1.0)
    x = np.zeros(full_model.padded_shape, np.float32)
    x[full_model.valid_slices] = truth
    data = (full_model.op.forward(x) + 1.0).astype(np.float32)
    solve = OT.richardson_lucy_solver(num_iter=10)
    full = solve(data, full_model)
    tiled = OT.process_tiles(data, psf, 1.0, solve, guard=6, tile_size=24)
    scale = float(np.max(full))
    np.testing.assert_allclose(tiled / scale, full / scale, atol=0.05)


def test_host_path_group_sizes():
    """Tiles solved together when the volume does not fit the device: bounded by half of
    the free device memory per tile's footprint, by the request and the tile count, and
    one at a time whenever a per-tile callback may stop the loop early."""
    gib = 1 << 30
    assert TP._host_group(20, 20, 2 * gib, 16 * gib, False) == 4
    assert TP._host_group(20, 3, 2 * gib, 16 * gib, False) == 3
    assert TP._host_group(2, 20, 1 * gib, 64 * gib, False) == 2
    assert TP._host_group(20, 20, 64 * gib, 16 * gib, False) == 1
    assert TP._host_group(20, 20, 1, 0, False) == 1
    assert TP._host_group(20, 20, 1 * gib, 64 * gib, True) == 1
