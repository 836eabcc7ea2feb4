"""The tiled path on the GPU: forward model, process_tiles, batched tile RL.

Checked against the oracle (oracle.tiles) on the same inputs: the composed
operator's forward/adjoint, the registration phase of even and odd PSFs,
the exact passthrough stitch, per-tile RL (restored and loss history, which
tests the data-box reduction of the fused iteration), the batched tile path
against the per-tile solve, and tiled versus full RL in the interior.
"""

import numpy as np
import pytest
import torch

from oracle import tiles as OT

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _host(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def _psf(shape, seed):
    rng = np.random.default_rng(seed)
    p = rng.random(shape) ** 4
    return (p / p.sum()).astype(np.float32)


@pytest.mark.parametrize("zoom", [1.0, (1.0, 1.25, 1.25), (1.0, 1.5, 1.5)])
def test_forward_model_matches_oracle(gpu, zoom):
    from deconlib_amd import ForwardModel, make_forward_model

    psf = _psf((3, 7, 6), 1)
    m = make_forward_model(psf, (4, 20, 20), zoom=zoom)
    o = OT.make_forward_model(psf, (4, 20, 20), zoom, dtype=np.float64)
    assert isinstance(m, ForwardModel)
    assert (m.data_shape, m.visible_shape, m.padded_shape, m.valid_slices) == \
        (o.data_shape, o.visible_shape, o.padded_shape, o.valid_slices)
    rng = np.random.default_rng(2)
    x = rng.random(m.padded_shape).astype(np.float32)
    y = rng.random(m.data_shape).astype(np.float32)
    fx = _host(m.op.forward(torch.from_numpy(x).cuda()))
    aty = _host(m.op.adjoint(torch.from_numpy(y).cuda()))
    assert fx.shape == m.data_shape and aty.shape == m.padded_shape
    assert rel_l2(fx, o.op.forward(x.astype(np.float64))) <= 1e-5
    assert rel_l2(aty, o.op.adjoint(y.astype(np.float64))) <= 1e-5


def test_registration_independent_of_psf_parity(gpu):
    from deconlib_amd import make_forward_model

    def peak(n):
        psf = np.zeros((n, n), np.float32)
        psf[0, 0] = 1.0
        m = make_forward_model(psf, (48, 48), zoom=1.25)
        vs = m.valid_slices
        x = torch.zeros(m.padded_shape, device="cuda")
        x[(vs[0].start + vs[0].stop) // 2, (vs[1].start + vs[1].stop) // 2] = 1.0
        y = _host(m.op.forward(x))
        py, px = np.unravel_index(np.argmax(y), y.shape)
        w = y[py - 1:py + 2, px]
        return (py, px), w / w.sum()

    ref_peak, ref_split = peak(25)
    for n in range(20, 31):
        p, split = peak(n)
        assert p == ref_peak
        np.testing.assert_allclose(split, ref_split, atol=1e-5)


def test_passthrough_stitches_input_exactly(gpu):
    from deconlib_amd import process_tiles

    data = np.random.default_rng(0).random((3, 61, 67)).astype(np.float32)
    psf = np.ones((3, 5, 5), np.float32) / 75

    def passthrough(tile, model):
        assert tile.shape == model.data_shape
        return tile

    out = process_tiles(data, psf, 1.0, passthrough, guard=4, tile_size=32)
    np.testing.assert_array_equal(out, data)


def test_model_built_once_and_early_stop(gpu):
    from deconlib_amd import process_tiles, richardson_lucy_solver

    seen = []

    def spy(tile, model):
        seen.append(model)
        return np.zeros(model.visible_shape, np.float32)

    psf = np.ones((3, 5, 5), np.float32) / 75
    process_tiles(np.ones((3, 48, 48), np.float32), psf, 1.0, spy, guard=4, tile_size=24)
    assert len(seen) > 1 and all(m is seen[0] for m in seen)
    data = np.random.default_rng(3).poisson(20, size=(4, 24, 24)).astype(np.float32)
    stops = []
    out = process_tiles(data, psf, 1.0, richardson_lucy_solver(num_iter=2), guard=2, tile_size=16,
                        on_tile_done=lambda spec, o: stops.append(spec.index) or len(stops) >= 2)
    assert len(stops) == 2 and out.dtype == np.float32


@pytest.mark.parametrize("zoom,shape", [(1.0, (4, 24, 24)), ((1.0, 1.5, 1.5), (4, 16, 16))])
def test_tiled_rl_matches_oracle(gpu, zoom, shape):
    """Tiled RL (batched fused tiles at zoom 1, the composed operator otherwise)
    against the oracle's tile-by-tile solve on the same data."""
    from deconlib_amd import process_tiles, richardson_lucy_solver

    data = np.random.default_rng(4).poisson(40, size=shape).astype(np.float32)
    psf = _psf((3, 5, 5), 5)
    for it, tol in ((1, 1e-5), (8, 1e-4)):
        got = process_tiles(data, psf, zoom, richardson_lucy_solver(num_iter=it), guard=2, tile_size=12)
        ref = OT.process_tiles(data, psf, zoom, OT.richardson_lucy_solver(num_iter=it), guard=2, tile_size=12)
        assert got.shape == ref.shape and np.all(got >= 0)
        assert rel_l2(got, ref) <= tol, (it, rel_l2(got, ref))


def test_batched_tiles_match_per_tile_solve_and_losses(gpu):
    """The lockstep tile jobs (private workspaces, data-box loss) against the
    composed-operator solve of each tile and the oracle's loss history."""
    from deconlib_amd import make_forward_model, richardson_lucy_solver
    from deconlib_amd.batch import TileRLJob, batched_richardson_lucy, fused_tiles_ok
    from oracle import deconv as O

    psf = _psf((5, 7, 7), 6)
    model = make_forward_model(psf, (12, 40, 40), zoom=1.0)
    assert fused_tiles_ok(model)
    rng = np.random.default_rng(7)
    tiles = [rng.poisson(30 * rng.random((12, 40, 40)) + 2).astype(np.float32) for _ in range(5)]
    solve = richardson_lucy_solver(num_iter=12, eval_interval=4)
    composed = [_host(solve.solve_composed(t, model)) for t in tiles]
    fused_one = [_host(solve(t, model)) for t in tiles]
    many = [_host(r) for r in solve.solve_batch(tiles, model)]
    for a, b, c in zip(many, fused_one, composed):
        assert np.array_equal(a, b)  # lockstep == alone, same kernels
        assert rel_l2(a, c) <= 1e-5
    # loss history of a tile job against the oracle's RL on the composed operator
    jobs = [TileRLJob(t, model, num_iter=12, eval_interval=4) for t in tiles[:2]]
    res = batched_richardson_lucy(jobs)
    om = OT.make_forward_model(psf, (12, 40, 40), 1.0)
    for t, r in zip(tiles[:2], res):
        ref = O.richardson_lucy(t, om.op, num_iter=12, eval_interval=4)
        assert rel_l2(_host(r.restored), ref.restored) <= 1e-4
        np.testing.assert_allclose(r.loss_history, ref.loss_history, rtol=1e-5)
        assert rel_l2(_host(r.pred), ref.pred) <= 1e-4


def test_tiled_matches_full_in_interior(gpu):
    from deconlib_amd import make_forward_model, process_tiles, richardson_lucy_solver

    truth = np.zeros((4, 40, 40), np.float32)
    truth[2, 10:30:6, 10:30:6] = 200.0
    psf = np.ones((3, 5, 5), np.float32) / 75
    full_model = make_forward_model(psf, truth.shape, zoom=1.0)
    x = torch.zeros(full_model.padded_shape, device="cuda")
    x[full_model.valid_slices] = torch.from_numpy(truth).cuda()
    data = _host(full_model.op.forward(x)) + 1.0
    solve = richardson_lucy_solver(num_iter=10)
    full = _host(solve(data, full_model))
    tiled = process_tiles(data, psf, 1.0, solve, guard=6, tile_size=24)
    scale = float(np.max(full))
    np.testing.assert_allclose(tiled / scale, full / scale, atol=0.05)


def test_data_box_outside_grid_rejected(gpu):
    import deconlib_amd._native as N
    from deconlib_amd import make_forward_model
    from deconlib_amd.batch import TileRLJob

    psf = np.ones((3, 5, 5), np.float32) / 75
    model = make_forward_model(psf, (8, 16, 16), zoom=1.0)
    job = TileRLJob(np.ones((8, 16, 16), np.float32), model, num_iter=2)
    item = job._rl_item()
    item.data_box[:] = [0, 99, 0, 4, 0, 4]
    arr = (N.RLItem * 1)(item)
    with pytest.raises(ValueError):
        N.call("dh_batch_rl_iteration", 1, arr, N.stream_ptr())
