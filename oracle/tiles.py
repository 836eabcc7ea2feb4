"""CPU oracle for the tiled path (TEST INFRASTRUCTURE, not the product).

Only ``tests/`` may import this module.  It restates, over the NumPy
operators of ``oracle.deconv``:

* ``compute_padded_shape`` / ``get_valid_slices``   (deconlib/deconvolution/shapes.py:92-214)
* ``make_forward_model``                            (deconlib/deconvolution/forward_model.py:53-121)
* ``optimal_tile_size`` / ``plan_tiles`` / ``process_tiles``
                                                    (deconlib/deconvolution/tile_processing.py:99-288)
* ``richardson_lucy_solver``                        (deconlib/deconvolution/rl_mlx.py:188-228)

Parity pins (tests/test_tiles.py): the reference tests' known answers and
invariants for the tiling (tests/test_tile_processing.py:76-150: balanced
cores, 5-smooth windows, uniform read shape, write partition, reads in
bounds), the passthrough stitch reproducing the input exactly (:157-173), the
delta-PSF registration phase independent of PSF parity
(tests/test_forward_model_registration.py:18-59), and the adjoint identity of
the composed operator.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import product

import numpy as np

from . import deconv as O


def compute_padded_shape(signal_shape, kernel_shape):
    padded, padding = [], []
    for n, m in zip(signal_shape, kernel_shape):
        total = int(m) - 1
        padding.append((total // 2, total - total // 2))
        padded.append(int(n) + total)
    return tuple(padded), tuple(padding)


def get_valid_slices(padded_shape, padding):
    return tuple(slice(b, p - a) for p, (b, a) in zip(padded_shape, padding))


@dataclass(frozen=True)
class ForwardModel:
    op: object
    data_shape: tuple
    visible_shape: tuple
    padded_shape: tuple
    valid_slices: tuple


def make_forward_model(psf, data_shape, zoom=1.0, dtype=np.float32):
    """convolve -> crop at the pad-before offset -> downsample (forward_model.py:53-121)."""
    nd = len(data_shape)
    zoom = (float(zoom),) * nd if np.isscalar(zoom) else tuple(float(z) for z in zoom)
    if any(z < 1.0 for z in zoom):
        raise ValueError("zoom must be >= 1")
    data_shape = tuple(int(d) for d in data_shape)
    visible = tuple(max(1, round(d * z)) for d, z in zip(data_shape, zoom))
    padded, padding = compute_padded_shape(visible, psf.shape)
    conv = O.LinearFFTConvolver(psf, padded, dtype=dtype)
    crop = O.Crop(padded, visible, start=tuple(b for b, _ in padding))
    if visible == data_shape:
        op = O.compose(crop, conv)
    else:
        down = O.FractionalAreaDownsample(tuple(v / d for v, d in zip(visible, data_shape)), in_shape=visible)
        op = O.compose(down, crop, conv)
    return ForwardModel(op, data_shape, visible, padded, get_valid_slices(padded, padding))


def next_smooth(n):
    return O.next_smooth_number(n)


def optimal_tile_size(N, guard, tile_size):
    if N <= tile_size:
        return N
    k = -(-N // tile_size)
    core = -(-N // k)
    if core + 2 * guard > N:
        raise ValueError("guard too large")
    return min(next_smooth(core + 2 * guard), N)


def axis_tiles(N, V, T, guard, zoom):
    if T >= N:
        return [(slice(0, N), slice(0, V), slice(0, V))]
    core = T - 2 * guard
    if core <= 0:
        raise ValueError("tile size must exceed 2 * guard")
    res = []
    for i in range(-(-N // core)):
        a, b = i * core, min(i * core + core, N)
        r0 = min(max(a - guard, 0), N - T)
        w0, w1, v0 = round(a * zoom), round(b * zoom), round(r0 * zoom)
        res.append((slice(r0, r0 + T), slice(w0, w1), slice(w0 - v0, w1 - v0)))
    return res


def plan_tiles(data_shape, zoom, guard, tile_size=512, min_z_slices=48):
    """Returns (tile_shape, visible_shape, [(read, write, crop)])."""
    nd = len(data_shape)
    zoom = (float(zoom),) * nd if np.isscalar(zoom) else tuple(float(z) for z in zoom)
    visible = tuple(max(1, round(n * z)) for n, z in zip(data_shape, zoom))
    tiled = {nd - 1, nd - 2} | ({nd - 3} if nd >= 3 and data_shape[nd - 3] > min_z_slices else set())
    shape, axes = [], []
    for i, (N, V, z) in enumerate(zip(data_shape, visible, zoom)):
        T = optimal_tile_size(N, guard, tile_size) if i in tiled else N
        shape.append(T)
        axes.append(axis_tiles(N, V, T, guard, z))
    tiles = [tuple(zip(*combo)) for combo in product(*axes)]
    return tuple(shape), visible, tiles


def process_tiles(data, psf, zoom, solve, guard=None, tile_size=512, min_z_slices=48, dtype=np.float32):
    if guard is None:
        guard = max(psf.shape[-2:]) // 2
    shape, visible, tiles = plan_tiles(data.shape, zoom, guard, tile_size, min_z_slices)
    model = make_forward_model(psf, shape, zoom, dtype=dtype)
    out = np.zeros(visible, dtype=np.float32)
    for read, write, crop in tiles:
        res = np.asarray(solve(np.asarray(data[read], dtype=np.float32), model))
        crop = tuple(slice(c.start, min(c.stop, res.shape[j])) for j, c in enumerate(crop))
        core = res[crop]
        out[tuple(slice(w.start, w.start + core.shape[j]) for j, w in enumerate(write))] = core
    return out


def richardson_lucy_solver(num_iter=50, background=0.0, init_value=None, eval_interval=10):
    def solve(data, model):
        init = None if init_value is None else np.full(model.padded_shape, float(init_value), np.float32)
        res = O.richardson_lucy(data, model.op, num_iter=num_iter, background=background, init=init,
                                eval_interval=eval_interval)
        return res.restored[model.valid_slices]

    return solve
