"""Tiled deconvolution of volumes too large for one pass (deconvolution/tile_processing.py).

Fixed-shape sliding windows (overlap-save): every tile reads a data window of
the same shape, so one ForwardModel serves all of them; interior tiles step by
the core (window minus a guard on each side), the last window along an axis is
shifted back inside the image, and each tile owns a disjoint interval of the
output, so every visible pixel is written exactly once.

``plan_tiles`` / ``optimal_tile_size`` are host arithmetic with the reference's
rounding.  ``process_tiles`` builds the one forward model and stitches the
owned cores; a solver that also offers ``solve.solve_batch(tiles, model)``
(``richardson_lucy_solver`` does) receives the tiles in groups (small tiles
advance in lockstep, one batched launch per pass for the group).  When the
volume and the output fit comfortably in device memory and no per-tile
callback is asked for, the volume is uploaded once, tiles are cut and cores
stitched on the device, and the output is downloaded once — for a batching
solver, and for any solver marked ``solve.accepts_device_tiles`` (the
package's NLCG and ER-Decon adapters), which then receives device tiles one at
a time.
"""

from __future__ import annotations

from dataclasses import dataclass
from itertools import product
from typing import Callable, List, Optional, Tuple, Union

import numpy as np

from .forward_model import ForwardModel, make_forward_model
from .linops import _next_smooth_number

Solver = Callable[[np.ndarray, ForwardModel], np.ndarray]


@dataclass(frozen=True)
class TileSpec:
    """read: data-space window; write: owned core in the output; crop: the core in the tile's result."""

    index: Tuple[int, ...]
    read: Tuple[slice, ...]
    write: Tuple[slice, ...]
    crop: Tuple[slice, ...]


@dataclass(frozen=True)
class TilePlan:
    """Uniform read shape tile_shape, stitched output shape visible_shape, one TileSpec per tile."""

    data_shape: Tuple[int, ...]
    tile_shape: Tuple[int, ...]
    visible_shape: Tuple[int, ...]
    tiles: List[TileSpec]


def _axis_tiles(N: int, V: int, T: int, guard: int, zoom: float):
    """(read, write, crop) slices along one axis (tile_processing.py:99-125).

    Ownership bounds are rounded from global positions, so rounding never drifts.
    """
    if T >= N:
        return [(slice(0, N), slice(0, V), slice(0, V))]
    core = T - 2 * guard
    if core <= 0:
        raise ValueError(f"tile size ({T}) must be > 2 * guard ({2 * guard})")
    out = []
    for i in range(-(-N // core)):
        own0 = i * core
        own1 = min(own0 + core, N)
        r0 = min(max(own0 - guard, 0), N - T)
        w0, w1 = round(own0 * zoom), round(own1 * zoom)
        v0 = round(r0 * zoom)
        out.append((slice(r0, r0 + T), slice(w0, w1), slice(w0 - v0, w1 - v0)))
    return out


def optimal_tile_size(N: int, guard: int, tile_size: int) -> int:
    """Balanced 5-smooth read window for one axis (tile_processing.py:128-165)."""
    if N <= tile_size:
        return N
    n_tiles = -(-N // tile_size)
    core = -(-N // n_tiles)
    window = core + 2 * guard
    if window > N:
        raise ValueError(
            f"guard ({guard}) too large for tile_size ({tile_size}) on an "
            f"axis of length {N}: balanced core {core} + 2*guard = {window} "
            f"would exceed the axis"
        )
    return min(_next_smooth_number(window), N)


def plan_tiles(data_shape: Tuple[int, ...], zoom: Union[float, Tuple[float, ...]], guard: int,
               tile_size: int = 512, min_z_slices: int = 48) -> TilePlan:
    """Y and X always tiled, Z only when longer than min_z_slices (tile_processing.py:168-227)."""
    ndim = len(data_shape)
    if ndim < 2:
        raise ValueError("data_shape must be at least 2D")
    if isinstance(zoom, (int, float)):
        zoom = (float(zoom),) * ndim
    else:
        zoom = tuple(float(z) for z in zoom)
    data_shape = tuple(int(n) for n in data_shape)
    visible_shape = tuple(max(1, round(n * z)) for n, z in zip(data_shape, zoom))
    tiled = {ndim - 1, ndim - 2}
    if ndim >= 3 and data_shape[ndim - 3] > min_z_slices:
        tiled.add(ndim - 3)
    per_axis, tile_shape = [], []
    for i, (N, V, z) in enumerate(zip(data_shape, visible_shape, zoom)):
        T = optimal_tile_size(N, guard, tile_size) if i in tiled else N
        tile_shape.append(T)
        per_axis.append(_axis_tiles(N, V, T, guard, z))
    tiles = [
        TileSpec(index=tuple(i for i, _ in combo), read=tuple(t[0] for _, t in combo),
                 write=tuple(t[1] for _, t in combo), crop=tuple(t[2] for _, t in combo))
        for combo in product(*(list(enumerate(a)) for a in per_axis))
    ]
    return TilePlan(data_shape=data_shape, tile_shape=tuple(tile_shape), visible_shape=visible_shape, tiles=tiles)


def _to_host(a) -> np.ndarray:
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def process_tiles(data, psf, zoom: Union[float, Tuple[float, ...]], solve: Solver, *,
                  guard: Optional[int] = None, tile_size: int = 512, min_z_slices: int = 48,
                  on_tile_done: Optional[Callable] = None, verbose: bool = False,
                  batch_tiles: Optional[int] = None) -> np.ndarray:
    """Deconvolve tile by tile with any solver; returns the stitched float32 volume (tile_processing.py:230-288).

    ``batch_tiles``: tiles handed to ``solve_batch`` at a time (default: all;
    the solver sizes its lockstep groups itself; 1 = one tile at a time).
    Results and callback order are those of the one-tile loop: tiles are
    stitched and ``on_tile_done`` called in plan order, and a truthy return
    stops before any later tile is written.
    """
    if guard is None:
        guard = max(psf.shape[-2:]) // 2
    plan = plan_tiles(data.shape, zoom, guard, tile_size, min_z_slices)
    model = make_forward_model(psf, plan.tile_shape, zoom)
    n = len(plan.tiles)
    batched = getattr(solve, "solve_batch", None)
    group = n if batch_tiles is None else max(1, int(batch_tiles))
    if batched is None:
        group = 1
    if on_tile_done is None and _fits_device(data, plan):
        if batched is not None:
            return _process_resident(data, plan, model, batched, group, verbose)
        if getattr(solve, "accepts_device_tiles", False):
            return _process_resident(data, plan, model, lambda tiles, m: [solve(tl, m) for tl in tiles], 1,
                                     verbose)
    output = np.zeros(plan.visible_shape, dtype=np.float32)
    if batched is not None and group > 1:
        group = _host_group(n, group, _tile_bytes(model), _free_device_bytes(), on_tile_done is not None)

    i = 0
    while i < n:
        specs = plan.tiles[i:i + group]
        tiles = [np.asarray(data[s.read], dtype=np.float32) for s in specs]
        results = batched(tiles, model) if (batched is not None and len(specs) > 1) else [solve(t, model)
                                                                                              for t in tiles]
        for spec, res in zip(specs, results):
            restored = _to_host(res)
            # a +-1 rounding guard: the crop is clipped to the result, the write matched to the core
            crop = tuple(slice(c.start, min(c.stop, restored.shape[j])) for j, c in enumerate(spec.crop))
            core = restored[crop]
            write = tuple(slice(w.start, w.start + core.shape[j]) for j, w in enumerate(spec.write))
            output[write] = core
            i += 1
            if verbose:
                print(f"  tile {i}/{n} {spec.index} done")
            if on_tile_done is not None and on_tile_done(spec, output):
                return output
    return output


def _host_group(n_tiles: int, requested: int, tile_bytes: int, free_bytes: int, has_callback: bool) -> int:
    """Tiles solved together on the host path (the volume does not fit the device):
    one at a time when a per-tile callback may stop early (so no later tile is
    solved in vain), else as many as the request allows within half of the free
    device memory (each tile's solver state is resident while its group runs)."""
    if has_callback:
        return 1
    fit = max(1, (free_bytes // 2) // max(int(tile_bytes), 1))
    return max(1, min(int(requested), int(n_tiles), int(fit)))


def _tile_bytes(model) -> int:
    """Device bytes one tile's solve holds (padded-domain state and plan workspace)."""
    from .batch import _tile_job_bytes, fused_tiles_ok

    if fused_tiles_ok(model):
        return _tile_job_bytes(model)
    pad = 1
    for v in model.padded_shape:
        pad *= int(v)
    return 8 * 4 * pad


def _free_device_bytes() -> int:
    import torch

    if not torch.cuda.is_available():
        return 0
    return int(torch.cuda.mem_get_info()[0])


def _fits_device(data, plan) -> bool:
    """Input and stitched output together take at most a quarter of free HBM."""
    import torch

    if not torch.cuda.is_available():
        return False
    vis = 1
    for v in plan.visible_shape:
        vis *= int(v)
    free, _ = torch.cuda.mem_get_info()
    return 4 * (int(np.prod(data.shape)) + vis) <= free // 4


def _process_resident(data, plan, model, batched, group, verbose) -> np.ndarray:
    """process_tiles with the volume uploaded once, tiles cut and cores stitched on
    the device, and the output downloaded once (no per-tile host transfers)."""
    import torch

    from ._tensor import device

    dev_data = torch.from_numpy(np.ascontiguousarray(data, dtype=np.float32)).to(device())
    out = torch.zeros(plan.visible_shape, dtype=torch.float32, device=device())
    n = len(plan.tiles)
    for s in range(0, n, group):
        specs = plan.tiles[s:s + group]
        results = batched([dev_data[sp.read].contiguous() for sp in specs], model)
        for j, (spec, res) in enumerate(zip(specs, results)):
            crop = tuple(slice(c.start, min(c.stop, res.shape[k])) for k, c in enumerate(spec.crop))
            core = res[crop]
            out[tuple(slice(w.start, w.start + core.shape[k]) for k, w in enumerate(spec.write))] = core
            if verbose:
                print(f"  tile {s + j + 1}/{n} {spec.index} done")
    return out.cpu().numpy()
