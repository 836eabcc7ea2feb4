"""Launch-shape autotuner for the convolution passes.

For one operator, each of the five passes is timed (HIP events around the
pass, ``dh_timing_*``) for every feasible workgroup tile (columns per
workgroup for the Y and Z passes, rows per workgroup for the X passes), each
launched either one workgroup per tile or as a persistent grid, and the
fastest is kept on the plan.  The choice is semantics-free: every launch shape
computes the same values.  Results are cached per (canvas, signal) shape in
memory and, optionally, in a JSON file.
"""

from __future__ import annotations

import ctypes as C
import json
import os
import sys
from typing import Dict, Optional

import torch

from . import _native as N
from ._tensor import empty_f32

_CACHE: Dict[str, Dict[int, int]] = {}


def _name(tile: int) -> str:
    return f"{tile & 0xFFFF}p" if tile & N.TILE_PERSISTENT else str(tile)


def _key(plan) -> str:
    i = plan.info
    return "s{}-k{}-p{}".format(tuple(i.signal_shape), tuple(i.kernel_shape), tuple(i.canvas_shape))


def _time_passes(op, x, reps: int):
    plan = op._plan
    out = empty_f32(x.shape)
    for _ in range(2):
        plan.apply(x, False, out)
        plan.apply(x, True, out)
    torch.cuda.synchronize()
    N.call("dh_timing_read", None, None, 1)
    N.call("dh_timing_enable", 1)
    try:
        for _ in range(reps):
            plan.apply(x, False, out)
            plan.apply(x, True, out)
        torch.cuda.synchronize()
    finally:
        N.call("dh_timing_enable", 0)
    ms = (C.c_double * 5)()
    cnt = (C.c_int64 * 5)()
    N.call("dh_timing_read", ms, cnt, 1)
    return [ms[i] / cnt[i] if cnt[i] else float("inf") for i in range(5)]


def autotune(op, reps: int = 10, candidates=(1, 2, 4, 8, 16, 32, 64), cache_file: Optional[str] = None,
             verbose: bool = False, passes=range(5)) -> Dict[int, int]:
    """Picks the fastest tile for every pass in ``passes`` (0..4: x_fwd, y_fwd, z_conv,
    y_inv, x_inv) of ``op`` (a LinearFFTConvolver/FFTConvolver); the others keep theirs."""
    plan = op._plan
    passes = tuple(sorted(set(int(p) for p in passes)))
    key = _key(plan) + ("" if passes == tuple(range(5)) else "-p" + "".join(map(str, passes)))
    if cache_file and key not in _CACHE and os.path.exists(cache_file):
        with open(cache_file) as f:
            for k, v in json.load(f).items():
                _CACHE[k] = {int(p): int(t) for p, t in v.items()}
    if key in _CACHE:
        for p, t in _CACHE[key].items():
            N.call("dh_conv_plan_set_tile", plan._h, p, t)
        return dict(_CACHE[key])
    x = torch.rand(tuple(op.in_shape), dtype=torch.float32, device="cuda")
    best: Dict[int, int] = {}
    for p in passes:
        timings = {}
        for t in candidates:
            for v in (t, t | N.TILE_PERSISTENT):
                if N.fns["dh_conv_plan_set_tile"](plan._h, p, v) != 0:
                    continue
                timings[v] = _time_passes(op, x, reps)[p]
        N.call("dh_conv_plan_set_tile", plan._h, p, 0)
        if timings:
            choice = min(timings, key=timings.get)
            default = _time_passes(op, x, reps)[p]
            if timings[choice] < default:
                N.call("dh_conv_plan_set_tile", plan._h, p, choice)
                best[p] = choice
            if verbose:
                print(f"[autotune] pass {N.PASS_NAMES[p]}: " +
                      ", ".join(f"{_name(t)}:{v * 1e3:.1f}us" for t, v in sorted(timings.items())) +
                      f" default {default * 1e3:.1f}us -> {_name(best[p]) if p in best else 'default'}",
                      file=sys.stderr)
    _CACHE[key] = best
    if cache_file:
        data = {}
        if os.path.exists(cache_file):
            with open(cache_file) as f:
                data = json.load(f)
        data[key] = {str(p): t for p, t in best.items()}
        with open(cache_file, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    return best
