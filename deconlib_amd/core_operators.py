"""Pad / Crop / FractionalAreaDownsample on MI355X (core_operators.py:29-352).

Same constructors, attributes, norms and ``ValueError`` checks as the
reference; the data movement runs in libdeconhip (``dh_box_copy`` for
pad/crop, ``dh_banded_apply`` for the banded fractional-area weights).
"""

from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Tuple, Union

import numpy as np
import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32


def _as3(shape):
    s = tuple(int(v) for v in shape)
    return (1,) * (3 - len(s)) + s


def _box(src: torch.Tensor, dst_shape, src_start, dst_start, extent) -> torch.Tensor:
    nd = src.dim()
    if nd > 3:
        raise ValueError("at most 3-D volumes are supported")
    out = empty_f32(dst_shape)
    a = lambda v, fill: (C.c_int64 * 3)(*((fill,) * (3 - nd) + tuple(int(t) for t in v)))
    N.call("dh_box_copy", src.data_ptr(), a(src.shape, 1), out.data_ptr(), a(dst_shape, 1),
           a(src_start, 0), a(dst_start, 0), a(extent, 1), N.stream_ptr())
    return out


class Pad:
    """Zero-padding; adjoint removes the padding (core_operators.py:29-59)."""

    def __init__(self, padding: Tuple[Tuple[int, int], ...]):
        self.padding = tuple((int(pb), int(pa)) for pb, pa in padding)
        self.operator_norm_sq = 1.0

    def forward(self, x) -> torch.Tensor:
        x = as_f32(x)
        out_shape = tuple(s + pb + pa for s, (pb, pa) in zip(x.shape, self.padding))
        return _box(x, out_shape, (0,) * x.dim(), [pb for pb, _ in self.padding], x.shape)

    def adjoint(self, y) -> torch.Tensor:
        y = as_f32(y)
        out_shape = tuple(s - pb - pa for s, (pb, pa) in zip(y.shape, self.padding))
        return _box(y, out_shape, [pb for pb, _ in self.padding], (0,) * y.dim(), out_shape)

    def __call__(self, x):
        return self.forward(x)


class Crop:
    """Crop to ``target_shape`` (centred unless ``start``); adjoint zero-pads (core_operators.py:62-137)."""

    def __init__(self, original_shape, target_shape, start=None):
        if len(original_shape) != len(target_shape):
            raise ValueError(
                f"original_shape {original_shape} and target_shape {target_shape} "
                "must have same number of dimensions"
            )
        self.original_shape = tuple(int(s) for s in original_shape)
        self.target_shape = tuple(int(s) for s in target_shape)
        self.in_shape = self.original_shape
        self.out_shape = self.target_shape
        self.operator_norm_sq = 1.0
        if start is not None and len(start) != len(self.original_shape):
            raise ValueError(
                f"start {start} must have same number of dimensions as "
                f"original_shape {self.original_shape}"
            )
        self._start = []
        self._padding = []
        for i, (orig, tgt) in enumerate(zip(self.original_shape, self.target_shape)):
            if tgt > orig:
                raise ValueError(f"Target size {tgt} > original size {orig}. Use Pad for upsizing.")
            s = (orig - tgt) // 2 if start is None else int(start[i])
            if start is not None and (s < 0 or s + tgt > orig):
                raise ValueError(
                    f"start {s} on axis {i} puts crop window [{s}, {s + tgt}) outside original size {orig}"
                )
            self._start.append(s)
            self._padding.append((s, orig - s - tgt))
        self._slices = tuple(slice(s, s + t) for s, t in zip(self._start, self.target_shape))
        self._padding = tuple(self._padding)

    def forward(self, x) -> torch.Tensor:
        x = as_f32(x)
        return _box(x, self.target_shape, self._start, (0,) * x.dim(), self.target_shape)

    def adjoint(self, y) -> torch.Tensor:
        y = as_f32(y)
        return _box(y, self.original_shape, (0,) * y.dim(), self._start, self.target_shape)

    def __call__(self, x):
        return self.forward(x)


def _banded_overlap_weights_1d(n_large: int, n_small: int):
    """Banded interval-overlap weights (core_operators.py:151-202), host float64 -> int32/float32."""
    scale = n_large / n_small
    w_f = min(n_large, int(math.ceil(scale)) + 2)
    i = np.arange(n_small)
    start_f = np.clip(np.floor(i * scale).astype(np.int64), 0, n_large - w_f)
    idx_f = start_f[:, None] + np.arange(w_f)[None, :]
    i0, i1 = (i * scale)[:, None], ((i + 1) * scale)[:, None]
    j0 = idx_f.astype(np.float64)
    wt_f = np.maximum(0.0, np.minimum(i1, j0 + 1) - np.maximum(i0, j0))
    w_a = min(n_small, int(math.ceil(1.0 / scale)) + 3)
    j = np.arange(n_large)
    start_a = np.clip(np.floor(j / scale).astype(np.int64) - 1, 0, n_small - w_a)
    idx_a = start_a[:, None] + np.arange(w_a)[None, :]
    a0 = idx_a.astype(np.float64) * scale
    a1 = (idx_a.astype(np.float64) + 1) * scale
    b0 = j[:, None].astype(np.float64)
    wt_a = np.maximum(0.0, np.minimum(a1, b0 + 1) - np.maximum(a0, b0))
    return (idx_f.astype(np.int32), wt_f.astype(np.float32), idx_a.astype(np.int32), wt_a.astype(np.float32))


def _apply_banded(x: torch.Tensor, idx: torch.Tensor, wt: torch.Tensor, axis: int) -> torch.Tensor:
    axis = axis % x.dim()
    shape = tuple(x.shape)
    outer = int(np.prod(shape[:axis], dtype=np.int64))
    inner = int(np.prod(shape[axis + 1:], dtype=np.int64))
    n_out, width = idx.shape
    out_shape = shape[:axis] + (n_out,) + shape[axis + 1:]
    out = empty_f32(out_shape)
    N.call("dh_banded_apply", x.data_ptr(), out.data_ptr(), outer, shape[axis], inner, n_out,
           idx.data_ptr(), wt.data_ptr(), width, N.stream_ptr())
    return out


class FractionalAreaDownsample:
    """Fractional-area fine -> coarse resampling (core_operators.py:222-352)."""

    def __init__(self, scale: Union[float, Tuple[float, ...]], axes: Tuple[int, ...] = None,
                 in_shape: Tuple[int, ...] = None):
        if isinstance(scale, (int, float)):
            scale = (float(scale),)
        self.scale = tuple(float(s) for s in scale)
        if any(s < 1.0 for s in self.scale):
            raise ValueError(
                f"FractionalAreaDownsample requires scale >= 1 (fine -> coarse); "
                f"got {self.scale}. Use FractionalAreaUpsample to upsample."
            )
        self.axes = axes
        self._band_cache: dict = {}
        self._block_cache: dict = {}
        self._axis_map: dict = {}
        self.in_shape = None
        self.out_shape = None
        if in_shape is not None:
            ax = axes if axes is not None else tuple(range(len(in_shape)))
            sc = self.scale * len(ax) if len(self.scale) == 1 and len(ax) > 1 else self.scale
            for a, s in zip(ax, sc):
                nl = in_shape[a]
                self._axis_map[a] = (nl, max(1, int(round(nl / s))))
            self.in_shape = tuple(int(s) for s in in_shape)
            o = list(self.in_shape)
            for a, (_, ns) in self._axis_map.items():
                o[a % len(o)] = ns
            self.out_shape = tuple(o)
        if self._axis_map:
            nsq = 1.0
            for nl, ns in self._axis_map.values():
                nsq *= nl / ns
        else:
            nsq = math.prod(self.scale)
        self.operator_norm_sq = float(nsq)

    def _get_band(self, n_large: int, n_small: int):
        key = (n_large, n_small)
        if key not in self._band_cache:
            i_f, w_f, i_a, w_a = _banded_overlap_weights_1d(n_large, n_small)
            dev = device()
            self._band_cache[key] = tuple(torch.from_numpy(a).to(dev) for a in (i_f, w_f, i_a, w_a))
        return self._band_cache[key]

    def _axes_scales(self, ndim):
        ax = self.axes if self.axes is not None else tuple(range(ndim))
        sc = self.scale * len(ax) if len(self.scale) == 1 and len(ax) > 1 else self.scale
        return ax, sc

    def axis_pairs(self, ndim, shape, forward=True):
        """[(axis, n_large, n_small)] for the non-identity axes of one application."""
        pairs = []
        cur = list(shape)
        for a, s in zip(*self._axes_scales(ndim)):
            if a in self._axis_map:
                nl, ns = self._axis_map[a]
            elif forward:
                nl = cur[a]
                ns = max(1, int(round(nl / s)))
            else:
                ns = cur[a]
                nl = max(1, int(round(ns * s)))
            if nl != ns:
                pairs.append((a, nl, ns))
            cur[a] = ns if forward else nl
        return pairs

    def _block_factors(self, pairs, ndim):
        """Per-axis integer factors (z, y, x of the volume promoted to 3-D) when every
        band table of ``pairs`` is a unit block sum -- then one dh_block_sum /
        dh_block_spread pass equals the banded passes bit for bit -- else None."""
        key = (ndim, tuple(pairs))
        if key in self._block_cache:
            return self._block_cache[key]
        f = [1, 1, 1]
        ok = ndim <= 3
        for a, nl, ns in pairs:
            if not ok:
                break
            if nl % ns:
                ok = False
                break
            fac = nl // ns
            i_f, w_f, i_a, w_a = _banded_overlap_weights_1d(nl, ns)
            o = np.arange(ns)[:, None]
            in_block = (i_f >= fac * o) & (i_f < fac * (o + 1))
            ok = bool(np.all(np.where(in_block, w_f == 1.0, w_f == 0.0)) and np.all(in_block.sum(1) == fac))
            j = np.arange(nl)[:, None]
            parent = i_a == j // fac
            ok = ok and bool(np.all(np.where(parent, w_a == 1.0, w_a == 0.0)) and np.all(parent.sum(1) == 1))
            f[(a % ndim) + 3 - ndim] = fac
        res = tuple(f) if ok else None
        self._block_cache[key] = res
        return res

    def forward(self, x) -> torch.Tensor:
        r = as_f32(x)
        pairs = self.axis_pairs(r.dim(), r.shape, True)
        f = self._block_factors(pairs, r.dim()) if pairs else None
        if f is not None:
            n = (1,) * (3 - r.dim()) + tuple(int(v) for v in r.shape)
            out = empty_f32(tuple(int(v) // int(q) for v, q in zip(r.shape, f[3 - r.dim():])))
            N.call("dh_block_sum", r.data_ptr(), out.data_ptr(), *n, *f, N.stream_ptr())
            return out
        for a, nl, ns in pairs:
            i_f, w_f, _, _ = self._get_band(nl, ns)
            r = _apply_banded(r, i_f, w_f, a)
        return r

    def sum_factors(self, shape) -> Optional[Tuple[int, int, int]]:
        """Integer factors (z, y, x promoted to 3-D) when the forward of a volume of
        ``shape`` is a unit block sum, else None."""
        shape = tuple(int(v) for v in shape)
        if len(shape) > 3:
            return None
        pairs = self.axis_pairs(len(shape), shape, True)
        return self._block_factors(pairs, len(shape)) if pairs else None

    def spread_factors(self, y) -> Optional[Tuple[int, int, int]]:
        """Integer factors (z, y, x of the volume promoted to 3-D) when the adjoint of
        ``y`` is a unit block spread with an x factor of 1 or 2 -- then a convolution can
        read ``y`` through its spread prologue instead of a materialised fine volume --
        else None."""
        if not isinstance(y, torch.Tensor) or y.dim() > 3:
            return None
        pairs = self.axis_pairs(y.dim(), y.shape, False)
        f = self._block_factors(pairs, y.dim()) if pairs else None
        return f if f is not None and f[2] <= 2 else None

    def adjoint(self, y) -> torch.Tensor:
        r = as_f32(y)
        pairs = self.axis_pairs(r.dim(), r.shape, False)
        f = self._block_factors(pairs, r.dim()) if pairs else None
        if f is not None:
            fine = tuple(int(v) * int(q) for v, q in zip(r.shape, f[3 - r.dim():]))
            out = empty_f32(fine)
            N.call("dh_block_spread", r.data_ptr(), out.data_ptr(), *((1,) * (3 - r.dim()) + fine), *f,
                   N.stream_ptr())
            return out
        for a, nl, ns in pairs:
            _, _, i_a, w_a = self._get_band(nl, ns)
            r = _apply_banded(r, i_a, w_a, a)
        return r

    def __call__(self, x):
        return self.forward(x)
