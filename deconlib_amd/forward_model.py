"""Canonical forward model: convolve -> crop -> downsample (deconvolution/forward_model.py).

The reconstruction lives on the padded visible domain (visible grid extended
by the PSF margins); ``op`` maps it to the data grid.  The crop to the visible
box happens on the fine grid at the exact (possibly asymmetric) pad-before
offset, before any downsampling, so even- and odd-sized PSFs register
identically (forward_model.py:96-113).  Every stage is a device operator of
this package: the convolution runs the pruned FFT passes, the crop a box copy,
the downsample the banded kernels.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional, Tuple, Union

from .composition import LinearOperator, compose
from .core_operators import Crop, FractionalAreaDownsample
from .linops import LinearFFTConvolver
from .shapes import compute_padded_shape, get_valid_slices


@dataclass(frozen=True)
class _Parts:
    """The stages of a ForwardModel's operator, for callers that fuse them."""

    conv: LinearFFTConvolver
    crop_start: Tuple[int, ...]
    downsample: Optional[FractionalAreaDownsample]


@dataclass(frozen=True)
class ForwardModel:
    """Operator plus shape bookkeeping (forward_model.py:31-50).

    op: padded visible -> data; data_shape, visible_shape, padded_shape and
    valid_slices (padded -> visible) as in the reference.
    """

    op: LinearOperator
    data_shape: Tuple[int, ...]
    visible_shape: Tuple[int, ...]
    padded_shape: Tuple[int, ...]
    valid_slices: Tuple[slice, ...]
    _parts: Any = field(default=None, repr=False, compare=False)


def make_forward_model(psf, data_shape: Tuple[int, ...],
                       zoom: Union[float, Tuple[float, ...]] = 1.0) -> ForwardModel:
    """convolve -> crop (fine grid) -> downsample (omitted when zoom maps to 1) (forward_model.py:53-121)."""
    ndim = len(data_shape)
    if isinstance(zoom, (int, float)):
        zoom = (float(zoom),) * ndim
    else:
        zoom = tuple(float(z) for z in zoom)
    if len(zoom) != ndim:
        raise ValueError(f"zoom has {len(zoom)} entries, expected {ndim}")
    if any(z < 1.0 for z in zoom):
        raise ValueError(f"zoom must be >= 1 (visible grid at least as fine as data); got {zoom}")
    psf_shape = tuple(int(s) for s in psf.shape)
    if len(psf_shape) != ndim:
        raise ValueError(f"psf is {len(psf_shape)}D but data_shape is {ndim}D")

    data_shape = tuple(int(d) for d in data_shape)
    visible_shape = tuple(max(1, round(d * z)) for d, z in zip(data_shape, zoom))
    padded_shape, padding = compute_padded_shape(visible_shape, psf_shape)
    valid_slices = get_valid_slices(padded_shape, visible_shape, padding)

    conv = LinearFFTConvolver(psf, signal_shape=padded_shape, normalize=True)
    crop_start = tuple(before for before, _ in padding)
    visible_crop = Crop(padded_shape, visible_shape, start=crop_start)
    if visible_shape == data_shape:
        op = compose(visible_crop, conv)
        down = None
    else:
        # the effective ratio, so that the visible shape lands on data_shape exactly
        eff_scale = tuple(v / d for v, d in zip(visible_shape, data_shape))
        down = FractionalAreaDownsample(scale=eff_scale, in_shape=visible_shape)
        op = compose(down, visible_crop, conv)
    return ForwardModel(op=op, data_shape=data_shape, visible_shape=visible_shape, padded_shape=padded_shape,
                        valid_slices=valid_slices, _parts=_Parts(conv, crop_start, down))
