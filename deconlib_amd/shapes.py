"""Shape and padding helpers of the three spaces (deconvolution/shapes.py).

Host arithmetic only: the visible shape from the data shape and a bin factor,
the wrap-free padded shape of a linear convolution with its (before, after)
padding, and the slices that crop a padded array back to the signal.  Same
names, arguments, rounding and ``ValueError`` messages as the reference.
"""

from __future__ import annotations

from typing import Optional, Tuple, Union


def compute_visible_shape(data_shape: Tuple[int, ...],
                          bin_factor: Union[float, Tuple[float, ...]] = 1.0) -> Tuple[int, ...]:
    """data_dim / bin_factor per axis, rounded to nearest (shapes.py:28-89)."""
    data_shape = tuple(int(s) for s in data_shape)
    ndim = len(data_shape)
    if isinstance(bin_factor, (int, float)):
        factors = (float(bin_factor),) * ndim
    else:
        factors = tuple(float(f) for f in bin_factor)
        if len(factors) != ndim:
            raise ValueError(
                f"bin_factor has {len(factors)} elements, "
                f"expected {ndim} to match data_shape"
            )
    return tuple(int(round(d / f)) for d, f in zip(data_shape, factors))


def compute_padded_shape(signal_shape: Tuple[int, ...], kernel_shape: Tuple[int, ...], *,
                         min_pad: Optional[Union[int, Tuple[Optional[int], ...]]] = None):
    """Signal padded by kernel-1 per axis (or min_pad), floor half before (shapes.py:92-160).

    Returns (padded_shape, ((before, after), ...)).
    """
    signal_shape = tuple(int(s) for s in signal_shape)
    kernel_shape = tuple(int(s) for s in kernel_shape)
    ndim = len(signal_shape)
    if len(kernel_shape) != ndim:
        raise ValueError(
            f"kernel_shape has {len(kernel_shape)} dimensions, "
            f"expected {ndim} to match signal_shape"
        )
    if min_pad is None:
        pads = (None,) * ndim
    elif isinstance(min_pad, int):
        pads = (min_pad,) * ndim
    else:
        pads = tuple(min_pad)
    padded, padding = [], []
    for n, m, p in zip(signal_shape, kernel_shape, pads):
        total = m - 1 if p is None else int(p)
        before = total // 2
        padding.append((before, total - before))
        padded.append(n + total)
    return tuple(padded), tuple(padding)


def get_valid_slices(padded_shape: Tuple[int, ...], signal_shape: Tuple[int, ...],
                     padding: Optional[Tuple[Tuple[int, int], ...]] = None) -> Tuple[slice, ...]:
    """Slices of the signal inside a padded array; symmetric padding when None (shapes.py:163-214)."""
    padded_shape = tuple(int(s) for s in padded_shape)
    signal_shape = tuple(int(s) for s in signal_shape)
    ndim = len(padded_shape)
    if len(signal_shape) != ndim:
        raise ValueError(
            f"padded_shape has {ndim} dimensions, "
            f"but signal_shape has {len(signal_shape)} dimensions. "
            f"Both must have the same ndim."
        )
    if padding is None:
        padding = tuple(((p - s) // 2, (p - s) - (p - s) // 2) for p, s in zip(padded_shape, signal_shape))
    return tuple(slice(before, p - after) for p, (before, after) in zip(padded_shape, padding))
