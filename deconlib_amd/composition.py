"""Operator protocol and composition (composition.py:24-122), unchanged semantics."""

from __future__ import annotations

from typing import Any, Protocol, runtime_checkable


@runtime_checkable
class LinearOperator(Protocol):
    """forward / adjoint / operator_norm_sq, optional in_shape / out_shape (composition.py:24-49)."""

    operator_norm_sq: float

    def forward(self, x: Any) -> Any: ...

    def adjoint(self, y: Any) -> Any: ...

    def __call__(self, x: Any) -> Any: ...


class Compose:
    """``Compose(A, B)(x) = A(B(x))``; shape-checked at construction (composition.py:52-106)."""

    def __init__(self, outer: LinearOperator, inner: LinearOperator):
        outer_in = getattr(outer, "in_shape", None)
        inner_out = getattr(inner, "out_shape", None)
        if outer_in is not None and inner_out is not None and tuple(outer_in) != tuple(inner_out):
            hint = ""
            if "LinearFFTConvolver" in (type(outer).__name__, type(inner).__name__):
                hint = (
                    " Note: LinearFFTConvolver pads internally for wrap-free "
                    "convolution — its input and output are both signal_shape; "
                    "do not pad its input or output yourself."
                )
            raise ValueError(
                f"Cannot compose {type(outer).__name__} after "
                f"{type(inner).__name__}: {type(outer).__name__} expects "
                f"input of shape {tuple(outer_in)} but "
                f"{type(inner).__name__} produces {tuple(inner_out)}.{hint}"
            )
        self.outer = outer
        self.inner = inner
        self.in_shape = getattr(inner, "in_shape", None)
        self.out_shape = getattr(outer, "out_shape", None)
        self.operator_norm_sq = float(outer.operator_norm_sq * inner.operator_norm_sq)

    def forward(self, x: Any) -> Any:
        # FractionalAreaDownsample(2) after a convolution: the 2x2x2 block sum runs in the
        # convolution's last pass (same values as the two steps)
        fac = getattr(self.outer, "sum_factors", None)
        fused = getattr(self.inner, "forward_block_sum", None)
        shape = getattr(self.inner, "out_shape", None)
        if fac is not None and fused is not None and shape is not None and fac(shape) == (2, 2, 2):
            r = fused(x)
            if r is not None:
                return r
        return self.outer.forward(self.inner.forward(x))

    def adjoint(self, y: Any) -> Any:
        # FractionalAreaDownsample after a convolution: the integer-factor spread runs
        # inside the correlation's first pass (same values as the two steps)
        spread = getattr(self.outer, "spread_factors", None)
        fused = getattr(self.inner, "adjoint_spread", None)
        if spread is not None and fused is not None:
            f = spread(y)
            if f is not None:
                return fused(y, f)
        return self.inner.adjoint(self.outer.adjoint(y))

    def __call__(self, x: Any) -> Any:
        return self.forward(x)


def compose(*ops: LinearOperator) -> LinearOperator:
    """``compose(A, B, C)(x) = A(B(C(x)))`` (composition.py:109-122)."""
    if not ops:
        raise ValueError("compose() requires at least one operator")
    if len(ops) == 1:
        return ops[0]
    result: LinearOperator = ops[-1]
    for op in reversed(ops[:-1]):
        result = Compose(op, result)
    return result
