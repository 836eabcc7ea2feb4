"""Operator-norm estimate by power iteration, with float64 sums on the GPU.

``power_iteration_norm_sq(op, shape)`` estimates ||A||^2 (the largest
eigenvalue of A^T A) of any operator with ``forward``/``adjoint`` -- a
``LinearFFTConvolver``, a ``Compose`` chain, a regularizer -- the way the
reference validates its closed-form ``operator_norm_sq`` values
(tests/test_operators_mlx_norms.py:21-53): v <- A^T A v / |A^T A v| from a
normalised Gaussian start, the Rayleigh quotient <v, A^T A v> as the estimate,
stop when it changes by less than ``tol`` relative.  The operator applications
run in float32; the Rayleigh quotient and the norm are float64 dot products
on the device (``dh_er_dot``), the only float64 work besides the PDHG step
sizes.  It is a diagnostic: the solvers use the operators' own
``operator_norm_sq``.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32

__all__ = ["power_iteration_norm_sq"]


def power_iteration_norm_sq(op, shape, num_iter: int = 100, tol: float = 1e-6, seed: int = 0,
                            return_history: bool = False):
    """Estimate ||A||^2 of ``op`` on inputs of ``shape`` (see the module docstring).

    Returns the estimate, or ``(estimate, history)`` with the Rayleigh quotient of
    every iteration when ``return_history`` is true.
    """
    shape = tuple(int(s) for s in shape)
    rng = np.random.default_rng(seed)
    v = as_f32(rng.standard_normal(shape).astype(np.float32))
    n = v.numel()
    dev = device()
    sc = torch.zeros(N.ER_NSLOTS, dtype=torch.float64, device=dev)
    part = torch.empty(int(N.fns["dh_er_partials"]()), dtype=torch.float64, device=dev)
    st = N.stream_ptr()

    def dot(a, b, slot):
        N.call("dh_er_dot", a.data_ptr(), b.data_ptr(), a.numel(), slot, sc.data_ptr(), part.data_ptr(), st)

    def scale(a, alpha):
        out = empty_f32(a.shape)
        N.call("dh_er_ew", N.ER_EW_SCALE, a.data_ptr(), None, float(alpha), 0.0, out.data_ptr(), None, a.numel(), st)
        return out

    dot(v, v, N.ER_DOT)
    v = scale(v, 1.0 / math.sqrt(float(sc[N.ER_DOT].item())))
    history = []
    prev = 0.0
    est = 0.0
    for _ in range(int(num_iter)):
        w = as_f32(op.adjoint(as_f32(op.forward(v))))
        if w.numel() != n:
            raise ValueError(f"adjoint(forward(x)) has {w.numel()} elements, x has {n}")
        dot(v, w, N.ER_DOT)
        dot(w, w, N.ER_GG)
        vals = sc[[N.ER_DOT, N.ER_GG]].cpu().numpy()
        est = float(vals[0])
        history.append(est)
        nw = math.sqrt(float(vals[1]))
        if nw == 0.0:
            break
        v = scale(w, 1.0 / nw)
        if abs(est - prev) < tol * abs(est):
            break
        prev = est
    return (est, history) if return_history else est
