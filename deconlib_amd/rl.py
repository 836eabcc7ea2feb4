"""Richardson–Lucy on MI355X (drop-in for rl_mlx.richardson_lucy_with_operator).

Semantics follow rl_mlx.py:83-185 step for step:

* ``x0 = max(A^T max(obs - bg, 1e-6), 1e-6)`` (or ``max(init, 1e-6)``);
* ``s = A^T 1``, ``floor = max(max(s) * 1e-6, 1e-10)``, off-support voxels get
  ``x = 0`` and ``s = +inf`` (so the update's ``x * (A^T r / s)`` is ``x * 0``
  there, exactly the reference's ``where(active, ., 0)``);
* per iteration ``r = obs / max(A x + bg, 1e-10)``, ``x = max(x * A^T r / s, 1e-10)``;
* loss (mean I-divergence of ``A x + bg``) at ``k % eval_interval == 0`` and at
  the last iteration.

For a ``LinearFFTConvolver`` or ``FFTConvolver`` the two elementwise steps
ride in the epilogues of the convolution passes, and the loss of iterate
``k+1`` is summed inside the forward pass that iteration ``k+1`` has to run
anyway, so one iteration is exactly two pruned convolutions and no host sync.
Any other operator (e.g. ``compose(Crop, FractionalAreaDownsample,
LinearFFTConvolver)``) goes through its own ``forward``/``adjoint`` plus the
library's elementwise kernels.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable, Optional, Tuple

import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32
from .linops import FFTConvolver, LinearFFTConvolver

_EPS = 1e-10


@dataclass
class RLResult:
    """Result of Richardson–Lucy (rl_mlx.py:35-55); arrays are device tensors."""

    restored: torch.Tensor
    pred: torch.Tensor
    iterations: int
    loss_history: list
    background: float = 0.0
    full_shape: Optional[Tuple[int, ...]] = None
    valid_slices: Optional[Tuple[slice, ...]] = None


def poisson_i_divergence(observed, model, eps: float = 1e-6) -> float:
    """Mean Poisson I-divergence (rl_mlx.py:58-80), summed in float64 on the GPU."""
    obs = as_f32(observed)
    mdl = as_f32(model)
    if tuple(obs.shape) != tuple(mdl.shape):
        raise ValueError(f"observed {tuple(obs.shape)} and model {tuple(mdl.shape)} differ in shape")
    n = obs.numel()
    part = torch.empty(N.fns["dh_reduce_partials"](n), dtype=torch.float64, device=device())
    out = torch.empty(1, dtype=torch.float64, device=device())
    N.call("dh_poisson_idiv", obs.data_ptr(), mdl.data_ptr(), n, float(eps), part.data_ptr(), out.data_ptr(),
           N.stream_ptr())
    return float(out.item()) / n


def _plan_of(op):
    if isinstance(op, (LinearFFTConvolver, FFTConvolver)):
        return op._plan, tuple(op.in_shape)
    return None, None


class _RLState:
    """Device-resident state of one RL solve (used by the batch/serving layers too)."""

    def __init__(self, observed, blur_op, background=0.0, init=None, num_iter=50, eval_interval=10):
        self.op = blur_op
        self.obs = as_f32(observed)
        self.bg = float(background)
        self.num_iter = int(num_iter)
        self.eval_interval = int(eval_interval)
        self.plan, shape = _plan_of(blur_op)
        dev = device()
        n_slots = self.num_iter // max(self.eval_interval, 1) + 3
        self.loss_slots = torch.zeros(n_slots, dtype=torch.float64, device=dev)
        self.slot_of = {}
        self.n_obs = self.obs.numel()
        # one-call fused iteration (ratio never in HBM) on specialised canvases
        self.fused_iter = self.plan is not None and int(self.plan.info.spectrum_block) > 0
        # the workspace the fused iterations use and carry the X spectrum in, with
        # its write counter: the plan's own unless a caller brings a private one
        self.ws = self.plan
        self.box = None  # I-divergence box of the signal grid (None: all of it)
        self._spec_gen = None
        if self.plan is not None:
            self._init_fused(init, shape)
        else:
            self._init_generic(init)

    # -- set-up ---------------------------------------------------------------
    def _init_fused(self, init, shape):
        plan = self.plan
        st = N.stream_ptr()
        if init is None:
            f = N.FuseArgs(prologue=N.PRO_SUB_FLOOR, pro_a=self.bg, pro_b=1e-6,
                           epilogue=N.EPI_FLOOR, epi_a=1e-6)
            self.x = plan.apply(self.obs, True, empty_f32(shape), f)
        else:
            ini = as_f32(init)
            self.x = empty_f32(ini.shape)
            N.call("dh_clamp_min", ini.data_ptr(), self.x.data_ptr(), ini.numel(), 1e-6, st)
        self.sens = plan.apply(None, True, empty_f32(shape), N.FuseArgs(prologue=N.PRO_ONES))
        self._support()
        self.ratio = empty_f32(self.obs.shape)

    def _init_generic(self, init):
        st = N.stream_ptr()
        if init is None:
            dmb = empty_f32(self.obs.shape)
            N.call("dh_sub_floor", self.obs.data_ptr(), dmb.data_ptr(), dmb.numel(), self.bg, 1e-6, st)
            a = as_f32(self.op.adjoint(dmb))
            self.x = empty_f32(a.shape)
            N.call("dh_clamp_min", a.data_ptr(), self.x.data_ptr(), a.numel(), 1e-6, st)
        else:
            ini = as_f32(init)
            self.x = empty_f32(ini.shape)
            N.call("dh_clamp_min", ini.data_ptr(), self.x.data_ptr(), ini.numel(), 1e-6, st)
        ones = torch.ones(self.obs.shape, dtype=torch.float32, device=device())
        self.sens = as_f32(self.op.adjoint(ones)).clone()
        self._support()
        self.ratio = empty_f32(self.obs.shape)
        self.part = torch.empty(N.fns["dh_reduce_partials"](self.n_obs), dtype=torch.float64,
                                device=device())

    def _support(self):
        n = self.sens.numel()
        part = torch.empty(N.fns["dh_reduce_partials"](n), dtype=torch.float64, device=device())
        mx = torch.empty(1, dtype=torch.float64, device=device())
        st = N.stream_ptr()
        N.call("dh_reduce", N.RED_MAX, self.sens.data_ptr(), n, part.data_ptr(), mx.data_ptr(), st)
        N.call("dh_rl_support", self.sens.data_ptr(), self.x.data_ptr(), n, mx.data_ptr(), st)
        self._keep = (part, mx)

    # -- iteration ------------------------------------------------------------
    def wants_loss(self, k: int) -> bool:
        return k % self.eval_interval == 0 or k == self.num_iter - 1

    def slot(self, k: int):
        if k not in self.slot_of:
            self.slot_of[k] = len(self.slot_of)
        return self.loss_slots[self.slot_of[k]:].data_ptr()

    def step(self, k: int) -> None:
        """x_k -> x_{k+1}; sums the loss of x_k when iteration k-1 asked for one."""
        red = self.slot(k - 1) if (k > 0 and self.wants_loss(k - 1)) else None
        st = N.stream_ptr()
        if self.fused_iter:
            plan, ws = self.plan, self.ws
            keep = k + 1 < self.num_iter
            ready = self._spec_gen is not None and ws.gen == self._spec_gen
            N.call("dh_rl_iteration", plan._h, plan.otf_buf.data_ptr(), ws.work.data_ptr(),
                   int(plan.info.work_bytes), self.x.data_ptr(), self.obs.data_ptr(), self.sens.data_ptr(),
                   self.bg, _EPS, red, int(ready), int(keep), st)
            ws.gen += 1
            self._spec_gen = ws.gen if keep else None
        elif self.plan is not None:
            f = N.FuseArgs(epilogue=N.EPI_RL_RATIO, epi_a=self.bg, epi_b=_EPS,
                           aux0=self.obs.data_ptr(), reduce_out=red)
            self.plan.apply(self.x, False, self.ratio, f)
            u = N.FuseArgs(epilogue=N.EPI_RL_UPDATE, epi_b=_EPS, aux0=self.x.data_ptr(),
                           aux1=self.sens.data_ptr())
            self.plan.apply(self.ratio, True, self.x, u)
        else:
            model = as_f32(self.op.forward(self.x))
            N.call("dh_rl_ratio", model.data_ptr(), self.obs.data_ptr(), self.ratio.data_ptr(),
                   self.n_obs, self.bg, _EPS, self.part.data_ptr(), red, st)
            corr = as_f32(self.op.adjoint(self.ratio))
            N.call("dh_rl_update", self.x.data_ptr(), corr.data_ptr(), self.sens.data_ptr(),
                   self.x.numel(), _EPS, st)

    def finish(self, k_last: int) -> torch.Tensor:
        """pred = A x + bg, summing the last pending loss."""
        red = self.slot(k_last) if self.wants_loss(k_last) else None
        if self.plan is not None:
            f = N.FuseArgs(epilogue=N.EPI_ADD, epi_a=self.bg, aux0=self.obs.data_ptr(), reduce_out=red)
            return self.plan.apply(self.x, False, empty_f32(self.obs.shape), f)
        model = as_f32(self.op.forward(self.x))
        pred = empty_f32(model.shape)
        N.call("dh_add_scalar", model.data_ptr(), self.obs.data_ptr(), pred.data_ptr(), model.numel(),
               self.bg, self.part.data_ptr(), red, N.stream_ptr())
        return pred

    def losses(self, k_last: int):
        vals = self.loss_slots.cpu().tolist()
        out = []
        for k in range(k_last + 1):
            if self.wants_loss(k):
                out.append(vals[self.slot_of[k]] / self.n_obs)
        return out


def richardson_lucy_with_operator(
    observed,
    blur_op,
    num_iter: int = 50,
    background: float = 0.0,
    init=None,
    callback: Optional[Callable[[int, torch.Tensor], bool]] = None,
    eval_interval: int = 10,
    verbose: bool = False,
) -> RLResult:
    """Richardson–Lucy with a pre-built positive operator (rl_mlx.py:83-185)."""
    st = _RLState(observed, blur_op, background, init, num_iter, eval_interval)
    k = None
    for k in range(num_iter):
        st.step(k)
        if verbose and st.wants_loss(k):
            print(f"  Iter {k:4d}: (loss reported at the end)")
        if callback is not None:
            torch.cuda.current_stream().synchronize()
            # a copy: the next iteration updates st.x in place (the reference hands over an immutable array)
            if callback(k, st.x.clone()):
                break
    if k is None:
        # the reference reads the loop variable after a zero-length loop
        raise NameError("name 'k' is not defined")
    pred = st.finish(k)
    loss_history = st.losses(k)
    return RLResult(
        restored=st.x,
        pred=pred,
        iterations=k + 1,
        loss_history=loss_history,
        background=float(background),
        full_shape=tuple(st.x.shape),
        valid_slices=None,
    )


def richardson_lucy_solver(num_iter: int = 50, background: float = 0.0, init_value: Optional[float] = None,
                           **rl_kwargs: Any) -> Callable:
    """RL adapted to the ``solve(data, model) -> visible array`` contract (rl_mlx.py:188-228).

    ``solve(data, model)`` runs Richardson–Lucy on ``model.op`` (initial
    estimate ``init_value`` everywhere on the padded domain when given) and
    returns ``restored[model.valid_slices]`` as a device tensor; a zoom-1 model
    runs the fused iteration on the padded domain (the crop folded into the
    data), any other ``richardson_lucy_with_operator`` on the composed operator.
    ``solve.solve_composed`` always takes the latter.  ``solve.solve_batch(tiles, model)`` solves several equally shaped
    data tiles against the one model and returns their results in order, each
    identical to ``solve(tile, model)``; ``process_tiles`` uses it.
    """

    def solve(data, model):
        from .batch import fused_tiles_ok, tiles_richardson_lucy

        if fused_tiles_ok(model) and not rl_kwargs.get("callback") and not rl_kwargs.get("verbose"):
            # zoom-1 model: the fused iteration on the padded domain
            return tiles_richardson_lucy([data], model, num_iter=num_iter, background=background,
                                         init_value=init_value,
                                         eval_interval=rl_kwargs.get("eval_interval", 10))[0]
        return solve_composed(data, model)

    def solve_composed(data, model):
        init = None
        if init_value is not None:
            init = torch.full(model.padded_shape, float(init_value), dtype=torch.float32, device=device())
        result = richardson_lucy_with_operator(observed=data, blur_op=model.op, num_iter=num_iter,
                                               background=background, init=init, **rl_kwargs)
        return result.restored[model.valid_slices]

    def solve_batch(tiles, model):
        from .batch import tiles_richardson_lucy

        return tiles_richardson_lucy(tiles, model, num_iter=num_iter, background=background,
                                     init_value=init_value, solve_one=solve_composed, **rl_kwargs)

    solve.solve_batch = solve_batch
    solve.solve_composed = solve_composed
    return solve
