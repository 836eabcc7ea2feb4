"""Batched Richardson–Lucy: many volumes of different shapes advanced in lockstep.

Each job keeps its own device-resident state (x, sensitivity, plan workspace,
OTF).  One iteration of the whole batch is one call of ``dh_batch_rl_iteration``
(the fused iteration of every job, stage by stage, volumes grouped by transform
length, so a stage over the whole batch is a handful of launches, not one per
volume); jobs whose canvas lacks specialised kernels use two calls of
``dh_batch_conv_apply`` instead.  Per-job results are bit-identical to running
the job alone (same kernels, same shapes, same reduction order).
"""

from __future__ import annotations

import ctypes as C
from typing import List, Optional

import torch

from . import _native as N
from ._tensor import as_f32, device, empty_f32
from .linops import FFTConvolver, LinearFFTConvolver
from .rl import RLResult, _RLState, _EPS


class RLJob:
    """One RL solve inside a batch (rl_mlx.richardson_lucy_with_operator semantics)."""

    def __init__(self, observed, blur_op, num_iter: int = 50, background: float = 0.0, init=None,
                 eval_interval: int = 10):
        if not isinstance(blur_op, (LinearFFTConvolver, FFTConvolver)):
            raise TypeError("batched RL needs LinearFFTConvolver / FFTConvolver operators")
        if num_iter < 1:
            raise ValueError("num_iter must be >= 1")
        self.state = _RLState(observed, blur_op, background, init, num_iter, eval_interval)
        self.num_iter = int(num_iter)
        self.k = 0          # iterations done
        self.result: Optional[RLResult] = None

    @property
    def voxels(self) -> int:
        return self.state.x.numel()

    @property
    def done(self) -> bool:
        return self.k >= self.num_iter

    def _item(self, forward: bool) -> N.BatchItem:
        st = self.state
        plan = st.plan
        if forward:
            red = st.slot(self.k - 1) if (self.k > 0 and st.wants_loss(self.k - 1)) else None
            f = N.FuseArgs(epilogue=N.EPI_RL_RATIO, epi_a=st.bg, epi_b=_EPS, aux0=st.obs.data_ptr(),
                           reduce_out=red)
            src, dst = st.x, st.ratio
        else:
            f = N.FuseArgs(epilogue=N.EPI_RL_UPDATE, epi_b=_EPS, aux0=st.x.data_ptr(),
                           aux1=st.sens.data_ptr())
            src, dst = st.ratio, st.x
        return N.BatchItem(plan=plan._h, otf=plan.otf_buf.data_ptr(), work=plan.work.data_ptr(),
                           work_bytes=int(plan.info.work_bytes), input=src.data_ptr(), out=dst.data_ptr(),
                           fuse=f)

    def _rl_item(self) -> N.RLItem:
        """This job's next fused iteration (dh_rl_iteration semantics)."""
        st = self.state
        plan = st.plan
        k = self.k
        red = st.slot(k - 1) if (k > 0 and st.wants_loss(k - 1)) else None
        keep = k + 1 < self.num_iter
        ready = st._spec_gen is not None and st.ws.gen == st._spec_gen
        item = N.RLItem(plan=plan._h, otf=plan.otf_buf.data_ptr(), work=st.ws.work.data_ptr(),
                        work_bytes=int(plan.info.work_bytes), x=st.x.data_ptr(), obs=st.obs.data_ptr(),
                        sens=st.sens.data_ptr(), background=st.bg, eps=_EPS, reduce_out=red,
                        spectrum_ready=int(ready), keep_spectrum=int(keep))
        if st.box is not None:
            item.data_box[:] = st.box
        return item

    def _after_fused(self) -> None:
        st = self.state
        st.ws.gen += 1
        st._spec_gen = st.ws.gen if self.k + 1 < self.num_iter else None

    def _steady(self) -> bool:
        """True when this job's next fused iteration is the steady-state call: its X
        spectrum is resident and kept for the following iteration, and no loss is
        due (the call's arguments are then the same every iteration)."""
        st = self.state
        k = self.k
        return (st.fused_iter and k >= 1 and k + 1 < self.num_iter and not st.wants_loss(k - 1)
                and st._spec_gen is not None and st.ws.gen == st._spec_gen)

    def finish(self) -> RLResult:
        st = self.state
        k_last = self.k - 1
        pred = st.finish(k_last)
        self.result = RLResult(restored=st.x, pred=pred, iterations=self.k, loss_history=st.losses(k_last),
                               background=st.bg, full_shape=tuple(st.x.shape), valid_slices=None)
        return self.result


_CAPTURE_STREAMS = {}


def _capture_stream():
    """One side stream per device for graph captures, created on first use (creating a
    stream costs milliseconds, more than a capture)."""
    dev = torch.cuda.current_device()
    if dev not in _CAPTURE_STREAMS:
        _CAPTURE_STREAMS[dev] = torch.cuda.Stream()
    return _CAPTURE_STREAMS[dev]


class BatchRunner:
    """Advances a changing set of RL jobs one iteration at a time (continuous batching).

    With ``graphs=True`` the steady-state iteration of an unchanged set of jobs
    (every job's spectrum resident, no loss due) is captured once as a HIP graph
    (``torch.cuda.CUDAGraph``) and replayed: one graph launch per iteration instead
    of every grouped kernel launch and its host-side argument packing.  Other
    iterations (first, loss-evaluation, last) and any change of the job set run
    eagerly; a replay executes the same kernels on the same buffers, so results are
    bit-identical to eager stepping.
    """

    def __init__(self, max_items: int = 256, graphs: bool = False):
        self.max_items = int(max_items)
        self.graphs = bool(graphs)
        self._graph = None
        self._graph_jobs = None
        self._stream = None
        self.replays = 0
        self.captures = 0

    def _try_graph(self, fused: List[RLJob]) -> bool:
        """Runs the iteration of ``fused`` as a graph replay when possible."""
        if not self.graphs or not fused or len(fused) > self.max_items or not all(j._steady() for j in fused):
            return False
        # the graph holds the jobs themselves: their buffers (the graph's kernel
        # arguments) stay alive while it exists, and identity cannot be mistaken
        same = (self._graph_jobs is not None and len(self._graph_jobs) == len(fused)
                and all(a is b for a, b in zip(self._graph_jobs, fused)))
        if not same:
            arr = (N.RLItem * len(fused))(*[j._rl_item() for j in fused])
            g = torch.cuda.CUDAGraph()
            # captured on a side stream ordered after the current one; unlike the
            # torch.cuda.graph context this neither synchronises the device nor
            # empties the caching allocator (the capture allocates nothing)
            cur = torch.cuda.current_stream()
            if self._stream is None:
                self._stream = _capture_stream()
            self._stream.wait_stream(cur)
            try:
                with torch.cuda.stream(self._stream):
                    g.capture_begin()
                    try:
                        N.call("dh_batch_rl_iteration", len(fused), arr, N.stream_ptr())
                    finally:
                        g.capture_end()
            except Exception:  # capture unsupported here: keep stepping eagerly
                cur.wait_stream(self._stream)
                self.graphs = False
                self._graph = self._graph_jobs = None
                return False
            cur.wait_stream(self._stream)
            self._graph, self._graph_jobs = g, list(fused)
            self.captures += 1
        self._graph.replay()
        self.replays += 1
        for j in fused:
            j._after_fused()
        return True

    def step(self, jobs: List[RLJob]) -> None:
        """One RL iteration for every job in ``jobs`` (all not yet done).

        Jobs on canvases with specialised kernels advance through the fused batched
        iteration (ratio kept on chip, each job's X spectrum carried between calls);
        the others through a batched forward and a batched adjoint application.
        """
        live = [j for j in jobs if not j.done]
        fused = [j for j in live if j.state.fused_iter]
        plain = [j for j in live if not j.state.fused_iter]
        if self._try_graph(fused):
            fused = []
        for s in range(0, len(fused), self.max_items):
            chunk = fused[s:s + self.max_items]
            arr = (N.RLItem * len(chunk))(*[j._rl_item() for j in chunk])
            N.call("dh_batch_rl_iteration", len(chunk), arr, N.stream_ptr())
            for j in chunk:
                j._after_fused()
        for j in plain:
            j.state.plan.gen += 1  # the batched launches overwrite each plan's workspace
        for s in range(0, len(plain), self.max_items):
            chunk = plain[s:s + self.max_items]
            for forward in (True, False):
                arr = (N.BatchItem * len(chunk))(*[j._item(forward) for j in chunk])
                N.call("dh_batch_conv_apply", len(chunk), arr, 0 if forward else 1, N.stream_ptr())
        for j in live:
            j.k += 1


def batched_richardson_lucy(jobs: List[RLJob], runner: Optional[BatchRunner] = None,
                            graphs: bool = False) -> List[RLResult]:
    """Runs every job to its own ``num_iter`` in lockstep; returns their RLResults.
    ``graphs``: replay the steady-state iteration as a HIP graph (see BatchRunner)."""
    runner = runner or BatchRunner(max(len(jobs), 1), graphs=graphs)
    while any(not j.done for j in jobs):
        runner.step(jobs)
        for j in jobs:
            if j.done and j.result is None:
                j.finish()
    return [j.result for j in jobs]


# ---- tiles of one forward model (tile_processing.process_tiles) ----

class _Workspace:
    """A private plan workspace with its own write counter (tiles share one plan)."""

    def __init__(self, plan):
        self.work = torch.empty(int(plan.info.work_bytes), dtype=torch.uint8, device=device())
        self.gen = 0


class _TileRLState(_RLState):
    """RL on ``crop . conv`` (a zoom-1 ForwardModel) run on the padded domain.

    The observed tile is embedded in the padded grid, zero outside the data
    box, so the ratio ``obs / max(conv x + bg, eps)`` is zero there: exactly the
    reference's ``crop^T`` of the data-space ratio, and the fused iteration of
    the convolution alone applies.  ``x0`` and ``sens = A^T 1`` come from the
    composed operator; the I-divergence sums only the data box.
    """

    def __init__(self, tile, model, background, init_value, num_iter, eval_interval):
        parts = model._parts
        conv = parts.conv
        plan = conv._plan
        dev = device()
        data = as_f32(tile)
        if tuple(data.shape) != tuple(model.data_shape):
            raise ValueError(f"tile shape {tuple(data.shape)} != model data_shape {model.data_shape}")
        self.op = model.op
        self.bg = float(background)
        self.num_iter = int(num_iter)
        self.eval_interval = int(eval_interval)
        self.plan = plan
        self.loss_slots = torch.zeros(self.num_iter // max(self.eval_interval, 1) + 3, dtype=torch.float64,
                                      device=dev)
        self.slot_of = {}
        self.n_obs = data.numel()
        self.fused_iter = True
        self.ws = _Workspace(plan)
        self._spec_gen = None
        start = parts.crop_start
        box = tuple(slice(s, s + n) for s, n in zip(start, data.shape))
        self.box = [v for s, n in zip(start, data.shape) for v in (s, s + n)]
        self.data = data
        self.obs = torch.zeros(model.padded_shape, dtype=torch.float32, device=dev)
        self.obs[box] = data
        st = N.stream_ptr()
        if init_value is None:
            dmb = empty_f32(data.shape)
            N.call("dh_sub_floor", data.data_ptr(), dmb.data_ptr(), dmb.numel(), self.bg, 1e-6, st)
            a = as_f32(model.op.adjoint(dmb))
        else:
            a = torch.full(model.padded_shape, float(init_value), dtype=torch.float32, device=dev)
        self.x = empty_f32(a.shape)
        N.call("dh_clamp_min", a.data_ptr(), self.x.data_ptr(), a.numel(), 1e-6, st)
        ones = torch.ones(data.shape, dtype=torch.float32, device=dev)
        self.sens = as_f32(model.op.adjoint(ones)).clone()
        self._support()
        self.ratio = None

    def finish(self, k_last: int) -> torch.Tensor:
        """pred = crop(conv x) + bg on the data grid, summing the last pending loss."""
        red = self.slot(k_last) if self.wants_loss(k_last) else None
        model_out = as_f32(self.op.forward(self.x))
        n = model_out.numel()
        pred = empty_f32(model_out.shape)
        part = torch.empty(N.fns["dh_reduce_partials"](n), dtype=torch.float64, device=device())
        N.call("dh_add_scalar", model_out.data_ptr(), self.data.data_ptr(), pred.data_ptr(), n, self.bg,
               part.data_ptr(), red, N.stream_ptr())
        return pred


class TileRLJob(RLJob):
    """One tile's RL solve against a shared zoom-1 ForwardModel (padded-domain form)."""

    def __init__(self, tile, model, num_iter: int = 50, background: float = 0.0,
                 init_value: Optional[float] = None, eval_interval: int = 10):
        if num_iter < 1:
            raise ValueError("num_iter must be >= 1")
        self.state = _TileRLState(tile, model, background, init_value, num_iter, eval_interval)
        self.num_iter = int(num_iter)
        self.k = 0
        self.result: Optional[RLResult] = None


def fused_tiles_ok(model) -> bool:
    """A ForwardModel whose tiles can run the batched fused RL iteration: no
    downsample stage (zoom 1) and a 3-D canvas with specialised kernels."""
    parts = getattr(model, "_parts", None)
    if parts is None or parts.downsample is not None or not isinstance(parts.conv, LinearFFTConvolver):
        return False
    return len(model.padded_shape) == 3 and int(parts.conv._plan.info.spectrum_block) > 0


def _tile_job_bytes(model) -> int:
    plan = model._parts.conv._plan
    pad = 1
    for n in model.padded_shape:
        pad *= int(n)
    data = 1
    for n in model.data_shape:
        data *= int(n)
    # workspace, obs (padded), x, sens, plus the data tile and set-up temporaries
    return int(plan.info.work_bytes) + 4 * 4 * pad + 3 * 4 * data


# Padded voxels per lockstep group: one volume of about this size already keeps
# every pass at its bandwidth, so larger tiles run one at a time (grouping them
# gains no launches back and spreads the working set over more of the cache).
_GROUP_VOXELS = 16 << 20


def tiles_richardson_lucy(tiles, model, num_iter: int = 50, background: float = 0.0,
                          init_value: Optional[float] = None, solve_one=None, eval_interval: int = 10,
                          callback=None, verbose: bool = False, max_tiles: Optional[int] = None):
    """RL on every tile against the one model; returns restored[valid_slices] per tile.

    Zoom-1 models on canvases with specialised kernels run as TileRLJobs through
    ``dh_batch_rl_iteration``: small tiles in lockstep groups of about
    16M padded voxels (bounded by free HBM and ``max_tiles``), large ones one
    at a time; anything else (a downsample stage, callbacks, verbose output)
    runs ``solve_one`` per tile.  Results agree with the composed operator's
    solve to float32 rounding (the fused iteration contracts multiply-adds
    differently from the separate kernels).
    """
    if not fused_tiles_ok(model) or callback is not None or verbose:
        return [solve_one(t, model) for t in tiles]
    free, _ = torch.cuda.mem_get_info()
    pad = 1
    for n in model.padded_shape:
        pad *= int(n)
    group = max(1, min(int(0.8 * free) // max(_tile_job_bytes(model), 1), -(-_GROUP_VOXELS // pad)))
    if max_tiles is not None:
        group = min(group, max(1, int(max_tiles)))
    out = []
    for s in range(0, len(tiles), group):
        jobs = [TileRLJob(t, model, num_iter, background, init_value, eval_interval) for t in tiles[s:s + group]]
        batched_richardson_lucy(jobs)
        # compact copies: the group's padded states are freed before the next group runs
        out.extend(j.state.x[model.valid_slices].clone() for j in jobs)
        del jobs
    return out
