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
from ._tensor import device
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
        ready = st._spec_gen is not None and plan.gen == st._spec_gen
        return N.RLItem(plan=plan._h, otf=plan.otf_buf.data_ptr(), work=plan.work.data_ptr(),
                        work_bytes=int(plan.info.work_bytes), x=st.x.data_ptr(), obs=st.obs.data_ptr(),
                        sens=st.sens.data_ptr(), background=st.bg, eps=_EPS, reduce_out=red,
                        spectrum_ready=int(ready), keep_spectrum=int(keep))

    def _after_fused(self) -> None:
        st = self.state
        st.plan.gen += 1
        st._spec_gen = st.plan.gen if self.k + 1 < self.num_iter else None

    def _steady(self) -> bool:
        """True when this job's next fused iteration is the steady-state call: its X
        spectrum is resident and kept for the following iteration, and no loss is
        due (the call's arguments are then the same every iteration)."""
        st = self.state
        k = self.k
        return (st.fused_iter and k >= 1 and k + 1 < self.num_iter and not st.wants_loss(k - 1)
                and st._spec_gen is not None and st.plan.gen == st._spec_gen)

    def finish(self) -> RLResult:
        st = self.state
        k_last = self.k - 1
        pred = st.finish(k_last)
        self.result = RLResult(restored=st.x, pred=pred, iterations=self.k, loss_history=st.losses(k_last),
                               background=st.bg, full_shape=tuple(st.x.shape), valid_slices=None)
        return self.result


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
                self._stream = torch.cuda.Stream()
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
