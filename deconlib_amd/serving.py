"""Deconvolution serving loop: HBM-budget admission and continuous batching.

Requests (RL or PDHG-Hessian jobs of different volume sizes) arrive over
time.  The server admits a request as soon as its device footprint fits the
HBM budget, builds its operator and state on the device (OTF, workspace,
iterate, duals, sensitivity: all resident until the job finishes), and then
advances every live job by one iteration per tick:

* all live RL jobs together through ``BatchRunner`` (the fused iteration of
  every job, volumes grouped per transform length);
* all live PDHG jobs together through ``pd_batch_step`` (their convolutions
  grouped per transform length, their scalars never leaving the GPU).

Nothing in the loop waits for the GPU except a bound on how far the host may
run ahead (``max_ahead`` ticks enqueued): admission builds operators with
asynchronous uploads and the OTF norm kept on the device, and a finished job
is noticed by polling the event recorded after its last kernel, so host-side
admission of new arrivals overlaps the device work of the jobs in flight.
A job's latency is its completion time on the device (that event) minus its
arrival time: queueing, set-up and iterations included.  The result of a job
does not depend on what else is in flight: the grouped launches compute each
volume exactly as a solo launch does.
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _native as N
from ._tensor import as_f32
from .batch import BatchRunner, RLJob
from .linops import LinearFFTConvolver
from .pdhg import _PDState, _nonneg, pd_batch_step


@dataclass
class Request:
    job_id: int
    arrival: float                 # seconds after the server starts
    kind: str                      # "rl" or "pdhg"
    observed: object               # float32 (z, y, x): host array or device tensor
    psf: object                    # float32, same shape, corner origin: host array or device tensor
    num_iter: int = 100
    alpha: float = 1e-3
    background: float = 0.0


@dataclass
class Completed:
    job_id: int
    kind: str
    shape: tuple
    iterations: int
    arrival: float
    admitted: float
    finished: float
    restored: Optional[torch.Tensor] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def plan_info(shape) -> N.PlanInfo:
    """Sizes of the plan a job of this shape (PSF of the same shape) gets; host only."""
    import ctypes as C

    shape = tuple(int(s) for s in shape)
    nd = len(shape)
    h = C.c_void_p()
    arr = (C.c_int64 * nd)(*shape)
    N.call("dh_conv_plan_create", nd, arr, arr, None, 0, C.byref(h))
    info = N.PlanInfo()
    N.call("dh_conv_plan_info", h, C.byref(info))
    N.fns["dh_conv_plan_destroy"](h)
    return info


def estimate_bytes(req: Request) -> int:
    """Device bytes a request holds while live (plan buffers plus solver state)."""
    shape = tuple(int(s) for s in req.observed.shape)
    info = plan_info(shape)
    vox = int(np.prod(shape))
    # RL: obs, x, sens, ratio, pred; PDHG: data, x, xp, x_new, y1, y2 and y2_new (6 each), Ax, Axp, g
    state = (5 if req.kind == "rl" else 21) * 4 * vox
    return int(info.otf_bytes + info.work_bytes + info.twiddle_bytes) + state + vox * 4  # + psf


class _Live:
    def __init__(self, req: Request, admitted: float):
        self.req = req
        self.admitted = admitted
        obs = as_f32(req.observed)
        op = LinearFFTConvolver(req.psf, tuple(obs.shape))
        if req.kind == "rl":
            self.job = RLJob(obs, op, num_iter=req.num_iter, background=req.background)
            self.pd = None
        else:
            self.job = None
            self.pd = _PDState(obs, op, _nonneg(obs), req.alpha, req.num_iter, req.background, tol=0.0)
        self.k = 0
        self.op = op

    @property
    def done(self) -> bool:
        return self.job.done if self.job is not None else self.k >= self.req.num_iter

    def restored(self) -> torch.Tensor:
        if self.job is not None:
            return self.job.finish().restored
        return self.pd.x


def pick_shortest(remaining: List[float], voxels: List[int], tick_voxels: int) -> List[int]:
    """Indices of the live jobs one tick advances under shortest-remaining-work-first:
    jobs in order of remaining work (ties in list order) until their voxels reach
    ``tick_voxels`` (at least one job), so a tick still fills the GPU while the jobs
    closest to completion finish first."""
    order = sorted(range(len(remaining)), key=lambda i: (remaining[i], i))
    out, acc = [], 0
    for i in order:
        if out and acc >= tick_voxels:
            break
        out.append(i)
        acc += voxels[i]
    return sorted(out)


# Relative cost of one iteration per voxel: a PDHG-Hessian iteration moves about 1.5x an RL
# iteration's bytes (two convolutions, the dual/primal sweep and its sums).
_PD_COST = 1.5


class DeconvServer:
    """Admits requests against an HBM budget and runs them with continuous batching.

    ``policy`` = "shortest" (default): each tick advances the live jobs with the least
    remaining work until ``tick_voxels`` voxels are in the tick (``pick_shortest``), which
    finishes short jobs early (lower median latency) at the same throughput once the
    tick fills the GPU; "all": every live job every tick (processor sharing).  Either way
    each job's iterations are the same computations, so its result is the same."""

    def __init__(self, hbm_budget_bytes: Optional[int] = None, reserve_bytes: int = 8 << 30,
                 keep_results: bool = False, max_ahead: int = 2, policy: str = "shortest",
                 tick_voxels: int = 1 << 25):
        total = torch.cuda.get_device_properties(torch.cuda.current_device()).total_memory
        self.budget = int(hbm_budget_bytes if hbm_budget_bytes is not None else total - reserve_bytes)
        self.keep_results = keep_results
        self.max_ahead = max(1, int(max_ahead))
        if policy not in ("shortest", "all"):
            raise ValueError(f"policy must be 'shortest' or 'all', got {policy!r}")
        self.policy = policy
        self.tick_voxels = int(tick_voxels)
        self.runner = BatchRunner(256)
        self.ticks = 0

    def _select(self, live: List["_Live"]) -> List["_Live"]:
        if self.policy == "all" or len(live) <= 1:
            return live
        vox = [int(l.op.signal_shape[0] * l.op.signal_shape[1] * l.op.signal_shape[2])
               if len(l.op.signal_shape) == 3 else int(np.prod(l.op.signal_shape)) for l in live]
        rem = [v * (l.req.num_iter - l.k) * (_PD_COST if l.pd is not None else 1.0) for v, l in zip(vox, live)]
        return [live[i] for i in pick_shortest(rem, vox, self.tick_voxels)]

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter) -> List[Completed]:
        queue = sorted(requests, key=lambda r: r.arrival)
        need = {r.job_id: estimate_bytes(r) for r in queue}
        for r in queue:
            if need[r.job_id] > self.budget:
                raise ValueError(f"request {r.job_id} needs {need[r.job_id]} B, more than the HBM budget")
        live: List[_Live] = []
        draining = []  # (live job, completion event, result)
        ticks = deque()
        used = 0
        done: List[Completed] = []
        # the start event is complete (its timestamp taken) before the host clock starts, so
        # device completion times and host times share one origin
        start = torch.cuda.Event(enable_timing=True)
        start.record()
        start.synchronize()
        t0 = clock()
        qi = 0

        def poll(block: bool = False) -> None:
            nonlocal used, draining
            keep = []
            for l, ev, out in draining:
                if (block and ev is draining[0][1]) or ev.query():
                    ev.synchronize()
                    t = start.elapsed_time(ev) / 1e3
                    done.append(Completed(l.req.job_id, l.req.kind, tuple(l.req.observed.shape), l.req.num_iter,
                                          l.req.arrival, l.admitted, t, out if self.keep_results else None))
                    used -= need[l.req.job_id]
                else:
                    keep.append((l, ev, out))
            draining = keep

        while qi < len(queue) or live or draining:
            now = clock() - t0
            # admission: in arrival order while memory allows
            while qi < len(queue) and queue[qi].arrival <= now and used + need[queue[qi].job_id] <= self.budget:
                req = queue[qi]
                live.append(_Live(req, clock() - t0))
                used += need[req.job_id]
                qi += 1
            if live:
                run = self._select(live)
                rl = [l.job for l in run if l.job is not None]
                if rl:
                    self.runner.step(rl)
                pd = [l for l in run if l.pd is not None]
                if pd:
                    pd_batch_step([l.pd for l in pd], [l.k for l in pd])
                for l in run:
                    l.k += 1
                for l in run:
                    if l.done:
                        out = l.restored()
                        ev = torch.cuda.Event(enable_timing=True)
                        ev.record()
                        draining.append((l, ev, out))
                live = [l for l in live if not l.done]
                self.ticks += 1
                tick = torch.cuda.Event()
                tick.record()
                ticks.append(tick)
                while len(ticks) > self.max_ahead:
                    ticks.popleft().synchronize()
                poll()
            elif draining:
                # nothing to enqueue: wait for the oldest job in flight (or an arrival)
                if qi < len(queue) and queue[qi].arrival <= clock() - t0 and used + need[queue[qi].job_id] > self.budget:
                    poll(block=True)
                else:
                    poll()
                    if draining:
                        time.sleep(2e-4)
            else:
                time.sleep(max(0.0, min(queue[qi].arrival - (clock() - t0), 0.01)))
        torch.cuda.current_stream().synchronize()
        return done


def run_plain(requests: List[Request]) -> float:
    """The plain single-volume loop: one job at a time, synchronously; returns wall seconds."""
    t0 = time.perf_counter()
    for r in requests:
        l = _Live(r, 0.0)
        if l.job is not None:
            from .rl import richardson_lucy_with_operator  # noqa: F401  (same kernels, one job)
            while not l.job.done:
                l.job.state.step(l.job.k)
                l.job.k += 1
                torch.cuda.current_stream().synchronize()
            l.job.finish()
        else:
            for k in range(r.num_iter):
                l.pd.step(k, False)
                torch.cuda.current_stream().synchronize()
        torch.cuda.current_stream().synchronize()
    return time.perf_counter() - t0


def latency_stats(done: List[Completed], makespan: float) -> dict:
    lat = np.array(sorted(c.latency for c in done))
    vox_it = sum(int(np.prod(c.shape)) * c.iterations for c in done)
    return {
        "jobs": len(done),
        "p50_s": float(np.percentile(lat, 50)) if len(lat) else math.nan,
        "p99_s": float(np.percentile(lat, 99)) if len(lat) else math.nan,
        "makespan_s": makespan,
        "voxel_iters_per_s": vox_it / makespan if makespan > 0 else math.nan,
    }
