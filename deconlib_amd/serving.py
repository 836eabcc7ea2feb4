"""Deconvolution serving loop: HBM-budget admission and continuous batching.

Requests (RL or PDHG-Hessian jobs of different volume sizes) arrive over
time.  The server admits a request as soon as its device footprint fits the
HBM budget, builds its operator and state on the device (OTF, workspace,
iterate, duals, sensitivity: all resident until the job finishes), and then
advances every live job by one iteration per tick:

* all live RL jobs together through ``BatchRunner`` (two batched convolution
  calls per tick, volumes grouped by transform length);
* each live PDHG job through its own device-resident iteration (its scalars
  never leave the GPU).

A job's latency is its completion time minus its arrival time (queueing,
set-up and iterations included).  The result of a job does not depend on what
else is in flight: the batched launches compute each volume exactly as a solo
launch does.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _native as N
from .batch import BatchRunner, RLJob
from .linops import LinearFFTConvolver
from .pdhg import _PDState, _nonneg


@dataclass
class Request:
    job_id: int
    arrival: float                 # seconds after the server starts
    kind: str                      # "rl" or "pdhg"
    observed: np.ndarray           # host float32 (z, y, x)
    psf: np.ndarray                # host float32, same shape, corner origin
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


def estimate_bytes(req: Request) -> int:
    """Device bytes a request holds while live (plan buffers plus solver state)."""
    import ctypes as C

    shape = tuple(int(s) for s in req.observed.shape)
    nd = len(shape)
    h = C.c_void_p()
    arr = (C.c_int64 * nd)(*shape)
    N.call("dh_conv_plan_create", nd, arr, arr, None, 0, C.byref(h))
    info = N.PlanInfo()
    N.call("dh_conv_plan_info", h, C.byref(info))
    N.fns["dh_conv_plan_destroy"](h)
    vox = int(np.prod(shape))
    # RL: obs, x, sens, ratio, pred; PDHG: data, x, xp, y1, y2 (6), Ax, Axp, g
    state = (5 if req.kind == "rl" else 14) * 4 * vox
    return int(info.otf_bytes + info.work_bytes + info.twiddle_bytes) + state + vox * 4  # + psf


class _Live:
    def __init__(self, req: Request, admitted: float):
        self.req = req
        self.admitted = admitted
        op = LinearFFTConvolver(req.psf, req.observed.shape)
        obs = torch.from_numpy(np.ascontiguousarray(req.observed, dtype=np.float32)).cuda()
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


class DeconvServer:
    """Admits requests against an HBM budget and runs them with continuous batching."""

    def __init__(self, hbm_budget_bytes: Optional[int] = None, reserve_bytes: int = 8 << 30,
                 keep_results: bool = False):
        total = torch.cuda.get_device_properties(torch.cuda.current_device()).total_memory
        self.budget = int(hbm_budget_bytes if hbm_budget_bytes is not None else total - reserve_bytes)
        self.keep_results = keep_results
        self.runner = BatchRunner(256)

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter) -> List[Completed]:
        queue = sorted(requests, key=lambda r: r.arrival)
        need = {r.job_id: estimate_bytes(r) for r in queue}
        for r in queue:
            if need[r.job_id] > self.budget:
                raise ValueError(f"request {r.job_id} needs {need[r.job_id]} B, more than the HBM budget")
        live: List[_Live] = []
        used = 0
        done: List[Completed] = []
        t0 = clock()
        qi = 0
        while qi < len(queue) or live:
            now = clock() - t0
            # admission: in arrival order while memory allows
            while qi < len(queue) and queue[qi].arrival <= now and used + need[queue[qi].job_id] <= self.budget:
                req = queue[qi]
                live.append(_Live(req, clock() - t0))
                used += need[req.job_id]
                qi += 1
            if not live:
                if qi < len(queue):
                    time.sleep(max(0.0, min(queue[qi].arrival - (clock() - t0), 0.01)))
                continue
            # one iteration of every live job
            rl = [l.job for l in live if l.job is not None and not l.job.done]
            if rl:
                self.runner.step(rl)
            for l in live:
                if l.pd is not None and l.k < l.req.num_iter:
                    l.pd.step(l.k, False)
                    l.k += 1
            finished = [l for l in live if l.done]
            if finished:
                outs = [l.restored() for l in finished]
                torch.cuda.current_stream().synchronize()
                t = clock() - t0
                for l, x in zip(finished, outs):
                    done.append(Completed(l.req.job_id, l.req.kind, tuple(l.req.observed.shape), l.req.num_iter,
                                          l.req.arrival, l.admitted, t, x if self.keep_results else None))
                    used -= need[l.req.job_id]
                live = [l for l in live if not l.done]
            else:
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
