"""Batched RL: every job's result is identical alone and inside a mixed batch."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _job_inputs(shape, seed):
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS

    rng = np.random.default_rng(seed)
    psf = widefield_psf(shape, 0.085, 0.1, WIDEFIELD_OPTICS)
    obs = rng.poisson(40 * rng.random(shape) + 3).astype(np.float32)
    return psf, obs


def test_batched_rl_equals_alone(gpu):
    from deconlib_amd import LinearFFTConvolver, richardson_lucy_with_operator
    from deconlib_amd.batch import RLJob, batched_richardson_lucy

    shapes = [(8, 32, 32), (16, 48, 48), (8, 32, 32), (4, 64, 40), (12, 20, 36), (16, 48, 48)]
    iters = [5, 7, 3, 6, 4, 7]
    inputs = [_job_inputs(s, 40 + i) for i, s in enumerate(shapes)]
    alone = [richardson_lucy_with_operator(obs, LinearFFTConvolver(psf, obs.shape), num_iter=n, eval_interval=2)
             for (psf, obs), n in zip(inputs, iters)]
    jobs = [RLJob(obs, LinearFFTConvolver(psf, obs.shape), num_iter=n, eval_interval=2)
            for (psf, obs), n in zip(inputs, iters)]
    mixed = batched_richardson_lucy(jobs)
    for a, m in zip(alone, mixed):
        assert m.iterations == a.iterations
        assert torch.equal(a.restored, m.restored)
        assert torch.equal(a.pred, m.pred)
        assert a.loss_history == m.loss_history


def test_batch_of_more_than_one_pack(gpu):
    """More volumes of one length than fit in one launch's argument pack."""
    from deconlib_amd import LinearFFTConvolver, richardson_lucy_with_operator
    from deconlib_amd.batch import RLJob, batched_richardson_lucy

    inputs = [_job_inputs((4, 16, 16), 60 + i) for i in range(20)]
    alone = [richardson_lucy_with_operator(obs, LinearFFTConvolver(psf, obs.shape), num_iter=3) for psf, obs in inputs]
    mixed = batched_richardson_lucy([RLJob(obs, LinearFFTConvolver(psf, obs.shape), num_iter=3)
                                     for psf, obs in inputs])
    for a, m in zip(alone, mixed):
        assert torch.equal(a.restored, m.restored)


def test_graph_replays_equal_eager(gpu):
    """Steady-state iterations replayed from a captured HIP graph, interleaved with
    eager iterations (first, loss, last), host synchronisations and job-set changes
    (recaptures), give the same bits as eager stepping."""
    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd.batch import BatchRunner, RLJob

    shapes = [(8, 32, 32), (16, 48, 48), (4, 64, 40), (12, 20, 36)]
    iters = [14, 9, 14, 11]
    inputs = [_job_inputs(s, 70 + i) for i, s in enumerate(shapes)]
    results = []
    for graphs in (False, True):
        print(f"batched RL, graphs={graphs}", flush=True)
        jobs = [RLJob(obs, LinearFFTConvolver(psf, obs.shape), num_iter=n, eval_interval=5)
                for (psf, obs), n in zip(inputs, iters)]
        runner = BatchRunner(len(jobs), graphs=graphs)
        while any(not j.done for j in jobs):
            runner.step(jobs)
            torch.cuda.synchronize()
            for j in jobs:
                if j.done and j.result is None:
                    j.finish()
        if graphs:
            # steady iterations of the whole live set: k = 2..5 and 7 (first capture),
            # 9 (after the 9-iteration job ends: recapture), 12 (after the
            # 11-iteration job ends: recapture); k = 6 and 11 follow a loss evaluation
            assert runner.replays == 7 and runner.captures == 3
        results.append([(j.result.restored.clone(), j.result.pred, j.result.loss_history) for j in jobs])
    for (x0, p0, l0), (x1, p1, l1) in zip(*results):
        assert torch.equal(x0, x1)
        assert torch.equal(p0, p1)
        assert l0 == l1
