"""Batched RL: every job's result is identical alone and inside a mixed batch."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _job_inputs(shape, seed):
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import CONFIG0_OPTICS

    rng = np.random.default_rng(seed)
    psf = widefield_psf(shape, 0.085, 0.1, CONFIG0_OPTICS)
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
