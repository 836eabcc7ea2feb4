"""Serving loop: admission by HBM budget, continuous batching, and every
result identical alone and in the mix."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _requests():
    from deconlib_amd.synth import serving_requests

    return serving_requests(6, 200.0, 77, z_range=(8, 16), xy_range=(24, 48), rl_iters=12, pdhg_iters=15,
                            rl_fraction=0.6)


def test_serving_results_equal_solo(gpu):
    from deconlib_amd import LinearFFTConvolver, richardson_lucy_with_operator, solve_pdhg_mlx
    from deconlib_amd.serving import DeconvServer, latency_stats

    reqs = _requests()
    assert {r.kind for r in reqs} == {"rl", "pdhg"}
    done = DeconvServer(keep_results=True).run(reqs)
    assert sorted(c.job_id for c in done) == [r.job_id for r in reqs]
    by_id = {c.job_id: c for c in done}
    for r in reqs:
        c = by_id[r.job_id]
        assert c.finished >= c.admitted >= 0 and c.latency >= 0
        if r.kind == "rl":
            solo = richardson_lucy_with_operator(r.observed, LinearFFTConvolver(r.psf, r.observed.shape),
                                                 num_iter=r.num_iter).restored
        else:
            solo = solve_pdhg_mlx(torch.from_numpy(r.observed).cuda(), r.psf, alpha=r.alpha, num_iter=r.num_iter,
                                  tol=0.0).restored
        assert torch.equal(c.restored, solo), r.job_id
    s = latency_stats(done, max(c.finished for c in done))
    assert s["jobs"] == 6 and s["p99_s"] >= s["p50_s"] > 0


def test_serving_budget_limits_concurrency(gpu):
    """A budget that fits one job at a time still completes every request, in arrival order."""
    from deconlib_amd.serving import DeconvServer, estimate_bytes

    reqs = _requests()
    budget = max(estimate_bytes(r) for r in reqs)
    done = DeconvServer(hbm_budget_bytes=budget).run(reqs)
    order = [c.job_id for c in sorted(done, key=lambda c: c.admitted)]
    assert order == [r.job_id for r in sorted(reqs, key=lambda r: r.arrival)]
    with pytest.raises(ValueError):
        DeconvServer(hbm_budget_bytes=budget // 100).run(reqs)


def test_serving_config_sizes_equal_solo(gpu):
    """Jobs at the serving configuration's sizes (32x128x128 .. 128x1024x1024, arbitrary
    integer extents, per-job aberrated PSFs, inputs made on the GPU), RL and PDHG mixed,
    admitted while others run: every result equals its solo run bit for bit."""
    from deconlib_amd import LinearFFTConvolver, richardson_lucy_with_operator, solve_pdhg_mlx
    from deconlib_amd.serving import DeconvServer
    from deconlib_amd.synth import serving_requests_device

    shapes = [(128, 1024, 1024), (37, 141, 141), (90, 700, 700), (32, 128, 128), (61, 333, 333)]
    kinds = ["rl", "pdhg", "rl", "pdhg", "rl"]
    print(f"serving at configuration sizes: {shapes}", flush=True)
    reqs = serving_requests_device(5, 50.0, 93, shapes=shapes, kinds=kinds, rl_iters=4, pdhg_iters=5)
    done = DeconvServer(keep_results=True).run(reqs)
    by_id = {c.job_id: c for c in done}
    assert sorted(by_id) == [r.job_id for r in reqs]
    for r in reqs:
        c = by_id[r.job_id]
        assert c.shape == tuple(r.observed.shape) and c.finished >= c.arrival >= 0
        if r.kind == "rl":
            solo = richardson_lucy_with_operator(r.observed, LinearFFTConvolver(r.psf, tuple(r.observed.shape)),
                                                 num_iter=r.num_iter).restored
        else:
            solo = solve_pdhg_mlx(r.observed, r.psf, alpha=r.alpha, num_iter=r.num_iter, tol=0.0).restored
        assert torch.equal(c.restored, solo), (r.job_id, r.kind, c.shape)
