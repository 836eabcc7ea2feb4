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
