"""Golden fixtures (tests/golden/golden_small.npz, written by make_golden.py).

CPU: the oracle still reproduces them.  GPU: the HIP path matches them to the
same bars as the parity tests.
"""

import os

import numpy as np
import pytest

from oracle import deconv as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz"))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_oracle_reproduces_golden():
    A = O.LinearFFTConvolver(G["conv_k"], G["conv_x"].shape)
    assert rel_l2(A.forward(G["conv_x"]), G["conv_fwd"]) < 1e-6
    assert rel_l2(A.adjoint(G["conv_x"]), G["conv_adj"]) < 1e-6
    H = O.Hessian3D(float(G["hess_r"]))
    assert rel_l2(H.forward(G["hess_f"]), G["hess_fwd"]) < 1e-6
    r = O.richardson_lucy(G["rl_obs"], O.LinearFFTConvolver(G["rl_psf"], G["rl_obs"].shape), num_iter=5,
                          background=0.5, eval_interval=2)
    assert rel_l2(r.restored, G["rl_x"]) < 1e-6
    p = O.solve_pdhg(G["rl_obs"], G["pd_psf"], alpha=1e-2, num_iter=5, eval_interval=2, tol=0.0)
    assert rel_l2(p.restored, G["pd_x"]) < 1e-6


@pytest.mark.gpu
def test_gpu_matches_golden(gpu):
    import torch

    from deconlib_amd import (Hessian3D, LinearFFTConvolver, richardson_lucy_with_operator,
                              solve_pdhg_mlx)

    A = LinearFFTConvolver(G["conv_k"], G["conv_x"].shape)
    assert rel_l2(A.forward(G["conv_x"]).cpu().numpy(), G["conv_fwd"]) < 2e-6
    assert rel_l2(A.adjoint(G["conv_x"]).cpu().numpy(), G["conv_adj"]) < 2e-6
    H = Hessian3D(float(G["hess_r"]))
    assert rel_l2(H.forward(G["hess_f"]).cpu().numpy(), G["hess_fwd"]) < 1e-6
    assert rel_l2(H.adjoint(G["hess_fwd"]).cpu().numpy(), G["hess_adj"]) < 1e-6
    r = richardson_lucy_with_operator(G["rl_obs"], LinearFFTConvolver(G["rl_psf"], G["rl_obs"].shape),
                                      num_iter=5, background=0.5, eval_interval=2)
    assert rel_l2(r.restored.cpu().numpy(), G["rl_x"]) < 1e-5
    np.testing.assert_allclose(r.loss_history, G["rl_loss"], rtol=1e-5)
    p = solve_pdhg_mlx(torch.from_numpy(G["rl_obs"]).cuda(), G["pd_psf"], alpha=1e-2, num_iter=5,
                       eval_interval=2, tol=0.0)
    assert rel_l2(p.restored.cpu().numpy(), G["pd_x"]) < 1e-5
    np.testing.assert_allclose(p.tau_history, G["pd_tau"], rtol=1e-5)
    np.testing.assert_allclose(p.loss_history, G["pd_loss"], rtol=1e-5)
