"""Write the small regression fixtures in tests/golden/ from the oracle.

The reference cannot run here (its hot path imports MLX, which is absent), so
these vectors come from the oracle, whose functions are pinned separately by
brute force and closed forms (tests/test_oracle_pins.py).  They freeze the
oracle's float32 outputs on tiny inputs: the CPU suite checks the oracle still
reproduces them, the GPU suite checks the HIP path against them.

Run:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import deconv as O  # noqa: E402


def main():
    rng = np.random.default_rng(2024)
    out = {}
    # linear convolution, odd and even kernels, 3-D
    k = (rng.random((5, 6, 7)) + 0.05).astype(np.float32)
    x = rng.random((9, 10, 12)).astype(np.float32)
    A = O.LinearFFTConvolver(k, x.shape)
    out.update(conv_k=k, conv_x=x, conv_fwd=A.forward(x), conv_adj=A.adjoint(x))
    # Hessian
    f = rng.standard_normal((6, 7, 8)).astype(np.float32)
    H = O.Hessian3D(0.7)
    out.update(hess_f=f, hess_r=np.float64(0.7), hess_fwd=H.forward(f), hess_adj=H.adjoint(H.forward(f)))
    # RL 5 iterations
    psf = (rng.random((4, 8, 8)) + 0.02).astype(np.float32)
    obs = rng.poisson(40 * rng.random((6, 12, 12)) + 2).astype(np.float32)
    r = O.richardson_lucy(obs, O.LinearFFTConvolver(psf, obs.shape), num_iter=5, background=0.5, eval_interval=2)
    out.update(rl_psf=psf, rl_obs=obs, rl_x=r.restored, rl_pred=r.pred, rl_loss=np.array(r.loss_history))
    # PDHG-Hessian 5 iterations
    pd_psf = np.pad(psf, ((1, 1), (2, 2), (2, 2)))
    p = O.solve_pdhg(obs, pd_psf, alpha=1e-2, num_iter=5, eval_interval=2, tol=0.0)
    out.update(pd_psf=pd_psf, pd_x=p.restored, pd_tau=np.array(p.tau_history),
               pd_sigma=np.array(p.sigma_history), pd_loss=np.array(p.loss_history))
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_small.npz"))


if __name__ == "__main__":
    main()
