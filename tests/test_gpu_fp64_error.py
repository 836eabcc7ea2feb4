"""Error of the GPU solvers against a float64 oracle run, per workload configuration.

On the same inputs each test computes the GPU result (float32 arithmetic), the
float32 oracle (the reference's own precision: MLX runs these solvers in float32)
and the float64 oracle, prints both float32 errors relative to float64, and
requires the GPU's to stay within the parity bar (relative L2 <= 1e-4) and within
10x what float32 arithmetic costs the oracle itself (plus 1e-6):

* RL, 64x256x256 (the RL configuration at full size and iteration count), 50 iterations;
* binned PDHG-Hessian (alpha 1e-3) at half size: object 64x256x256 -> data
  32x128x128 through FractionalAreaDownsample(2) o LinearFFTConvolver, PSF at
  0.0425/0.05 um, peak ~500 counts, 20 iterations;
* batched RL: two volume shapes of the batched configuration (16x384x384 and
  48x192x192) with their own aberrated PSFs, run together in one batch, 50 iterations;
* the large-volume configuration's solver (PDHG-Hessian alpha 5e-4), 20 iterations, on
  a 64x256x256 crop and (with DECONLIB_AMD_LONG_TESTS=1: the float64 oracle takes
  several minutes of host time there, about 3.5 on 16 cores) on the configuration's 128x512x512 crop:
  blobs at the configuration's density plus thin filaments, the configuration's
  vectorial PSF (isotropic emitter, NA 1.4) with IndexMismatch depth 10 um.

The serving configuration runs the same RL and PDHG kernels as these.
"""

import os

import numpy as np
import pytest
import torch

from oracle import deconv as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _report(name, gpu, ref32, ref64):
    e_gpu, e_ref = rel_l2(gpu, ref64), rel_l2(ref32, ref64)
    print(f"fp64 error {name}: GPU {e_gpu:.3e}, float32 oracle {e_ref:.3e}", flush=True)
    assert e_gpu <= 1e-4
    assert e_gpu <= 10 * e_ref + 1e-6
    return e_gpu, e_ref


def _poisson(clean, seed, peak=None):
    clean = np.maximum(np.asarray(clean, np.float64), 0.0)
    if peak is not None:
        clean = clean * (peak / clean.max())
    return np.random.default_rng(seed).poisson(clean).astype(np.float32)


def test_rl_config_fp64_error(gpu):
    from deconlib_amd import LinearFFTConvolver, richardson_lucy_with_operator
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS, blob_volume

    shape, it = (64, 256, 256), 50
    print(f"RL {shape}, {it} iterations", flush=True)
    psf = widefield_psf(shape, 0.085, 0.1, WIDEFIELD_OPTICS).astype(np.float32)
    obj = blob_volume(shape, 200, np.random.default_rng(1000))
    obs = _poisson(O.LinearFFTConvolver(psf, shape, dtype=np.float64).forward(obj), 1000)
    gpu = richardson_lucy_with_operator(obs, LinearFFTConvolver(psf, shape), num_iter=it).restored.cpu().numpy()
    r32 = O.richardson_lucy(obs, O.LinearFFTConvolver(psf, shape), num_iter=it).restored
    r64 = O.richardson_lucy(obs, O.LinearFFTConvolver(psf, shape, dtype=np.float64), num_iter=it,
                            dtype=np.float64).restored
    _report(f"RL 64x256x256, {it} it", gpu, r32, r64)


def test_binned_pdhg_config_fp64_error(gpu):
    from deconlib_amd import solve_pdhg_mlx
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS, blob_volume

    hi, it = (64, 256, 256), 20
    print(f"binned PDHG {hi} -> half, {it} iterations", flush=True)
    psf = widefield_psf(hi, 0.0425, 0.05, WIDEFIELD_OPTICS).astype(np.float32)
    obj = blob_volume(hi, 250, np.random.default_rng(1001))
    A = O.compose(O.FractionalAreaDownsample(2.0), O.LinearFFTConvolver(psf, hi, dtype=np.float64))
    obs = _poisson(A.forward(obj), 1001, peak=500.0)
    kw = dict(alpha=1e-3, num_iter=it, bin_factors=2, tol=0.0)
    gpu = solve_pdhg_mlx(torch.from_numpy(obs).cuda(), psf, **kw)
    r32 = O.solve_pdhg(obs, psf, **kw)
    r64 = O.solve_pdhg(obs, psf, dtype=np.float64, **kw)
    assert gpu.iterations == r32.iterations == r64.iterations == it
    _report(f"binned PDHG-Hessian 64x256x256 -> 32x128x128, {it} it", gpu.restored.cpu().numpy(), r32.restored,
            r64.restored)


def test_batched_config_fp64_error(gpu):
    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd.batch import RLJob, batched_richardson_lucy
    from deconlib_amd.synth import aberrated_psf, blob_volume

    rng = np.random.default_rng(1002)
    it = 50
    jobs, refs = [], []
    for shape in ((16, 384, 384), (48, 192, 192)):
        print(f"batched RL volume {shape}, {it} iterations", flush=True)
        psf = aberrated_psf(shape, rng).astype(np.float32)
        obj = blob_volume(shape, max(1, round(200 * np.prod(shape) / 4194304)), rng)
        obs = _poisson(O.LinearFFTConvolver(psf, shape, dtype=np.float64).forward(obj), int(rng.integers(1 << 30)))
        jobs.append(RLJob(obs, LinearFFTConvolver(psf, shape), num_iter=it))
        refs.append((O.richardson_lucy(obs, O.LinearFFTConvolver(psf, shape), num_iter=it).restored,
                     O.richardson_lucy(obs, O.LinearFFTConvolver(psf, shape, dtype=np.float64), num_iter=it,
                                       dtype=np.float64).restored))
    out = batched_richardson_lucy(jobs)
    for job, res, (r32, r64) in zip(jobs, out, refs):
        _report(f"batched RL {tuple(res.restored.shape)}, {it} it", res.restored.cpu().numpy(), r32, r64)


LONG = os.environ.get("DECONLIB_AMD_LONG_TESTS") == "1"


@pytest.mark.timeout(900)
@pytest.mark.parametrize("shape,it", [
    ((64, 256, 256), 20),
    pytest.param((128, 512, 512), 20, marks=pytest.mark.skipif(
        not LONG, reason="the float64 oracle needs several minutes of host time at this size; "
                         "run with DECONLIB_AMD_LONG_TESTS=1")),
])
def test_large_volume_crop_fp64_error(gpu, shape, it):
    from deconlib_amd import solve_pdhg_mlx
    from deconlib_amd.psf import IndexMismatch, widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS, blob_volume, filaments

    print(f"large-volume solver on a {shape} crop, {it} iterations", flush=True)
    rng = np.random.default_rng(1004)
    psf = widefield_psf(shape, 0.085, 0.1, WIDEFIELD_OPTICS, aberrations=[IndexMismatch(depth=10.0)],
                        dipole="isotropic")
    # 1e5 blobs over 256x2048x2048 voxels: the same density on the crop
    obj = blob_volume(shape, round(1e5 * np.prod(shape) / (256 * 2048 * 2048)), rng)
    obj = np.maximum(obj, filaments(shape, max(3, round(24 * np.prod(shape) / (128 * 512 * 512))), rng))
    obs = _poisson(O.LinearFFTConvolver(psf, shape, dtype=np.float64).forward(obj), 1004)
    kw = dict(alpha=5e-4, num_iter=it, tol=0.0)
    gpu = solve_pdhg_mlx(torch.from_numpy(obs).cuda(), psf, **kw)
    def progress(name):
        return lambda k, x: print(f"  {name} oracle iteration {k + 1}/{it}", flush=True)

    r32 = O.solve_pdhg(obs, psf, callback=progress("float32"), **kw)
    r64 = O.solve_pdhg(obs, psf, dtype=np.float64, callback=progress("float64"), **kw)
    assert gpu.iterations == r32.iterations == r64.iterations == it
    _report(f"PDHG-Hessian {'x'.join(map(str, shape))} crop, {it} it", gpu.restored.cpu().numpy(), r32.restored,
            r64.restored)
