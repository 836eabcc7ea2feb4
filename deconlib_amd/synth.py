"""Seeded synthetic workloads: blob objects, Poisson observations, config recipes.

The objects follow the workload descriptions: isotropic Gaussian blobs with
centres uniform in the volume, sigma ~ U[1, 3] px, peak ~ U[100, 1000], plus a
constant background.  Observations are Poisson(blur(object)) drawn on the
host with numpy's PCG64 stream.
"""

from __future__ import annotations

import numpy as np

from .psf import Optics, widefield_psf

CONFIG0_OPTICS = Optics(wavelength=0.525, na=1.4, ni=1.515, ns=1.334)


def blob_volume(shape, n_blobs: int, rng: np.random.Generator, background: float = 10.0,
                sigma=(1.0, 3.0), peak=(100.0, 1000.0)) -> np.ndarray:
    """Sum of separable Gaussian blobs (each evaluated on a +-4 sigma box) plus a constant."""
    shape = tuple(int(s) for s in shape)
    vol = np.full(shape, background, dtype=np.float64)
    centres = rng.uniform(0, 1, size=(n_blobs, 3)) * np.array(shape)
    sig = rng.uniform(sigma[0], sigma[1], size=n_blobs)
    amp = rng.uniform(peak[0], peak[1], size=n_blobs)
    for c, s, a in zip(centres, sig, amp):
        r = int(np.ceil(4 * s))
        lo = [max(0, int(np.floor(c[d])) - r) for d in range(3)]
        hi = [min(shape[d], int(np.floor(c[d])) + r + 1) for d in range(3)]
        if any(h <= l for l, h in zip(lo, hi)):
            continue
        g = [np.exp(-0.5 * ((np.arange(lo[d], hi[d]) - c[d]) / s) ** 2) for d in range(3)]
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += a * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    return vol


def rl_workload(shape=(64, 256, 256), n_blobs=200, seed=1000, lateral=0.085, axial=0.1, blur=None):
    """PSF (shape-sized, corner origin), object and Poisson observation of the RL recipe.

    ``blur(psf, obj) -> blurred`` computes the noiseless observation; it defaults
    to a float64 host FFT linear convolution.
    """
    rng = np.random.default_rng(seed)
    psf = widefield_psf(shape, lateral, axial, CONFIG0_OPTICS)
    obj = blob_volume(shape, n_blobs, rng)
    if blur is None:
        blur = host_linear_blur
    clean = np.maximum(blur(psf, obj), 0.0)
    prng = np.random.default_rng(np.random.SeedSequence(seed).spawn(1)[0])
    obs = prng.poisson(clean).astype(np.float32)
    return psf.astype(np.float32), obj.astype(np.float32), obs


def host_linear_blur(psf: np.ndarray, obj: np.ndarray) -> np.ndarray:
    """Wrap-free convolution of obj with a corner-origin psf (float64, numpy FFT)."""
    s = obj.shape
    m = psf.shape
    P = tuple(n + k - 1 for n, k in zip(s, m))
    k = np.zeros(P)
    idx = []
    for mm, pp in zip(m, P):
        npos = (mm + 1) // 2
        d = np.arange(mm)
        d[npos:] = pp - (mm - npos) + np.arange(mm - npos)
        idx.append(d)
    k[np.ix_(*idx)] = psf / psf.sum()
    out = np.fft.irfftn(np.fft.rfftn(obj, P) * np.fft.rfftn(k), P)
    return out[: s[0], : s[1], : s[2]]
