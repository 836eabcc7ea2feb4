"""Seeded synthetic workloads: blob objects, Poisson observations, config recipes.

The objects follow the workload descriptions: isotropic Gaussian blobs with
centres uniform in the volume, sigma ~ U[1, 3] px, peak ~ U[100, 1000], plus a
constant background.  Observations are Poisson(blur(object)) drawn on the
host with numpy's PCG64 stream.
"""

from __future__ import annotations

import numpy as np

from .psf import Optics, widefield_psf

WIDEFIELD_OPTICS = Optics(wavelength=0.525, na=1.4, ni=1.515, ns=1.334)


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
    psf = widefield_psf(shape, lateral, axial, WIDEFIELD_OPTICS)
    obj = blob_volume(shape, n_blobs, rng)
    if blur is None:
        blur = host_linear_blur
    clean = np.maximum(blur(psf, obj), 0.0)
    prng = np.random.default_rng(np.random.SeedSequence(seed).spawn(1)[0])
    obs = prng.poisson(clean).astype(np.float32)
    return psf.astype(np.float32), obj.astype(np.float32), obs


def aberrated_psf(shape, rng: np.random.Generator, lateral=0.085, axial=0.1, device: bool = False):
    """The widefield optics (WIDEFIELD_OPTICS) with IndexMismatch depth ~ U[0, 20] um and Zernike spherical ~ U[-0.5, 0.5] rad
    (``device=True``: synthesised on the GPU, a float32 device tensor)."""
    from .psf import IndexMismatch, ZernikeAberration

    ab = [IndexMismatch(depth=float(rng.uniform(0.0, 20.0))),
          ZernikeAberration({ZernikeAberration.SPHERICAL: float(rng.uniform(-0.5, 0.5))})]
    return widefield_psf(shape, lateral, axial, WIDEFIELD_OPTICS, aberrations=ab, device=device)


def filaments(shape, n: int, rng: np.random.Generator, intensity: float = 300.0, vertices: int = 6) -> np.ndarray:
    """Thin filaments: n random 3D polylines one voxel wide (each segment sampled every
    0.25 voxel and rounded), value `intensity` on the line, 0 elsewhere."""
    vol = np.zeros(tuple(int(s) for s in shape), np.float32)
    hi = np.array(vol.shape, np.float64) - 1
    for _ in range(n):
        pts = rng.uniform(0, 1, size=(vertices, 3)) * hi
        idx = []
        for a, b in zip(pts[:-1], pts[1:]):
            k = int(np.ceil(np.abs(b - a).max() * 4)) + 1
            idx.append(np.rint(a + np.linspace(0, 1, k)[:, None] * (b - a)).astype(np.int64))
        idx = np.unique(np.concatenate(idx), axis=0)
        vol[idx[:, 0], idx[:, 1], idx[:, 2]] = intensity
    return vol


def serving_requests(n_jobs: int, rate_per_s: float, seed: int, z_range=(16, 48), xy_range=(64, 256),
                     rl_fraction=0.7, rl_iters=100, pdhg_iters=200, density=200 / (64 * 256 * 256),
                     blur=None):
    """Poisson arrivals of mixed RL / PDHG-Hessian jobs with log-uniform volume sizes.

    Sizes are rounded to multiples of 8; each job has its own aberrated PSF,
    a blob object at the RL workload's blob density (200 per 64x256x256) and a Poisson observation.
    ``blur(psf, obj)`` computes the noiseless image (default: host float64
    FFT).  Returns a list of ``serving.Request``.
    """
    from .serving import Request

    blur = blur or host_linear_blur
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for j in range(n_jobs):
        t += float(rng.exponential(1.0 / rate_per_s))
        z = int(round(np.exp(rng.uniform(np.log(z_range[0]), np.log(z_range[1]))) / 8)) * 8
        xy = int(round(np.exp(rng.uniform(np.log(xy_range[0]), np.log(xy_range[1]))) / 8)) * 8
        shape = (max(z, 8), max(xy, 8), max(xy, 8))
        psf = aberrated_psf(shape, rng)
        obj = blob_volume(shape, max(1, int(round(density * np.prod(shape)))), rng)
        clean = np.maximum(blur(psf, obj), 0.0)
        obs = rng.poisson(clean).astype(np.float32)
        if rng.uniform() < rl_fraction:
            out.append(Request(j, t, "rl", obs, psf, num_iter=rl_iters))
        else:
            alpha = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
            out.append(Request(j, t, "pdhg", obs, psf, num_iter=pdhg_iters, alpha=alpha))
    return out


def blob_volume_device(shape, n_blobs: int, rng: np.random.Generator, background: float = 10.0,
                       sigma=(1.0, 3.0), peak=(100.0, 1000.0)):
    """``blob_volume`` rendered into a float32 device tensor (same parameter law and
    draw order; separable Gaussians on +-4 sigma boxes, summed in float32)."""
    import torch

    shape = tuple(int(s) for s in shape)
    vol = torch.full(shape, float(background), dtype=torch.float32, device="cuda")
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
        box = (a * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]).astype(np.float32)
        vol[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] += torch.from_numpy(box).to("cuda", non_blocking=False)
    return vol


def _synth_side(n: int) -> int:
    """Largest length of the form 2^a * {1, 3, 5} (<= 4096) not above n: the sides the
    GPU PSF synthesis computes."""
    best = 1
    for base in (1, 3, 5):
        v = base
        while v * 2 <= min(n, 4096):
            v *= 2
        if v <= n and v > best:
            best = v
    return best


def serving_requests_device(n_jobs: int, rate_per_s: float, seed: int, z_range=(32, 128), xy_range=(128, 1024),
                            rl_fraction=0.7, rl_iters=100, pdhg_iters=200, alpha_range=(1e-4, 1e-2),
                            density=200 / (64 * 256 * 256), shapes=None, kinds=None):
    """The serving configuration's request stream with inputs resident on the GPU.

    Exponential inter-arrival times (rate ``rate_per_s``); z and y = x drawn
    independently log-uniform over ``z_range`` / ``xy_range`` (any integer size);
    70 % RL (``rl_iters``) / 30 % PDHG-Hessian (``pdhg_iters``, alpha log-uniform over
    ``alpha_range``); each job with its own aberrated PSF (synthesised on the GPU, on the
    largest 2^a * {1, 3, 5} grid not above the volume per axis: a linear convolution
    takes any kernel extent up to the volume's),
    a blob object at that blob density rendered on the GPU and a Poisson
    observation of its blurred image drawn on the GPU.  ``shapes`` / ``kinds``
    (lists of n_jobs) replace the drawn sizes / job kinds.
    """
    import torch

    from .linops import LinearFFTConvolver
    from .serving import Request

    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for j in range(n_jobs):
        t += float(rng.exponential(1.0 / rate_per_s))
        z = int(round(np.exp(rng.uniform(np.log(z_range[0]), np.log(z_range[1])))))
        xy = int(round(np.exp(rng.uniform(np.log(xy_range[0]), np.log(xy_range[1])))))
        shape = tuple(shapes[j]) if shapes is not None else (z, xy, xy)
        psf = aberrated_psf(tuple(_synth_side(n) for n in shape), rng, device=True)
        obj = blob_volume_device(shape, max(1, int(round(density * np.prod(shape)))), rng)
        clean = LinearFFTConvolver(psf, shape).forward(obj).clamp_min_(0.0)
        del obj
        gen = torch.Generator(device="cuda")
        gen.manual_seed(int(rng.integers(1 << 31)))
        obs = torch.poisson(clean, generator=gen)
        del clean
        is_rl = rng.uniform() < rl_fraction
        if kinds is not None:
            is_rl = kinds[j] == "rl"
        if is_rl:
            out.append(Request(j, t, "rl", obs, psf, num_iter=rl_iters))
        else:
            alpha = float(np.exp(rng.uniform(np.log(alpha_range[0]), np.log(alpha_range[1]))))
            out.append(Request(j, t, "pdhg", obs, psf, num_iter=pdhg_iters, alpha=alpha))
    return out


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
    ax = tuple(range(len(P)))
    out = np.fft.irfftn(np.fft.rfftn(obj, P, axes=ax) * np.fft.rfftn(k, axes=ax), P, axes=ax)
    return out[: s[0], : s[1], : s[2]]
