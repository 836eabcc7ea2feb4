"""CPU oracle for PSF synthesis (TEST INFRASTRUCTURE, not the product).

Only ``tests/`` may import this module.  It restates, in complex128 NumPy:

* ``pupil_to_psf`` (deconlib/psf/widefield.py:18-73): per z-plane
  ``|ifft2(pupil * exp(2 pi i kz z))|^2``, DC at the corner, normalised to sum 1;
* ``pupil_to_vectorial_psf`` (deconlib/psf/widefield.py:110-234) with its
  Fresnel coefficients, dipole-to-field factors and aplanatic apodization
  (deconlib/psf/pupil.py:48-75,156-262).

Parity pins (tests/test_psf.py): Parseval's identity per plane
(sum |ifft2(Q)|^2 = sum |Q|^2 / (ny nx) for a unit-modulus defocus, for the scalar
PSF and for the weighted vectorial channels), a brute-force 2-D inverse DFT on
tiny pupils, the flat plane of a single-sample pupil, the z-mirror symmetry of a
real pupil, and the dipole relations (isotropic = mean of the axis dipoles; a
fixed dipole along an axis equals that axis's PSF).
"""

from __future__ import annotations

import numpy as np


def pupil_to_psf(pupil, kz, z, normalize=True):
    """(nz, ny, nx) float64 intensity PSF from a (ny, nx) complex pupil (widefield.py:57-72)."""
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    defocus = np.exp(2j * np.pi * np.asarray(kz, np.float64) * z[:, None, None])
    amp = np.fft.ifft2(np.asarray(pupil, np.complex128) * defocus, axes=(-2, -1))
    psf = np.abs(amp) ** 2
    if normalize:
        total = psf.sum()
        if total > 0:
            psf = psf / total
    return psf


def idft2_brute(q):
    """Direct 2-D inverse DFT with numpy's 1/(ny nx) normalisation (O(N^2) per plane)."""
    ny, nx = q.shape
    y = np.arange(ny)
    x = np.arange(nx)
    ey = np.exp(2j * np.pi * np.outer(y, y) / ny)
    ex = np.exp(2j * np.pi * np.outer(x, x) / nx)
    return ey @ q @ ex.T / (ny * nx)


def vectorial_factors(cos_t, sin_t, phi, mask, ni, ns):
    """(3, 2, ny, nx) dipole-to-field factors with Fresnel transmission (psf/pupil.py:156-262)."""
    s2 = np.clip((ni / ns) * sin_t, 0.0, 1.0)
    c2 = np.sqrt(1.0 - s2 * s2)
    eps = np.finfo(np.float64).eps
    m = mask.astype(np.float64)
    ts = 2.0 * ns * c2 / np.maximum(ns * c2 + ni * cos_t, eps) * m
    tp = 2.0 * ns * c2 / np.maximum(ni * c2 + ns * cos_t, eps) * m
    cp, sp = np.cos(phi), np.sin(phi)
    p = (tp * cos_t * cp, tp * cos_t * sp)
    s = (-ts * sp, ts * cp)
    f = np.zeros((3, 2) + cos_t.shape)
    for c in range(2):
        f[0, c] = cp * p[c] - sp * s[c]
        f[1, c] = sp * p[c] + cp * s[c]
    f[2, 0] = -sin_t * cp * tp
    f[2, 1] = -sin_t * sp * tp
    return f


def vectorial_psf(pupil, kz, cos_t, sin_t, phi, mask, ni, ns, z, dipole="isotropic", normalize=True):
    """Incoherent sum over dipoles / coherent sum for a fixed one (psf/widefield.py:163-232)."""
    f = vectorial_factors(cos_t, sin_t, phi, mask, ni, ns)
    apod = 1.0 / np.sqrt(np.maximum(cos_t, np.finfo(np.float64).eps)) * mask.astype(np.float64)
    z = np.atleast_1d(np.asarray(z, np.float64))
    pd = (np.asarray(pupil, np.complex128) * apod) * np.exp(2j * np.pi * np.asarray(kz) * z[:, None, None])

    def inten(mx, my):
        ex = np.fft.ifft2(pd * mx, axes=(-2, -1))
        ey = np.fft.ifft2(pd * my, axes=(-2, -1))
        return np.abs(ex) ** 2 + np.abs(ey) ** 2

    if dipole == "isotropic":
        psf = sum(inten(f[d, 0], f[d, 1]) for d in range(3)) / 3.0
    elif dipole in ("x", "y", "z"):
        d = "xyz".index(dipole)
        psf = inten(f[d, 0], f[d, 1])
    else:
        th, ph = dipole
        mu = (np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th))
        psf = inten(sum(mu[d] * f[d, 0] for d in range(3)), sum(mu[d] * f[d, 1] for d in range(3)))
    if normalize and psf.sum() > 0:
        psf = psf / psf.sum()
    return psf
