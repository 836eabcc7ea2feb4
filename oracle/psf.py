"""CPU oracle for PSF synthesis (TEST INFRASTRUCTURE, not the product).

Only ``tests/`` may import this module.  It restates ``pupil_to_psf``
(deconlib/psf/widefield.py:18-73) in complex128 NumPy: per z-plane
``|ifft2(pupil * exp(2 pi i kz z))|^2``, DC at the corner, normalised to sum 1.

Parity pins (tests/test_psf.py): Parseval's identity per plane
(sum |ifft2(Q)|^2 = sum |Q|^2 / (ny nx) for a unit-modulus defocus), a brute-force
2-D inverse DFT on tiny pupils, the flat plane of a single-sample pupil, and the
z-mirror symmetry of a real, point-symmetric pupil.
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
