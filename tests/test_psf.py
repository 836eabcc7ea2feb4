"""PSF synthesis: the oracle pinned, and the host restatement against it (CPU).

The oracle (oracle/psf.py) is pinned by Parseval's identity per plane, a
brute-force inverse DFT, the flat plane of a single-sample pupil and the
z-mirror symmetry of a real point-symmetric pupil; the product's host
``pupil_to_psf`` (used to synthesise the workloads) is checked against it.
The GPU synthesis is tested in tests/test_gpu_psf.py.
"""

import numpy as np
import pytest

from oracle import psf as OP


def _random_pupil(ny, nx, seed):
    rng = np.random.default_rng(seed)
    pupil = (rng.random((ny, nx)) + 1j * rng.random((ny, nx))) * (rng.random((ny, nx)) > 0.3)
    kz = rng.random((ny, nx)) * 3.0
    return pupil, kz


def test_oracle_parseval_per_plane():
    pupil, kz = _random_pupil(12, 10, 0)
    z = np.array([-1.3, 0.0, 0.4, 2.7])
    psf = OP.pupil_to_psf(pupil, kz, z, normalize=False)
    expect = float(np.sum(np.abs(pupil) ** 2)) / (12 * 10)  # |exp(i phi)| = 1 leaves it unchanged
    for plane in psf:
        assert abs(plane.sum() - expect) <= 1e-12 * expect
    assert abs(OP.pupil_to_psf(pupil, kz, z).sum() - 1.0) < 1e-12


def test_oracle_matches_brute_force_dft():
    pupil, kz = _random_pupil(6, 5, 1)
    z = np.array([0.7, -0.2])
    psf = OP.pupil_to_psf(pupil, kz, z, normalize=False)
    for i, zi in enumerate(z):
        a = OP.idft2_brute(pupil * np.exp(2j * np.pi * kz * zi))
        np.testing.assert_allclose(psf[i], np.abs(a) ** 2, rtol=1e-12, atol=1e-15)


def test_oracle_single_sample_pupil_is_flat():
    pupil = np.zeros((8, 6), complex)
    pupil[3, 2] = 2.0 - 1.0j
    psf = OP.pupil_to_psf(pupil, np.ones((8, 6)), [0.0, 0.31], normalize=False)
    np.testing.assert_allclose(psf, abs(2.0 - 1.0j) ** 2 / (8 * 6) ** 2, rtol=1e-12)


def test_oracle_z_mirror_symmetry():
    # a real pupil: psf(-z)[y, x] = psf(z)[-y mod ny, -x mod nx]
    n = 8
    k = np.minimum(np.arange(n), n - np.arange(n))
    kz = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2) * 0.37
    pupil = (kz < 1.2).astype(float)
    a = OP.pupil_to_psf(pupil, kz, [0.8], normalize=False)[0]
    b = OP.pupil_to_psf(pupil, kz, [-0.8], normalize=False)[0]
    mirrored = np.roll(a[::-1, ::-1], (1, 1), axis=(0, 1))
    np.testing.assert_allclose(b, mirrored, rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("shape", [(6, 16, 16), (5, 12, 20)])
def test_host_pupil_to_psf_matches_oracle(shape):
    from deconlib_amd.psf import (Optics, ZernikeAberration, apply_aberrations, fft_coords, make_geometry,
                                  make_pupil, pupil_to_psf)

    nz, ny, nx = shape
    optics = Optics(wavelength=0.525, na=1.4, ni=1.515, ns=1.334)
    geom = make_geometry((ny, nx), 0.085, optics)
    pupil = apply_aberrations(make_pupil(geom), geom, optics, [ZernikeAberration({12: 0.4})])
    z = fft_coords(nz, 0.1)
    got = pupil_to_psf(pupil, geom, z)
    ref = OP.pupil_to_psf(pupil, geom.kz, z)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-16)


def _geom(ny, nx):
    from deconlib_amd.psf import Optics, make_geometry

    optics = Optics(wavelength=0.525, na=1.4, ni=1.515, ns=1.334)
    return make_geometry((ny, nx), 0.085, optics), optics


def _vec(g, o, pupil, z, dipole, normalize=False):
    return OP.vectorial_psf(pupil, g.kz, g.cos_theta, g.sin_theta, g.phi, g.mask, o.ni, o.ns, z, dipole, normalize)


def test_oracle_vectorial_dipole_relations():
    """isotropic = mean of the axis dipoles; a fixed dipole along an axis is that axis."""
    g, o = _geom(16, 20)
    pupil = g.support_weight.astype(complex)
    z = np.array([-0.3, 0.0, 0.2])
    px, py, pz = (_vec(g, o, pupil, z, d) for d in "xyz")
    np.testing.assert_allclose(_vec(g, o, pupil, z, "isotropic"), (px + py + pz) / 3.0, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(_vec(g, o, pupil, z, (0.0, 0.0)), pz, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(_vec(g, o, pupil, z, (np.pi / 2, 0.0)), px, rtol=1e-10, atol=1e-16)
    np.testing.assert_allclose(_vec(g, o, pupil, z, (np.pi / 2, np.pi / 2)), py, rtol=1e-10, atol=1e-16)


def test_oracle_vectorial_parseval():
    """Each raw plane holds sum_c w_c sum |P_c|^2 / (ny nx), independent of z."""
    g, o = _geom(12, 16)
    pupil = g.support_weight.astype(complex)
    f = OP.vectorial_factors(g.cos_theta, g.sin_theta, g.phi, g.mask, o.ni, o.ns)
    apod = 1.0 / np.sqrt(np.maximum(g.cos_theta, np.finfo(float).eps)) * g.mask
    energy = sum(np.sum(np.abs(pupil * apod * f[d, c]) ** 2) for d in range(3) for c in range(2)) / 3.0
    psf = _vec(g, o, pupil, [-1.0, 0.0, 0.7], "isotropic")
    for plane in psf:
        assert abs(plane.sum() - energy / (12 * 16)) <= 1e-12 * energy


@pytest.mark.parametrize("dipole", ["isotropic", "y", (0.7, 1.1)])
def test_host_vectorial_matches_oracle(dipole):
    from deconlib_amd.psf import IndexMismatch, apply_aberrations, fft_coords, make_pupil, pupil_to_vectorial_psf

    g, o = _geom(16, 24)
    pupil = apply_aberrations(make_pupil(g), g, o, [IndexMismatch(4.0)])
    z = fft_coords(5, 0.1)
    got = pupil_to_vectorial_psf(pupil, g, o, z, dipole)
    ref = _vec(g, o, pupil, z, dipole, normalize=True)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-16)
