"""GPU PSF synthesis (dh_psf_synth) against the oracle and closed forms."""

import numpy as np
import pytest

from oracle import psf as OP

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _pupil(ny, nx, lateral=0.085, zernike=0.4, depth=None):
    from deconlib_amd.psf import (IndexMismatch, Optics, ZernikeAberration, apply_aberrations, make_geometry,
                                  make_pupil)

    optics = Optics(wavelength=0.525, na=1.4, ni=1.515, ns=1.334)
    geom = make_geometry((ny, nx), lateral, optics)
    ab = [ZernikeAberration({12: zernike})] + ([IndexMismatch(depth)] if depth else [])
    return apply_aberrations(make_pupil(geom), geom, optics, ab), geom


@pytest.mark.parametrize("shape", [(8, 32, 32), (5, 48, 40), (3, 20, 12), (4, 1, 16), (6, 96, 80)])
def test_device_psf_matches_oracle(gpu, shape):
    from deconlib_amd.psf import fft_coords, pupil_to_psf_device

    nz, ny, nx = shape
    pupil, geom = _pupil(ny, nx, depth=5.0)
    z = fft_coords(nz, 0.1)
    for norm in (True, False):
        got = pupil_to_psf_device(pupil, geom, z, normalize=norm).cpu().numpy()
        ref = OP.pupil_to_psf(pupil, geom.kz, z, normalize=norm)
        assert got.shape == (nz, ny, nx) and got.dtype == np.float32
        assert rel_l2(got, ref) <= 1e-5, (shape, norm, rel_l2(got, ref))
    assert abs(float(got.sum(dtype=np.float64)) - float(ref.sum())) <= 1e-5 * float(ref.sum())


def test_device_psf_parseval(gpu):
    """Every raw plane sums to sum|P|^2 / (ny nx), whatever the defocus."""
    from deconlib_amd.psf import pupil_to_psf_device

    rng = np.random.default_rng(3)
    ny, nx = 64, 48
    pupil = (rng.random((ny, nx)) + 1j * rng.random((ny, nx))) * (rng.random((ny, nx)) > 0.4)

    class G:
        kz = rng.random((ny, nx)) * 2.9

    z = np.linspace(-6.0, 6.0, 9)
    psf = pupil_to_psf_device(pupil, G, z, normalize=False).cpu().numpy().astype(np.float64)
    expect = float(np.sum(np.abs(pupil) ** 2)) / (ny * nx)
    for plane in psf:
        assert abs(plane.sum() - expect) <= 1e-5 * expect


def test_device_psf_config_shape_matches_host(gpu):
    """The 64x256x256 widefield PSF of the RL workload, GPU against the host synthesis."""
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS

    host = widefield_psf((64, 256, 256), 0.085, 0.1, WIDEFIELD_OPTICS, dtype=np.float64)
    dev = widefield_psf((64, 256, 256), 0.085, 0.1, WIDEFIELD_OPTICS, device=True).cpu().numpy()
    assert rel_l2(dev, host) <= 1e-5
    assert abs(float(dev.sum(dtype=np.float64)) - 1.0) <= 1e-5


def test_device_psf_rejects_unsupported_sides(gpu):
    from deconlib_amd.psf import pupil_to_psf_device

    class G:
        kz = np.zeros((7, 9))

    with pytest.raises(NotImplementedError):
        pupil_to_psf_device(np.ones((7, 9), complex), G, [0.0])


@pytest.mark.parametrize("dipole", ["isotropic", "z", (0.6, 2.0)])
def test_device_vectorial_psf_matches_oracle(gpu, dipole):
    from deconlib_amd.psf import (IndexMismatch, Optics, apply_aberrations, fft_coords, make_geometry, make_pupil,
                                  pupil_to_vectorial_psf_device)

    optics = Optics(wavelength=0.525, na=1.4, ni=1.515, ns=1.334)
    g = make_geometry((32, 48), 0.085, optics)
    pupil = apply_aberrations(make_pupil(g), g, optics, [IndexMismatch(10.0)])
    z = fft_coords(6, 0.1)
    for norm in (True, False):
        got = pupil_to_vectorial_psf_device(pupil, g, optics, z, dipole, normalize=norm).cpu().numpy()
        ref = OP.vectorial_psf(pupil, g.kz, g.cos_theta, g.sin_theta, g.phi, g.mask, optics.ni, optics.ns, z,
                               dipole, norm)
        assert rel_l2(got, ref) <= 1e-5, (dipole, norm, rel_l2(got, ref))
