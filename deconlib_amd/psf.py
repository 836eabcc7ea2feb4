"""Widefield PSF synthesis on the host (the input producer of the solvers).

Restates the reference's scalar PSF model so that workloads can be generated
without the reference installed:

* ``Optics``                      psf/optics.py:10-53
* ``make_geometry``               psf/optics.py:96-225 (anti-aliased support weight)
* ``make_pupil``                  psf/pupil.py:17-45 (uniform pupil, no apodization)
* ``fft_coords``                  utils/fourier.py:11-36
* ``pupil_to_psf``                psf/widefield.py:18-73 (DC at corner, normalised)
* ``IndexMismatch``               psf/aberrations/geometric.py:11-61 (Gibson–Lanni OPD)
* ``ZernikeAberration``           psf/aberrations/zernike.py:55-127, utils/zernike.py:21-120
* ``apply_aberrations``           psf/aberrations/base.py:40-71

``pupil_to_psf`` works z-plane by z-plane (the reference forms the whole
(nz, ny, nx) complex128 stack at once), so large PSFs fit in host memory.
``pupil_to_psf_device`` is the same computation on the GPU (``dh_psf_synth``:
defocus formed on load, float32 transforms, float64 normalisation); the pupil,
geometry and aberration phases are 2-D and stay on the host.
"""

from __future__ import annotations

from dataclasses import dataclass
from math import factorial
from typing import Dict, Tuple, Union

import numpy as np


@dataclass(frozen=True)
class Optics:
    wavelength: float
    na: float
    ni: float
    ns: float = None

    def __post_init__(self):
        if self.ns is None:
            object.__setattr__(self, "ns", self.ni)
        if self.na > self.ni:
            raise ValueError(f"NA ({self.na}) cannot exceed immersion refractive index ({self.ni})")
        if self.wavelength <= 0:
            raise ValueError(f"Wavelength must be positive, got {self.wavelength}")

    @property
    def k_cutoff(self) -> float:
        return self.na / self.wavelength

    @property
    def k_immersion(self) -> float:
        return self.ni / self.wavelength


@dataclass(frozen=True)
class Geometry:
    kx: np.ndarray
    ky: np.ndarray
    kz: np.ndarray
    rho: np.ndarray
    phi: np.ndarray
    mask: np.ndarray
    support_weight: np.ndarray
    cos_theta: np.ndarray
    sin_theta: np.ndarray

    @property
    def shape(self) -> Tuple[int, int]:
        return self.mask.shape


def make_geometry(shape, spacing, optics: Optics, oversample: int = 8) -> Geometry:
    ny, nx = shape
    if isinstance(spacing, (int, float)):
        dy = dx = float(spacing)
    else:
        dy, dx = spacing
    if dy <= 0 or dx <= 0:
        raise ValueError(f"Spacing must be positive, got ({dy}, {dx})")
    if oversample < 1:
        raise ValueError(f"oversample must be >= 1, got {oversample}")
    kx1 = np.fft.fftfreq(nx, dx)
    ky1 = np.fft.fftfreq(ny, dy)
    kx, ky = np.meshgrid(kx1, ky1, indexing="xy")
    kr = np.sqrt(kx ** 2 + ky ** 2)
    mask = kr <= optics.k_cutoff
    dkx, dky = 1.0 / (nx * dx), 1.0 / (ny * dy)
    kc2 = optics.k_cutoff ** 2
    if oversample == 1:
        w = mask.astype(np.float64)
    else:
        sub = (np.arange(oversample) + 0.5) / oversample - 0.5
        w = np.zeros_like(kr)
        for jx in range(oversample):
            kxs2 = (kx + sub[jx] * dkx) ** 2
            for jy in range(oversample):
                kys = ky + sub[jy] * dky
                w += kxs2 + kys * kys <= kc2
        w /= oversample * oversample
    rho = np.zeros_like(kr)
    rho[mask] = kr[mask] / optics.k_cutoff
    phi = np.arctan2(ky, kx)
    sin_t = np.clip(optics.wavelength * kr / optics.ni, 0.0, 1.0)
    cos_t = np.sqrt(1.0 - sin_t ** 2)
    kz = np.sqrt(np.maximum(0.0, optics.k_immersion ** 2 - kr ** 2))
    return Geometry(kx=kx, ky=ky, kz=kz, rho=rho, phi=phi, mask=mask, support_weight=w,
                    cos_theta=cos_t, sin_theta=sin_t)


def make_pupil(geom: Geometry) -> np.ndarray:
    return geom.support_weight.astype(np.complex128)


def fft_coords(n: int, spacing: float = 1.0) -> np.ndarray:
    return np.fft.fftfreq(n) * n * spacing


def pupil_to_psf(pupil: np.ndarray, geom: Geometry, z: np.ndarray, normalize: bool = True,
                 dtype=np.float64) -> np.ndarray:
    """|IFFT2(P exp(2 pi i kz z))|^2 per z-plane, DC at the corner."""
    z = np.atleast_1d(z)
    out = np.empty((len(z),) + pupil.shape, dtype=dtype)
    total = 0.0
    for i, zi in enumerate(z):
        a = np.fft.ifft2(pupil * np.exp(2j * np.pi * geom.kz * zi))
        plane = a.real ** 2 + a.imag ** 2
        total += float(plane.sum())
        out[i] = plane
    if normalize and total > 0:
        out /= total
    return out


class _PsfPlan:
    """A bound dh_psf_plan for one pupil shape (its twiddle rows on the device)."""

    def __init__(self, ny: int, nx: int):
        import ctypes as C

        import torch

        from . import _native as N
        from ._tensor import device

        h = C.c_void_p()
        N.call("dh_psf_plan_create", int(ny), int(nx), C.byref(h))
        self._h = h
        self._destroy = N.fns["dh_psf_plan_destroy"]  # bound now: no imports at interpreter exit
        tb = C.c_int64()
        N.call("dh_psf_plan_info", h, 1, 1, C.byref(tb), None)
        self.twiddles = torch.empty(int(tb.value), dtype=torch.uint8, device=device())
        N.call("dh_psf_plan_bind", h, self.twiddles.data_ptr(), N.stream_ptr())

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)
            self._h = None


_PSF_PLANS: Dict[Tuple[int, int], _PsfPlan] = {}


def _synth_device(channels, weights, geom: Geometry, z, normalize: bool):
    """sum_c w_c |ifft2(channel_c * defocus)|^2 on the GPU (dh_psf_synth)."""
    import ctypes as C

    import torch

    from . import _native as N
    from ._tensor import device

    dev = device()
    if isinstance(channels, torch.Tensor):
        pt = channels.to(device=dev, dtype=torch.complex64).contiguous()
    else:
        pt = torch.from_numpy(np.ascontiguousarray(channels, dtype=np.complex64)).to(dev)
    nchan, ny, nx = (int(v) for v in pt.shape)
    zz = np.atleast_1d(np.asarray(z, dtype=np.float64))
    nz = len(zz)
    key = (ny, nx)
    if key not in _PSF_PLANS:
        _PSF_PLANS[key] = _PsfPlan(ny, nx)
    plan = _PSF_PLANS[key]
    kz = torch.from_numpy(np.ascontiguousarray(geom.kz, dtype=np.float64)).to(dev)
    zt = torch.from_numpy(zz).to(dev)
    w = (C.c_double * nchan)(*[float(v) for v in weights])
    wb = C.c_int64()
    N.call("dh_psf_plan_info", plan._h, nz, nchan, None, C.byref(wb))
    work = torch.empty(int(wb.value), dtype=torch.uint8, device=dev)
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    N.call("dh_psf_synth", plan._h, nchan, pt.data_ptr(), w, kz.data_ptr(), zt.data_ptr(), nz,
           int(bool(normalize)), out.data_ptr(), work.data_ptr(), int(wb.value), None, N.stream_ptr())
    return out


def pupil_to_psf_device(pupil, geom: Geometry, z, normalize: bool = True):
    """pupil_to_psf (psf/widefield.py:18-73) on the GPU; a (nz, ny, nx) float32 device tensor.

    ``pupil`` (ny, nx) complex, numpy or a device tensor; ``geom.kz`` and ``z`` as
    for ``pupil_to_psf``.  Pupil sides must be 2^a * {1, 3, 5} <= 4096.
    """
    p = pupil[None] if not hasattr(pupil, "unsqueeze") else pupil.unsqueeze(0)
    return _synth_device(p, [1.0], geom, z, normalize)


# ---- vectorial PSF (psf/pupil.py:48-75,156-262; psf/widefield.py:110-234) ----

def aplanatic_apodization(geom: Geometry, eps: float = None) -> np.ndarray:
    """1 / sqrt(cos theta) on the binary NA mask (pupil.py:48-75)."""
    eps = np.finfo(np.float64).eps if eps is None else eps
    return 1.0 / np.sqrt(np.maximum(geom.cos_theta, eps)) * geom.mask.astype(np.float64)


def compute_fresnel_coefficients(geom: Geometry, optics: Optics):
    """(t_s, t_p) transmission sample -> immersion, on the binary NA mask (pupil.py:156-202)."""
    c1, s1 = geom.cos_theta, geom.sin_theta
    s2 = np.clip((optics.ni / optics.ns) * s1, 0.0, 1.0)
    c2 = np.sqrt(1.0 - s2 ** 2)
    num = 2.0 * optics.ns * c2
    eps = np.finfo(np.float64).eps
    mask = geom.mask.astype(np.float64)
    t_s = num / np.maximum(optics.ns * c2 + optics.ni * c1, eps) * mask
    t_p = num / np.maximum(optics.ni * c2 + optics.ns * c1, eps) * mask
    return t_s, t_p


def compute_vectorial_factors(geom: Geometry, optics: Optics) -> np.ndarray:
    """(3 dipoles x, y, z; 2 fields Ex, Ey; ny, nx) dipole-to-field factors (pupil.py:205-262)."""
    t_s, t_p = compute_fresnel_coefficients(geom, optics)
    c1, s1 = geom.cos_theta, geom.sin_theta
    cp, sp = np.cos(geom.phi), np.sin(geom.phi)
    px, py = t_p * c1 * cp, t_p * c1 * sp  # p polarisation (plane of incidence)
    sx, sy = -t_s * sp, t_s * cp           # s polarisation
    f = np.empty((3, 2) + geom.shape, dtype=np.float64)
    f[0, 0], f[0, 1] = cp * px - sp * sx, cp * py - sp * sy
    f[1, 0], f[1, 1] = sp * px + cp * sx, sp * py + cp * sy
    f[2, 0], f[2, 1] = -s1 * cp * t_p, -s1 * sp * t_p
    return f


def _vectorial_channels(pupil, geom: Geometry, optics: Optics, dipole):
    """(channels (K, ny, nx), weights) of the incoherent sum for a dipole mode (widefield.py:172-224)."""
    f = compute_vectorial_factors(geom, optics)
    pa = np.asarray(pupil) * aplanatic_apodization(geom)
    if isinstance(dipole, str) and dipole == "isotropic":
        return np.stack([pa * f[d, c] for d in range(3) for c in range(2)]), [1.0 / 3.0] * 6
    if isinstance(dipole, str) and dipole in ("x", "y", "z"):
        d = {"x": 0, "y": 1, "z": 2}[dipole]
        return np.stack([pa * f[d, 0], pa * f[d, 1]]), [1.0, 1.0]
    if isinstance(dipole, tuple):
        th, ph = dipole
        mu = (np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th))
        mx = mu[0] * f[0, 0] + mu[1] * f[1, 0] + mu[2] * f[2, 0]
        my = mu[0] * f[0, 1] + mu[1] * f[1, 1] + mu[2] * f[2, 1]
        return np.stack([pa * mx, pa * my]), [1.0, 1.0]
    raise ValueError(f"Unknown dipole type: {dipole}")


def pupil_to_vectorial_psf(pupil, geom: Geometry, optics: Optics, z, dipole="isotropic",
                           normalize: bool = True) -> np.ndarray:
    """Vectorial PSF on the host (widefield.py:110-234), plane by plane in complex128."""
    ch, w = _vectorial_channels(pupil, geom, optics, dipole)
    z = np.atleast_1d(z)
    out = np.zeros((len(z),) + geom.shape, dtype=np.float64)
    for i, zi in enumerate(z):
        d = np.exp(2j * np.pi * geom.kz * zi)
        for c in range(len(w)):
            a = np.fft.ifft2(ch[c] * d)
            out[i] += (a.real ** 2 + a.imag ** 2) * w[c]
    if normalize:
        total = out.sum()
        if total > 0:
            out /= total
    return out


def pupil_to_vectorial_psf_device(pupil, geom: Geometry, optics: Optics, z, dipole="isotropic",
                                  normalize: bool = True):
    """pupil_to_vectorial_psf on the GPU (dh_psf_synth over the field channels)."""
    ch, w = _vectorial_channels(pupil, geom, optics, dipole)
    return _synth_device(ch, w, geom, z, normalize)


class IndexMismatch:
    """Gibson–Lanni depth aberration: exp(2 pi i d (n2 cos t2 - n1 cos t1) / lambda)."""

    def __init__(self, depth: float):
        self.depth = depth

    def __call__(self, geom: Geometry, optics: Optics) -> np.ndarray:
        n1, n2 = optics.ni, optics.ns
        s2 = np.clip((n1 / n2) * geom.sin_theta, 0.0, 1.0)
        c2 = np.sqrt(1.0 - s2 ** 2)
        opd = self.depth * (n2 * c2 - n1 * geom.cos_theta)
        return np.exp(1j * 2.0 * np.pi * opd / optics.wavelength)


def _ansi_to_nm(j: int):
    n = int(np.ceil((-3 + np.sqrt(9 + 8 * j)) / 2))
    return n, 2 * j - n * (n + 2)


def zernike_polynomial(j: int, rho: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Orthonormal OSA/ANSI Zernike Z_j."""
    n, m = _ansi_to_nm(j)
    am = abs(m)
    R = np.zeros_like(rho)
    for k in range((n - am) // 2 + 1):
        c = (-1) ** k * factorial(n - k) / (
            factorial(k) * factorial((n + am) // 2 - k) * factorial((n - am) // 2 - k))
        R += c * rho ** (n - 2 * k)
    norm = np.sqrt(2.0 * (n + 1) / (1.0 + float(m == 0)))
    return norm * R * np.cos(m * phi) if m >= 0 else -norm * R * np.sin(am * phi)


class ZernikeAberration:
    SPHERICAL = 12

    def __init__(self, coefficients: Union[Dict[int, float], np.ndarray]):
        self.coefficients = dict(coefficients) if isinstance(coefficients, dict) else {
            j: c for j, c in enumerate(np.asarray(coefficients))}

    def __call__(self, geom: Geometry, optics: Optics) -> np.ndarray:
        phase = np.zeros_like(geom.rho)
        for j, c in self.coefficients.items():
            if c != 0:
                phase += c * zernike_polynomial(int(j), geom.rho, geom.phi)
        return np.exp(1j * phase * geom.mask)


def apply_aberrations(pupil, geom, optics, aberrations) -> np.ndarray:
    out = pupil.copy()
    for a in aberrations:
        out = out * a(geom, optics)
    return out


def widefield_psf(shape, lateral: float, axial: float, optics: Optics = None, aberrations=(),
                  dtype=np.float32, device: bool = False, dipole=None):
    """(nz, ny, nx) corner-origin PSF normalised to sum 1 (the configs' recipe).

    ``dipole=None``: the scalar widefield PSF (pupil_to_psf, widefield.py:18-73);
    ``dipole="isotropic"`` (or "x"/"y"/"z"/(theta, phi)): the vectorial PSF
    (pupil_to_vectorial_psf, widefield.py:110-234).
    ``device=True``: synthesised on the GPU, returned as a float32 device tensor.
    """
    optics = optics or Optics(wavelength=0.525, na=1.4, ni=1.515, ns=1.334)
    nz, ny, nx = shape
    geom = make_geometry((ny, nx), lateral, optics)
    pupil = make_pupil(geom)
    if aberrations:
        pupil = apply_aberrations(pupil, geom, optics, list(aberrations))
    z = fft_coords(nz, axial)
    if dipole is not None:
        if device:
            return pupil_to_vectorial_psf_device(pupil, geom, optics, z, dipole=dipole, normalize=True)
        return pupil_to_vectorial_psf(pupil, geom, optics, z, dipole=dipole, normalize=True).astype(dtype)
    if device:
        return pupil_to_psf_device(pupil, geom, z, normalize=True)
    psf = pupil_to_psf(pupil, geom, z, normalize=True)
    return psf.astype(dtype)
