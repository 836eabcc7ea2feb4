"""``.psf.h5`` PSF files (deconlib/io.py:37-227): the reference's HDF5 layout, written and
read through the HDF5 C library (``_hdf5``), so the files are interchangeable with the
reference's h5py-written ones.

Layout: root attributes ``format`` = "deconlib_psf", ``version`` = 1, ``source``,
``created`` (ISO-8601 UTC), optional ``pupil_ref`` and free-form metadata; dataset ``psf``
(float32, gzip level 3); group ``optics`` (attributes wavelength, na, ni, ns); group
``sampling`` (attribute ``pixel_size``, optional dataset ``z_planes``); optional group
``distillation`` (positions, amplitudes and the three histories as datasets).
"""

from __future__ import annotations

import datetime as _dt
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Optional, Tuple, Union

import numpy as np

from . import _hdf5 as H
from .psf import Optics

_PSF_FORMAT = "deconlib_psf"
_FORMAT_VERSION = 1


@dataclass
class Psf:
    """A 2-D or 3-D PSF image (corner origin) with the optics and voxel sampling it was
    computed at (io.py:37-78).  ``pixel_size`` is (dz, dy, dx) or (dy, dx) in um;
    ``z_planes`` an explicit z grid for non-uniform sampling; ``source`` one of
    "theoretical", "distilled", "pupil_sampled"."""

    psf: np.ndarray
    optics: Optics
    pixel_size: Tuple[float, ...]
    z_planes: Optional[np.ndarray] = None
    source: str = "theoretical"
    pupil_ref: Optional[str] = None
    distillation_diagnostics: Optional[dict] = None

    def __post_init__(self) -> None:
        if self.psf.ndim not in (2, 3):
            raise ValueError(f"psf must be 2D or 3D, got shape {self.psf.shape}")
        if len(self.pixel_size) != self.psf.ndim:
            raise ValueError(
                f"pixel_size length {len(self.pixel_size)} does not match psf ndim {self.psf.ndim}"
            )

    @classmethod
    def from_distillation(cls, result: Any, *, optics: Optics, pixel_size: Tuple[float, ...]) -> "Psf":
        """From a PSF-distillation result (any object with ``psf``, ``positions``,
        ``amplitudes`` and the ``chi2`` / ``psf_change`` / ``amp_change`` histories,
        io.py:80-97)."""
        return cls(
            psf=np.asarray(result.psf).astype(np.float32, copy=False),
            optics=optics,
            pixel_size=tuple(pixel_size),
            source="distilled",
            distillation_diagnostics={
                "positions": np.asarray(result.positions, dtype=np.int32),
                "amplitudes": np.asarray(result.amplitudes, dtype=np.float64),
                "chi2_history": list(result.chi2_history),
                "psf_change_history": list(result.psf_change_history),
                "amp_change_history": list(result.amp_change_history),
            },
        )


def _now_iso() -> str:
    return _dt.datetime.now(_dt.timezone.utc).isoformat(timespec="seconds")


def save_psf(filepath: Union[str, Path], psf: Psf, *, metadata: Optional[dict] = None) -> None:
    """Write ``psf`` to ``filepath`` as ``.psf.h5`` (io.py:140-192).  Metadata values that
    are not numbers, number vectors or strings are skipped, as the reference's best-effort
    writer skips what h5py refuses."""
    with H.File(str(Path(filepath)), "w") as f:
        root = f.id
        H.write_attr(root, "format", _PSF_FORMAT)
        H.write_attr(root, "version", np.int64(_FORMAT_VERSION))
        H.write_attr(root, "source", str(psf.source))
        H.write_attr(root, "created", _now_iso())
        if psf.pupil_ref:
            H.write_attr(root, "pupil_ref", str(psf.pupil_ref))
        for k, v in (metadata or {}).items():
            try:
                H.write_attr(root, str(k), v)
            except (TypeError, ValueError):
                pass
        H.write_dataset(root, "psf", np.asarray(psf.psf).astype(np.float32, copy=False), gzip=3)
        og = f.create_group(root, "optics")
        for key in ("wavelength", "na", "ni", "ns"):
            H.write_attr(og, key, np.float64(getattr(psf.optics, key)))
        sg = f.create_group(root, "sampling")
        H.write_attr(sg, "pixel_size", np.asarray(psf.pixel_size, dtype=np.float64))
        if psf.z_planes is not None:
            H.write_dataset(sg, "z_planes", np.asarray(psf.z_planes, dtype=np.float64))
        if psf.distillation_diagnostics is not None:
            dg = f.create_group(root, "distillation")
            d = psf.distillation_diagnostics
            if "positions" in d:
                H.write_dataset(dg, "positions", np.asarray(d["positions"], dtype=np.int32))
            if "amplitudes" in d:
                H.write_dataset(dg, "amplitudes", np.asarray(d["amplitudes"], dtype=np.float64))
            for key in ("chi2_history", "psf_change_history", "amp_change_history"):
                if key in d:
                    H.write_dataset(dg, key, np.asarray(d[key], dtype=np.float64))


def load_psf(filepath: Union[str, Path]) -> Psf:
    """Read a ``.psf.h5`` file (io.py:195-227); ``ValueError`` when its ``format``
    attribute is not "deconlib_psf"."""
    path = Path(filepath)
    with H.File(str(path), "r") as f:
        root = f.id
        fmt = H.read_attr(root, "format") if H.has_attr(root, "format") else ""
        if fmt != _PSF_FORMAT:
            raise ValueError(f"{path}: not a deconlib psf file (format={fmt!r})")
        source = str(H.read_attr(root, "source")) if H.has_attr(root, "source") else "theoretical"
        pupil_ref = str(H.read_attr(root, "pupil_ref")) if H.has_attr(root, "pupil_ref") else None
        og = f.open_group(root, "optics")
        optics = Optics(**{k: float(H.read_attr(og, k)) for k in ("wavelength", "na", "ni", "ns")})
        sg = f.open_group(root, "sampling")
        pixel_size = tuple(float(v) for v in np.atleast_1d(H.read_attr(sg, "pixel_size")))
        z_planes = H.read_dataset(sg, "z_planes") if H.exists(sg, "z_planes") else None
        psf_arr = H.read_dataset(root, "psf").astype(np.float32, copy=False)
        diag = None
        if H.exists(root, "distillation"):
            dg = f.open_group(root, "distillation")
            diag = {key: H.read_dataset(dg, key) for key in H.members(dg)}
    return Psf(psf=psf_arr, optics=optics, pixel_size=pixel_size, z_planes=z_planes, source=source,
               pupil_ref=pupil_ref, distillation_diagnostics=diag)
