"""``.psf.h5`` I/O (deconlib/io.py:37-227): the reference's round-trip tests
(tests/test_io.py:160-248 of the reference, PSF part), and the file layout checked with the
HDF5 distribution's own ``h5dump`` (an independent reader): attribute types as h5py writes
them (variable-length UTF-8 strings, little-endian int64 / float64), the gzip-3 ``psf``
dataset, groups and values."""

import os
import re
import shutil
import subprocess
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from deconlib_amd import _hdf5 as H
from deconlib_amd.io import Psf, load_psf, save_psf
from deconlib_amd.psf import Optics

try:
    H.lib()
    _HAVE_HDF5 = True
except RuntimeError:
    _HAVE_HDF5 = False

pytestmark = pytest.mark.skipif(not _HAVE_HDF5, reason="no HDF5 C library in this environment")


def _h5dump():
    for c in (shutil.which("h5dump"), "/opt/conda/bin/h5dump"):
        if c and os.path.exists(c):
            return c
    pytest.skip("h5dump not available")


@pytest.fixture
def optics() -> Optics:
    return Optics(wavelength=0.525, na=1.4, ni=1.515, ns=1.334)


def test_psf_3d_round_trip(tmp_path: Path, optics: Optics) -> None:
    rng = np.random.default_rng(0)
    psf_arr = rng.random((32, 64, 64)).astype(np.float32)
    psf = Psf(psf=psf_arr, optics=optics, pixel_size=(0.1, 0.085, 0.085), source="theoretical")
    path = tmp_path / "theory.psf.h5"
    save_psf(path, psf)
    loaded = load_psf(path)
    np.testing.assert_array_equal(loaded.psf, psf_arr)
    assert loaded.psf.dtype == np.float32
    assert loaded.optics == optics
    assert loaded.pixel_size == (0.1, 0.085, 0.085)
    assert loaded.source == "theoretical"
    assert loaded.z_planes is None
    assert loaded.pupil_ref is None
    assert loaded.distillation_diagnostics is None


def test_psf_2d_round_trip(tmp_path: Path, optics: Optics) -> None:
    psf_arr = np.ones((32, 32), dtype=np.float32)
    psf = Psf(psf=psf_arr, optics=optics, pixel_size=(0.085, 0.085), source="theoretical")
    path = tmp_path / "psf2d.psf.h5"
    save_psf(path, psf)
    loaded = load_psf(path)
    assert loaded.psf.shape == psf_arr.shape
    assert loaded.pixel_size == (0.085, 0.085)


def test_psf_with_z_planes(tmp_path: Path, optics: Optics) -> None:
    z_planes = np.array([-1.0, -0.3, 0.0, 0.3, 1.0], dtype=np.float64)
    psf = Psf(psf=np.zeros((5, 16, 16), dtype=np.float32), optics=optics, pixel_size=(0.3, 0.085, 0.085),
              z_planes=z_planes, source="pupil_sampled", pupil_ref="parent.pupil.h5")
    path = tmp_path / "psf_sampled.psf.h5"
    save_psf(path, psf)
    loaded = load_psf(path)
    np.testing.assert_array_equal(loaded.z_planes, z_planes)
    assert loaded.source == "pupil_sampled"
    assert loaded.pupil_ref == "parent.pupil.h5"


def test_psf_from_distillation(tmp_path: Path, optics: Optics) -> None:
    psf_arr = np.zeros((16, 32, 32), dtype=np.float32)
    psf_arr[0, 0, 0] = 1.0
    result = SimpleNamespace(
        psf=psf_arr,
        positions=np.array([[8, 16, 16], [10, 12, 18]], dtype=np.int32),
        amplitudes=np.array([1.5, 0.8], dtype=np.float64),
        chi2_history=[10.0, 5.0, 2.0],
        psf_change_history=[1.0, 0.3, 0.05],
        amp_change_history=[0.5, 0.1, 0.01],
    )
    psf = Psf.from_distillation(result, optics=optics, pixel_size=(0.2, 0.1, 0.1))
    path = tmp_path / "distilled.psf.h5"
    save_psf(path, psf)
    loaded = load_psf(path)
    assert loaded.source == "distilled"
    d = loaded.distillation_diagnostics
    assert sorted(d) == ["amp_change_history", "amplitudes", "chi2_history", "positions", "psf_change_history"]
    np.testing.assert_array_equal(d["positions"], result.positions)
    assert d["positions"].dtype == np.int32
    np.testing.assert_array_equal(d["amplitudes"], result.amplitudes)
    np.testing.assert_array_equal(d["chi2_history"], result.chi2_history)


def test_psf_load_rejects_wrong_format(tmp_path: Path) -> None:
    path = tmp_path / "bogus.h5"
    with H.File(str(path), "w") as f:
        H.write_attr(f.id, "format", "deconlib_pupil")
    with pytest.raises(ValueError, match="not a deconlib psf"):
        load_psf(path)


def test_psf_pixel_size_ndim_mismatch_rejected(optics: Optics) -> None:
    with pytest.raises(ValueError, match="does not match"):
        Psf(psf=np.zeros((8, 16, 16), dtype=np.float32), optics=optics, pixel_size=(0.1, 0.1))
    with pytest.raises(ValueError, match="2D or 3D"):
        Psf(psf=np.zeros(8, dtype=np.float32), optics=optics, pixel_size=(0.1,))


def test_metadata_attributes(tmp_path: Path, optics: Optics) -> None:
    psf = Psf(psf=np.ones((4, 8, 8), np.float32), optics=optics, pixel_size=(0.1, 0.1, 0.1))
    path = tmp_path / "meta.psf.h5"
    save_psf(path, psf, metadata={"objective": "100x/1.4", "n_beads": 12, "snr": 31.5,
                                  "flags": np.array([1, 2, 3], np.int32), "skipped": {"a": 1}})
    with H.File(str(path), "r") as f:
        assert H.read_attr(f.id, "objective") == "100x/1.4"
        assert H.read_attr(f.id, "n_beads") == 12
        assert H.read_attr(f.id, "snr") == 31.5
        np.testing.assert_array_equal(H.read_attr(f.id, "flags"), [1, 2, 3])
        assert not H.has_attr(f.id, "skipped")
        assert H.read_attr(f.id, "version") == 1


def test_layout_seen_by_h5dump(tmp_path: Path, optics: Optics) -> None:
    """The HDF5 distribution's reader sees the reference's layout and encodings."""
    exe = _h5dump()
    rng = np.random.default_rng(5)
    psf_arr = rng.random((6, 20, 24)).astype(np.float32)
    z = np.linspace(-0.5, 0.5, 6)
    psf = Psf(psf=psf_arr, optics=optics, pixel_size=(0.2, 0.085, 0.085), z_planes=z, source="pupil_sampled",
              pupil_ref="p.pupil.h5")
    path = tmp_path / "dump.psf.h5"
    save_psf(path, psf)
    hdr = subprocess.run([exe, "-p", "-A", str(path)], capture_output=True, text=True, check=True).stdout
    flat = " ".join(hdr.split())
    # string attributes: variable-length UTF-8, scalar (h5py's encoding of str)
    m = re.search(r'ATTRIBUTE "format" \{ DATATYPE H5T_STRING \{ STRSIZE H5T_VARIABLE; STRPAD H5T_STR_NULLTERM; '
                  r'CSET H5T_CSET_UTF8; CTYPE H5T_C_S1; \} DATASPACE SCALAR DATA \{ \(0\): "deconlib_psf" \}', flat)
    assert m, flat[:800]
    assert re.search(r'ATTRIBUTE "version" \{ DATATYPE H5T_STD_I64LE DATASPACE SCALAR DATA \{ \(0\): 1 \}', flat)
    assert re.search(r'ATTRIBUTE "pupil_ref" \{ DATATYPE H5T_STRING \{ STRSIZE H5T_VARIABLE;.*?"p.pupil.h5"', flat)
    assert re.search(r'DATASET "psf" \{ DATATYPE H5T_IEEE_F32LE DATASPACE SIMPLE \{ \( 6, 20, 24 \) / \( 6, 20, 24 \) \}',
                     flat)
    assert re.search(r'COMPRESSION DEFLATE \{ LEVEL 3 \}', flat)
    assert re.search(r'GROUP "optics" \{ ATTRIBUTE "na" \{ DATATYPE H5T_IEEE_F64LE DATASPACE SCALAR DATA \{ \(0\): 1.4 \}',
                     flat)
    assert re.search(r'ATTRIBUTE "pixel_size" \{ DATATYPE H5T_IEEE_F64LE DATASPACE SIMPLE \{ \( 3 \) / \( 3 \) \} '
                     r'DATA \{ \(0\): 0.2, 0.085, 0.085 \}', flat)
    assert re.search(r'DATASET "z_planes" \{ DATATYPE H5T_IEEE_F64LE DATASPACE SIMPLE \{ \( 6 \) / \( 6 \) \}', flat)
    # the psf values through h5dump's binary export
    out = tmp_path / "psf.bin"
    subprocess.run([exe, "-d", "/psf", "-b", "LE", "-o", str(out), str(path)], capture_output=True, check=True)
    np.testing.assert_array_equal(np.fromfile(out, dtype="<f4").reshape(psf_arr.shape), psf_arr)
