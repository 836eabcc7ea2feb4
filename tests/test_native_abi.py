"""The C-ABI library loads, exports every declared symbol, and its host-side
plan logic (no GPU call) matches the reference's sizing and error behaviour."""

import ctypes as C
import re
from pathlib import Path

import pytest

from deconlib_amd import _native as N
from oracle import deconv as O

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    txt = (ROOT / "include" / "deconhip.h").read_text()
    return sorted(set(re.findall(r"\b(dh_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    names = _declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(N.lib, n), n
        assert n in N.SYMBOLS, f"{n} has no ctypes signature"


def test_abi_version():
    assert N.fns["dh_abi_version"]() == N.ABI_VERSION


def _plan(sig, ker, fft=None, circular=0):
    nd = len(sig)
    h = C.c_void_p()
    rc = N.fns["dh_conv_plan_create"](nd, (C.c_int64 * nd)(*sig), (C.c_int64 * nd)(*ker),
                                      (C.c_int64 * nd)(*fft) if fft else None, circular, C.byref(h))
    return rc, h


@pytest.mark.parametrize("sig,ker", [((64, 256, 256), (64, 256, 256)), ((128, 512, 512), (128, 512, 512)),
                                     ((101, 589, 549), (101, 256, 256)), ((17, 33), (5, 9)), ((100,), (7,))])
def test_plan_canvas_matches_fast_padded_shape(sig, ker):
    rc, h = _plan(sig, ker)
    assert rc == 0
    info = N.PlanInfo()
    assert N.fns["dh_conv_plan_info"](h, C.byref(info)) == 0
    nd = len(sig)
    ref = O.fast_padded_shape(sig, ker)
    assert tuple(info.fft_shape)[3 - nd:] == ref
    canvas = tuple(info.canvas_shape)[3 - nd:]
    assert all(c >= r for c, r in zip(canvas, ref))
    assert canvas[-1] % 2 == 0
    P = tuple(info.canvas_shape)
    assert info.otf_bytes == P[0] * P[1] * (P[2] // 2 + 1) * 8
    N.fns["dh_conv_plan_destroy"](h)


def test_plan_errors_map_to_reference_exceptions():
    rc, _ = _plan((8, 8), (3, 3), fft=(9, 9))
    assert rc == N.DH_EINVAL
    with pytest.raises(ValueError):
        N.check(rc)
    rc, _ = _plan((8, 8), (4, 4), circular=1)
    assert rc == N.DH_EINVAL
    rc, _ = _plan((9, 9), (9, 9), circular=1)       # odd last axis: not supported
    assert rc == N.DH_EUNSUPPORTED
    with pytest.raises(NotImplementedError):
        N.check(rc)


def _null_args(sig):
    """Null pointers, zero sizes and zero scalars for a ctypes argument list."""
    out = []
    for a in sig:
        if a in (C.c_float, C.c_double):
            out.append(0.0)
        elif a in (C.c_int, C.c_int32, C.c_int64):
            out.append(0)
        else:
            out.append(None)
    return out


@pytest.mark.parametrize("prefix", ["dh_nlcg_", "dh_er_", "dh_jn_"])
def test_solver_entry_points_reject_bad_arguments(prefix):
    """Every solver entry point validates its arguments before any device work: null
    pointers and empty sizes return the invalid-argument status with a message (the
    Python layer raises ValueError), and the partials sizes are positive."""
    names = [n for n in N.SYMBOLS if n.startswith(prefix)]
    assert len(names) >= 9
    for name in names:
        if name.endswith("_partials"):
            assert N.fns[name]() > 0
            continue
        rc = N.fns[name](*_null_args(N.SYMBOLS[name][1:]))
        assert rc == N.DH_EINVAL, name
        with pytest.raises(ValueError, match="arguments"):
            N.check(rc, name)



def test_pd_dual_primal_rejects_overlapping_buffers():
    """The fused PDHG sweep re-reads x, xp, y2 and g around every tile while it writes:
    outputs that overlap an input (or each other) are refused before any device work.
    The pointers are never dereferenced on this path."""
    fn = N.fns["dh_pd_dual_primal"]
    nz, ny, nx = 4, 5, 6
    n = nz * ny * nx
    w = (C.c_float * 6)(1, 1, 1, 1, 1, 1)
    base = 1 << 20
    x, xp, g = base, base + 4 * n, base + 8 * n
    y2, y2n, xn = base + 12 * n, base + 40 * n, base + 70 * n
    sc, part = base + 80 * n, base + 90 * n

    def call(y2n_, xn_, x_=x, xp_=xp, y2_=y2, g_=g):
        return fn(y2n_, xn_, x_, xp_, y2_, g_, nz, ny, nx, sc, w, 1.0, 0, part, None)

    for bad in [call(y2, xn), call(y2n, x), call(y2n, xp), call(y2n, g), call(y2n, y2 + 4),
                call(y2n, y2n + 8), call(x, xn), call(y2n, xn, x_=xn)]:
        assert bad == N.DH_EINVAL
    assert fn(None, xn, x, xp, y2, g, nz, ny, nx, sc, w, 1.0, 0, part, None) == N.DH_EINVAL
    assert fn(y2n, xn, x, xp, y2, g, 0, ny, nx, sc, w, 1.0, 0, part, None) == N.DH_EINVAL
