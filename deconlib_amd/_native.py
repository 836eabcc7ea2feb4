"""ctypes binding of libdeconhip.so (the C ABI declared in include/deconhip.h).

The HIP library is the only compute path: if it is missing this module raises
at import time, so nothing can silently fall back to a CPU or PyTorch route.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch  # loads the HIP runtime first, so the library shares it

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libdeconhip.so"

if not _LIB_PATH.exists():
    raise ImportError(
        f"libdeconhip.so not found at {_LIB_PATH}; build it with "
        "`python -m deconlib_amd._build` (hipcc, gfx950)"
    )

lib = C.CDLL(str(_LIB_PATH))

ABI_VERSION = 11

DH_OK, DH_EINVAL, DH_EHIP, DH_EWORKSPACE, DH_EUNSUPPORTED = 0, -1, -2, -3, -4

PRO_NONE, PRO_ONES, PRO_SUB_FLOOR, PRO_SPREAD, PRO_PD_Y1 = 0, 1, 2, 3, 4
EPI_NONE, EPI_FLOOR, EPI_RL_RATIO, EPI_RL_UPDATE, EPI_ADD, EPI_SQDIFF, EPI_BLOCKSUM2 = 0, 1, 2, 3, 4, 5, 6
RED_SUM, RED_MAX, RED_SUMSQ = 0, 1, 2
REG_HESSIAN, REG_GRADIENT, REG_IDENTITY = 0, 1, 2
STATIC_SCRATCH_BYTES = 128 << 20  # DH_STATIC_SCRATCH_BYTES
TILE_PERSISTENT = 0x10000  # dh_conv_plan_set_tile flag: persistent grid with next-tile prefetch


class PlanInfo(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("signal_shape", C.c_int32 * 3),
        ("kernel_shape", C.c_int32 * 3),
        ("fft_shape", C.c_int32 * 3),
        ("canvas_shape", C.c_int32 * 3),
        ("otf_bytes", C.c_int64),
        ("work_bytes", C.c_int64),
        ("twiddle_bytes", C.c_int64),
        ("spectrum_block", C.c_int32),
    ]


class FuseArgs(C.Structure):
    _fields_ = [
        ("prologue", C.c_int32),
        ("pro_a", C.c_float),
        ("pro_b", C.c_float),
        ("epilogue", C.c_int32),
        ("epi_a", C.c_float),
        ("epi_b", C.c_float),
        ("aux0", C.c_void_p),
        ("aux1", C.c_void_p),
        ("reduce_out", C.c_void_p),
        ("spread", C.c_int32 * 3),
        ("pro_loss", C.c_int32),
        ("pro_aux", C.c_void_p * 3),
        ("pro_out", C.c_void_p),
        ("pro_scalars", C.c_void_p),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("plan", C.c_void_p),
        ("otf", C.c_void_p),
        ("work", C.c_void_p),
        ("work_bytes", C.c_int64),
        ("input", C.c_void_p),
        ("out", C.c_void_p),
        ("fuse", FuseArgs),
    ]


class RLItem(C.Structure):
    _fields_ = [
        ("plan", C.c_void_p),
        ("otf", C.c_void_p),
        ("work", C.c_void_p),
        ("work_bytes", C.c_int64),
        ("x", C.c_void_p),
        ("obs", C.c_void_p),
        ("sens", C.c_void_p),
        ("background", C.c_float),
        ("eps", C.c_float),
        ("reduce_out", C.c_void_p),
        ("spectrum_ready", C.c_int32),
        ("keep_spectrum", C.c_int32),
        ("data_box", C.c_int32 * 6),
    ]


class PdParams(C.Structure):
    _fields_ = [
        ("delta", C.c_double),
        ("eta", C.c_double),
        ("tol", C.c_double),
        ("min_iter", C.c_int32),
        ("patience", C.c_int32),
    ]


PD_NSLOTS = 16

# NLCG float64 scalar block slots (DH_NL_* in deconhip.h)
NL_LAM, NL_NEWTON_OK, NL_DPHI1, NL_DPHI2, NL_PHI0, NL_PHI_TRIAL, NL_ACCEPT = 0, 1, 2, 3, 4, 5, 6
NL_TRIAL_DATA, NL_TRIAL_LOSS, NL_MDATA, NL_LOSS, NL_IDIV, NL_SUMF, NL_RR = 7, 8, 9, 10, 11, 12, 13
NL_GAMMA, NL_LAM_ACCEPTED, NL_A0, NL_NSLOTS = 14, 15, 16, 32

# ER-Decon float64 scalar block slots and elementwise ops (DH_ER_* in deconhip.h)
ER_RS, ER_RS_THR, ER_DENOM, ER_BETA, ER_STOP, ER_GG, ER_DOT = 0, 1, 2, 3, 4, 5, 6
ER_DATA, ER_IDIV, ER_LOGQ, ER_Q, ER_NSLOTS = 7, 8, 9, 10, 16
ER_EW_DIV, ER_EW_SQUARE_SCALE, ER_EW_AXPY_SQUARE, ER_EW_NEG, ER_EW_MUL, ER_EW_SCALE = 0, 1, 2, 3, 4, 5
ER_EW_AFFINE, ER_EW_MAX_DIV, ER_EW_AXPY = 6, 7, 8

# JetNewton float64 scalar block slots (DH_JN_* in deconhip.h)
JN_EPS, JN_PG, JN_KKT2, JN_ACTIVE, JN_RZ, JN_TOL_K, JN_STOP, JN_DENOM, JN_BETA = 0, 1, 2, 3, 4, 5, 6, 7, 8
JN_NCG, JN_HITNEG, JN_USE_RHS, JN_ND, JN_DIRECTIONAL, JN_LOGU, JN_USUM, JN_R0, JN_NSLOTS = 9, 10, 11, 12, 13, 14, 15, 16, 24

_vp, _i64, _int, _f, _d = C.c_void_p, C.c_int64, C.c_int, C.c_float, C.c_double


def _sig(name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


# Every symbol the header declares; a missing one fails the import.
SYMBOLS = {
    "dh_abi_version": (_int,),
    "dh_last_error": (C.c_char_p,),
    "dh_conv_plan_create": (_int, _int, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64), _int, C.POINTER(_vp)),
    "dh_conv_plan_info": (_int, _vp, C.POINTER(PlanInfo)),
    "dh_conv_plan_bind": (_int, _vp, _vp, _vp),
    "dh_conv_plan_destroy": (_int, _vp),
    "dh_conv_plan_set_tile": (_int, _vp, _int, _int),
    "dh_conv_plan_get_tile": (_int, _vp, _int),
    "dh_otf_build": (_int, _vp, _vp, _int, _vp, _vp, _i64, C.POINTER(_d), _vp),
    "dh_otf_build_dev": (_int, _vp, _vp, _int, _vp, _vp, _i64, _vp, _vp),
    "dh_pd_init_scalars": (_int, _vp, _vp, _d, _vp),
    "dh_conv_apply": (_int, _vp, _int, _vp, _vp, _vp, _vp, _i64, C.POINTER(FuseArgs), _vp),
    "dh_reduce_partials": (_i64, _i64),
    "dh_reduce": (_int, _int, _vp, _i64, _vp, _vp, _vp),
    "dh_rl_support": (_int, _vp, _vp, _i64, _vp, _vp),
    "dh_clamp_min": (_int, _vp, _vp, _i64, _f, _vp),
    "dh_sub_floor": (_int, _vp, _vp, _i64, _f, _f, _vp),
    "dh_rl_ratio": (_int, _vp, _vp, _vp, _i64, _f, _f, _vp, _vp, _vp),
    "dh_rl_update": (_int, _vp, _vp, _vp, _i64, _f, _vp),
    "dh_add_scalar": (_int, _vp, _vp, _vp, _i64, _f, _vp, _vp, _vp),
    "dh_box_copy": (_int, _vp, C.POINTER(_i64), _vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64),
                    C.POINTER(_i64), _vp),
    "dh_block_sum": (_int, _vp, _vp, _int, _int, _int, _int, _int, _int, _vp),
    "dh_block_spread": (_int, _vp, _vp, _int, _int, _int, _int, _int, _int, _vp),
    "dh_banded_apply": (_int, _vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp, _int, _vp),
    "dh_pd_partials": (_i64, _int, _int, _int),
    "dh_pd_y1": (_int, _vp, _vp, _vp, _vp, _i64, _vp, _f, _int, _vp),
    "dh_pd_dual": (_int, _vp, _vp, _vp, _int, _int, _int, _vp, C.POINTER(_f), _f, _int, _vp),
    "dh_pd_primal": (_int, _vp, _vp, _vp, _vp, _int, _int, _int, _vp, C.POINTER(_f), _vp, _vp),
    "dh_pd_dual_primal": (_int, _vp, _vp, _vp, _vp, _vp, _vp, _int, _int, _int, _vp, C.POINTER(_f), _f, _int, _vp,
                          _vp),
    "dh_pd_hess_sums": (_int, _vp, _vp, _int, _int, _int, C.POINTER(_f), _vp, _vp, _int, _vp, _vp),
    "dh_pd_data_sums": (_int, _vp, _vp, _vp, _i64, _f, _int, _vp, _vp, _vp, _vp),
    "dh_pd_scalars": (_int, _vp, _vp, _int, C.POINTER(PdParams), _vp),
    "dh_hessian_forward": (_int, _vp, _vp, _int, _int, _int, C.POINTER(_f), _vp),
    "dh_hessian_adjoint": (_int, _vp, _vp, _int, _int, _int, C.POINTER(_f), _vp),
    "dh_reg_components": (_int, _int),
    "dh_pd_reg_dual": (_int, _int, _vp, _vp, _vp, _int, _int, _int, _vp, C.POINTER(_f), _f, _int, _vp),
    "dh_pd_reg_primal": (_int, _int, _vp, _vp, _vp, _vp, _int, _int, _int, _vp, C.POINTER(_f), _vp, _vp),
    "dh_pd_reg_sums": (_int, _int, _vp, _vp, _int, _int, _int, C.POINTER(_f), _vp, _vp, _int, _vp, _vp),
    "dh_reg_forward": (_int, _int, _vp, _vp, _int, _int, _int, C.POINTER(_f), _vp),
    "dh_reg_adjoint": (_int, _int, _vp, _vp, _int, _int, _int, C.POINTER(_f), _vp),
    "dh_poisson_idiv": (_int, _vp, _vp, _i64, _f, _vp, _vp, _vp),
    "dh_conv_apply_static": (_int, _vp, _int, _vp, _vp, _vp, C.POINTER(FuseArgs), _vp),
    "dh_batch_conv_apply": (_int, _int, C.POINTER(BatchItem), _int, _vp),
    "dh_rl_iteration": (_int, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _f, _f, _vp, _int, _int, _vp),
    "dh_batch_rl_iteration": (_int, _int, C.POINTER(RLItem), _vp),
    "dh_psf_plan_create": (_int, _int, _int, C.POINTER(C.c_void_p)),
    "dh_psf_plan_info": (_int, _vp, _int, _int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)),
    "dh_psf_plan_bind": (_int, _vp, _vp, _vp),
    "dh_psf_synth": (_int, _vp, _int, _vp, C.POINTER(_d), _vp, _vp, _int, _int, _vp, _vp, _i64, _vp, _vp),
    "dh_psf_plan_destroy": (_int, _vp),
    "dh_nlcg_partials": (_i64,),
    "dh_nlcg_init": (_int, _vp, _vp, _vp, _i64, _f, _vp),
    "dh_nlcg_model": (_int, _vp, _vp, _vp, _i64, _f, _vp, _vp, _vp),
    "dh_nlcg_begin_step": (_int, _vp, _vp, _vp, _vp, _i64, _vp, _int, _vp),
    "dh_nlcg_reg_dots": (_int, _vp, _vp, _vp, _i64, _vp, _vp, _vp),
    "dh_nlcg_newton": (_int, _vp, _vp, _vp, _vp, _i64, _f, _d, _d, _vp, _vp, _vp),
    "dh_nlcg_trial": (_int, _vp, _vp, _vp, _vp, _i64, _f, _d, _d, _vp, _vp, _vp, _vp, _vp),
    "dh_nlcg_update": (_int, _vp, _vp, _vp, _i64, _d, _vp, _vp, _vp),
    "dh_nlcg_gradient": (_int, _vp, _vp, _vp, _f, _vp, _i64, _int, _d, _vp, _vp, _vp),
    "dh_nlcg_direction": (_int, _vp, _vp, _i64, _vp, _vp),
    "dh_nlcg_quad": (_int, _vp, _vp, _vp, _vp, _vp, _i64, _d, _vp, _vp, _vp),
    "dh_nlcg_idiv": (_int, _vp, _vp, _i64, _d, _vp, _vp, _vp),
    "dh_er_partials": (_i64,),
    "dh_er_hess_weights": (_int, _vp, _int, _int, _int, C.POINTER(_f), _f, _f, _f, _f, _vp, _vp, _vp, _vp, _vp),
    "dh_er_weights_stacked": (_int, _vp, _int, _vp, _i64, _f, _f, _f, _f, _vp, _vp, _vp, _vp, _vp),
    "dh_er_hess_apply_w": (_int, _vp, _vp, _int, _int, _int, C.POINTER(_f), _vp, _vp),
    "dh_er_scale_stacked": (_int, _vp, _vp, _int, _i64, _vp, _vp),
    "dh_er_data": (_int, _vp, _vp, _i64, _int, _vp, _vp, _vp, _vp, _vp),
    "dh_er_gradient": (_int, _vp, _vp, _vp, _vp, _vp, _f, _i64, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp),
    "dh_er_cg_ap": (_int, _vp, _vp, _vp, _vp, _vp, _f, _vp, _vp, _i64, _vp, _vp, _vp),
    "dh_er_cg_update": (_int, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp),
    "dh_er_dot": (_int, _vp, _vp, _i64, _int, _vp, _vp, _vp),
    "dh_er_ew": (_int, _int, _vp, _vp, _f, _f, _vp, _vp, _i64, _vp),
    "dh_jn_partials": (_i64,),
    "dh_jn_penalty": (_int, _vp, _int, _int, _int, C.POINTER(_f), _f, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp),
    "dh_jn_penalty_stacked": (_int, _vp, _int, _i64, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp),
    "dh_jn_hvp_operand": (_int, _vp, _vp, _vp, _vp, _int, _int, _int, C.POINTER(_f), _vp, _vp),
    "dh_jn_hvp_operand_stacked": (_int, _vp, _vp, _vp, _vp, _int, _i64, _vp, _vp),
    "dh_jn_kkt": (_int, _vp, _vp, _i64, _d, _vp, _vp, _vp),
    "dh_jn_active": (_int, _vp, _vp, _i64, _f, _f, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp),
    "dh_jn_cg_ap": (_int, _vp, _f, _vp, _vp, _vp, _vp, _i64, _d, _vp, _vp, _vp),
    "dh_jn_cg_update": (_int, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp),
    "dh_jn_direction": (_int, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp),
    "dh_jn_trial": (_int, _vp, _vp, _f, _vp, _vp, _i64, _vp, _vp, _vp),
    "dh_timing_enable": (_int, _int),
    "dh_timing_select": (_int, _int),
    "dh_timing_read": (_int, C.POINTER(_d), C.POINTER(_i64), _int),
}

PASS_NAMES = ("x_fwd", "y_fwd", "z_conv", "y_inv", "x_inv")

fns = {name: _sig(name, sig[0], *sig[1:]) for name, sig in SYMBOLS.items()}

if fns["dh_abi_version"]() != ABI_VERSION:
    raise ImportError(
        f"libdeconhip.so ABI {fns['dh_abi_version']()} != expected {ABI_VERSION}; rebuild it"
    )


_VALUE_RETURNING = {"dh_reduce_partials", "dh_abi_version", "dh_pd_partials", "dh_conv_plan_get_tile",
                    "dh_nlcg_partials", "dh_er_partials", "dh_jn_partials", "dh_reg_components"}


class NativeError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    """Map a status code to the reference's exception types."""
    if rc == DH_OK:
        return
    msg = fns["dh_last_error"]().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == DH_EINVAL:
        raise ValueError(text)
    if rc == DH_EUNSUPPORTED:
        raise NotImplementedError(text)
    raise NativeError(f"{text} (status {rc})")


def call(name: str, *args) -> int:
    rc = fns[name](*args)
    if isinstance(rc, int) and name not in _VALUE_RETURNING:
        check(rc, name)
    return rc


def stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()
