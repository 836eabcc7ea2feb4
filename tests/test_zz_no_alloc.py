"""dh_conv_apply_static (the entry point that allocates nothing) against dh_conv_apply.

Through the C ABI: two calls in a row on one GPU, the second on a smaller
volume, each bit-identical to dh_conv_apply with a caller workspace on the same
plan (same kernels and launch shapes); a forward with the RL-ratio epilogue and
its float64 reduction, and an adjoint.  A plan whose workspace exceeds the static
scratch is rejected with DH_EWORKSPACE before anything is launched.
"""

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _apply_both(conv, x, adjoint, fuse_kind=None, obs=None):
    from deconlib_amd import _native as N

    p = conv._plan
    st = N.stream_ptr()
    outs, reds = [], []
    for static in (False, True):
        out = torch.empty_like(x)
        red = torch.zeros(1, dtype=torch.float64, device=x.device)
        fuse = None
        if fuse_kind == "ratio":
            fuse = N.FuseArgs(epilogue=N.EPI_RL_RATIO, epi_a=1.5, epi_b=1e-10, aux0=obs.data_ptr(),
                              reduce_out=red.data_ptr())
        fp = C.byref(fuse) if fuse is not None else None
        if static:
            rc = N.fns["dh_conv_apply_static"](p._h, int(adjoint), x.data_ptr(), out.data_ptr(),
                                                p.otf_buf.data_ptr(), fp, st)
        else:
            rc = N.fns["dh_conv_apply"](p._h, int(adjoint), x.data_ptr(), out.data_ptr(), p.otf_buf.data_ptr(),
                                         p.work.data_ptr(), int(p.info.work_bytes), fp, st)
        N.check(rc, "static" if static else "workspace")
        outs.append(out)
        reds.append(red)
    torch.cuda.synchronize()
    return outs, reds


def test_static_scratch_matches_workspace_calls(gpu):
    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd import _native as N

    rng = np.random.default_rng(90)
    # call 1: a 64x256x256 volume with a full-size kernel (canvas 128x512x512, ~101 MB workspace)
    shape = (64, 256, 256)
    print(f"no-alloc call 1: {shape}", flush=True)
    psf = rng.random(shape).astype(np.float32) + 0.01
    conv = LinearFFTConvolver(psf, shape)
    assert int(conv._plan.info.work_bytes) <= N.STATIC_SCRATCH_BYTES
    x = torch.from_numpy(rng.random(shape).astype(np.float32)).cuda()
    obs = torch.from_numpy(rng.poisson(20, size=shape).astype(np.float32)).cuda()
    (a, b), (ra, rb) = _apply_both(conv, x, False, "ratio", obs)
    assert torch.equal(a, b)
    assert ra.item() == rb.item() and ra.item() != 0.0
    (a, b), _ = _apply_both(conv, x, True)
    assert torch.equal(a, b)
    # call 2: a smaller volume on the same scratch (its earlier contents must not matter)
    shape2 = (12, 40, 56)
    print(f"no-alloc call 2: {shape2}", flush=True)
    psf2 = rng.random((5, 9, 9)).astype(np.float32) + 0.01
    conv2 = LinearFFTConvolver(psf2, shape2)
    x2 = torch.from_numpy(rng.random(shape2).astype(np.float32)).cuda()
    (a, b), _ = _apply_both(conv2, x2, False)
    assert torch.equal(a, b)
    (a, b), _ = _apply_both(conv2, x2, True)
    assert torch.equal(a, b)


def test_static_scratch_rejects_large_plans(gpu):
    from deconlib_amd import _native as N

    h = C.c_void_p()
    sig = (C.c_int64 * 3)(64, 512, 512)
    N.check(N.fns["dh_conv_plan_create"](3, sig, sig, None, 0, C.byref(h)))
    try:
        info = N.PlanInfo()
        N.check(N.fns["dh_conv_plan_info"](h, C.byref(info)))
        assert info.work_bytes > N.STATIC_SCRATCH_BYTES
        # rejected on the host before any pointer is used or any kernel launched
        rc = N.fns["dh_conv_apply_static"](h, 0, None, None, None, None, N.stream_ptr())
        assert rc == N.DH_EWORKSPACE
        assert "static scratch" in N.fns["dh_last_error"]().decode()
    finally:
        N.fns["dh_conv_plan_destroy"](h)
