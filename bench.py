"""Benchmark: Richardson–Lucy voxel-iterations/s at 64x512x512 float32 on one MI355X.

A step is one RL iteration (forward convolution, ratio, adjoint correlation,
multiplicative update: one dh_rl_iteration call, the ratio kept on chip) on a
64x512x512 volume with a full-size widefield PSF (FFT canvas 128x1024x1024).
Inputs are resident in HBM before the timed region starts.  Set-up runs untimed
iterations for 0.25 s on a separate state (the GPU leaves its idle clocks), then
the W warm-up iterations, then the K timed ones; HIP events bracket only the Z
pass in the timed region, and a further K iterations give the per-pass breakdown.

Prints one JSON line (see the keys below).  ``--gpus N`` with N > 1 exits:
every run here uses one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "voxel-iters/s (RL, PDHG-Hessian) at 64x512x512 fp32; serving p50/p99 latency"
SHAPE = (64, 512, 512)
LATERAL, AXIAL = 0.085, 0.1
N_BLOBS = 800          # the 64x256x256 RL workload's blob density (200 per volume) on a 64x512x512 volume
PEAK_HBM_GBS = 8000.0
# HBM bytes per k_z<128> launch at this workload, from rocprofv3 FETCH_SIZE (x2) and
# WRITE_SIZE (x1), both scaled by what they read for a device copy of known size
# (profiles/fft_passes_64x512x512_hbm_counters.md): (394075.8 * 2 + 262656.0) KiB.
Z_TRAFFIC_BYTES = int((394084.0 * 2 + 262656.0) * 1024)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def conv_pass_bytes(S, P, epi_operand_bytes=0):
    """Algorithmic HBM bytes of the five passes of one pruned convolution (complex64 = 8 B).

    ``epi_operand_bytes``: bytes per output voxel the fused epilogue of the X inverse
    pass reads besides the spectrum (RL: obs in the forward, x and the sensitivity in
    the adjoint, 6 B on average per launch)."""
    s0, s1, s2 = S
    p0, p1, p2 = P
    H = p2 // 2 + 1
    V = s0 * s1 * s2
    return {
        "x_fwd": 4 * V + 8 * s0 * s1 * H,
        "y_fwd": 8 * s0 * s1 * H + 8 * s0 * p1 * H,
        "z_conv": 8 * s0 * p1 * H + 8 * p0 * p1 * H + 8 * s0 * p1 * H,
        "y_inv": 8 * s0 * p1 * H + 8 * s0 * s1 * H,
        "x_inv": 8 * s0 * s1 * H + 4 * V + epi_operand_bytes * V,
    }


def rl_iteration_bytes(S, P, fused=True):
    """Algorithmic HBM bytes of one RL iteration.

    Two-application form: two convolutions plus the epilogue operands (forward: obs;
    adjoint: x, sensitivity).  Fused form (dh_rl_iteration): the ratio is never
    written or read back and the updated x is transformed where it is produced, so
    12 B per voxel less (ratio write + read, x re-read)."""
    core = sum(conv_pass_bytes(S, P).values())
    V = int(np.prod(S))
    return 2 * core + 12 * V - (12 * V if fused else 0)


def round_trip_bytes(S, P):
    """Mean algorithmic bytes of the two X round trips of a fused RL iteration:
    spectrum in and out (8 B x s0*s1*H each) plus obs (ratio) or x, sensitivity in
    and x out (update): 16*s0*s1*H + 8*V on average."""
    s0, s1, _ = S
    H = P[2] // 2 + 1
    return 16 * s0 * s1 * H + 8 * int(np.prod(S))


def make_inputs(torch):
    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS, blob_volume

    t0 = time.time()
    psf = widefield_psf(SHAPE, LATERAL, AXIAL, WIDEFIELD_OPTICS)
    obj = blob_volume(SHAPE, N_BLOBS, np.random.default_rng(1000)).astype(np.float32)
    blur = LinearFFTConvolver(psf, SHAPE)
    clean = np.maximum(blur.forward(torch.from_numpy(obj).cuda()).cpu().numpy(), 0.0)
    prng = np.random.default_rng(np.random.SeedSequence(1000).spawn(1)[0])
    obs = prng.poisson(clean).astype(np.float32)
    log(f"inputs ready in {time.time() - t0:.1f}s")
    return psf, obs, blur


def bench_pdhg(torch, N, obs, blur, steps, warmup):
    """PDHG-Hessian (alpha 1e-3, L1, Poisson) iterations on the same volume and operator."""
    from deconlib_amd.pdhg import _PDState, _nonneg

    x0 = _nonneg(torch.from_numpy(obs).cuda())
    st = _PDState(obs, blur, x0, 1e-3, warmup + steps, background=0.0, r=1.0, tol=0.0)
    for k in range(warmup):
        st.step(k, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        st.step(k, False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    vox = int(np.prod(obs.shape))
    return {
        "metric": "voxel-iters/s (PDHG-Hessian) at 64x512x512 fp32",
        "value": vox * steps / dt,
        "ms_per_step": dt * 1e3 / steps,
        "steps": steps,
        "warmup": warmup,
        "finite": bool(torch.isfinite(st.x).all().item()),
        "tau_last": float(st.hist[3 * (warmup + steps - 1)].item()),
    }


def bench_pdhg_binned(torch, steps=10, warmup=2, seed=1001, tune=False, tune_cache=None):
    """The binned PDHG configuration at full size: object 128x512x512 (2000 blobs + 10),
    widefield PSF of the same optics at 0.0425/0.05 um (synthesised on the GPU), data
    64x256x256 = FractionalAreaDownsample(2) o LinearFFTConvolver (canvas 256x1024x1024),
    Poisson at peak ~500 counts; PDHG-Hessian (alpha 1e-3, L1) iterations from the start
    solve_pdhg_mlx(bin_factors=2) uses.  Voxel-iterations count the 128x512x512 object."""
    from deconlib_amd import FractionalAreaDownsample, LinearFFTConvolver, compose
    from deconlib_amd import _native as N
    from deconlib_amd.pdhg import _PDState, _nonneg
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS, blob_volume

    hi = (128, 512, 512)
    psf = widefield_psf(hi, 0.0425, 0.05, WIDEFIELD_OPTICS, device=True)
    obj = blob_volume(hi, 2000, np.random.default_rng(seed)).astype(np.float32)
    conv = LinearFFTConvolver(psf, hi)
    tiles = {}
    if tune:  # launch shapes of this canvas's passes, as for the metric shape
        from deconlib_amd.autotune import autotune

        tiles = autotune(conv, verbose=True, cache_file=tune_cache)
    A = compose(FractionalAreaDownsample(scale=(2.0, 2.0, 2.0)), conv)
    clean = A.forward(torch.from_numpy(obj).cuda()).clamp_min(0.0)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    obs = torch.poisson(clean * (500.0 / float(clean.max())), generator=gen)
    st = _PDState(obs, A, _nonneg(A.adjoint(obs)), 1e-3, warmup + steps, background=0.0, r=1.0, tol=0.0)
    for k in range(warmup):
        st.step(k, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        st.step(k, False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    vox = int(np.prod(hi))
    return {
        "metric": "voxel-iters/s (binned PDHG-Hessian) on the 128x512x512 object grid, data 64x256x256",
        "value": vox * steps / dt,
        "ms_per_step": dt * 1e3 / steps,
        "steps": steps,
        "warmup": warmup,
        "canvas": list(conv.canvas_shape),
        "tiles": {N.PASS_NAMES[p]: t for p, t in tiles.items()},
        "data_shape": list(obs.shape),
        "finite": bool(torch.isfinite(st.x).all().item()),
    }


def bench_large_volume(torch, steps=3, warmup=1, seed=1004, tune=False):
    """The large single-volume configuration at full size: 256x2048x2048 float32 on a
    512x4096x4096 canvas, PDHG-Hessian (alpha 5e-4, L1) iterations.  Object: 1e5 blobs
    + 10 and 64 thin filaments (intensity 300); vectorial PSF of an isotropic emitter
    (pupil_to_vectorial_psf) of the volume's size with IndexMismatch depth 10 um,
    synthesised on the GPU; Poisson data.  Skipped below
    200 GB of free device memory (about 150 GB stay resident).  With ``tune`` the two Y
    passes' launch shapes are autotuned (the 4096-point Y passes' best shape differs
    from box to box; a whole-canvas tuning of every pass would take minutes)."""
    import ctypes as C

    from deconlib_amd import LinearFFTConvolver
    from deconlib_amd import _native as N
    from deconlib_amd.pdhg import _PDState, _nonneg
    from deconlib_amd.psf import IndexMismatch, widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS, blob_volume, filaments

    free, _ = torch.cuda.mem_get_info()
    if free < 200e9:
        return {"skipped": f"needs 200 GB free device memory, {free / 1e9:.0f} GB free"}
    shape = (256, 2048, 2048)
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    psf = widefield_psf(shape, LATERAL, AXIAL, WIDEFIELD_OPTICS, aberrations=[IndexMismatch(depth=10.0)],
                        device=True, dipole="isotropic")
    op = LinearFFTConvolver(psf, shape)
    del psf
    tiles = {}
    if tune:
        from deconlib_amd.autotune import autotune

        tiles = autotune(op, reps=2, candidates=(1, 2), passes=(1, 3))
        torch.cuda.empty_cache()
    obj = blob_volume(shape, 100000, rng)
    obj = torch.from_numpy(np.maximum(obj, filaments(shape, 64, rng)).astype(np.float32)).cuda()
    clean = op.forward(obj).clamp_min_(0.0)
    del obj
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    obs = torch.poisson(clean, generator=gen)
    del clean
    st = _PDState(obs, op, _nonneg(obs), 5e-4, warmup + steps, background=0.0, r=1.0, tol=0.0)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    for k in range(warmup):
        st.step(k, False)
    torch.cuda.synchronize()
    N.call("dh_timing_read", None, None, 1)
    N.call("dh_timing_enable", 1)
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        st.step(k, False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    N.call("dh_timing_enable", 0)
    ms = (C.c_double * 5)()
    cnt = (C.c_int64 * 5)()
    N.call("dh_timing_read", ms, cnt, 1)
    pb = conv_pass_bytes(shape, op.canvas_shape)
    passes = {name: {"avg_ms": ms[i] / cnt[i], "gbs": pb[name] / (ms[i] / cnt[i] * 1e-3) / 1e9,
                     "launches": int(cnt[i])} for i, name in enumerate(N.PASS_NAMES) if cnt[i]}
    vox = int(np.prod(shape))
    out = {
        "metric": "voxel-iters/s (PDHG-Hessian) at 256x2048x2048 fp32, canvas 512x4096x4096",
        "psf": "vectorial, isotropic dipole, NA 1.4, IndexMismatch depth 10 um (GPU synthesis)",
        "value": vox * steps / dt,
        "ms_per_step": dt * 1e3 / steps,
        "steps": steps,
        "warmup": warmup,
        "setup_s": setup_s,
        "canvas": list(op.canvas_shape),
        "tiles": {N.PASS_NAMES[p]: t for p, t in tiles.items()},
        "device_gb_resident": torch.cuda.memory_allocated() / 1e9,
        "passes": passes,
        "finite": bool(torch.isfinite(st.x).all().item()),
    }
    del st, op, obs
    torch.cuda.empty_cache()
    return out


def nlcg_iteration_bytes(S, P):
    """Algorithmic HBM bytes of one NLCG iteration: three plain convolutions (K(s d),
    K(d d), K^T w) plus the elementwise passes, in float32 words per voxel: line-search
    set-up 4 (s, d in; s d, d d out), three Newton passes 4 each (m, Ksd, Kdd, g), trial
    6 (those four in; m_new, w out), update 5 (s, d, f in; s, f out), gradient 3 (s, KTw
    in; r out), direction 3 (r, d in; d out): 33 words."""
    return 3 * sum(conv_pass_bytes(S, P).values()) + 33 * 4 * int(np.prod(S))


def bench_nlcg(torch, obs, blur, steps, warmup):
    """NLCG iterations (3 Newton refinements, no regulariser) on the same volume and
    operator: each is the solver's loop body, line search, one host read of the accept
    flag, accepted step, gradient and direction."""
    from deconlib_amd import _native as N
    from deconlib_amd.nlcg import _NLCGState

    st = _NLCGState(obs, blur, 0.0, None, 0.0, None, 3)
    fallbacks = 0

    def iteration():
        nonlocal fallbacks
        st.line_search()
        if st.S.read()[N.NL_ACCEPT] == 0.0:
            fallbacks += 1
            st.fall_back_to_steepest_descent()
            st.line_search()
            if st.S.read()[N.NL_ACCEPT] == 0.0:
                return False
        st.accept(restart=False)
        return True

    for _ in range(warmup):
        iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    for _ in range(steps):
        if not iteration():
            break
        done += 1
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    vox = int(np.prod(obs.shape))
    it_bytes = nlcg_iteration_bytes(obs.shape, blur.canvas_shape)
    v = st.S.read()
    return {
        "metric": "voxel-iters/s (NLCG, 3 Newton refinements) at 64x512x512 fp32",
        "value": vox * done / dt,
        "ms_per_step": dt * 1e3 / max(done, 1),
        "steps": done,
        "warmup": warmup,
        "steepest_descent_fallbacks": fallbacks,
        "algorithmic_bytes_per_iteration": it_bytes,
        "effective_gbs": it_bytes * done / dt / 1e9,
        "mean_idiv_last": v[N.NL_LOSS] / vox,
        "finite": bool(torch.isfinite(st.f).all().item()),
    }


def nlcg_vs_rl(torch, obs, blur, rl_iters=30, max_nlcg=30):
    """Wall time for NLCG to reach the mean I-divergence RL reaches after `rl_iters`
    iterations on the same volume and operator (both from their default starts;
    RL's loss summed in every iteration's fused forward pass, NLCG's read with its
    accept flag)."""
    from deconlib_amd import _native as N
    from deconlib_amd.nlcg import _NLCGState
    from deconlib_amd.rl import _RLState

    vox = int(np.prod(obs.shape))
    st = _RLState(obs, blur, background=0.0, num_iter=rl_iters + 1, eval_interval=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(rl_iters + 1):
        st.step(k)  # step k sums the loss of iterate k - 1
    torch.cuda.synchronize()
    rl_s = (time.perf_counter() - t0) * rl_iters / (rl_iters + 1)
    target = float(st.loss_slots[st.slot_of[rl_iters - 1]].item()) / vox
    del st
    ns = _NLCGState(obs, blur, 0.0, None, 0.0, None, 3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reached = None
    for k in range(max_nlcg):
        ns.line_search()
        if ns.S.read()[N.NL_ACCEPT] == 0.0:
            break
        ns.accept(restart=False)
        if ns.S.read()[N.NL_LOSS] / vox <= target:
            reached = k + 1
            break
    torch.cuda.synchronize()
    nl_s = time.perf_counter() - t0
    return {"target_mean_idiv": target, "rl_iterations": rl_iters, "rl_seconds": rl_s,
            "nlcg_iterations": reached, "nlcg_seconds": nl_s if reached else None,
            "speedup": (rl_s / nl_s) if reached else None}


def _time_cg_steps(torch, step, read, steps, warmup):
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        read()  # the solvers read the CG flags once per step
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_erdecon(torch, obs, blur, steps=10, warmup=2):
    """ER-Decon (Hessian3D, Gaussian data term) on the same volume and operator: the
    inner Gauss-Newton CG step (two convolutions, the weighted Hessian stencil and its
    adjoint, the CG updates; the relative-residual stop disabled so every step runs in
    full), and one whole outer iteration through erdecon_with_operator."""
    from deconlib_amd import _native as N
    from deconlib_amd import erdecon_with_operator
    from deconlib_amd.erdecon import _ERState
    from deconlib_amd.linops import Hessian3D

    f = torch.from_numpy(obs).cuda() / float(obs.max())
    st = _ERState(f, blur, Hessian3D(), 0.0, 0.05, 1e-2, "gaussian", 0.0, 0.0, None)
    st.gradient(0.1)
    st.S.sc[N.ER_RS_THR] = 0.0

    def step():
        kt, htu = st.hvp_parts()
        N.call("dh_er_cg_ap", st.s.data_ptr(), kt.data_ptr(), htu.data_ptr(), st.w.data_ptr(), st.sv.data_ptr(),
               0.0, st.p.data_ptr(), st.ap.data_ptr(), st.n, st.S.sc.data_ptr(), st.S.part.data_ptr(),
               N.stream_ptr())
        N.call("dh_er_cg_update", st.x.data_ptr(), st.r.data_ptr(), st.p.data_ptr(), st.ap.data_ptr(),
               st.s.data_ptr(), st.sv.data_ptr(), st.n, st.S.sc.data_ptr(), st.S.part.data_ptr(), N.stream_ptr())

    per_step = _time_cg_steps(torch, step, st.S.read, steps, warmup)
    del st
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = erdecon_with_operator(obs, blur, num_iter=1, newton_tol=0.0)
    torch.cuda.synchronize()
    outer = time.perf_counter() - t0
    vox = int(np.prod(obs.shape))
    return {
        "cg_step_ms": per_step * 1e3,
        "voxel_cg_steps_per_s": vox / per_step,
        "one_outer_iteration_s": outer,
        "note": "one outer iteration = set-up, gradient, up to 25 CG steps, Armijo search, result",
        "finite": bool(torch.isfinite(res.restored).all().item()),
    }


def bench_jetnewton(torch, obs, blur, steps=10, warmup=2):
    """JetNewton (AnisotropicHessian3D, Poisson) on the same volume and operator: the
    inner reduced CG step with the exact Hessian product (forcing stop disabled), and one
    whole outer iteration through jetnewton_with_operator."""
    from deconlib_amd import AnisotropicHessian3D
    from deconlib_amd import _native as N
    from deconlib_amd import jetnewton_with_operator
    from deconlib_amd.jetnewton import _JNState

    s0 = float(np.sqrt(max(float(obs.mean()), 1.0)))
    H = AnisotropicHessian3D((0.85, 1.0, 1.0))
    f = torch.from_numpy(obs).cuda()
    st = _JNState(f, blur, H, s0, 10.0, 1.0, 1.0, "poisson", None)
    st.gradient()
    st.active_set(1e-2, 1e-3, 1e-6)
    st.B.jn[N.JN_TOL_K] = 0.0

    def step():
        st.hvp_apply()
        N.call("dh_jn_cg_update", st.xcg.data_ptr(), st.rcg.data_ptr(), st.pcg.data_ptr(), st.ap.data_ptr(), st.n,
               st.B.jn.data_ptr(), st.B.part.data_ptr(), N.stream_ptr())

    per_step = _time_cg_steps(torch, step, st.B.read, steps, warmup)
    del st
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = jetnewton_with_operator(obs, blur, H, s0, background=10.0, num_iter=1, newton_tol=0.0)
    torch.cuda.synchronize()
    outer = time.perf_counter() - t0
    vox = int(np.prod(obs.shape))
    return {
        "cg_step_ms": per_step * 1e3,
        "voxel_cg_steps_per_s": vox / per_step,
        "one_outer_iteration_s": outer,
        "cg_steps_in_outer_iteration": res.cg_iterations_history,
        "finite": bool(torch.isfinite(res.restored).all().item()),
    }


def stream_ceiling(torch, mib=1024, reps=10):
    """Device streaming rates this GPU reaches with library kernels (no FFT structure):
    a device copy (1 read + 1 write) and an elementwise add (2 reads + 1 write), both
    timed with HIP events over `reps` launches; context for the roofline fraction."""
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b, c = torch.empty_like(a), torch.empty_like(a)
    out = {}
    for name, fn, words in (("copy", lambda: b.copy_(a), 2), ("add", lambda: torch.add(a, b, out=c), 3)):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name + "_gbs"] = words * n * 4 / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    del a, b, c
    return out


def bench_host_to_host(torch, psf, obs, iters=50):
    """One whole RL solve from host arrays to a host array: operator construction (OTF
    build), PCIe upload, `iters` iterations, PCIe download.  Reported beside `value`
    (which times device-resident iterations only)."""
    from deconlib_amd import LinearFFTConvolver, richardson_lucy_with_operator

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op = LinearFFTConvolver(psf, obs.shape)
    res = richardson_lucy_with_operator(obs, op, num_iter=iters)
    out = res.restored.cpu().numpy()
    dt = time.perf_counter() - t0
    x = torch.from_numpy(obs).pin_memory()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    d = x.cuda(non_blocking=True)
    torch.cuda.synchronize()
    up = time.perf_counter() - t1
    t1 = time.perf_counter()
    d.cpu()
    down = time.perf_counter() - t1
    vox = int(np.prod(obs.shape))
    return {
        "iterations": iters,
        "seconds": dt,
        "voxel_iters_per_s": vox * iters / dt,
        "h2d_GBps_pinned": obs.nbytes / up / 1e9,
        "d2h_GBps": obs.nbytes / down / 1e9,
        "finite": bool(np.isfinite(out).all()),
    }


def bench_batched_small(torch, n_vol=32, shape=(16, 64, 64), iters=50, seed=1005):
    """Many small RL volumes (canvas 32x128x128): the plain one-job-at-a-time loop
    against the batched fused iteration, eager and with the steady-state iteration
    replayed from a HIP graph (launch overhead dominates at this size)."""
    from deconlib_amd import LinearFFTConvolver, richardson_lucy_with_operator
    from deconlib_amd.batch import RLJob, batched_richardson_lucy
    from deconlib_amd.synth import blob_volume

    rng = np.random.default_rng(seed)
    z = np.minimum(np.arange(shape[0]), shape[0] - np.arange(shape[0]))[:, None, None]
    y = np.minimum(np.arange(shape[1]), shape[1] - np.arange(shape[1]))[None, :, None]
    x = np.minimum(np.arange(shape[2]), shape[2] - np.arange(shape[2]))[None, None, :]
    psf = np.exp(-0.5 * ((z / 2.0) ** 2 + (y / 1.5) ** 2 + (x / 1.5) ** 2)).astype(np.float32)
    vols = []
    for _ in range(n_vol):
        obj = blob_volume(shape, 6, rng).astype(np.float32)
        clean = LinearFFTConvolver(psf, shape).forward(obj).cpu().numpy()
        vols.append(rng.poisson(np.maximum(clean, 0.0)).astype(np.float32))
    vox_it = n_vol * int(np.prod(shape)) * iters

    def timed(fn):
        fn()  # warm-up (plans, occupancy queries, graph capture path)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return vox_it / (time.perf_counter() - t0)

    ops = [LinearFFTConvolver(psf, shape) for _ in vols]
    plain = timed(lambda: [richardson_lucy_with_operator(v, o, num_iter=iters) for v, o in zip(vols, ops)])
    eager = timed(lambda: batched_richardson_lucy([RLJob(v, o, num_iter=iters) for v, o in zip(vols, ops)]))
    graph = timed(lambda: batched_richardson_lucy([RLJob(v, o, num_iter=iters) for v, o in zip(vols, ops)],
                                                  graphs=True))

    # steady state: per-iteration time of a running batch once the graph exists (its
    # capture is a one-time cost that the end-to-end figures above include), the same
    # iterations eager; every 10th iteration evaluates the loss and runs eagerly in both
    from deconlib_amd.batch import BatchRunner

    def steady(graphs: bool, n: int = 40) -> float:
        jobs = [RLJob(v, o, num_iter=n + 10) for v, o in zip(vols, ops)]
        runner = BatchRunner(len(jobs), graphs=graphs)
        for _ in range(5):
            runner.step(jobs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            runner.step(jobs)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / n

    steady_eager, steady_graph = steady(False), steady(True)
    return {
        "workload": f"{n_vol} volumes {shape[0]}x{shape[1]}x{shape[2]}, RL {iters} it (canvas 32x128x128)",
        "plain_loop_voxel_iters_per_s": plain,
        "batched_voxel_iters_per_s": eager,
        "batched_graph_voxel_iters_per_s": graph,
        "graph_speedup_vs_plain_loop": graph / plain,
        "steady_iteration_ms": {"eager": steady_eager, "graph": steady_graph,
                                "note": "one batched iteration of all volumes after 5 warm-up iterations "
                                        "(graph captured), mean of 40"},
    }


def bench_batched_config(torch, n_vol=64, iters=100, seed=1002, tune=False):
    """The batched configuration at full size: 64 volumes with z in {16,32,48,64} and
    y = x in {128,192,256,384,512}, each with its own aberrated PSF (IndexMismatch depth
    U[0,20] um, spherical U[-0.5,0.5] rad, synthesised on the GPU) and blob object
    (200 blobs per 64x256x256 voxels + 10), Poisson observations resident on the
    device, RL 100 iterations each.  The plain loop solves the volumes one after
    another; the batched run advances all of them together, each pass one grouped
    launch per transform length and 16 volumes.  Both timings include every job's
    set-up (sensitivity) and result; the batched results must equal the plain ones
    bit for bit.  With ``tune`` the launch shapes come from the autotuner, one tuning
    per shape class (reported as ``autotune_s``, outside both timings)."""
    from deconlib_amd import LinearFFTConvolver, richardson_lucy_with_operator
    from deconlib_amd.batch import RLJob, batched_richardson_lucy
    from deconlib_amd.synth import aberrated_psf, blob_volume

    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    ops, obs = [], []
    for _ in range(n_vol):
        z = int(rng.choice([16, 32, 48, 64]))
        xy = int(rng.choice([128, 192, 256, 384, 512]))
        shape = (z, xy, xy)
        psf = aberrated_psf(shape, rng, device=True)
        obj = blob_volume(shape, max(1, round(200 * np.prod(shape) / 4194304)), rng)
        op = LinearFFTConvolver(psf, shape)
        clean = op.forward(obj.astype(np.float32)).clamp_min(0.0)
        gen = torch.Generator(device=clean.device)
        gen.manual_seed(int(rng.integers(1 << 30)))
        obs.append(torch.poisson(clean, generator=gen))
        ops.append(op)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    vox = sum(int(o.numel()) for o in obs)
    tune_s = None
    if tune:  # launch shapes per shape class (plans of one shape share the tuned result)
        from deconlib_amd.autotune import autotune

        t0 = time.perf_counter()
        for op in ops:
            autotune(op, reps=5)
        tune_s = time.perf_counter() - t0

    def plain(n):
        return [richardson_lucy_with_operator(o, op, num_iter=n) for o, op in zip(obs, ops)]

    def batched(n):
        return batched_richardson_lucy([RLJob(o, op, num_iter=n) for o, op in zip(obs, ops)])

    def timed(fn):
        torch.cuda.synchronize()
        s = time.perf_counter()
        out = fn(iters)
        torch.cuda.synchronize()
        return out, time.perf_counter() - s

    plain(2), batched(2)  # first launches of every plan (occupancy queries, kernel loads)
    rp, tp = timed(plain)
    rb, tb = timed(batched)
    same = all(torch.equal(a.restored, b.restored) and a.loss_history == b.loss_history for a, b in zip(rp, rb))
    finite = all(bool(torch.isfinite(b.restored).all()) for b in rb)
    del rp, rb
    return {
        "workload": f"{n_vol} volumes z in {{16,32,48,64}}, y=x in {{128..512}}, per-volume aberrated PSFs, "
                    f"RL {iters} it (seed {seed})",
        "voxels": vox,
        "shapes": sorted({tuple(int(s) for s in o.shape) for o in obs}),
        "setup_s": setup_s,
        "autotune_s": tune_s,
        "plain_loop_s": tp,
        "plain_loop_voxel_iters_per_s": vox * iters / tp,
        "batched_s": tb,
        "batched_voxel_iters_per_s": vox * iters / tb,
        "speedup_vs_plain_loop": tp / tb,
        "bit_identical_to_plain": same,
        "finite": finite,
    }


def bench_tiled(torch, shape=(48, 1024, 1024), psf_shape=(32, 64, 64), tile_size=256, iters=20, seed=1006):
    """process_tiles with RL on a volume larger than one tile: the reference's
    per-tile loop through the composed operator (crop . conv) against the fused
    tile path (padded-domain fused iteration; small tiles in lockstep groups)."""
    from deconlib_amd import plan_tiles, process_tiles, richardson_lucy_solver
    from deconlib_amd.synth import blob_volume

    rng = np.random.default_rng(seed)
    ax = [np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64) for n in psf_shape]
    psf = np.exp(-0.5 * ((ax[0][:, None, None] / 3.0) ** 2 + (ax[1][None, :, None] / 2.0) ** 2
                         + (ax[2][None, None, :] / 2.0) ** 2)).astype(np.float32)
    data = rng.poisson(np.maximum(blob_volume(shape, 400, rng), 0.0) * 0.1 + 5.0).astype(np.float32)
    guard = max(psf_shape[-2:]) // 2
    plan = plan_tiles(shape, 1.0, guard, tile_size)
    vox_it = int(np.prod(shape)) * iters
    solve = richardson_lucy_solver(num_iter=iters)

    def loop(t, m):
        # the reference's execution: one tile after another through the composed operator
        return solve.solve_composed(t, m)

    res = {}
    for name, fn in (("per_tile_loop", loop), ("fused_tiles", solve)):
        process_tiles(data[:, :tile_size + 2 * guard, :tile_size + 2 * guard], psf, 1.0, fn, tile_size=tile_size)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = process_tiles(data, psf, 1.0, fn, tile_size=tile_size)
        torch.cuda.synchronize()
        res[name] = {"seconds": time.perf_counter() - t0, "voxel_iters_per_s": vox_it / (time.perf_counter() - t0)}
        res[name]["finite"] = bool(np.isfinite(out).all())
    return {
        "workload": (f"RL {iters} it on {shape[0]}x{shape[1]}x{shape[2]} float32, Gaussian PSF "
                     f"{psf_shape[0]}x{psf_shape[1]}x{psf_shape[2]}, {len(plan.tiles)} tiles of "
                     f"{plan.tile_shape[0]}x{plan.tile_shape[1]}x{plan.tile_shape[2]} (guard {guard})"),
        **res,
        "speedup": res["per_tile_loop"]["seconds"] / res["fused_tiles"]["seconds"],
        "note": ("numpy volume in and out; the per-tile loop moves every tile and result over PCIe, the fused "
                 "path uploads the volume once, cuts and stitches tiles on the device and downloads once"),
    }


def bench_psf_synthesis(torch, shape=SHAPE):
    """The RL workload's widefield PSF (pupil_to_psf): GPU synthesis (dh_psf_synth)
    against the host restatement (numpy complex128, plane by plane)."""
    from deconlib_amd.psf import widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS

    widefield_psf((2,) + tuple(shape[1:]), LATERAL, AXIAL, WIDEFIELD_OPTICS, device=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev = widefield_psf(shape, LATERAL, AXIAL, WIDEFIELD_OPTICS, device=True)
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    host = widefield_psf(shape, LATERAL, AXIAL, WIDEFIELD_OPTICS)
    t_host = time.perf_counter() - t0
    d = dev.cpu().numpy().astype(np.float64)
    err = float(np.linalg.norm(d - host) / np.linalg.norm(host))
    # the vectorial PSF (isotropic dipole: six field channels), with index mismatch
    from deconlib_amd.psf import (IndexMismatch, apply_aberrations, fft_coords, make_geometry, make_pupil,
                                  pupil_to_vectorial_psf, pupil_to_vectorial_psf_device)

    g = make_geometry(tuple(shape[1:]), LATERAL, WIDEFIELD_OPTICS)
    pupil = apply_aberrations(make_pupil(g), g, WIDEFIELD_OPTICS, [IndexMismatch(10.0)])
    z = fft_coords(shape[0], AXIAL)
    pupil_to_vectorial_psf_device(pupil, g, WIDEFIELD_OPTICS, z[:2])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vdev = pupil_to_vectorial_psf_device(pupil, g, WIDEFIELD_OPTICS, z)
    torch.cuda.synchronize()
    tv_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    vhost = pupil_to_vectorial_psf(pupil, g, WIDEFIELD_OPTICS, z)
    tv_host = time.perf_counter() - t0
    verr = float(np.linalg.norm(vdev.cpu().numpy().astype(np.float64) - vhost) / np.linalg.norm(vhost))
    return {"shape": list(shape),
            "scalar": {"gpu_seconds": t_dev, "host_seconds": t_host, "speedup": t_host / t_dev, "rel_l2_vs_host": err},
            "vectorial_isotropic": {"gpu_seconds": tv_dev, "host_seconds": tv_host, "speedup": tv_host / tv_dev,
                                    "rel_l2_vs_host": verr},
            "note": "end to end: pupil geometry and aberrations on the host, uploads, synthesis, normalisation"}


def pdhg_iteration_bytes(S, P):
    """Algorithmic HBM bytes of one PDHG-Hessian iteration on a plain convolution: two
    convolutions plus the dual/primal kernels in float32 words per voxel: y1 prox 5
    (A x, A x_prev, data, y1 in; y1 out), Hessian dual step 14 (x, x_prev, y2[6] in;
    y2[6] out), primal step 8 (x, g, y2[6] in; x_new out), data sums 2, Hessian sums 2."""
    return 2 * sum(conv_pass_bytes(S, P).values()) + 31 * 4 * int(np.prod(S))


def bench_serving(torch, n_jobs, rate, seed=1003):
    """The serving configuration: Poisson arrivals at ``rate`` jobs/s of 70 % RL (100 it) /
    30 % PDHG-Hessian (200 it, alpha log-uniform 1e-4..1e-2) jobs, z log-uniform 32..128 and
    y = x log-uniform 128..1024 (independent, any integer), per-job aberrated PSFs, blob
    objects and Poisson observations made on the GPU before the run.  The server admits
    by HBM budget (free device memory at the start less 16 GiB) and batches iterations
    across live jobs; the plain loop runs the same jobs one at a time, synchronously.
    Reports p50/p99 latency, aggregate voxel-iterations/s, the plain loop's rate and the
    fraction of the HBM roofline (the jobs' algorithmic bytes / makespan / 8 TB/s)."""
    from deconlib_amd.serving import DeconvServer, latency_stats, plan_info, run_plain
    from deconlib_amd.synth import serving_requests_device

    t0 = time.time()
    reqs = serving_requests_device(n_jobs, rate, seed)
    torch.cuda.synchronize()
    log(f"serving: {n_jobs} requests synthesised on the GPU in {time.time() - t0:.1f}s")
    warm = serving_requests_device(3, 1000.0, seed + 1, z_range=(32, 40), xy_range=(128, 160), rl_iters=3,
                                   pdhg_iters=3, kinds=["rl", "pdhg", "rl"])
    DeconvServer().run(warm)
    run_plain(warm)
    free, _ = torch.cuda.mem_get_info()
    srv = DeconvServer(hbm_budget_bytes=int(free - (16 << 30)))
    t0 = time.perf_counter()
    done = srv.run(reqs)
    makespan = time.perf_counter() - t0
    stats = latency_stats(done, makespan)
    log(f"serving: {json.dumps(stats)}")
    plain_s = run_plain(reqs)
    vox_it = sum(int(np.prod(r.observed.shape)) * r.num_iter for r in reqs)
    nbytes = 0
    for r in reqs:
        S = tuple(int(v) for v in r.observed.shape)
        P = tuple(int(v) for v in plan_info(S).canvas_shape)
        nbytes += r.num_iter * (rl_iteration_bytes(S, P, fused=True) if r.kind == "rl" else pdhg_iteration_bytes(S, P))
    sizes = [int(np.prod(r.observed.shape)) for r in reqs]
    stats.update({
        "workload": f"{n_jobs} jobs, Poisson arrivals {rate}/s, 70% RL 100 it / 30% PDHG-Hessian 200 it "
                    "(alpha log-uniform 1e-4..1e-2), z log-uniform 32..128, y=x log-uniform 128..1024, "
                    "per-job aberrated PSFs, inputs on the GPU",
        "jobs_rl": sum(r.kind == "rl" for r in reqs),
        "voxels_min_median_max": [min(sizes), int(np.median(sizes)), max(sizes)],
        "arrival_span_s": max(r.arrival for r in reqs),
        "ticks": srv.ticks,
        "plain_loop_s": plain_s,
        "plain_loop_voxel_iters_per_s": vox_it / plain_s,
        "speedup_vs_plain_loop": plain_s / makespan,
        "algorithmic_bytes": nbytes,
        "hbm_frac": nbytes / makespan / (PEAK_HBM_GBS * 1e9),
        "plain_loop_hbm_frac": nbytes / plain_s / (PEAK_HBM_GBS * 1e9),
    })
    del reqs
    torch.cuda.empty_cache()
    return stats


def cpu_baseline(psf, obs, iters):
    """The oracle's RL iteration (fp32 NumPy/SciPy restatement) timed on the host cores."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import deconv as O

    cores = len(os.sched_getaffinity(0))
    A = O.LinearFFTConvolver(psf, obs.shape)
    t = np.float32
    x = np.maximum(A.adjoint(np.maximum(obs, t(1e-6))), t(1e-6))
    s = A.adjoint(np.ones_like(obs))
    act = s > t(max(float(s.max()) * 1e-6, 1e-10))
    s = np.where(act, s, t(1.0))
    eps = t(1e-10)

    def step(x):
        r = obs / np.maximum(A.forward(x), eps)
        return np.maximum(x * np.where(act, A.adjoint(r) / s, t(0.0)), eps)

    x = step(x)  # warm-up iteration
    t0 = time.perf_counter()
    for _ in range(iters):
        x = step(x)
    dt = time.perf_counter() - t0
    return {
        "value": float(np.prod(obs.shape) * iters / dt),
        "unit": "voxel-iters/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{iters} RL iterations of the fp32 NumPy/SciPy restatement at 64x512x512 "
                  f"(scipy.fft workers={cores}), after 1 warm-up iteration; {dt:.1f}s",
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _timed_oracle_iterations(run, vox, n_iter):
    """Runs an oracle solve with a per-iteration callback and times iterations 2..n
    (the first is the warm-up); returns (voxel-iterations/s, seconds timed)."""
    stamps = []
    run(lambda k, x: stamps.append(time.perf_counter()))
    dt = stamps[-1] - stamps[0]
    return vox * (len(stamps) - 1) / dt, dt


def cpu_baseline_configs(seed=2000):
    """The oracle (fp32 NumPy/SciPy restatement) timed on the host cores for each workload
    configuration, on bounded samples (about 10-30 s of CPU work each, inputs made on the
    host with the configurations' recipes):

    * RL 64x256x256 (canvas 128x512x512): 20 iterations after 1 warm-up;
    * binned PDHG-Hessian, object 128x512x512 -> data 64x256x256 (canvas 256x1024x1024):
      3 iterations after 1 warm-up, voxel-iterations on the object grid;
    * batched RL: 6 volumes drawn from the batched configuration's shapes with aberrated
      PSFs, 6 iterations each after 1 warm-up, one after another (the reference's loop);
    * the large-volume solver (PDHG-Hessian alpha 5e-4) on its 128x512x512 crop with the
      vectorial PSF: 3 iterations after 1 warm-up."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import deconv as O

    from deconlib_amd.psf import IndexMismatch, widefield_psf
    from deconlib_amd.synth import WIDEFIELD_OPTICS, aberrated_psf, blob_volume

    rng = np.random.default_rng(seed)
    out = {}

    def poisson(clean):
        return rng.poisson(np.maximum(clean, 0.0)).astype(np.float32)

    # RL, 64x256x256
    shape = (64, 256, 256)
    psf = widefield_psf(shape, LATERAL, AXIAL, WIDEFIELD_OPTICS)
    A = O.LinearFFTConvolver(psf, shape)
    obs = poisson(A.forward(blob_volume(shape, 200, rng).astype(np.float32)))
    rate, dt = _timed_oracle_iterations(
        lambda cb: O.richardson_lucy(obs, A, num_iter=21, callback=cb), int(np.prod(shape)), 21)
    out["rl_64x256x256"] = {"value": rate, "unit": "voxel-iters/s", "sample": f"20 RL iterations, {dt:.1f}s"}
    log(f"cpu baseline RL 64x256x256: {rate:.3e} voxel-iters/s")

    # binned PDHG-Hessian
    hi = (128, 512, 512)
    psf = widefield_psf(hi, 0.0425, 0.05, WIDEFIELD_OPTICS)
    Ab = O.compose(O.FractionalAreaDownsample(scale=(2.0, 2.0, 2.0)), O.LinearFFTConvolver(psf, hi))
    clean = Ab.forward(blob_volume(hi, 2000, rng).astype(np.float32))
    obs = poisson(clean * (500.0 / float(clean.max())))
    rate, dt = _timed_oracle_iterations(
        lambda cb: O.solve_pdhg(obs, psf, alpha=1e-3, num_iter=4, bin_factors=2, tol=0.0, callback=cb),
        int(np.prod(hi)), 4)
    out["pdhg_binned_128x512x512"] = {"value": rate, "unit": "voxel-iters/s",
                                      "sample": f"3 PDHG-Hessian iterations on the object grid, {dt:.1f}s"}
    log(f"cpu baseline binned PDHG: {rate:.3e} voxel-iters/s")

    # batched RL, the reference's one-volume-at-a-time loop
    total_vox_it, total_dt = 0, 0.0
    shapes = []
    for _ in range(6):
        z = int(rng.choice([16, 32, 48, 64]))
        xy = int(rng.choice([128, 192, 256, 384, 512]))
        sh = (z, xy, xy)
        shapes.append(sh)
        p = aberrated_psf(sh, rng)
        Av = O.LinearFFTConvolver(p, sh)
        ob = poisson(Av.forward(blob_volume(sh, max(1, round(200 * np.prod(sh) / 4194304)), rng).astype(np.float32)))
        r, d = _timed_oracle_iterations(lambda cb: O.richardson_lucy(ob, Av, num_iter=7, callback=cb),
                                        int(np.prod(sh)), 7)
        total_vox_it += r * d
        total_dt += d
    rate = total_vox_it / total_dt
    out["batched_rl"] = {"value": rate, "unit": "voxel-iters/s",
                         "sample": f"6 RL iterations on each of 6 volumes {shapes}, {total_dt:.1f}s"}
    log(f"cpu baseline batched RL: {rate:.3e} voxel-iters/s")

    # large-volume solver on its 128x512x512 crop
    crop = (128, 512, 512)
    psf = widefield_psf(crop, LATERAL, AXIAL, WIDEFIELD_OPTICS, aberrations=[IndexMismatch(depth=10.0)],
                        dipole="isotropic")
    Ac = O.LinearFFTConvolver(psf, crop)
    obs = poisson(Ac.forward(blob_volume(crop, 3125, rng).astype(np.float32)))
    rate, dt = _timed_oracle_iterations(
        lambda cb: O.solve_pdhg(obs, psf, alpha=5e-4, num_iter=4, tol=0.0, callback=cb), int(np.prod(crop)), 4)
    out["pdhg_large_volume_crop_128x512x512"] = {"value": rate, "unit": "voxel-iters/s",
                                                 "sample": f"3 PDHG-Hessian iterations, vectorial PSF, {dt:.1f}s"}
    log(f"cpu baseline large-volume crop: {rate:.3e} voxel-iters/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-iters", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-configs", action="store_true", help="CPU baseline at the metric shape only")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tune-cache", default=None, help="JSON file the autotuner reads/writes")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the host-to-host, batching, tiled and PSF synthesis legs (profiled runs)")
    ap.add_argument("--serving-jobs", type=int, default=60)
    ap.add_argument("--serving-rate", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import torch

    from deconlib_amd import _native as N
    from deconlib_amd.rl import _RLState

    psf, obs, blur = make_inputs(torch)
    tiles = {}
    if not args.no_tune:
        from deconlib_amd.autotune import autotune

        t0 = time.time()
        tiles = autotune(blur, verbose=True, cache_file=args.tune_cache)
        log(f"autotune {time.time() - t0:.1f}s: {tiles}")
    import ctypes as C

    st = _RLState(obs, blur, background=0.0, num_iter=args.warmup + 2 * args.steps + 1, eval_interval=10)
    # set-up: untimed RL iterations on a separate state for at least 0.25 s, right before
    # the W warm-up iterations, so that the GPU has left its idle clock state
    ramp = _RLState(obs, blur, background=0.0, num_iter=1 << 30, eval_interval=1 << 30)
    torch.cuda.synchronize()
    t_ramp = time.perf_counter()
    k_ramp = 0
    while time.perf_counter() - t_ramp < 0.25:
        for _ in range(5):
            ramp.step(k_ramp)
            k_ramp += 1
        torch.cuda.synchronize()
    for k in range(args.warmup):
        st.step(k)
    torch.cuda.synchronize()
    # timed region: HIP events (fence-free, on the launch stream) bracket the Z pass only,
    # the kernel the roofline reports, so the measurement adds two event records per
    # application
    N.call("dh_timing_read", None, None, 1)
    N.call("dh_timing_select", 1 << N.PASS_NAMES.index("z_conv"))
    N.call("dh_timing_enable", 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        st.step(k)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    N.call("dh_timing_enable", 0)
    del ramp
    zms = (C.c_double * 5)()
    zcnt = (C.c_int64 * 5)()
    N.call("dh_timing_read", zms, zcnt, 1)
    # per-pass breakdown: the next K iterations with every pass bracketed (not timed)
    N.call("dh_timing_select", (1 << len(N.PASS_NAMES)) - 1)
    N.call("dh_timing_enable", 1)
    for k in range(args.warmup + args.steps, args.warmup + 2 * args.steps):
        st.step(k)
    torch.cuda.synchronize()
    N.call("dh_timing_enable", 0)
    ms = (C.c_double * 5)()
    cnt = (C.c_int64 * 5)()
    N.call("dh_timing_read", ms, cnt, 1)
    vox = int(np.prod(SHAPE))
    value = vox * args.steps / dt
    P = blur.canvas_shape
    pb = conv_pass_bytes(SHAPE, P, epi_operand_bytes=6)
    if st.fused_iter:  # pass 4 is the X round trip (inverse + epilogue + forward)
        pb["x_inv"] = round_trip_bytes(SHAPE, P)
    it_bytes = rl_iteration_bytes(SHAPE, P, fused=st.fused_iter)
    per_pass = {}
    for i, name in enumerate(N.PASS_NAMES):
        if cnt[i]:
            avg = ms[i] / cnt[i]
            key = "x_round_trip" if (name == "x_inv" and st.fused_iter) else name
            per_pass[key] = {"avg_ms": avg, "gbs": pb[name] / (avg * 1e-3) / 1e9, "launches": int(cnt[i])}
    # the roofline kernel over the timed region itself
    zi = N.PASS_NAMES.index("z_conv")
    z_avg = zms[zi] / zcnt[zi]
    z = {"avg_ms": z_avg, "gbs": pb["z_conv"] / (z_avg * 1e-3) / 1e9, "launches": int(zcnt[zi])}
    log(f"per-pass: {json.dumps(per_pass)}")
    log(f"z_conv over the timed region: {json.dumps(z)}")
    ok = bool(torch.isfinite(st.x).all().item())
    pdhg = bench_pdhg(torch, N, obs, blur, args.steps, args.warmup)
    log(f"pdhg: {json.dumps(pdhg)}")
    pdhg["binned_configuration"] = bench_pdhg_binned(torch, tune=not args.no_tune, tune_cache=args.tune_cache)
    log(f"pdhg binned configuration: {json.dumps(pdhg['binned_configuration'])}")
    pdhg["large_volume"] = bench_large_volume(torch, tune=not args.no_tune)
    log(f"pdhg large volume: {json.dumps(pdhg['large_volume'])}")
    nlcg = bench_nlcg(torch, obs, blur, args.steps, args.warmup)
    nlcg["vs_rl_to_equal_fit"] = nlcg_vs_rl(torch, obs, blur)
    log(f"nlcg: {json.dumps(nlcg)}")
    erdecon = bench_erdecon(torch, obs, blur)
    log(f"erdecon: {json.dumps(erdecon)}")
    jetnewton = bench_jetnewton(torch, obs, blur)
    log(f"jetnewton: {json.dumps(jetnewton)}")
    host = small = batched_cfg = tiled = psf_synth = None
    if not args.no_extras:
        host = bench_host_to_host(torch, psf, obs)
        log(f"host-to-host solve: {json.dumps(host)}")
        small = bench_batched_small(torch)
        log(f"batched small volumes: {json.dumps(small)}")
        batched_cfg = bench_batched_config(torch, tune=not args.no_tune)
        log(f"batched configuration: {json.dumps(batched_cfg)}")
        tiled = bench_tiled(torch)
        log(f"tiled: {json.dumps(tiled)}")
        psf_synth = bench_psf_synthesis(torch)
        log(f"psf synthesis: {json.dumps(psf_synth)}")
    if args.copy_probe:
        # a plain device copy of known size: the reference point for the HBM counters
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 3 x 268435456-byte device-to-device copies")
    ceiling = stream_ceiling(torch)
    log(f"stream ceiling (1 GiB device copy / add): {json.dumps(ceiling)}")
    serving = None
    if args.serving_jobs > 0:
        serving = bench_serving(torch, args.serving_jobs, args.serving_rate)
        log(f"serving: {json.dumps(serving)}")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(psf, obs, args.cpu_iters)
        cpu["cpu_model"] = _cpu_model()
        if not args.no_cpu_configs:
            cpu["per_config"] = cpu_baseline_configs()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "Richardson-Lucy, 64x512x512 float32, widefield PSF 64x512x512 "
                        "(NA 1.4, 0.525 um, 0.085/0.1 um), FFT canvas "
                        f"{P[0]}x{P[1]}x{P[2]}, {N_BLOBS} Gaussian blobs + 10, Poisson",
            "volume": list(SHAPE),
            "canvas": list(P),
            "solver": "richardson_lucy_with_operator (fused iteration, dh_rl_iteration)"
                      if st.fused_iter else "richardson_lucy_with_operator (two fused applications)",
            "tiles": {N.PASS_NAMES[p]: t for p, t in tiles.items()},
            "finite": ok,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "z_conv (fused Z FFT x OTF x inverse Z FFT)",
            "achieved": z["gbs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": z["gbs"] / PEAK_HBM_GBS,
            "traffic": Z_TRAFFIC_BYTES,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated "
                              "(profiles/fft_passes_64x512x512_hbm_counters.md)",
            "algorithmic_bytes": pb["z_conv"],
            "avg_ms": z["avg_ms"],
            "avg_ms_source": "HIP events (no system-scope fence) around every Z launch of the timed "
                             "iterations, on the launch stream",
            "stream_ceiling": ceiling,
        },
        "passes": per_pass,
        "passes_source": "the next K iterations after the timed ones, every pass bracketed by HIP events",
        "iteration": {"algorithmic_bytes": it_bytes,
                      "hbm_frac": it_bytes / (dt / args.steps) / (PEAK_HBM_GBS * 1e9)},
        "pdhg": pdhg,
        "nlcg": nlcg,
        "erdecon": erdecon,
        "jetnewton": jetnewton,
        "host_to_host": host,
        "batched_small": small,
        "batched_config": batched_cfg,
        "tiled": tiled,
        "psf_synthesis": psf_synth,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
