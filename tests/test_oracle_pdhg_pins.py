"""Pins of the oracle's PDHG pieces beyond the first iteration (CPU only).

What each check pins, and the slip it is chosen to catch:

* the Malitsky-Pock step rule (pdhg_mlx.py:761-797): hand-derived numbers for
  each branch (growth-bound, ratio-bound, backtracking, tau floor, theta floor,
  sigma ceiling), and along real runs its invariants with dx^2 and dK^2
  recomputed by four separate operator applications; a test plants slips in
  a copy of the rule and checks that the invariants reject each one;
* the patience/min_iter stop (pdhg_mlx.py:805-835): hand-built rel-change
  sequences with the stop iteration worked out by hand, and the loop's own
  stop against its recorded rel-change history;
* convergence: a tiny problem run to a fixed point satisfies the KKT
  conditions of min F(Ax) + alpha |Lx|_1 s.t. x >= 0, computed independently;
* the loss formulas (rl_mlx.py:58-80, pdhg_mlx.py:860-881): hand-computed
  values on a few numbers with sum(obs - model) != 0;
* the gradient / 1-D and 2-D Hessian / identity regularizers
  (linops_core_mlx.py:25-80, linops_mlx.py:46-179, pdhg_mlx.py:241-456):
  ramp and quadratic known answers, a dense brute-force transpose, the
  reference's power-iteration norm checks and closed-form norms.
"""

import math

import numpy as np
import pytest

from oracle import deconv as O


# ---------------------------------------------------------------------------
# step rule: hand-derived cases
# ---------------------------------------------------------------------------

def test_step_rule_growth_bound_no_backtrack():
    # tau = sigma = theta = step = 1, dx2 = 2, dK2 = 1:
    #   ratio = 2 / (2*1*1) = 1, delta*sqrt(ratio) = 0.99 < sqrt(2) -> candidate 0.99
    #   0.99*1*1 = 0.99 <= delta*dx2 = 1.98 -> no backtracking, tau' = 0.99
    #   theta' = 0.99, sigma' = 1/0.99
    tn, sn, thn, back = O.malitsky_pock_step(1.0, 1.0, 1.0, 2.0, 1.0, 1.0)
    assert not back
    assert tn == pytest.approx(0.99, rel=1e-15)
    assert thn == pytest.approx(0.99, rel=1e-15)
    assert sn == pytest.approx(1.0 / 0.99, rel=1e-15)


def test_step_rule_growth_term_binds():
    # dK2 = 0: ratio = dx2 / (2 sigma 1e-12) is huge, the growth term sqrt(1 + theta) tau binds:
    #   tau = 0.5, theta = 1 -> tau' = 0.5*sqrt(2) = 0.7071067811865476; theta' = sqrt(2);
    #   sigma' = 0.8 / sqrt(2) = 0.565685424949238
    tn, sn, thn, back = O.malitsky_pock_step(0.5, 0.8, 1.0, 3.0, 0.0, 1.0)
    assert not back
    assert tn == pytest.approx(0.7071067811865476, rel=1e-15)
    assert thn == pytest.approx(1.4142135623730951, rel=1e-15)
    assert sn == pytest.approx(0.565685424949238, rel=1e-14)


def test_step_rule_backtracks():
    # tau = sigma = theta = step = 1, dx2 = 1, dK2 = 8:
    #   ratio = 1/16, candidate = 0.99/4 = 0.2475 (< sqrt 2)
    #   0.2475*1*8 = 1.98 > 0.99*1 -> backtrack: tau' = 0.5*0.2475 = 0.12375
    #   theta' = 0.12375 (>= 0.1), sigma' = 1/0.12375 = 8.080808...
    tn, sn, thn, back = O.malitsky_pock_step(1.0, 1.0, 1.0, 1.0, 8.0, 1.0)
    assert back
    assert tn == pytest.approx(0.12375, rel=1e-14)
    assert thn == pytest.approx(0.12375, rel=1e-14)
    assert sn == pytest.approx(8.080808080808081, rel=1e-14)


def test_step_rule_theta_floor_and_tau_floor():
    # dx2 = 1e-4, dK2 = 1: candidate = 0.99*sqrt(5e-5) = 0.0070003571337468,
    # 0.0070003571*1 > 0.99e-4 -> backtrack: tau' = 0.0035001785668734
    # theta' = 0.0035 < 0.1 -> sigma' = sigma / 0.1 = 10
    tn, sn, thn, back = O.malitsky_pock_step(1.0, 1.0, 1.0, 1e-4, 1.0, 1.0)
    assert back
    assert tn == pytest.approx(0.0035001785668734103, rel=1e-13)
    assert sn == pytest.approx(10.0, rel=1e-15)
    # step = 100: the tau floor 1e-4*100 = 0.01 binds; theta' = 0.01 -> sigma' = 1/0.1
    tn, sn, thn, back = O.malitsky_pock_step(1.0, 1.0, 1.0, 1e-4, 1.0, 100.0)
    assert tn == pytest.approx(0.01, rel=1e-15)
    assert thn == pytest.approx(0.01, rel=1e-15)
    assert sn == pytest.approx(10.0, rel=1e-15)


def test_step_rule_sigma_clamps():
    # step = 1e-3 -> sigma in [1e-7, 10]; sigma = 5 with theta' = 0.0035 -> 5/0.1 = 50 -> 10
    tn, sn, thn, back = O.malitsky_pock_step(1.0, 5.0, 1.0, 1e-4, 0.2, 1e-3)
    assert sn == pytest.approx(10.0, rel=1e-15)
    # lower clamp: step = 1 -> sigma >= 1e-4; sigma = 1e-6, theta' = sqrt 2 -> 7.07e-7 -> 1e-4
    tn, sn, thn, back = O.malitsky_pock_step(1.0, 1e-6, 1.0, 3.0, 0.0, 1.0)
    assert sn == pytest.approx(1e-4, rel=1e-15)


# ---------------------------------------------------------------------------
# step rule: invariants along real runs
# ---------------------------------------------------------------------------

def _problem(reg="hessian", ndim=3, seed=0, gain=None):
    """A tiny PDHG problem.  gain: A = gain * (identity + 1 % spread along z),
    unnormalised, so that |A|^2 dominates |K|^2 and the step rule backtracks."""
    rng = np.random.default_rng(seed)
    if gain is not None:
        psf = np.zeros((3, 3, 3))
        psf[0, 0, 0], psf[1, 0, 0] = gain, 0.01 * gain
        xt = rng.random((6, 8, 8)) * 30
        A = O.LinearFFTConvolver(psf, xt.shape, normalize=False, dtype=np.float64)
        obs = rng.poisson(A.forward(xt) + 2.0).astype(np.float64)
        return A, O.create_regularizer(reg, 3, 1.0, "L1"), obs
    if ndim == 3:
        psf = rng.random((3, 5, 5)) + 0.05
        xt = rng.random((6, 8, 8)) * 30
    else:
        psf = np.fft.ifftshift(np.exp(-0.5 * ((np.arange(7) - 3) / 1.2) ** 2))
        xt = np.zeros(20)
        xt[[4, 11, 12]] = [60.0, 90.0, 30.0]
    A = O.LinearFFTConvolver(psf, xt.shape, dtype=np.float64)
    obs = rng.poisson(A.forward(xt) + 2.0).astype(np.float64)
    L = O.create_regularizer(reg, ndim, 1.0, "L1")
    return A, L, obs


def _step_rule_violations(trace, A, L, step, delta=0.99, eta=0.5):
    """Checks each traced iteration against pdhg_mlx.py:761-797 with dx^2 and dK^2
    recomputed by four separate applications A x', A x, L x', L x."""
    bad = []
    for k, t in enumerate(trace):
        x, xn = t["x"], t["x_new"]
        dx2 = float(np.sum((xn - x) ** 2))
        dK2 = float(np.sum((A.forward(xn) - A.forward(x)) ** 2) + np.sum((L.forward(xn) - L.forward(x)) ** 2))
        tau, sigma, theta = t["tau"], t["sigma"], t["theta"]
        tn, sn, thn = t["tau_new"], t["sigma_new"], t["theta_new"]
        floor_tau = tn <= step * 1e-4 * (1 + 1e-12)
        if abs(thn - tn / tau) > 1e-14 * thn:
            bad.append((k, "theta' != tau'/tau"))
        if not floor_tau and tn > math.sqrt(1.0 + theta) * tau * (1 + 1e-12):
            bad.append((k, "tau' above sqrt(1+theta) tau"))
        cand = min(math.sqrt(1.0 + theta) * tau, delta * math.sqrt(dx2 / (2.0 * sigma * max(dK2, 1e-12))))
        back = cand * sigma * dK2 > delta * dx2
        want = max(eta * cand if back else cand, step * 1e-4)
        if abs(tn - want) > 1e-9 * want:
            bad.append((k, "tau' does not follow the backtracking branch"))
        free = step * 1e-4 * (1 + 1e-12) < sn < step * 1e4 * (1 - 1e-12)
        if thn >= 0.1 and free and abs(sn * tn - sigma * tau) > 1e-12 * sigma * tau:
            bad.append((k, "sigma' tau' != sigma tau"))
        if thn < 0.1 and free and abs(sn - sigma / 0.1) > 1e-12 * sn:
            bad.append((k, "sigma' != sigma / 0.1 under the theta floor"))
        if not (step * 1e-4 * (1 - 1e-12) <= sn <= step * 1e4 * (1 + 1e-12)):
            bad.append((k, "sigma' outside its clamp"))
    return bad


def _run_traced(reg, ndim, iters=40, seed=0, gain=None):
    A, L, obs = _problem(reg, ndim, seed, gain)
    tr = []
    x0 = np.maximum(obs, 0.0) * (0.2 if gain else 1.0)
    O.pdhg_loop(obs, A, L, x0, 0.05, iters, 2.0, tol=0.0, dtype=np.float64, trace=tr)
    step = 0.99 / math.sqrt(A.operator_norm_sq + L.operator_norm_sq)
    return tr, A, L, step


RUNS = [("hessian", 3, None), ("gradient", 3, None), ("identity", 3, None), ("hessian", 1, None),
        ("hessian", 3, 10.0), ("identity", 3, 10.0)]


@pytest.mark.parametrize("reg,ndim,gain", RUNS)
def test_step_rule_invariants_along_runs(reg, ndim, gain):
    tr, A, L, step = _run_traced(reg, ndim, gain=gain)
    assert len(tr) == 40
    assert _step_rule_violations(tr, A, L, step) == []


def test_step_rule_runs_exercise_both_branches():
    seen_back, seen_grow, seen_ratio = False, False, False
    for reg, ndim, gain in RUNS:
        tr, A, L, step = _run_traced(reg, ndim, gain=gain)
        for t in tr:
            seen_back |= bool(t["backtracked"])
            g = math.sqrt(1.0 + t["theta"]) * t["tau"]
            seen_grow |= (not t["backtracked"]) and abs(t["tau_new"] - g) < 1e-12 * g
            seen_ratio |= (not t["backtracked"]) and t["tau_new"] < g * (1 - 1e-9)
    assert seen_back and seen_grow and seen_ratio


def _slipped(kind):
    def rule(tau, sigma, theta, dx2, dK2, step, delta=0.99, eta=0.5):
        ratio = dx2 / (2.0 * sigma * max(dK2, 1e-12))
        grow = (1.0 + theta) * tau if kind == "growth" else math.sqrt(1.0 + theta) * tau
        tc = min(grow, delta * math.sqrt(ratio))
        back = tc * sigma * dK2 > delta * dx2
        if kind == "branch":
            back = not back
        tn = eta * tc if back else tc
        tn = max(tn, step * 1e-4)
        thn = tau / tn if kind == "theta" else tn / tau
        sn = sigma * max(thn, 0.1) if kind == "sigma" else sigma / max(thn, 0.1)
        sn = min(max(sn, step * 1e-4), step * 1e4)
        return tn, sn, thn, back
    return rule


@pytest.mark.parametrize("kind", ["growth", "branch", "theta", "sigma"])
def test_step_rule_checks_reject_planted_slips(monkeypatch, kind):
    monkeypatch.setattr(O, "malitsky_pock_step", _slipped(kind))
    found = []
    for reg, ndim, gain in RUNS:
        tr, A, L, step = _run_traced(reg, ndim, iters=20, gain=gain)
        found += _step_rule_violations(tr, A, L, step)
    assert found, f"slip {kind!r} not detected"


# ---------------------------------------------------------------------------
# stopping rule
# ---------------------------------------------------------------------------

def _stop_index(seq, min_iter, tol, patience):
    count = 0
    for k, rel in enumerate(seq):
        count, stop = O.convergence_step(k, rel, count, min_iter, tol, patience)
        if stop:
            return k
    return None


def test_stop_rule_hand_built_sequences():
    # min_iter 3, tol 0.1, patience 2: k = 0..2 ignored although below tol;
    # k=3 0.5 resets, k=4 counts 1, k=5 0.2 resets, k=6 counts 1, k=7 counts 2 -> stop at k=7
    seq = [0.01, 0.01, 0.01, 0.5, 0.05, 0.2, 0.05, 0.05, 0.05]
    assert _stop_index(seq, 3, 0.1, 2) == 7
    # rel == tol does not count (strict <): never stops
    assert _stop_index([0.1] * 10, 0, 0.1, 1) is None
    # patience 1 stops at the first qualifying iteration at or after min_iter
    assert _stop_index([0.01] * 10, 4, 0.1, 1) == 4
    # min_iter 0, patience 3: 0.05, 0.05, 0.3 (reset), 0.05 x3 -> stop at k=5
    assert _stop_index([0.05, 0.05, 0.3, 0.05, 0.05, 0.05], 0, 0.1, 3) == 5


def test_stop_rule_drives_the_loop():
    A, L, obs = _problem("hessian", 1, seed=3)
    tr = []
    res = O.pdhg_loop(obs, A, L, np.maximum(obs, 0.0), 0.5, 3000, 2.0, tol=1e-4, min_iter=20, patience=5,
                      dtype=np.float64, trace=tr)
    assert res.converged
    k = _stop_index([t["rel"] for t in tr], 20, 1e-4, 5)
    assert k is not None and res.iterations == k + 1 == len(tr)
    # the stop returns x' of the stopping iteration, without over-relaxation
    np.testing.assert_array_equal(res.restored, tr[-1]["x_new"])
    assert len(res.tau_history) == res.iterations
    # the relative change is |x' - x| / max(|x|, |x'|), computed independently
    t = tr[-1]
    rel = np.linalg.norm(t["x_new"] - t["x"]) / max(np.linalg.norm(t["x"]), np.linalg.norm(t["x_new"]))
    assert t["rel"] == pytest.approx(rel, rel=1e-9)


# ---------------------------------------------------------------------------
# convergence: KKT conditions of the converged solution
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("reg", ["hessian", "gradient"])
def test_converged_solution_satisfies_kkt(reg):
    A, L, obs = _problem(reg, 1, seed=5)
    alpha, bg = 0.5, 2.0
    res = O.pdhg_loop(obs, A, L, np.maximum(obs, 0.0), alpha, 6000, bg, tol=0.0, dtype=np.float64)
    x, y1, y2 = res.restored, res.metadata["y1"], res.metadata["y2"]
    m = A.forward(x) + bg
    # data dual = gradient of F(u) = sum(u + b - D log(u + b)) at A x
    assert np.max(np.abs(y1 - (1.0 - obs / m))) < 1e-4
    # regularizer dual in alpha * subdifferential of |.|_1 at L x
    Lx = L.forward(x)
    assert np.max(np.abs(y2)) <= alpha * (1 + 1e-12)
    assert np.max(np.abs(y2 * Lx - alpha * np.abs(Lx))) < 1e-4 * alpha * (1 + np.max(np.abs(Lx)))
    # stationarity with the nonnegativity multiplier
    g = A.adjoint(y1) + L.adjoint(y2)
    on = x > 1e-6
    assert on.any() and (~on).any()
    assert np.max(np.abs(g[on])) < 1e-4
    assert np.min(g[~on]) > -1e-4


# ---------------------------------------------------------------------------
# loss formulas on hand-computed values
# ---------------------------------------------------------------------------

def test_i_divergence_hand_values():
    # mean over [2 ln(2/1) - 2 + 1, 0 + 3, 0, ln(1/4) - 1 + 4] = (0.3863 + 3 + 0 + 1.6137)/4 = 1.25
    obs = np.array([2.0, 0.0, 5.0, 1.0])
    mdl = np.array([1.0, 3.0, 5.0, 4.0])
    assert np.sum(obs - mdl) == -5.0
    assert O.poisson_i_divergence(obs, mdl) == pytest.approx(1.25, rel=1e-12)
    # model floored at eps: 1*ln(1/1e-6) - (1 - 1e-6)
    assert O.poisson_i_divergence(np.array([1.0]), np.array([0.0])) == pytest.approx(12.815511557964273, rel=1e-12)
    # another eps: 1*ln(1/1e-3) - (1 - 1e-3)
    assert O.poisson_i_divergence(np.array([1.0]), np.array([0.0]), eps=1e-3) == pytest.approx(
        5.908755278982137, rel=1e-12)


def test_pdhg_objective_hand_values():
    data = np.array([2.0, 0.0, 5.0])
    fm = np.array([1.0, 3.0, 5.0])
    Lx = np.array([[1.0, -2.0, 0.0], [3.0, 0.0, 0.0]])
    # Poisson: 2 ln(2.000001) - 1 + (0 + 3) + 5 ln(1.0000002) = 3.38629636111954
    # L1 reg: 0.1 * (1 + 2 + 3) = 0.6;  L1_2 reg: 0.1 * (sqrt(10) + 2 + 1e-6) (the 1e-12 floor of an empty voxel)
    assert O.pdhg_objective(data, fm, Lx, 0.1, "L1", "poisson") == pytest.approx(3.38629636111954 + 0.6, rel=1e-12)
    assert O.pdhg_objective(data, fm, Lx, 0.1, "L1_2", "poisson") == pytest.approx(
        3.38629636111954 + 0.1 * (math.sqrt(10.0) + 2.0 + 1e-6), rel=1e-12)
    # Gaussian: 0.5 * (1 + 9 + 0) = 5
    assert O.pdhg_objective(data, fm, Lx, 0.1, "L1", "gaussian") == pytest.approx(5.6, rel=1e-12)


def test_loop_records_the_objective():
    A, L, obs = _problem("hessian", 3, seed=1)
    tr = []
    res = O.pdhg_loop(obs, A, L, np.maximum(obs, 0.0), 0.05, 6, 2.0, tol=0.0, eval_interval=3,
                      dtype=np.float64, trace=tr)
    assert len(res.loss_history) == 2
    for j, k in enumerate((2, 5)):
        x = tr[k]["x_new"]
        m = A.forward(x) + 2.0
        Lx = L.forward(x)
        want = float(np.sum(obs * np.log((obs + 1e-6) / np.maximum(m, 1e-6)) - (obs - m))
                     + 0.05 * np.sum(np.abs(Lx)))
        assert res.loss_history[j] == pytest.approx(want, rel=1e-12)


# ---------------------------------------------------------------------------
# first-order and 1-D / 2-D regularizers
# ---------------------------------------------------------------------------

def _dense(op, shape):
    n = int(np.prod(shape))
    cols = [np.asarray(op(e.reshape(shape))).ravel() for e in np.eye(n)]
    return np.stack(cols, axis=1)


def test_forward_difference_known_answers():
    f = 3.0 * np.arange(6) + 1.0
    np.testing.assert_array_equal(O.d1_fwd(f, 0), [3, 3, 3, 3, 3, 0])
    # transpose by brute force for n >= 2
    for n in (2, 3, 7):
        D = _dense(lambda v: O.d1_fwd(v, 0), (n,))
        Dt = _dense(lambda v: O.d1_fwd_adj(v, 0), (n,))
        np.testing.assert_array_equal(Dt, D.T)
    # n == 1: the reference's concatenation gives [-g0]
    np.testing.assert_array_equal(O.d1_fwd_adj(np.array([2.5]), 0), [-2.5])


def test_gradient3d_ramps_and_transpose():
    z, y, x = np.meshgrid(np.arange(4), np.arange(5), np.arange(6), indexing="ij")
    f = 2.0 * z - 1.0 * y + 0.5 * x
    G = O.Gradient3D(r=1.5).forward(f)
    np.testing.assert_allclose(G[0][:-1], 1.5 * 2.0)
    np.testing.assert_allclose(G[0][-1], 0.0)
    np.testing.assert_allclose(G[1][:, :-1], -1.0)
    np.testing.assert_allclose(G[2][:, :, :-1], 0.5)
    for op in (O.GradientRegularizer(3, 0.7), O.GradientRegularizer(2), O.GradientRegularizer(1),
               O.HessianRegularizer(2), O.HessianRegularizer(1), O.IdentityRegularizer()):
        shape = {3: (3, 4, 5), 2: (4, 5), 1: (6,)}[getattr(op, "ndim", 2)]
        M = _dense(op.forward, shape)
        out_shape = (op.output_components,) + shape
        Mt = _dense(op.adjoint, out_shape)
        np.testing.assert_allclose(Mt, M.T, atol=1e-13)
        assert op.forward(np.zeros(shape)).shape == out_shape


def test_hessian2d_quadratic_and_hessian1d():
    y, x = np.meshgrid(np.arange(8.0), np.arange(9.0), indexing="ij")
    f = 0.5 * y * y + 1.5 * x * x - 0.25 * x * y
    H = O.Hessian2D().forward(f)
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(H[0][inner], 1.0)
    np.testing.assert_allclose(H[1][inner], 3.0)
    np.testing.assert_allclose(H[2][inner], math.sqrt(2.0) * -0.25, rtol=1e-12)
    np.testing.assert_allclose(O.Hessian1D().forward(np.arange(7.0) ** 2)[1:-1], 2.0)


def _power_norm(op, shape, iters=300, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(shape)
    v /= np.linalg.norm(v)
    s = 0.0
    for _ in range(iters):
        w = op.adjoint(op.forward(v))
        s = float(np.sum(v * w))
        v = w / np.linalg.norm(w)
    return s


@pytest.mark.parametrize("op,shape,exact", [
    (O.Gradient1D(), (64,), 4.0),
    (O.Gradient2D(), (48, 48), 8.0),
    (O.Gradient3D(1.0), (24, 24, 24), 12.0),
    (O.Gradient3D(3.0), (16, 24, 24), 44.0),
    (O.Hessian1D(), (64,), 16.0),
    (O.Hessian2D(), (48, 48), 48.0),
])
def test_regularizer_norm_bounds(op, shape, exact):
    assert op.operator_norm_sq == exact
    est = _power_norm(op, shape)
    assert est <= exact * 1.0001
    # first differences and d2 reach their bounds at high frequency
    if not isinstance(op, O.Hessian2D):
        assert est >= 0.97 * exact
    else:
        assert est >= 20.0
    assert O.IdentityRegularizer().operator_norm_sq == 1.0
