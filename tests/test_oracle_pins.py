"""Closed-form pins of oracle functions that convergence and consistency
properties alone would not catch.

* Backtracking CG (CGMethod.hpp:238-373, 1279-1573; CGMethod.cpp:43-360) on
  the reference's 10-D weighted quadratic (TestCGMethod.cpp:20-56). Along an
  exact quadratic every quantity of an iteration is a closed form: with
  lambda* = s.F / s.W.s the energy change at step lambda is
  -lambda s.F + lambda^2 s.W.s / 2, so the Armijo test with slope c accepts
  exactly lambda <= 2 (1 - c) lambda* (1.2 lambda* for c = 0.4, 2 lambda*
  in safe mode); the trials are lambda0 = min(LAMBDAMAX, MAXDIST / max|F|)
  halved until accepted; the new force is F - lambda W s; beta is
  max(0, (F'.F' - F'.F) / F.F) (Polak-Ribiere), F'.F' / F.F
  (Fletcher-Reeves) or 0 (steepest descent); the direction resets to F'
  when s'.F' <= 0. The model below advances x, F and s with these closed
  forms, never by comparing computed energies, and every energy call of the
  oracle (the line search's current point and each trial) and its trace are
  checked against it. The runs are checked to contain steps that a wrong
  slope (0.3 or 0.5), halving factor or initial-lambda cap would decide
  differently.
* energyN / forceN (CylinderExclVolRepulsion.cpp:884-1254): the node and
  weight tables at N = 2 and N = 4, written out by hand from the loops (each
  abscissa advanced before its first use, weights 4 at odd and 2 at even loop
  index, corners weight 1, scale 1/(9 N1 N2)); forceN equals minus the
  gradient of the same quadrature sum, differentiated here by hand.
* The 0.999 clamp of the cosine bending forces (FilamentBendingCosine.cpp:
  287-297, BranchingBendingCosine.cpp:298-303): within 1e-3 of |cos| = 1 and
  with theta0 != 0 the force is the true gradient (central differences of the
  unclamped energy) times sin(t' - theta0) sin(t) / (sin(t') sin(t - theta0)),
  t' = acos(0.999 cos t); outside the window it is the true gradient.
"""
import math

import numpy as np
import pytest

from medyan_amd import _abi

OPT = np.array([-500.0, -400.0, -300.0, -200.0, -100.0, 0.0, 100.0, 200.0, 300.0, 400.0])
W = np.array([3, 1, 0.1, 0.5, 1, 10, 20, 50, 100, 0.1])


def quad_energy(x):
    return float(np.sum(0.5 * W * (x[:10] - OPT) ** 2))


def quad_force(x, f):
    f[:10] = -W * (x[:10] - OPT)


# ---------------------------------------------------------------------------
# CG with backtracking on the exact quadratic
# ---------------------------------------------------------------------------
def cg_model(method, x0, max_dist, lambda_max, grad_tol, n_iter_max=100000, slope=0.4, reduce=0.5):
    """Closed-form model of the minimization. Returns per iteration the
    current point, direction, trial lambdas, accepted lambda, beta and reset
    flag, plus the number of steps that a different slope or halving factor
    would have decided differently."""
    x = np.array(x0, float)
    F = -W * (x - OPT)
    s = F.copy()
    last_ff = F @ F
    maxf = np.abs(F).max()
    its = []
    disc = dict(slope_03=0, slope_05=0, cap_dist=0, cap_lmax=0, rejections=0)
    while not maxf < grad_tol and len(its) < n_iter_max:
        lam0 = min(lambda_max, max_dist / maxf)
        disc["cap_dist" if max_dist / maxf < lambda_max else "cap_lmax"] += 1
        sf = s @ F
        lstar = sf / (s @ (W * s))
        thr = 2.0 * (1.0 - slope) * lstar
        trials = []
        lam = lam0
        while True:
            trials.append(lam)
            assert abs(lam - thr) > 1e-9 * thr, "trial on the Armijo boundary: ambiguous case"
            if lstar < lam <= thr:
                disc["slope_05"] += 1          # c = 0.5 accepts only lambda <= lambda*
            if thr < lam <= 1.4 * lstar:
                disc["slope_03"] += 1          # c = 0.3 accepts up to 1.4 lambda*
            if lam <= thr:
                break
            lam *= reduce
            disc["rejections"] += 1
            assert lam > 1e-8
        x = x + lam * s
        Fn = F - lam * W * s
        ffn = Fn @ Fn
        if method == _abi.MDY_CG_POLAK_RIBIERE:
            beta = max(0.0, (ffn - Fn @ F) / last_ff)
        elif method == _abi.MDY_CG_FLETCHER_RIEVES:
            beta = ffn / last_ff
        else:
            beta = 0.0
        s_new = Fn + beta * s
        reset = bool(s_new @ Fn <= 0)
        if reset:
            s_new = Fn.copy()
        its.append(dict(x_before=x - lam * s, s=s.copy(), trials=trials, lam=lam, beta=beta, reset=reset,
                        maxf=np.abs(Fn).max(), x_after=x.copy()))
        F, s, last_ff, maxf = Fn, s_new, ffn, np.abs(Fn).max()
    return its, disc


def run_logged(oracle, p, x0, iter_cap=0):
    calls = []

    def energy(x):
        calls.append(np.array(x, float))
        return quad_energy(x)

    x, f, res, tr = oracle.minimize(p, x0, 10, energy, quad_force, trace_cap=100000, iter_cap=iter_cap)
    return x, f, res, tr, calls


# Iterations compared. On this problem the accepted steps of the slow
# components make the stiff ones (w = 50, 100) grow by |1 - lambda w| > 1
# per step once they have converged, so any two correct codes that round
# differently drift apart after 80-200 iterations; the first 60 agree to
# rounding in every case below.
N_CMP = 60


CG_CASES = {
    # the reference test's settings (TestCGMethod.cpp:57-92) from its third start
    "ref_input_2": (np.array([1000.0, 900, 800, 700, 600, 500, 400, 300, 200, 100]), 10.0, 1.0, 0.1),
    "ref_input_0": (np.full(10, 10.0), 10.0, 1.0, 0.1),
    # a smaller LAMBDAMAX that binds within the compared iterations (both
    # branches of min(LAMBDAMAX, MAXDIST / maxF))
    "small_lmax": (np.array([1000.0, 900, 800, 700, 600, 500, 400, 300, 200, 100]), 10.0, 0.02, 0.1),
}
METHODS = {"PR": _abi.MDY_CG_POLAK_RIBIERE, "FR": _abi.MDY_CG_FLETCHER_RIEVES, "SD": _abi.MDY_CG_STEEPEST}


@pytest.mark.parametrize("case", list(CG_CASES))
@pytest.mark.parametrize("mname", list(METHODS))
def test_backtracking_cg_matches_closed_forms(oracle, case, mname):
    x0, max_dist, lambda_max, grad_tol = CG_CASES[case]
    method = METHODS[mname]
    p = _abi.default_cg_params(0, grad_tol=grad_tol, max_dist=max_dist, lambda_max=lambda_max, method=method,
                               try_to_recover=0, force_call_limit=100000, energy_rel_tol=0.0)
    its, disc = cg_model(method, x0, max_dist, lambda_max, grad_tol, n_iter_max=N_CMP)
    x, f, res, tr, calls = run_logged(oracle, p, x0, iter_cap=N_CMP)
    assert res.status == 0 and res.error_flags == 0
    assert res.num_iter == len(its) == N_CMP
    # energy calls: energiesBefore, then per iteration the current point and each trial
    assert res.num_energy_call >= 1 + sum(1 + len(it["trials"]) for it in its)
    assert len(calls) == res.num_energy_call
    assert res.num_force_call == 1 + len(its)
    scale = np.abs(OPT).max() + np.abs(x0).max()
    k = 1
    for n, it in enumerate(its):
        np.testing.assert_allclose(calls[k], it["x_before"], rtol=0, atol=1e-11 * scale, err_msg=f"iteration {n}")
        k += 1
        s = it["s"]
        for j, lam in enumerate(it["trials"]):
            got = (calls[k] - it["x_before"]) @ s / (s @ s)
            assert got == pytest.approx(lam, rel=1e-9), (n, j)
            if j:
                # LAMBDAREDUCE: exactly one halving between trials (bits: a power-of-two scale)
                prev = (calls[k - 1] - it["x_before"]) @ s / (s @ s)
                assert got / prev == pytest.approx(0.5, rel=1e-9)
            k += 1
        assert tr["lambda_"][n] == pytest.approx(it["lam"], rel=1e-9)
        assert tr["beta"][n] == pytest.approx(it["beta"], rel=1e-7, abs=1e-12)
        assert tr["maxf"][n] == pytest.approx(it["maxf"], rel=1e-9, abs=1e-12)
        assert tr["safe"][n] == 0
    np.testing.assert_allclose(x, its[-1]["x_after"], rtol=0, atol=1e-9 * scale)


def test_backtracking_runs_discriminate_the_constants(oracle):
    """The runs above contain steps that a slope of 0.3 or 0.5 would decide
    differently, rejected trials (halving), and both branches of the initial
    lambda cap; so the constants are pinned, not only convergence."""
    tot = dict(slope_03=0, slope_05=0, cap_dist=0, cap_lmax=0, rejections=0)
    for x0, max_dist, lambda_max, grad_tol in CG_CASES.values():
        for method in METHODS.values():
            _, disc = cg_model(method, x0, max_dist, lambda_max, grad_tol, n_iter_max=N_CMP)
            for key in tot:
                tot[key] += disc[key]
    print("discriminating steps:", tot)
    assert all(v > 0 for v in tot.values()), tot


@pytest.mark.parametrize("ratio,accepted", [(1.95, True), (2.05, False)])
def test_safe_mode_accepts_up_to_twice_the_line_minimizer(oracle, ratio, accepted):
    """Safe mode (Armijo factor 0, CGMethod.hpp:1300-1360): after a failed
    line search the next one accepts any step that does not raise the
    energy, i.e. exactly lambda <= 2 lambda* on a quadratic. The energy
    rejects every trial of the first search (it fails, lambda 0, safe mode
    on, direction reset to F); the second search runs on the true quadratic
    from the same point with LAMBDAMAX = ratio * lambda*."""
    x0 = np.array([1000.0, 900, 800, 700, 600, 500, 400, 300, 200, 100])
    F0 = -W * (x0 - OPT)
    lstar = (F0 @ F0) / (F0 @ (W * F0))
    seen = dict(x0=0)
    calls = []

    def energy(x):
        calls.append(np.array(x, float))
        if np.array_equal(x, x0):
            seen["x0"] += 1
        elif seen["x0"] <= 2:
            return 1e300                      # the first line search finds no lower energy
        return quad_energy(x)

    p = _abi.default_cg_params(0, grad_tol=0.1, max_dist=1e9, lambda_max=ratio * lstar,
                               method=_abi.MDY_CG_POLAK_RIBIERE, try_to_recover=1, force_call_limit=100000,
                               energy_rel_tol=0.0)
    _, _, res, tr = oracle.minimize(p, x0, 10, energy, quad_force, trace_cap=8, iter_cap=2)
    assert tr["lambda_"][0] == 0.0
    assert res.num_safe_mode == 1
    k = [i for i, c in enumerate(calls) if np.array_equal(c, x0)][2]     # the second search's current point
    trials = [(c - x0) @ F0 / (F0 @ F0) for c in calls[k + 1:]]
    assert trials[0] == pytest.approx(ratio * lstar, rel=1e-12)
    if accepted:
        assert tr["lambda_"][1] == pytest.approx(ratio * lstar, rel=1e-12)
    else:
        assert trials[1] == pytest.approx(0.5 * ratio * lstar, rel=1e-12)
        assert tr["lambda_"][1] == pytest.approx(0.5 * ratio * lstar, rel=1e-12)


# ---------------------------------------------------------------------------
# energyN / forceN node tables
# ---------------------------------------------------------------------------
def simpson_table(n):
    """(s, t, weight) of the reference rule at N1 = N2 = n, written out."""
    if n == 2:
        # loop index 1 only: abscissa 2/2 = 1, weight 4
        return [(0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 1),      # corners
                (1, 0, 4), (1, 1, 4),                            # s-edge terms, s = 1
                (0, 1, 4), (1, 1, 4),                            # t-edge terms, t = 1
                (1, 1, 16)]                                      # interior, s = t = 1
    if n == 4:
        # loop indices 1, 2, 3 -> abscissae 2/4, 3/4, 4/4 with weights 4, 2, 4
        ab = [(0.5, 4), (0.75, 2), (1.0, 4)]
        tab = [(0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 1)]
        tab += [(0.5, 0, 4), (0.5, 1, 4), (0.75, 0, 2), (0.75, 1, 2), (1.0, 0, 4), (1.0, 1, 4)]
        tab += [(0, 0.5, 4), (1, 0.5, 4), (0, 0.75, 2), (1, 0.75, 2), (0, 1.0, 4), (1, 1.0, 4)]
        tab += [(s, t, ws * wt) for (s, ws) in ab for (t, wt) in ab]
        return tab
    raise ValueError(n)


def simpson_mixed_table():
    """N1 = 2 along the first cylinder, N2 = 4 along the second."""
    s_ab = [(1.0, 4)]
    t_ab = [(0.5, 4), (0.75, 2), (1.0, 4)]
    tab = [(0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 1)]
    tab += [(s, e, w) for (s, w) in s_ab for e in (0, 1)]
    tab += [(e, t, w) for (t, w) in t_ab for e in (0, 1)]
    tab += [(s, t, ws * wt) for (s, ws) in s_ab for (t, wt) in t_ab]
    return tab


GEOMS = [
    ([10, 0, 0], [-10, 0, 0], [0, 10, 5], [3, -10, 5]),
    ([0, 0, 0], [100, 0, 0], [20, 30, 5], [90, 50, -20]),
    ([5, 5, 5], [60, -10, 30], [40, 40, 40], [-20, 35, 70]),
]


def table_energy_force(tab, c1, c2, c3, c4, kr, n1, n2):
    """Quadrature sum of kR / r^4 over the table and its gradient, by hand:
    r(s, t) = c1 + s (c2 - c1) - c3 - t (c4 - c3), d(r^-4)/dr = -4 r^-6 r."""
    c1, c2, c3, c4 = (np.asarray(v, float) for v in (c1, c2, c3, c4))
    h = 1.0 / (9.0 * n1 * n2)
    e = 0.0
    g = np.zeros((4, 3))
    for s, t, w in tab:
        r = c1 + s * (c2 - c1) - c3 - t * (c4 - c3)
        r2 = r @ r
        e += w / r2 ** 2
        dr = -4.0 * w / r2 ** 3 * r            # d(w r^-4)/dr
        g[0] += (1 - s) * dr
        g[1] += s * dr
        g[2] -= (1 - t) * dr
        g[3] -= t * dr
    return kr * h * e, -kr * h * g


@pytest.mark.parametrize("geom", range(len(GEOMS)))
@pytest.mark.parametrize("n1,n2", [(2, 2), (4, 4), (2, 4)])
def test_simpson_node_tables(oracle, geom, n1, n2):
    c = [np.array(v, float) for v in GEOMS[geom]]
    kr = 1e5 * 108.0 * 108.0
    tab = simpson_mixed_table() if (n1, n2) == (2, 4) else simpson_table(n1)
    e_ref, f_ref = table_energy_force(tab, *c, kr, n1, n2)
    e = oracle.ev_pair_energy_simpson(*c, kr, n1, n2)
    assert e == pytest.approx(e_ref, rel=1e-15)
    f = np.asarray(oracle.ev_pair_force_simpson(*c, kr, n1, n2)).reshape(4, 3)
    np.testing.assert_allclose(f, f_ref, rtol=0, atol=1e-13 * np.abs(f_ref).max())


def test_simpson_fallback_orders(oracle):
    """The energy fallback of an invalid analytic pair is energyN at
    N1 = N2 = 400 (CylinderExclVolRepulsion.cpp:929-930): exactly parallel
    cylinders give 0/0 analytically, and the pair energy is that rule's
    value bit for bit."""
    x = np.array([0, 0, 0, 100, 0, 0, 0, 20, 0, 100, 20, 0], float)
    e = oracle.ev_energy(x, [0, 3, 6, 9], [1.0], [1.0, 1.0])
    c = [x[i:i + 3] for i in (0, 3, 6, 9)]
    assert e == oracle.ev_pair_energy_simpson(*c, 1.0, 400, 400)
    assert e != oracle.ev_pair_energy_simpson(*c, 1.0, 200, 200)


# ---------------------------------------------------------------------------
# the 0.999 clamp of the cosine bending forces
# ---------------------------------------------------------------------------
def bend_geometry(cos_t, L1=108.0, L2=95.0):
    """Three beads with the angle between (r2 - r1) and (r3 - r2) of cosine cos_t."""
    t = math.acos(cos_t)
    r1 = np.array([1.0, 2.0, 3.0])
    r2 = r1 + L1 * np.array([1.0, 0.0, 0.0])
    r3 = r2 + L2 * np.array([math.cos(t), math.sin(t) * 0.6, math.sin(t) * 0.8])
    return np.concatenate([r1, r2, r3])


def clamp_factor(cos_t, theta0):
    """Force / true gradient force for the reference's clamped k_eff."""
    t = math.acos(cos_t)
    if abs(abs(cos_t) - 1.0) < 1e-3:
        tp = math.acos(0.999 * cos_t)
        return (math.sin(tp - theta0) / math.sin(tp)) / (math.sin(t - theta0) / math.sin(t))
    return 1.0


def central_diff(fn, x, h):
    g = np.zeros_like(x)
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (fn(xp) - fn(xm)) / (2 * h)
    return g


@pytest.mark.parametrize("cos_t", [0.9995, 0.99905, 0.99895, -0.9995, -0.99895])
@pytest.mark.parametrize("theta0", [0.5, 2.0])
def test_filament_bending_clamp(oracle, cos_t, theta0):
    x = bend_geometry(cos_t)
    bs, k, th = [0, 3, 6], [672.0], [theta0]
    f = oracle.ff_forces("fb", x, bs, k, th)
    g = central_diff(lambda y: oracle.ff_energy("fb", y, bs, k, th), x, 1e-4)
    fac = clamp_factor(cos_t, theta0)
    if abs(abs(cos_t) - 1.0) < 1e-3:
        assert abs(fac - 1.0) > 0.05          # the clamp changes the force visibly here
    np.testing.assert_allclose(f, -fac * g, rtol=0, atol=2e-6 * np.abs(g).max())


@pytest.mark.parametrize("cos_t", [0.9995, -0.9995, 0.99895])
def test_branching_bending_clamp(oracle, cos_t):
    """BranchingBendingCosine: the angle between (r2 - r1) on the mother and
    (r4 - r3) on the daughter, theta0 = 70 deg (never zero, so the clamp is
    always armed)."""
    theta0 = 70.0 * math.pi / 180.0
    t = math.acos(cos_t)
    r1 = np.array([1.0, 2.0, 3.0])
    r2 = r1 + 108.0 * np.array([1.0, 0.0, 0.0])
    r3 = r1 + np.array([40.0, 7.0, -3.0])
    r4 = r3 + 90.0 * np.array([math.cos(t), math.sin(t) * 0.6, math.sin(t) * 0.8])
    x = np.concatenate([r1, r2, r3, r4])
    bs, k, th = [0, 3, 6, 9], [10.0], [theta0]
    f = oracle.ff_forces("bb", x, bs, k, th)
    g = central_diff(lambda y: oracle.ff_energy("bb", y, bs, k, th), x, 1e-4)
    fac = clamp_factor(cos_t, theta0)
    np.testing.assert_allclose(f, -fac * g, rtol=0, atol=2e-6 * np.abs(g).max())
