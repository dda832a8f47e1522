"""Pins of the CPU oracle's force fields (the checker for the HIP kernels).

Each potential is pinned by what fixes it independently of the restatement:
closed forms at hand-computed geometries, central differences of the energy,
Newton's third law, rigid-motion invariance, and, for excluded volume, the
reference's own test geometries and equalities
(src/TESTS/Mechanics/ForceField/Volume/TestCylinderExclVolume.cpp,
src/TESTS/Mechanics/ForceField/Boundary/TestBoundaryBeadRepulsion.cpp,
src/TESTS/test_boundaries.cpp:45-49) plus an independent Gauss-Legendre
quadrature of the excluded-volume integral.
"""
import math

import functools

import numpy as np
import pytest

RNG = np.random.default_rng(20261202)


def rot(a, b, c):
    ca, sa, cb, sb, cc, sc = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(c), math.sin(c)
    rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return rz @ ry @ rx


def central_diff(efn, x, h):
    g = np.zeros_like(x)
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (efn(xp) - efn(xm)) / (2 * h)
    return g


def ff_case(name):
    """One or two interactions of each bonded / cross-link force field, on
    generic (non-degenerate) geometry."""
    if name == "fs":
        x = np.array([0, 0, 0, 3, 4, 0, 10, 1, -2], float) + RNG.normal(0, 0.1, 9)
        return x, np.array([0, 3, 3, 6]), (np.array([100.0, 50.0]), np.array([4.0, 7.0]))
    if name == "fb":
        x = np.array([0, 0, 0, 100, 0, 0, 190, 40, 10, 250, 120, 60], float) + RNG.normal(0, 1, 12)
        return x, np.array([0, 3, 6, 3, 6, 9]), (np.array([672.0, 300.0]), np.array([0.0, 0.5]))
    if name == "ls":
        x = np.array([0, 0, 0, 100, 0, 0, 20, 60, 5, 110, 80, -10], float) + RNG.normal(0, 1, 12)
        return x, np.array([0, 3, 6, 9]), (np.array([8.0]), np.array([40.0]), np.array([0.25]), np.array([0.625]))
    if name == "bs":
        x = np.array([0, 0, 0, 100, 0, 0, 40, 9, 3], float) + RNG.normal(0, 0.5, 9)
        return x, np.array([0, 3, 6]), (np.array([100.0]), np.array([6.0]), np.array([0.375]))
    if name == "bb":
        x = np.array([0, 0, 0, 100, 0, 0, 40, 6, 0, 80, 100, 30], float) + RNG.normal(0, 1, 12)
        return x, np.array([0, 3, 6, 9]), (np.array([10.0]), np.array([70.0 * math.pi / 180]))
    raise KeyError(name)


# --------------------------------------------------------------------------
# closed forms
# --------------------------------------------------------------------------
def test_fs_closed_form(oracle):
    x = np.array([0, 0, 0, 3, 4, 0], float)
    bs, k, l0 = np.array([0, 3]), np.array([100.0]), np.array([4.0])
    assert oracle.ff_energy("fs", x, bs, k, l0) == pytest.approx(0.5 * 100 * 1.0, rel=1e-15)
    f = oracle.ff_forces("fs", x, bs, k, l0)
    # |F| = k (d - l0) = 100 along the bond, pulling the beads together
    np.testing.assert_allclose(f, [60, 80, 0, -60, -80, 0], rtol=1e-14)


@pytest.mark.parametrize("theta_deg,theta0_deg", [(0, 0), (90, 0), (180, 0), (70, 70), (90, 70), (30, 70)])
def test_bending_closed_form(oracle, theta_deg, theta0_deg):
    """U = k(1 - cos(theta - theta0)), theta the angle between consecutive bonds
    (FilamentBendingCosine.cpp:199-243); for theta0 = 0, U = k(1 - cos theta)."""
    t, t0, k = math.radians(theta_deg), math.radians(theta0_deg), 672.0
    x = np.array([-108, 0, 0, 0, 0, 0, 108 * math.cos(t), 108 * math.sin(t), 0], float)
    u = oracle.ff_energy("fb", x, np.array([0, 3, 6]), np.array([k]), np.array([t0]))
    assert u == pytest.approx(k * (1 - math.cos(t - t0)), rel=1e-12, abs=1e-9)
    # branching bending: the same angle between (c2-c1) and (c4-c3)
    x4 = np.array([-108, 0, 0, 0, 0, 0, 5, 5, 5, 5 + 108 * math.cos(t), 5 + 108 * math.sin(t), 5], float)
    u4 = oracle.ff_energy("bb", x4, np.array([0, 3, 6, 9]), np.array([k]), np.array([t0]))
    assert u4 == pytest.approx(k * (1 - math.cos(t - t0)), rel=1e-12, abs=1e-9)


def test_linker_closed_form(oracle):
    x = np.array([0, 0, 0, 100, 0, 0, 0, 30, 0, 100, 30, 0], float)
    # points at 25 % of cylinder 1 and 75 % of cylinder 2: (25,0,0) and (75,30,0)
    d = math.hypot(50, 30)
    u = oracle.ff_energy("ls", x, np.array([0, 3, 6, 9]), np.array([8.0]), np.array([40.0]), np.array([0.25]),
                         np.array([0.75]))
    assert u == pytest.approx(0.5 * 8 * (d - 40) ** 2, rel=1e-14)
    f = oracle.ff_forces("ls", x, np.array([0, 3, 6, 9]), np.array([8.0]), np.array([40.0]), np.array([0.25]),
                         np.array([0.75]))
    pull = 8 * (d - 40) * np.array([50, 30, 0]) / d      # force on point 1 toward point 2
    np.testing.assert_allclose(f[0:3], 0.75 * pull, rtol=1e-13)
    np.testing.assert_allclose(f[3:6], 0.25 * pull, rtol=1e-13)
    np.testing.assert_allclose(f[6:9], -0.25 * pull, rtol=1e-13)
    np.testing.assert_allclose(f[9:12], -0.75 * pull, rtol=1e-13)


def test_branch_stretch_closed_form(oracle):
    x = np.array([0, 0, 0, 100, 0, 0, 37.5, 10, 0], float)
    u = oracle.ff_energy("bs", x, np.array([0, 3, 6]), np.array([100.0]), np.array([6.0]), np.array([0.375]))
    assert u == pytest.approx(0.5 * 100 * 4.0 ** 2, rel=1e-14)


def test_boundary_reference_fixture(oracle):
    """TestBoundaryBeadRepulsion.cpp:23-40: plane through (1,2,3) with normal
    +z; beads on, 10 above and 10 below it; U = K exp(-r/lambda)."""
    plane = np.array([0.0, 0.0, 1.0, -3.0])
    x = np.array([5.0, -100.0, 3.0, 0.0, 0.0, 13.0, -1.0, -1e8, -7.0])
    krep, slen = np.array([1e5, 1e4, 1e3]), np.array([1.0, 1.5, 2.0])
    nn, bs = np.array([3]), np.array([0, 3, 6])
    expect = 1e5 + 1e4 * math.exp(-10 / 1.5) + 1e3 * math.exp(10 / 2.0)
    assert oracle.bd_energy(x, plane, nn, bs, krep, slen) == pytest.approx(expect, rel=1e-14)
    f = oracle.bd_forces(x, plane, nn, bs, krep, slen).reshape(3, 3)
    for i, r in enumerate([0.0, 10.0, -10.0]):
        np.testing.assert_allclose(f[i], [0, 0, krep[i] / slen[i] * math.exp(-r / slen[i])], rtol=1e-14)


def test_plane_distance_kat(oracle):
    """test_boundaries.cpp:45-49: plane through (10,10,10) with normal +x;
    (15,10,10) is 5 above and (5,10,10) 5 below (K = lambda = 1)."""
    plane = np.array([1.0, 0.0, 0.0, -10.0])
    x = np.array([15.0, 10, 10, 5, 10, 10])
    e_above = oracle.bd_energy(x, plane, np.array([1]), np.array([0]), np.array([1.0]), np.array([1.0]))
    e_below = oracle.bd_energy(x, plane, np.array([1]), np.array([3]), np.array([1.0]), np.array([1.0]))
    assert e_above == pytest.approx(math.exp(-5.0), rel=1e-15)
    assert e_below == pytest.approx(math.exp(5.0), rel=1e-15)


# --------------------------------------------------------------------------
# energy / force consistency, Newton's third law, rigid motions
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["fs", "fb", "ls", "bs", "bb"])
def test_bonded_force_is_minus_gradient(oracle, name):
    x, bs, params = ff_case(name)
    f = oracle.ff_forces(name, x, bs, *params)
    g = central_diff(lambda y: oracle.ff_energy(name, y, bs, *params), x, 1e-5)
    np.testing.assert_allclose(f, -g, rtol=2e-6, atol=2e-6 * np.abs(f).max())


@pytest.mark.parametrize("name", ["fs", "fb", "ls", "bs", "bb"])
def test_bonded_newton_third_law_and_invariance(oracle, name):
    x, bs, params = ff_case(name)
    f = oracle.ff_forces(name, x, bs, *params).reshape(-1, 3)
    assert np.abs(f.sum(axis=0)).max() <= 1e-10 * np.abs(f).max()
    R = rot(0.3, -1.1, 2.0)
    y = (x.reshape(-1, 3) @ R.T + np.array([13.0, -7.0, 101.0])).ravel()
    e0, e1 = oracle.ff_energy(name, x, bs, *params), oracle.ff_energy(name, y, bs, *params)
    assert e1 == pytest.approx(e0, rel=1e-9, abs=1e-9)


def test_boundary_force_is_minus_gradient(oracle):
    plane = np.array([0.0, 0.0, 1.0, -3.0])
    x = np.array([5.0, -100.0, 3.5, 0.0, 0.0, 6.0, -1.0, 2.0, 1.0])
    args = (plane, np.array([3]), np.array([0, 3, 6]), np.array([41.0, 41.0, 41.0]), np.array([2.7, 2.7, 2.7]))
    f = oracle.bd_forces(x, *args)
    g = central_diff(lambda y: oracle.bd_energy(y, *args), x, 1e-6)
    np.testing.assert_allclose(f, -g, rtol=1e-7, atol=1e-9)


# --------------------------------------------------------------------------
# excluded volume
# --------------------------------------------------------------------------
EV_SYMMETRIC = dict(bs=[0, 3, 6, 9], k=[0.25], eql=[20.0, 20.0],
                    x=[10, 0, 0, -10, 0, 0, 0, 10, 5, 0, -10, 5])
EV_GENERAL = dict(bs=[0, 3, 6, 9, 0, 3, 12, 15], k=[100.0, 2e3], eql=[20.0, 20.0, 20.0, 100.0],
                  x=[10, 0, 0, -10, 0, 0, 0, 10, 5, 3, -10, 5, 3, 70, 30, 2, -20, 0])


@functools.lru_cache(maxsize=8)
def _leggauss01(n):
    t, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (t + 1), 0.5 * w


def gauss_ev(c1, c2, c3, c4, kr, n=400):
    """kr * Int_0^1 Int_0^1 ds dt / |P(s) - Q(t)|^4 by tensor Gauss-Legendre:
    an independent quadrature of the integral the analytic form evaluates."""
    t, w = _leggauss01(n)
    P = c1[None, :] + t[:, None] * (c2 - c1)[None, :]
    Q = c3[None, :] + t[:, None] * (c4 - c3)[None, :]
    d2 = ((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1)
    return kr * float(w @ (1.0 / d2 ** 2) @ w)


def test_ev_scaling_fixture(oracle):
    """TestCylinderExclVolume.cpp:13-47: the energy of two cylinders equals the
    sum over their sub-segments (k scaled by the segment lengths)."""
    coord = np.array([-60.0, 0, 0, 0, 0, 0, 60, 0, 0, -40, -40, 10, -28, -24, 10, -10, 0, 10, 20, 40, 10])
    e0 = oracle.ev_energy(coord, [0, 6, 9, 18], [1e-2], [100.0, 100.0])
    e1 = oracle.ev_energy(coord, [0, 3, 9, 12, 0, 3, 12, 15, 0, 3, 15, 18, 3, 6, 9, 12, 3, 6, 12, 15, 3, 6, 15, 18],
                          [1e-2] * 6, [50.0, 20.0, 50.0, 30.0, 50.0, 50.0, 50.0, 20.0, 50.0, 30.0, 50.0, 50.0])
    # the reference requires Approx equality; the integral is additive, so it holds tightly
    assert e1 == pytest.approx(e0, rel=1e-10)


def test_ev_monomer_consistency_fixture(oracle):
    """TestCylinderExclVolume.cpp:49-100: the integral formula agrees with the
    monomer-sum formula (40 monomers each, interval 1 and 4) within 1e-2."""
    x = np.array([10.0, 0, 0, -10, 0, 0, 0, 10, 5, 0, -10, 5])
    eq, kvol = 20.0, 0.25
    e_int = oracle.ev_energy(x, [0, 3, 6, 9], [kvol], [eq, eq])
    c11, c12, c21, c22 = x[0:3], x[3:6], x[6:9], x[9:12]

    def mons(lo, hi, step):
        start = -((-lo) // step) * step     # ceildiv(lo, step) * step
        return list(range(start, hi, step))

    n1 = n2 = 40
    coeff = kvol * (1 * eq / 40) * (4 * eq / 40)
    e_mon = 0.0
    for m1 in mons(0, 40, 1):
        p = c11 + (m1 - 0 + 0.5) / n1 * (c12 - c11)
        for m2 in mons(-9, 31, 4):
            q = c21 + (m2 + 9 + 0.5) / n2 * (c22 - c21)
            d2 = ((p - q) ** 2).sum()
            e_mon += coeff / d2 ** 2
    assert abs(e_int - e_mon) <= min(abs(e_int), abs(e_mon)) * 1e-2


@pytest.mark.parametrize("case", [EV_SYMMETRIC, EV_GENERAL])
def test_ev_against_independent_quadrature(oracle, case):
    x = np.array(case["x"], float)
    bs, k, eql = case["bs"], case["k"], case["eql"]
    total = 0.0
    for i in range(len(k)):
        c = [x[bs[4 * i + j]: bs[4 * i + j] + 3] for j in range(4)]
        kr = k[i] * eql[2 * i] * eql[2 * i + 1]
        total += gauss_ev(*c, kr)
    assert oracle.ev_energy(x, bs, k, eql) == pytest.approx(total, rel=1e-9)


@pytest.mark.parametrize("case", [EV_SYMMETRIC, EV_GENERAL])
def test_ev_force_is_minus_gradient(oracle, case):
    """TestCylinderExclVolume.cpp:104-243 (energy / force consistency), here
    per coordinate by central differences."""
    x = np.array(case["x"], float)
    bs, k, eql = case["bs"], case["k"], case["eql"]
    f = oracle.ev_forces(x, bs, k, eql)
    g = central_diff(lambda y: oracle.ev_energy(y, bs, k, eql), x, 1e-5)
    np.testing.assert_allclose(f, -g, rtol=1e-6, atol=1e-7 * np.abs(f).max())
    assert np.abs(f.reshape(-1, 3).sum(axis=0)).max() <= 1e-10 * np.abs(f).max()


def test_ev_swap_symmetry_and_rigid_invariance(oracle):
    x = np.array(EV_GENERAL["x"], float)
    a = oracle.ev_energy(x, [0, 3, 6, 9], [100.0], [20.0, 20.0])
    b = oracle.ev_energy(x, [6, 9, 0, 3], [100.0], [20.0, 20.0])
    assert b == pytest.approx(a, rel=1e-12)
    c = oracle.ev_energy(x, [3, 0, 9, 6], [100.0], [20.0, 20.0])   # both cylinders reversed
    assert c == pytest.approx(a, rel=1e-12)
    R = rot(1.0, 0.2, -0.7)
    y = (x.reshape(-1, 3) @ R.T + 50.0).ravel()
    assert oracle.ev_energy(y, [0, 3, 6, 9], [100.0], [20.0, 20.0]) == pytest.approx(a, rel=1e-10)


def test_ev_simpson_fallback_converges_first_order(oracle):
    """energyN / forceN (CylinderExclVolRepulsion.cpp:884-1254). The reference
    advances each abscissa before its first use, so interior nodes sit one
    step late and the rules converge to the integral (and to minus its
    gradient) at first order in 1/N: the error halves when N doubles. The
    exact outputs of the reference's own run stay unpinned."""
    c = [np.array(v, float) for v in ([10, 0, 0], [-10, 0, 0], [0, 10, 5], [3, -10, 5])]
    kr = 100.0 * 20 * 20
    exact = gauss_ev(*c, kr)
    err = {n: oracle.ev_pair_energy_simpson(*c, kr, n, n) / exact - 1 for n in (200, 400)}
    assert abs(err[400]) < 1e-3
    assert 1.8 < err[200] / err[400] < 2.3
    x = np.concatenate(c)
    g = central_diff(lambda y: oracle.ev_energy(y, [0, 3, 6, 9], [100.0], [20.0, 20.0]), x, 1e-5).reshape(4, 3)
    ferr = {n: np.abs(oracle.ev_pair_force_simpson(*c, kr, n, n) + g).max() / np.abs(g).max() for n in (100, 200)}
    assert ferr[200] < 2e-3
    assert 1.8 < ferr[100] / ferr[200] < 2.3


def test_ev_parallel_pair_uses_quadrature(oracle):
    """Exactly parallel cylinders make the analytic form 0/0; the energy then
    comes from the Simpson rule (CylinderExclVolRepulsion.cpp:442-481)."""
    x = np.array([0, 0, 0, 100, 0, 0, 0, 20, 0, 100, 20, 0], float)
    e = oracle.ev_energy(x, [0, 3, 6, 9], [1.0], [1.0, 1.0])
    c = [x[i:i + 3] for i in (0, 3, 6, 9)]
    assert np.isfinite(e)
    assert e == pytest.approx(gauss_ev(*c, 1.0), rel=1e-6)


def test_ev_coplanar_perturbation_invariant(oracle):
    """Coplanar pairs: the reference force is the analytic force of the
    geometry with c2 moved 0.01 nm along the plane normal
    (CylinderExclVolRepulsion.cpp:610-626; MathFunctions.h:1221-1270)."""
    c1, c2, c3, c4 = (np.array(v, float) for v in ([0, 0, 0], [100, 0, 0], [20, 30, 0], [90, 50, 0]))
    assert oracle.are_in_plane(c1, c2, c3, c4)
    x = np.concatenate([c1, c2, c3, c4])
    f_ref = oracle.ev_forces(x, [0, 3, 6, 9], [1.0], [1.0, 1.0], coplanar_mode=0)
    nrm = np.cross(c2 - c1, c3 - c1)
    nrm /= np.linalg.norm(nrm)
    y = np.concatenate([c1, c2 + 0.01 * nrm, c3, c4])
    f_moved = oracle.ev_forces(y, [0, 3, 6, 9], [1.0], [1.0, 1.0], coplanar_mode=1)
    np.testing.assert_allclose(f_ref, f_moved, rtol=1e-12, atol=1e-14 * np.abs(f_moved).max())
