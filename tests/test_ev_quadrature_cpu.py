"""Pins tests/ev_quadrature.py, the independent evaluation that the device's
quadrature mode is checked against: on well-conditioned pairs its energy and
forces equal the oracle's closed form (CylinderExclVolRepulsion.cpp) to
rounding, its forces are minus the central differences of its energy, and
its H v equals central differences of its forces. No GPU."""
import numpy as np
import pytest

import ev_quadrature as evq


def random_pairs(n, seed, coplanar_offset=None):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        c1 = rng.uniform(0, 200, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        c2 = c1 + 108.0 * u
        c3 = c1 + rng.uniform(-60, 160) * u + rng.uniform(15, 90) * np.cross(u, v) / max(np.linalg.norm(np.cross(u, v)), 1e-3)
        c4 = c3 + 108.0 * v
        if coplanar_offset is not None:   # move the second cylinder into (near) the first's plane with v
            n_ = np.cross(u, v)
            n_ /= np.linalg.norm(n_)
            off = (c3 - c1) @ n_
            c3 = c3 - (off - coplanar_offset) * n_
            c4 = c4 - (off - coplanar_offset) * n_
        if evq.segment_distance(c1, c2, c3, c4) > 8.0:
            out.append((c1, c2, c3, c4))
    return out


def test_energy_and_forces_equal_closed_form_on_well_conditioned_pairs(oracle):
    oracle.set_exact(1)
    try:
        for c1, c2, c3, c4 in random_pairs(12, 3):
            if evq.triple_q(c1, c2, c3, c4) < 1e-2:
                continue
            k, l1, l2 = 1e5, 108.0, 108.0
            E, F = evq.pair(c1, c2, c3, c4, k * l1 * l2)
            x = np.concatenate([c1, c2, c3, c4])
            eo = oracle.ev_energy(x, [0, 3, 6, 9], [k], [l1, l2])
            fo = oracle.ev_forces(x, [0, 3, 6, 9], [k], [l1, l2], coplanar_mode=1).reshape(4, 3)
            assert E == pytest.approx(eo, rel=1e-12)
            assert np.abs(F - fo).max() <= 1e-10 * np.abs(fo).max()
    finally:
        oracle.set_exact(0)


@pytest.mark.parametrize("offset", [None, 1e-3])
def test_forces_and_hvp_are_derivatives(offset):
    rng = np.random.default_rng(9)
    for c in random_pairs(4, 5, coplanar_offset=offset):
        kr = 1e5 * 108.0 * 108.0
        v = rng.normal(size=(4, 3))
        E, F, Hv = evq.pair(*c, kr, v=v)
        h = 1e-4
        # forces: -dE/dc by central differences, one bead coordinate at a time
        for b in range(4):
            for a in range(3):
                cp = [x.copy() for x in c]
                cm = [x.copy() for x in c]
                cp[b][a] += h
                cm[b][a] -= h
                fd = -(evq.pair(*cp, kr)[0] - evq.pair(*cm, kr)[0]) / (2 * h)
                assert fd == pytest.approx(F[b, a], rel=1e-6, abs=1e-7 * np.abs(F).max())
        cp = [x + h * v[i] for i, x in enumerate(c)]
        cm = [x - h * v[i] for i, x in enumerate(c)]
        fd = -(evq.pair(*cp, kr)[1] - evq.pair(*cm, kr)[1]) / (2 * h)
        assert np.abs(fd - Hv).max() <= 1e-6 * np.abs(Hv).max()


def test_near_coplanar_closed_form_drifts_quadrature_does_not(oracle):
    """Where the flag applies (q < 1e-3) the closed form's rounding is
    amplified: at q ~ 1e-12 its forces move by far more than the geometry's
    own change under a 1e-12 nm displacement; the quadrature's do not."""
    c1, c2, c3, c4 = random_pairs(1, 11, coplanar_offset=1e-4)[0]
    assert evq.triple_q(c1, c2, c3, c4) < 1e-3
    k = 1e5
    x = np.concatenate([c1, c2, c3, c4])
    rng = np.random.default_rng(2)
    oracle.set_exact(1)
    try:
        f0 = oracle.ev_forces(x, [0, 3, 6, 9], [k], [108.0, 108.0], coplanar_mode=1)
        _, q0 = evq.pair(c1, c2, c3, c4, k * 108.0 * 108.0)
        dx = rng.normal(size=12) * 1e-12
        f1 = oracle.ev_forces(x + dx, [0, 3, 6, 9], [k], [108.0, 108.0], coplanar_mode=1)
        y = (x + dx).reshape(4, 3)
        _, q1 = evq.pair(y[0], y[1], y[2], y[3], k * 108.0 * 108.0)
    finally:
        oracle.set_exact(0)
    scale = np.abs(q0).max()
    assert np.abs(q1 - q0).max() <= 1e-12 * scale
    assert np.abs(f1 - f0).max() > 1e3 * np.abs(q1 - q0).max()
