"""Branching dihedral and position force fields (cosine forms) in the oracle,
pinned by closed forms and by energy-force consistency, as the reference's
own tests do (TestBranchingDihedralCosine.cpp, TestBranchingPositionCosine.cpp:
the same input geometries, central differences of the energy against the
forces, including the branch point at the mother's plus end):
* position: U = k (1 - sin theta) for the angle theta between the mother's
  plus direction and the daughter, measured at the branch point;
* dihedral: U = k (1 - cos a) when the daughter's far bead turns by a about
  the daughter axis out of the plane of mother and daughter."""
import numpy as np
import pytest

REF_POSITION = [  # (c1, c2, c3, pos) from TestBranchingPositionCosine.cpp:26-58
    ([0.0, -1.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0], 0.6),
    ([0.0, -1.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0], 0.5),
    ([0.0, -1.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0], 0.4),
    ([0.0, -1.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0], 1.0),
]


def test_position_closed_form(oracle):
    for c1, c2, c3, p in REF_POSITION + [([0, -1, 0], [0, 1, 0], [1.0, 0.7, 0.2], 0.3)]:
        x = np.array(c1 + c2 + c3, float)
        u = oracle.ff_energy("bp", x, [0, 3, 6], [2.0], [p])
        c1, c2, c3 = map(np.array, (c1, c2, c3))
        if p == 1.0:
            c2 = 2.0 * (c2 - 0.5 * c1)
            p = 0.5
        mp = c1 * (1 - p) + p * c2
        a, b = c2 - mp, c3 - mp
        cos = a @ b / np.linalg.norm(a) / np.linalg.norm(b)
        assert np.isclose(u, 2.0 * (1 - np.sqrt(1 - cos * cos)), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("ang", [0.0, 0.3, np.pi / 2, 2.0])
def test_dihedral_closed_form(oracle, ang):
    c1, c2, c3 = [0.0, -1.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 0.0]
    c4 = [1.0, np.cos(ang), np.sin(ang)]
    x = np.array(c1 + c2 + c3 + c4, float)
    u = oracle.ff_energy("bdh", x, [0, 3, 6, 9], [3.0], [0.5])
    assert np.isclose(u, 3.0 * (1 - np.cos(ang)), rtol=1e-12, atol=1e-14)


def _consistency(oracle, name, x, bs, params, h=1e-6):
    f = oracle.ff_forces(name, x, bs, *params)
    cd = np.zeros_like(x)
    for i in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        cd[i] = -(oracle.ff_energy(name, xp, bs, *params) - oracle.ff_energy(name, xm, bs, *params)) / (2 * h)
    return f, cd


@pytest.mark.parametrize("case", range(4))
def test_position_forces_are_energy_gradient(oracle, case):
    c1, c2, c3, p = REF_POSITION[case]
    rng = np.random.default_rng(case)
    x = np.array(c1 + c2 + c3, float) + rng.normal(0, 0.05, 9)   # generic geometry near the reference input
    f, cd = _consistency(oracle, "bp", x, [0, 3, 6], [[1.0], [p]])
    np.testing.assert_allclose(f, cd, rtol=0, atol=1e-7 * max(1.0, np.abs(cd).max()))
    assert np.abs(f.reshape(-1, 3).sum(0)).max() < 1e-12          # internal forces sum to zero


@pytest.mark.parametrize("pos", [0.3, 0.5, 0.8, 1.0])
def test_dihedral_forces_are_energy_gradient(oracle, pos):
    rng = np.random.default_rng(int(pos * 10))
    x = np.array([0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.2, 0.4, 0.7]) + rng.normal(0, 0.05, 12)
    f, cd = _consistency(oracle, "bdh", x, [0, 3, 6, 9], [[2.0], [pos]])
    np.testing.assert_allclose(f, cd, rtol=0, atol=1e-7 * max(1.0, np.abs(cd).max()))
    assert np.abs(f.reshape(-1, 3).sum(0)).max() < 1e-12
