"""Pins of the oracle's quadratic branching dihedral (BRDIHEDRALFFTYPE
QUADRATIC, BranchingDihedralQuadratic.cpp:26-398), E = kdih theta^2, and of
the shared acos it uses in exact mode.

* The reference's own energy-force consistency inputs
  (TestBranchingDihedralQuadratic.cpp:21-75: small, two near-90 degree and a
  large dihedral, and the branch point at the mother's plus end) under its
  test (random central differences, TestFFCommon.hpp:137-170), with a
  tolerance 1e-7 (of the summed magnitudes of the terms, which partly
  cancel) instead of its 5e-4 relative, every repetition required to pass.
* Closed forms: the daughter's far bead turned by a about the daughter axis
  out of the mother-daughter plane gives E = k a^2 and the force
  -2 k a t on the far bead (t the unit tangent of the turn).
* Internal forces: they sum to zero and exert no net torque.
* Per-coordinate central differences on random geometries and positions.
* mdy_acos against the C library's acos (within 2 ulp; exact at +-1).
"""
import numpy as np
import pytest

from medyan_amd import _abi, network

REF_INPUT = dict(  # TestBranchingDihedralQuadratic.cpp:21-75
    x=[0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.5, 1.0, 0.1,
       0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.5, 0.1, 1.0,
       0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.5, -0.1, 1.0,
       0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.5, -1.0, 0.1,
       0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.5, -1.0, 0.1],
    pos=[0.6, 0.5, 0.5, 0.4, 1.0], kdih=[1.0] * 5)


@pytest.fixture(params=[0, 1], ids=["reference", "exact"])
def mode(oracle, request):
    oracle.set_exact(request.param)
    yield oracle
    oracle.set_exact(0)


def test_acos_against_c_library(oracle):
    rng = np.random.default_rng(3)
    x = np.concatenate([np.linspace(-1, 1, 200001), rng.uniform(-1, 1, 200000),
                        1 - np.logspace(-17, -1, 2000), -1 + np.logspace(-17, -1, 2000),
                        np.logspace(-300, -1, 2000), [-1.0, -0.5, 0.0, 0.5, 1.0, -0.0]])
    x = x[(x >= -1) & (x <= 1)]
    y, ref = oracle.xmath_acos(x), np.arccos(x)
    ulp = np.abs(y - ref) / np.spacing(ref)
    assert ulp.max() <= 2.0, (ulp.max(), x[np.argmax(ulp)])
    assert oracle.xmath_acos([1.0])[0] == 0.0 and oracle.xmath_acos([-1.0])[0] == np.pi


def test_reference_energy_force_consistency(mode):
    x0 = np.array(REF_INPUT["x"])
    bs = 3 * np.arange(20)
    args = (bs, REF_INPUT["kdih"], REF_INPUT["pos"])
    f = mode.ff_forces("bdhq", x0, *args)
    rng = np.random.default_rng(11)
    for _ in range(10):
        dc = rng.normal(0.0, 1e-5, x0.shape)
        de = mode.ff_energy("bdhq", x0 + dc, *args) - mode.ff_energy("bdhq", x0 - dc, *args)
        expected = -2 * float(dc @ f)
        # relative to the size of the terms: the five contributions partly cancel
        assert abs(de - expected) <= 1e-7 * 2 * float(np.abs(dc * f).sum())


def turned(a, pos=0.5):
    """Mother (0,-1,0)-(0,1,0), branch point at the origin for pos 0.5,
    daughter from (1,0,0) with its far bead turned by a about the x axis."""
    return np.array([0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.0, np.cos(a), np.sin(a)])


@pytest.mark.parametrize("a", [0.0, 0.3, 1.0, np.pi / 2, 2.0, 3.0])
def test_closed_form(mode, a):
    k = 3.0
    x = turned(a)
    u = mode.ff_energy("bdhq", x, [0, 3, 6, 9], [k], [0.5])
    assert u == pytest.approx(k * a * a, rel=1e-12, abs=1e-30)
    f = mode.ff_forces("bdhq", x, [0, 3, 6, 9], [k], [0.5]).reshape(4, 3)
    t = np.array([0.0, -np.sin(a), np.cos(a)])
    np.testing.assert_allclose(f[3], -2 * k * a * t, rtol=0, atol=1e-12 * max(1.0, 2 * k * a))
    np.testing.assert_allclose(f[0] + f[1] + f[2] + f[3], 0.0, atol=1e-12 * max(1.0, 2 * k * a))


def rand_geometry(rng):
    x = np.array([0.0, -1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 1.2, 0.4, 0.7]) + rng.normal(0, 0.3, 12)
    return x


@pytest.mark.parametrize("pos", [0.3, 0.5, 0.8, 1.0])
def test_forces_are_energy_gradient(mode, pos):
    rng = np.random.default_rng(int(pos * 10) + 40)
    for _ in range(4):
        x = rand_geometry(rng)
        args = ([0, 3, 6, 9], [2.0], [pos])
        f = mode.ff_forces("bdhq", x, *args)
        h = 1e-6
        cd = np.array([-(mode.ff_energy("bdhq", x + h * e, *args) - mode.ff_energy("bdhq", x - h * e, *args)) / (2 * h)
                       for e in np.eye(12)])
        np.testing.assert_allclose(f, cd, rtol=0, atol=1e-7 * max(1.0, np.abs(cd).max()))
        F = f.reshape(4, 3)
        R = x.reshape(4, 3)
        scale = np.abs(F).max() * np.abs(R).max()
        assert np.abs(F.sum(0)).max() <= 1e-12 * max(1.0, np.abs(F).max())
        assert np.abs(np.cross(R, F).sum(0)).max() <= 1e-11 * max(1.0, scale)   # no net torque


def test_modes_differ_by_rounding_only(oracle):
    rng = np.random.default_rng(8)
    xs = [rand_geometry(rng) for _ in range(50)]
    out = {}
    for m in (0, 1):
        oracle.set_exact(m)
        try:
            out[m] = [(oracle.ff_energy("bdhq", x, [0, 3, 6, 9], [2.0], [0.7]),
                       oracle.ff_forces("bdhq", x, [0, 3, 6, 9], [2.0], [0.7])) for x in xs]
        finally:
            oracle.set_exact(0)
    for (e0, f0), (e1, f1) in zip(out[0], out[1]):
        assert abs(e0 - e1) <= 1e-13 * abs(e0)
        np.testing.assert_allclose(f1, f0, rtol=0, atol=1e-12 * np.abs(f0).max())


def quadratic(net):
    return net.with_branching_dihedral("QUADRATIC")


def test_dihedral_type_names(lib):
    net = network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, seed=4)
    assert net.br_dihedral_type == _abi.MDY_BRDIH_COSINE
    q = net.with_branching_dihedral("QUADRATIC")
    assert q.br_dihedral_type == _abi.MDY_BRDIH_QUADRATIC and net.br_dihedral_type == _abi.MDY_BRDIH_COSINE
    assert q.with_branching_dihedral("COSINE").br_dihedral_type == _abi.MDY_BRDIH_COSINE
    assert q.system().br_dihedral_type == _abi.MDY_BRDIH_QUADRATIC
    for v2 in ("COSINEV2", "QUADRATICV2"):
        with pytest.raises(NotImplementedError):
            net.with_branching_dihedral(v2)
    with pytest.raises(ValueError, match="not recognized"):
        net.with_branching_dihedral("HARMONIC")


def test_system_uses_the_selected_form(oracle, lib):
    base = network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, seed=4,
                            k_br_dihedral=10.0, k_br_position=20.0)
    net = quadratic(base)
    assert len(net.bdh_k) > 0
    e_cos = oracle.system(base).energy(base.coord, individual=True)[1]
    s = oracle.system(net)
    e, per = s.energy(net.coord, individual=True)
    args = (net.bdh_beadset, net.bdh_k, net.bdh_pos)
    assert per[_abi.MDY_FF_BDH] == oracle.ff_energy("bdhq", net.coord, *args)
    assert per[_abi.MDY_FF_BDH] != e_cos[_abi.MDY_FF_BDH]
    others = [i for i in range(_abi.MDY_NFF) if i != _abi.MDY_FF_BDH]
    np.testing.assert_array_equal(per[others], e_cos[others])
    f_q = oracle.ff_forces("bdhq", net.coord, *args)
    f_c = oracle.ff_forces("bdh", net.coord, *args)
    np.testing.assert_allclose(s.forces(net.coord) - oracle.system(base).forces(base.coord), f_q - f_c,
                               rtol=0, atol=1e-9 * np.abs(f_q).max())


def test_system_minimizes(oracle, lib):
    net = quadratic(network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, seed=4,
                                     k_br_dihedral=10.0, k_br_position=20.0))
    s = oracle.system(net)
    p = _abi.default_cg_params(net.n_beads)
    x, f, res, _ = s.minimize(p)
    assert res.status == 0 and res.error_flags == 0
    assert res.grad_inf_norm < p.grad_tol or res.energy_rel_change < p.energy_rel_tol
    np.testing.assert_array_equal(f, s.forces(x))
