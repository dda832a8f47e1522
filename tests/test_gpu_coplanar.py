"""The reference's near-coplanar stress input on the device
(TestCylinderExclVolume.cpp:166-188, the failure point of a simulation of
Sep 1, 2021; tests/coplanar_fixture.py): its 140 cylinders as filaments,
with an excluded-volume cutoff of 150 nm so that the neighbor list holds all
197 of the fixture's interactions (the one the reference singles out,
coplanar to a triple product of 9e-8, among them).

* neighbor list and tables bit-exact; energies and forces bit for bit
  against the exact-mode oracle, in both coplanar modes;
* against the reference-order oracle: forces within 1e-10 of max|F| outside
  ill-conditioned pairs; the energy differs only by the unstable pair's own
  difference between the two arithmetic modes (last-ulp amplification, shown
  on the CPU in test_oracle_coplanar.py);
* a minimization from this state to 1 pN: bitwise trajectory against the
  exact-mode oracle.
"""
import numpy as np
import pytest

from coplanar_fixture import as_network, load
from medyan_amd import _abi, mechanics
from test_gpu_parity import bits_equal, ill_conditioned_beads

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF


@pytest.fixture(scope="module")
def net():
    n = as_network()
    n.exvol_cutoff = 150.0
    return n


def ctx_for(net, coplanar=_abi.MDY_COPLANAR_REFERENCE):
    ctx = mechanics.Context(0, coplanar_mode=coplanar)
    ctx.set_network(net)
    ctx.build_neighbors()
    return ctx


def test_fixture_pairs_all_listed(oracle, net):
    bs, _, _, _, bad = load()
    ctx = ctx_for(net)
    np.testing.assert_array_equal(ctx.pairs(), oracle.nl_pairs(net))
    got = {tuple(r) for r in ctx.exvol_table()[0]}
    fix = [tuple(r) for r in bs.reshape(-1, 4)]
    assert all(p in got or (p[2], p[3], p[0], p[1]) in got for p in fix)
    ctx.close()


@pytest.mark.parametrize("coplanar", [_abi.MDY_COPLANAR_REFERENCE, _abi.MDY_COPLANAR_EXACT])
def test_energy_forces_bitwise_exact_mode(oracle, net, coplanar):
    ctx = ctx_for(net, coplanar)
    s = oracle.system(net, coplanar_mode=coplanar)
    oracle.set_exact(1)
    try:
        e, _ = ctx.energy(net.coord)
        eo, _ = s.energy()
        assert bits_equal([e], [eo])
        assert bits_equal(ctx.forces(net.coord), s.forces())
    finally:
        oracle.set_exact(0)
    ctx.close()


def test_against_reference_order(oracle, net):
    bs, eql, x, k, bad = load()
    ctx = ctx_for(net)
    s = oracle.system(net)
    e, _ = ctx.energy(net.coord)
    eo, _ = s.energy()
    f, fo = ctx.forces(net.coord), s.forces()
    ok = np.ones(net.n_beads, bool)
    ill = ill_conditioned_beads(net, s)
    ok[ill] = False
    diff = np.abs(f - fo).reshape(-1, 3).max(1)
    assert diff[ok].max() <= 1e-10 * np.abs(fo).max()
    # the energy difference is the unstable pair's own mode difference
    X = x.reshape(-1, 3)
    a, b, c, d = bs.reshape(-1, 4)[bad] // 3
    kr = k * eql[2 * bad] * eql[2 * bad + 1]
    oracle.set_exact(1)
    try:
        ex = oracle.ev_pair_energy(X[a], X[b], X[c], X[d], kr)
    finally:
        oracle.set_exact(0)
    d68 = abs(ex - oracle.ev_pair_energy(X[a], X[b], X[c], X[d], kr))
    print(f"[gpu] energy vs reference order: {abs(e - eo):.3e}; unstable pair's mode difference {d68:.3e}; "
          f"{len(ill)} beads in ill-conditioned pairs")
    assert abs(e - eo) <= d68 + 1e-10 * abs(eo)
    ctx.close()


def test_minimization_from_the_failure_point(oracle, net):
    ctx = ctx_for(net)
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0)
    x, f, res = ctx.minimize(p, trace_cap=100000)
    tr = ctx.trace()
    oracle.set_exact(1)
    try:
        s = oracle.system(net)
        xo, fo, reso, tro = s.minimize(p, trace_cap=100000)
    finally:
        oracle.set_exact(0)
    assert res.status == reso.status and res.num_iter == reso.num_iter
    for key in ("lambda_", "energy", "maxf", "beta"):
        assert bits_equal(tr[key], tro[key]), key
    assert bits_equal(x, xo) and bits_equal(f, fo)
    print(f"[gpu] from the failure point: {res.num_iter} iterations, status {res.status}, "
          f"max|F| {res.grad_inf_norm:.3f}")
    ctx.close()
