"""The quadratic branching dihedral (BRDIHEDRALFFTYPE QUADRATIC,
BranchingDihedralQuadratic.cpp:26-398) on the device against the oracle,
whose restatement is pinned in test_oracle_bdh_quadratic.py: energies and
forces bit for bit against the exact-mode oracle and within 1e-10 of the
reference-order oracle, bit-identical minimization trajectories, the
mixed-precision form, branch-point stretch forces and Hessian-vector
products."""
import ctypes as C

import numpy as np
import pytest
import torch

from medyan_amd import _abi, autodiff, mechanics, network
from test_gpu_parity import bits_equal, ctx_for, exact  # noqa: F401  (exact is a fixture)

pytestmark = pytest.mark.gpu


def quadratic(net):
    return net.with_branching_dihedral("QUADRATIC")


@pytest.fixture(scope="module")
def qnets(lib):
    return {
        "branched_full": quadratic(network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8,
                                                    seed=4, k_br_dihedral=10.0, k_br_position=20.0)),
        "many_branches": quadratic(network.generate(lib, "net500", n_filaments=60, cyl_per_filament=4,
                                                    n_branches=20, seed=6, k_br_dihedral=10.0, k_br_position=20.0)),
        # with the monomer-sum excluded volume as well (every optional form at once)
        "with_monomer_exvol": quadratic(network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5,
                                                         n_branches=8, seed=4, k_br_dihedral=10.0,
                                                         k_br_position=20.0).with_monomer_exvol(2)),
    }


CASES = ["branched_full", "many_branches", "with_monomer_exvol"]


@pytest.mark.parametrize("case", CASES)
def test_energies_forces_bitwise(exact, qnets, case):
    net = qnets[case]
    print(f"[gpu] quadratic branching dihedral, {case}", flush=True)
    ctx = ctx_for(net)
    s = exact.system(net)
    rng = np.random.default_rng(5)
    for x in (net.coord, net.coord + rng.normal(0, 0.5, net.coord.shape)):
        e, per = ctx.energy(x)
        eo, pero = s.energy(x)
        assert bits_equal([e], [eo])
        np.testing.assert_allclose(per, pero, rtol=1e-13, atol=0)
        assert per[_abi.MDY_FF_BDH] > 0
        assert bits_equal(ctx.forces(x), s.forces(x))


@pytest.mark.parametrize("case", CASES)
def test_reference_order_within_1e10(oracle, qnets, case):
    net = qnets[case]
    ctx = ctx_for(net)
    s = oracle.system(net)
    e, _ = ctx.energy(net.coord)
    eo, _ = s.energy(net.coord)
    assert abs(e - eo) <= 1e-10 * abs(eo)
    f, fo = ctx.forces(net.coord), s.forces(net.coord)
    assert np.abs(f - fo).max() <= 1e-10 * np.abs(fo).max()


@pytest.mark.parametrize("case", CASES)
def test_minimization_trajectory_bitwise(exact, qnets, case):
    net = qnets[case]
    ctx = ctx_for(net)
    s = exact.system(net)
    p = _abi.default_cg_params(net.n_beads)
    x, f, res = ctx.minimize(p, trace_cap=100000)
    tr = ctx.trace()
    xo, fo, reso, tro = s.minimize(p, trace_cap=100000)
    assert res.status == 0 and reso.status == 0 and res.error_flags == 0
    assert (res.num_iter, res.num_energy_call, res.num_force_call) == (
        reso.num_iter, reso.num_energy_call, reso.num_force_call)
    for key in ("lambda_", "energy", "maxf", "beta"):
        assert bits_equal(tr[key], tro[key]), key
    assert bits_equal(x, xo) and bits_equal(f, fo)
    print(f"[gpu] {case}: {res.num_iter} iterations, max|F| {res.grad_inf_norm:.3f}")


def test_mixed_precision_form(exact, qnets):
    net = qnets["many_branches"]
    ctx = mechanics.Context(precision=_abi.MDY_PREC_MIXED)
    ctx.set_network(net)
    ctx.build_neighbors()
    s = exact.system(net)
    fo = s.forces(net.coord)
    assert np.abs(ctx.forces(net.coord) - fo).max() <= 1e-6 * np.abs(fo).max()
    p = _abi.default_cg_params(net.n_beads)
    x, f, res = ctx.minimize(p)
    assert res.status == 0 and res.error_flags == 0
    assert res.grad_inf_norm < p.grad_tol or res.energy_rel_change < p.energy_rel_tol


def test_branch_stretch_forces(oracle, qnets):
    net = qnets["branched_full"]
    ctx = ctx_for(net)
    x, _, res = ctx.minimize(_abi.default_cg_params(net.n_beads, energy_rel_tol=0.0))
    assert res.status == 0
    _, _, br = ctx.stretch_forces(x)
    oracle.set_exact(1)
    try:
        _, _, obr = oracle.system(net).stretch_forces(x)
    finally:
        oracle.set_exact(0)
    assert bits_equal(br.ravel(), obr)


def test_hvp(oracle, qnets):
    """H v of the quadratic form minus H v of the cosine form on the same
    network isolates the dihedral's Hessian (every other force field's part
    cancels); checked against a Richardson-extrapolated central difference
    of the oracle's dihedral forces (truncation error ~1e-10 here)."""
    q = qnets["branched_full"]
    c = q.copy()
    c.br_dihedral_type = _abi.MDY_BRDIH_COSINE
    rng = np.random.default_rng(7)
    v = rng.normal(size=3 * q.n_beads)
    dev = lambda a: torch.tensor(np.ascontiguousarray(a, np.float64), device="cuda")
    hv = []
    for net in (q, c):
        ctx = mechanics.Context(coplanar_mode=_abi.MDY_COPLANAR_EXACT)
        ctx.set_network(net)
        ctx.build_neighbors()
        hv.append(autodiff.hvp(ctx, dev(net.coord), dev(v)).cpu().numpy())
    args = (q.bdh_beadset, q.bdh_k, q.bdh_pos)
    F = lambda x: oracle.ff_forces("bdhq", x, *args) - oracle.ff_forces("bdh", x, *args)
    cd = lambda h: -(F(q.coord + h * v) - F(q.coord - h * v)) / (2 * h)
    ref = (4 * cd(5e-4) - cd(1e-3)) / 3
    err = np.abs(hv[0] - hv[1] - ref).max() / np.abs(ref).max()
    print(f"[gpu] quadratic dihedral hvp: max |dHv - CD| / max|CD| = {err:.2e}")
    assert err < 1e-7


def test_batch_mixes_dihedral_forms(qnets):
    """One batched launch over networks of both dihedral forms: each equals
    its own single-network minimization bit for bit."""
    q = qnets["many_branches"]
    c = q.copy()
    c.br_dihedral_type = _abi.MDY_BRDIH_COSINE
    nets = [q, c, qnets["branched_full"]]
    singles = [ctx_for(n).minimize(_abi.default_cg_params(n.n_beads)) for n in nets]
    ctxs = [ctx_for(n) for n in nets]
    xs = [torch.tensor(n.coord, device="cuda") for n in nets]
    res, _, _ = mechanics.minimize_batch(ctxs, [_abi.default_cg_params(n.n_beads) for n in nets],
                                         [x.data_ptr() for x in xs], threads_per_block=128)
    torch.cuda.synchronize()
    for i, (x1, _, r1) in enumerate(singles):
        assert res[i].status == 0 and res[i].num_iter == r1.num_iter
        assert bits_equal(xs[i].cpu().numpy(), x1)
    assert singles[0][2].num_iter != singles[1][2].num_iter or not bits_equal(singles[0][0], singles[1][0])


def test_energy_jvp(oracle, qnets):
    """d E_k / dx . w per force field; the dihedral's entry against a central
    difference of the oracle's quadratic-form energy."""
    net = qnets["branched_full"]
    ctx = mechanics.Context(coplanar_mode=_abi.MDY_COPLANAR_EXACT)
    ctx.set_network(net)
    ctx.build_neighbors()
    rng = np.random.default_rng(9)
    w = rng.normal(size=3 * net.n_beads)
    dev = lambda a: torch.tensor(np.ascontiguousarray(a, np.float64), device="cuda")
    jv = autodiff.energy_jvp(ctx, dev(net.coord), dev(w)).cpu().numpy()
    args = (net.bdh_beadset, net.bdh_k, net.bdh_pos)
    E = lambda x: oracle.ff_energy("bdhq", x, *args)
    cd = lambda h: (E(net.coord + h * w) - E(net.coord - h * w)) / (2 * h)
    ref = (4 * cd(5e-5) - cd(1e-4)) / 3
    assert abs(jv[_abi.MDY_FF_BDH] - ref) <= 1e-7 * max(1.0, abs(ref))


def test_unknown_form_rejected(qnets):
    """BranchingFF.h:84-88 throws for an unrecognized dihedral type; the
    ABI returns MDY_ERR_ARG and keeps no network."""
    n = qnets["branched_full"].copy()
    n.br_dihedral_type = 2
    ctx = mechanics.Context(0)
    with pytest.raises(mechanics.MdyError, match="br_dihedral_type") as e:
        ctx.set_network(n)
    assert e.value.code == _abi.MDY_ERR_ARG
    with pytest.raises(mechanics.MdyError):
        ctx.energy(n.coord)


def test_rejected_network_drops_the_previous_one(qnets):
    """A context that held a network and then gets a rejected one holds no
    network at all: the library answers MDY_ERR_STATE, never the old tables
    with the new coordinates."""
    good = qnets["branched_full"]
    bad = good.copy()
    bad.br_dihedral_type = 2
    ctx = mechanics.Context(0)
    ctx.set_network(good)
    ctx.build_neighbors()
    e_ok, _ = ctx.energy(good.coord)
    assert np.isfinite(e_ok)
    with pytest.raises(mechanics.MdyError, match="br_dihedral_type"):
        ctx.set_network(bad)
    # straight through the C ABI, with a buffer of the old network's size
    x = np.ascontiguousarray(good.coord, dtype=np.float64)
    tot = C.c_double()
    rc = ctx.lib.mdy_energy(ctx.h, x.ctypes.data_as(C.POINTER(C.c_double)), C.byref(tot), None)
    assert rc == _abi.MDY_ERR_STATE
    p = _abi.default_cg_params(good.n_beads)
    res = _abi.MinResult()
    rc = ctx.lib.mdy_minimize(ctx.h, C.byref(p), x.ctypes.data_as(C.POINTER(C.c_double)), None, C.byref(res))
    assert rc == _abi.MDY_ERR_STATE
    with pytest.raises(mechanics.MdyError):
        ctx.energy(good.coord)
    # a valid network afterwards works again
    ctx.set_network(good)
    ctx.build_neighbors()
    assert ctx.energy(good.coord)[0] == e_ok
    ctx.close()
