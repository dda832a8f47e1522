"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same
seeded inputs.

* neighbor lists, excluded-volume and boundary tables: bit-exact;
* energies and forces: bit-exact against the oracle in exact mode (the
  shared atan/exp and exact sums); against the oracle in reference mode (C
  library atan/exp, sequential sums) energies within 1e-10 relative and
  forces within 1e-10 of max|F| outside ill-conditioned excluded-volume
  pairs, where a last-ulp atan difference is amplified;
* minimization: bit-identical trajectories (every lambda, energy, max force
  and beta of the trace) against the oracle in exact mode at net500 in full
  and for the first iterations at net40k; at full size the end state is
  checked through size-independent properties (the returned forces are the
  oracle's forces at the returned coordinates; the force criterion holds).
"""
import numpy as np
import pytest

from medyan_amd import _abi, mechanics, network

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.fixture
def exact(oracle):
    oracle.set_exact(1)
    yield oracle
    oracle.set_exact(0)


def small_networks(lib):
    return {
        "net500": network.generate(lib, "net500"),
        "branched": network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, seed=2),
        "dense": network.generate(lib, "net500", n_filaments=60, cyl_per_filament=3, box=500.0, n_linkers=40,
                                  n_motors=4, seed=3),
        # all four branching force fields (dihedral k 10, position k 20 as in the shipped examples)
        "branched_full": network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, seed=4,
                                          k_br_dihedral=10.0, k_br_position=20.0),
    }


@pytest.fixture(scope="module")
def nets(lib):
    d = small_networks(lib)
    d["net40k"] = network.generate(lib, "net40k")
    return d


CASES = ["net500", "branched", "dense", "branched_full", "net40k"]


def ctx_for(net):
    print(f"[gpu] context for {net.n_cyl} cylinders", flush=True)
    ctx = mechanics.Context(0)
    ctx.set_network(net)
    ctx.build_neighbors()
    return ctx


@pytest.mark.parametrize("case", CASES)
def test_neighbor_tables_bit_exact(oracle, nets, case):
    net = nets[case]
    ctx = ctx_for(net)
    np.testing.assert_array_equal(ctx.pairs(), oracle.nl_pairs(net))
    s = oracle.system(net)
    bs, kr = ctx.exvol_table()
    obs, okr, oeq = s.ev_table
    np.testing.assert_array_equal(bs, obs)
    assert bits_equal(kr, okr * oeq[:, 0] * oeq[:, 1])
    nn, bbs = ctx.boundary_table()
    onn, obbs = s.bd_table
    np.testing.assert_array_equal(nn, onn)
    np.testing.assert_array_equal(bbs, obbs)


@pytest.mark.parametrize("case", CASES)
def test_energy_forces_exact_mode_bitwise(exact, nets, case):
    net = nets[case]
    ctx = ctx_for(net)
    s = exact.system(net)
    rng = np.random.default_rng(5)
    for x in (net.coord, net.coord + rng.normal(0, 0.5, net.coord.shape)):
        e, per = ctx.energy(x)
        eo, pero = s.energy(x)
        assert bits_equal([e], [eo])
        np.testing.assert_allclose(per, pero, rtol=1e-13, atol=0)
        f = ctx.forces(x)
        assert bits_equal(f, s.forces(x))


def ill_conditioned_beads(net, s, tau=2e-3, sin_min=0.05):
    """Beads of excluded-volume pairs whose analytic force is ill-conditioned:
    nearly coplanar (normalized triple product of the areInPlane test,
    MathFunctions.h:1101-1133, below tau) or nearly parallel cylinders. There
    a last-ulp difference in atan is amplified by up to ~1e8."""
    bs, _, _ = s.ev_table
    x = net.coord.reshape(-1, 3)
    c = x[bs // 3]
    v1, v2, v3 = c[:, 1] - c[:, 0], c[:, 3] - c[:, 2], c[:, 2] - c[:, 0]
    n = np.cross(v1, v2)
    nn = np.linalg.norm(n, axis=1)
    trip = np.abs((n * v3).sum(1)) / (nn * np.linalg.norm(v3, axis=1))
    sin = nn / (np.linalg.norm(v1, axis=1) * np.linalg.norm(v2, axis=1))
    bad = (trip < tau) | (sin < sin_min)
    return np.unique(bs[bad] // 3)


@pytest.mark.parametrize("case", CASES)
def test_energy_forces_reference_mode_1e10(oracle, nets, case):
    """Against the oracle with the C library's atan/exp and sequential sums:
    energies within 1e-10 relative; forces within 1e-10 of max|F| on every
    bead outside ill-conditioned excluded-volume pairs; on those, the whole
    difference is the oracle's own exact-mode minus reference-mode
    difference (the GPU equals exact mode bit for bit)."""
    net = nets[case]
    ctx = ctx_for(net)
    s = oracle.system(net)
    e, per = ctx.energy(net.coord)
    eo, pero = s.energy()
    assert abs(e - eo) <= 1e-10 * abs(eo)
    np.testing.assert_allclose(per, pero, rtol=1e-10, atol=1e-10 * abs(eo))
    f, fo = ctx.forces(net.coord), s.forces()
    fmax = np.abs(fo).max()
    diff = np.abs(f - fo).reshape(-1, 3).max(1)
    ill = ill_conditioned_beads(net, s)
    ok = np.ones(net.n_beads, bool)
    ok[ill] = False
    assert diff[ok].max() <= 1e-10 * fmax
    oracle.set_exact(1)
    try:
        fe = s.forces()
    finally:
        oracle.set_exact(0)
    assert bits_equal(f, fe)
    print(f"[gpu] {case}: {len(ill)} beads in ill-conditioned pairs, max rel diff there "
          f"{diff[~ok].max() / fmax if (~ok).any() else 0:.2e}")


def test_forces_accumulate(nets):
    """mdy_forces has ForceField::computeForces semantics: it adds."""
    net = nets["net500"]
    ctx = ctx_for(net)
    base = np.linspace(-1, 1, 3 * net.n_beads)
    f = ctx.forces(net.coord, base)
    np.testing.assert_allclose(f - base, ctx.forces(net.coord), rtol=0, atol=1e-9)


@pytest.mark.parametrize("case", ["net500", "branched", "dense", "branched_full"])
def test_minimization_trajectory_bitwise(exact, nets, case):
    net = nets[case]
    ctx = ctx_for(net)
    s = exact.system(net)
    p = _abi.default_cg_params(net.n_beads)
    x, f, res = ctx.minimize(p, trace_cap=100000)
    tr = ctx.trace()
    xo, fo, reso, tro = s.minimize(p, trace_cap=100000)
    assert res.status == 0 and reso.status == 0
    assert res.num_iter == reso.num_iter and res.num_energy_call == reso.num_energy_call
    assert res.num_force_call == reso.num_force_call and res.error_flags == reso.error_flags
    for key in ("lambda_", "energy", "maxf", "beta"):
        assert bits_equal(tr[key], tro[key]), key
    np.testing.assert_array_equal(tr["safe"], tro["safe"])
    assert bits_equal(x, xo) and bits_equal(f, fo)
    assert bits_equal(res.energies_after[NFF:NFF + 1], reso.energies_after[NFF:NFF + 1])
    assert bits_equal(res.energies_before[NFF:NFF + 1], reso.energies_before[NFF:NFF + 1])


@pytest.mark.parametrize("method", [_abi.MDY_CG_STEEPEST, _abi.MDY_CG_FLETCHER_RIEVES])
def test_other_directions_trajectory_bitwise(exact, nets, method):
    """CGMethod::minimize<steepestDescent> and <FletcherRieves>
    (CGMethod.hpp:1746-1759): bit-identical trajectories to the oracle."""
    net = nets["branched_full"]
    ctx = ctx_for(net)
    s = exact.system(net)
    p = _abi.default_cg_params(net.n_beads, method=method, force_call_limit=3000)
    x, f, res = ctx.minimize(p, trace_cap=100000)
    tr = ctx.trace()
    xo, fo, reso, tro = s.minimize(p, trace_cap=100000)
    assert res.status == 0 and reso.status == 0
    assert (res.num_iter, res.num_energy_call, res.num_force_call, res.error_flags) == (
        reso.num_iter, reso.num_energy_call, reso.num_force_call, reso.error_flags)
    for key in ("lambda_", "energy", "maxf", "beta"):
        assert bits_equal(tr[key], tro[key]), key
    assert bits_equal(x, xo) and bits_equal(f, fo)
    print(f"[gpu] method {method}: {res.num_iter} iterations, max|F| {res.grad_inf_norm:.3f}")


@pytest.mark.parametrize("case", ["net500", "branched_full"])
def test_quadratic_line_search_trajectory_bitwise(exact, nets, case):
    """lineSearchAlgorithm QUADRATIC (quadraticLineSearchV2 and
    quadraticoptimization, CGMethod.hpp:522-765): bit-identical trajectories
    to the oracle, with parabola refinements among the accepted steps."""
    from test_oracle_cg import refined_steps
    net = nets[case]
    ctx = ctx_for(net)
    s = exact.system(net)
    p = _abi.default_cg_params(net.n_beads, line_search=_abi.MDY_LS_QUADRATIC, force_call_limit=3000)
    x, f, res = ctx.minimize(p, trace_cap=100000)
    tr = ctx.trace()
    xo, fo, reso, tro = s.minimize(p, trace_cap=100000)
    assert res.status == 0 and reso.status == 0
    assert (res.num_iter, res.num_energy_call, res.num_force_call, res.error_flags, res.num_safe_mode) == (
        reso.num_iter, reso.num_energy_call, reso.num_force_call, reso.error_flags, reso.num_safe_mode)
    for key in ("lambda_", "energy", "maxf", "beta", "safe"):
        assert bits_equal(tr[key], tro[key]), key
    assert bits_equal(x, xo) and bits_equal(f, fo)
    assert bits_equal(res.energies_after[NFF:NFF + 1], reso.energies_after[NFF:NFF + 1])
    nref = refined_steps(tro, p)
    print(f"[gpu] quadratic {case}: {res.num_iter} iterations, {nref} refined steps, "
          f"max|F| {res.grad_inf_norm:.3f}")
    assert nref > 0


def test_net40k_first_iterations_bitwise_and_end_state(exact, nets):
    net = nets["net40k"]
    ctx = ctx_for(net)
    s = exact.system(net)
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0)
    x, f, res = ctx.minimize(p, trace_cap=100000)
    tr = ctx.trace()
    assert res.status == 0 and res.error_flags == 0
    assert res.grad_inf_norm < 1.0                      # minimized to 1 pN
    _, _, _, tro = s.minimize(p, iter_cap=25, trace_cap=100)
    for key in ("lambda_", "energy", "maxf", "beta"):
        assert bits_equal(tr[key][:25], tro[key]), key
    # size-independent end-state checks: the returned forces are the oracle's
    # forces at the returned coordinates, and the force criterion holds
    assert bits_equal(f, s.forces(x))
    assert np.abs(f).max() == res.grad_inf_norm
    eo, _ = s.energy(x, individual=True)
    assert bits_equal([res.energies_after[NFF]], [eo])


def test_call_limit_and_line_search_flags(exact, nets):
    net = nets["net500"]
    ctx = ctx_for(net)
    s = exact.system(net)
    p = _abi.default_cg_params(net.n_beads, force_call_limit=40, energy_rel_tol=0.0)
    p.n_beads_limit = 1
    x, f, res = ctx.minimize(p)
    xo, fo, reso, _ = s.minimize(p)
    assert res.num_force_call == reso.num_force_call == 41
    assert bits_equal(x, xo)
    # without recovery, a line search that cannot move sets the error flag
    p2 = _abi.default_cg_params(net.n_beads, try_to_recover=0, lambda_max=1e-9)
    x2, f2, res2 = ctx.minimize(p2)
    xo2, fo2, reso2, _ = s.minimize(p2)
    assert res2.error_flags == reso2.error_flags == _abi.MDY_ERROR_LINESEARCH
    assert res2.num_iter == reso2.num_iter
    assert bits_equal(x2, xo2)


@pytest.mark.parametrize("report", [1, 0])
def test_call_limit_report_flag(exact, nets, report):
    """report_on_call_limit reaches the device: a run stopped by the call
    limit after taking lowest-energy backups ends with the oracle's
    coordinates, forces and energies for either value of the flag."""
    net = nets["net500"]
    ctx = ctx_for(net)
    s = exact.system(net)
    p = _abi.default_cg_params(net.n_beads, force_call_limit=400, energy_rel_tol=0.0, grad_tol=0.05,
                               report_on_call_limit=report)
    p.n_beads_limit = 1
    x, f, res = ctx.minimize(p)
    xo, fo, reso, _ = s.minimize(p)
    assert res.num_force_call == reso.num_force_call == 401
    assert bits_equal(x, xo) and bits_equal(f, fo)
    # the total is an exact sum; per-force-field reports are plain double sums
    assert bits_equal(res.energies_after[NFF:NFF + 1], reso.energies_after[NFF:NFF + 1])
    np.testing.assert_allclose(res.energies_after[:NFF], reso.energies_after[:NFF], rtol=1e-13, atol=0)


def test_search_direction_after_minimization(exact, nets):
    """mdy_get_search_dir: CGMethod::searchDir after a converged run (the
    final shift still pending on the device), a call-limit stop and a
    line-search error, against the oracle's (-0 and +0 compare equal)."""
    net = nets["net500"]
    ctx = ctx_for(net)
    s = exact.system(net)
    limit = _abi.default_cg_params(net.n_beads, force_call_limit=40, energy_rel_tol=0.0)
    limit.n_beads_limit = 1
    cases = {"converged": _abi.default_cg_params(net.n_beads), "call_limit": limit,
             "line_search_error": _abi.default_cg_params(net.n_beads, try_to_recover=0, lambda_max=1e-9)}
    for name, p in cases.items():
        _, _, res = ctx.minimize(p)
        d = ctx.search_dir()
        _, _, reso, _ = s.minimize(p)
        do = exact.last_search_dir()
        assert res.num_iter == reso.num_iter, name
        assert np.array_equal(d, do), name
        print(f"[gpu] search direction after {name}: {res.num_iter} iterations, max|d| {np.abs(d).max():.3g}")


def test_parallel_pair_quadrature_fallback(exact):
    """Exactly parallel cylinders: the analytic excluded-volume energy is 0/0
    and the Simpson rule takes over (energy path); the force path perturbs
    the coplanar pair like the reference."""
    f1 = np.array([[200.0, 200.0, 200.0 + 108.0 * i] for i in range(4)])
    f2 = f1 + np.array([30.0, 0.0, 0.0])
    net = network.from_filaments([f1, f2], 1000.0)
    ctx = ctx_for(net)
    s = exact.system(net)
    assert len(ctx.exvol_table()[1]) > 0
    e, _ = ctx.energy(net.coord)
    eo, _ = s.energy()
    assert np.isfinite(e) and bits_equal([e], [eo])
    assert bits_equal(ctx.forces(net.coord), s.forces())


def test_disabled_force_fields(exact, lib):
    net = network.generate(lib, "net500", n_filaments=20, seed=9)
    net.exvol_enabled = 0
    net.boundary_enabled = 0
    ctx = ctx_for(net)
    s = exact.system(net)
    e, per = ctx.energy(net.coord)
    eo, pero = s.energy()
    assert bits_equal([e], [eo]) and per[_abi.MDY_FF_EV] == 0.0 and per[_abi.MDY_FF_BD] == 0.0
    assert bits_equal(ctx.forces(net.coord), s.forces())


def test_single_filament_hairpin(exact):
    """One filament folded back on itself: pairs on the same filament count
    only when their positions differ by more than 2, and the exactly parallel
    legs go through the quadrature fallback."""
    f1 = np.array([[300.0, 300, 300], [400, 300, 300], [500, 300, 300], [500, 330, 300], [400, 330, 300],
                   [300, 330, 300]])
    net = network.from_filaments([f1], 1000.0)
    ctx = ctx_for(net)
    s = exact.system(net)
    np.testing.assert_array_equal(ctx.pairs(), [[3, 0], [4, 0], [4, 1]])
    assert len(s.ev_table[0]) == 3
    p = _abi.default_cg_params(net.n_beads)
    x, f, res = ctx.minimize(p)
    xo, fo, reso, _ = s.minimize(p)
    assert res.status == reso.status == 0
    assert bits_equal(x, xo) and res.num_iter == reso.num_iter


def test_line_search_argument_checked(nets):
    net = nets["net500"]
    ctx = ctx_for(net)
    with pytest.raises(mechanics.MdyError) as e:
        ctx.minimize(_abi.default_cg_params(net.n_beads, line_search=7))
    assert e.value.code == _abi.MDY_ERR_ARG


def test_error_paths(lib, nets):
    ctx = mechanics.Context(0)
    with pytest.raises(mechanics.MdyError) as e:
        ctx.minimize(_abi.default_cg_params(10), np.zeros(30))
    assert e.value.code in (_abi.MDY_ERR_STATE, _abi.MDY_ERR_ARG)
    bad = nets["net500"].copy()
    bad.fs_beadset = bad.fs_beadset.copy()
    bad.fs_beadset[0] = 1                      # not a multiple of 3
    with pytest.raises(mechanics.MdyError) as e:
        ctx.set_network(bad)
    assert e.value.code == _abi.MDY_ERR_ARG
