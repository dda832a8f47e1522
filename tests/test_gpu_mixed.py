"""The mixed-precision form: FP32 derivative blocks for well-conditioned
excluded-volume pairs, FP32 force records, FP64 energies, sums, dot
products and directions. Its error against the FP64 oracle is stated here:
forces within 1e-5 of max|F| (measured ~3e-7), energies bit-identical to
the FP64 form (they are FP64)."""
import numpy as np
import pytest

from medyan_amd import _abi, mechanics, network

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF


@pytest.fixture
def exact(oracle):
    oracle.set_exact(1)
    yield oracle
    oracle.set_exact(0)


@pytest.mark.parametrize("workload,kw", [("net500", {}), ("net500", dict(n_filaments=40, cyl_per_filament=5,
                                                                        n_branches=8, seed=2)),
                                         ("net40k", {})])
def test_mixed_forces_and_energies(exact, lib, workload, kw):
    net = network.generate(lib, workload, **kw)
    print(f"[gpu] mixed form, {net.n_cyl} cylinders", flush=True)
    ctx = mechanics.Context(precision=_abi.MDY_PREC_MIXED)
    ctx.set_network(net)
    ctx.build_neighbors()
    s = exact.system(net)
    e, per = ctx.energy(net.coord)
    eo, _ = s.energy()
    assert e == eo                                     # energies stay FP64, bit for bit
    f, fo = ctx.forces(net.coord), s.forces()
    err = np.abs(f - fo).max() / np.abs(fo).max()
    print(f"[gpu] mixed force error {err:.2e} of max|F|")
    assert err <= 1e-5


def test_mixed_minimization_reaches_tolerance(exact, lib):
    net = network.generate(lib, "net500")
    ctx = mechanics.Context(precision=_abi.MDY_PREC_MIXED)
    ctx.set_network(net)
    ctx.build_neighbors()
    p = _abi.default_cg_params(net.n_beads, energy_rel_tol=0.0)
    x, f, res = ctx.minimize(p)
    assert res.status == 0 and res.error_flags == 0 and res.grad_inf_norm < 1.0
    s = exact.system(net)
    eo, _ = s.energy(x, individual=True)
    assert res.energies_after[NFF] == eo              # FP64 energy of the returned coordinates
    fo = s.forces(x)
    assert abs(np.abs(fo).max() - res.grad_inf_norm) <= 1e-3   # FP32 record rounding only
    assert res.energies_after[NFF] < res.energies_before[NFF]
