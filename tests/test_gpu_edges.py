"""Edge cases of the device path against the oracle: a network of one
cylinder (no neighbor pairs), a start that already meets the force
tolerance (no CG iteration, energiesAfter still computed), filaments too far
apart for any excluded-volume pair (empty dynamic tables) and rejected
monomer tables."""
import numpy as np
import pytest

from medyan_amd import _abi, mechanics, network
from test_gpu_parity import bits_equal, ctx_for, exact  # noqa: F401  (exact is a fixture)

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF


def test_one_cylinder(exact):
    f = np.array([[400.0, 500.0, 500.0], [515.0, 504.0, 497.0]])       # stretched to ~115 nm
    net = network.from_filaments([f], 1000.0)
    print(f"[gpu] one cylinder, {net.n_beads} beads", flush=True)
    ctx = ctx_for(net)
    s = exact.system(net)
    assert len(ctx.pairs()) == 0 and len(ctx.exvol_table()[1]) == 0
    e, per = ctx.energy(net.coord)
    eo, _ = s.energy()
    assert bits_equal([e], [eo]) and per[_abi.MDY_FF_FS] > 0
    p = _abi.default_cg_params(net.n_beads)
    x, fo, res = ctx.minimize(p)
    xo, foo, reso, _ = s.minimize(p)
    assert res.status == reso.status == 0 and res.num_iter == reso.num_iter > 0
    assert bits_equal(x, xo) and bits_equal(fo, foo)
    assert abs(np.linalg.norm(x[3:] - x[:3]) - 108.0) < 0.02               # back near the rest length


def test_start_already_within_tolerance(exact):
    f = np.array([[400.0, 500.0, 500.0], [508.0, 500.0, 500.0]])       # exactly the rest length
    net = network.from_filaments([f], 1000.0)
    ctx = ctx_for(net)
    s = exact.system(net)
    p = _abi.default_cg_params(net.n_beads)
    x, fo, res = ctx.minimize(p, trace_cap=10)
    xo, foo, reso, _ = s.minimize(p)
    assert res.num_iter == reso.num_iter == 0 and res.status == 0
    assert (res.num_energy_call, res.num_force_call) == (reso.num_energy_call, reso.num_force_call)
    assert bits_equal(x, net.coord) and bits_equal(res.energies_after[:NFF + 1], reso.energies_after[:NFF + 1])
    assert len(ctx.trace()["lambda_"]) == 0


def test_no_excluded_volume_pairs(exact):
    fils = [np.array([[200.0 + 400 * i + 3.0 * (k % 2), 300.0, 300.0 + 112 * k] for k in range(4)]) for i in range(3)]
    net = network.from_filaments(fils, 1500.0)
    ctx = ctx_for(net)
    s = exact.system(net)
    assert len(ctx.exvol_table()[1]) == 0 and len(s.ev_table[1]) == 0
    e, per = ctx.energy(net.coord)
    assert bits_equal([e], [s.energy()[0]]) and per[_abi.MDY_FF_EV] == 0.0
    assert bits_equal(ctx.forces(net.coord), s.forces())
    p = _abi.default_cg_params(net.n_beads)
    x, _, res = ctx.minimize(p)
    xo, _, reso, _ = s.minimize(p)
    assert res.num_iter == reso.num_iter > 0 and bits_equal(x, xo)


@pytest.mark.parametrize("field,value", [("cyl_mon_interval", 0), ("cyl_mon_plus", -100)])
def test_monomer_tables_rejected(lib, field, value):
    net = network.generate(lib, None, seed=5, n_filaments=3, cyl_per_filament=3, box=700.0, n_linkers=2,
                           n_motors=0, n_branches=0).with_monomer_exvol(4)
    bad = getattr(net, field).copy()
    bad[1] = value
    setattr(net, field, bad)
    ctx = mechanics.Context()
    with pytest.raises(mechanics.MdyError) as e:
        ctx.set_network(net)
    assert e.value.code == _abi.MDY_ERR_ARG
