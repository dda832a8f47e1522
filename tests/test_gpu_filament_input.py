"""A network read from MEDYAN's FILAMENT input format (the reader of
Parser.cpp:3060-3094 as medyan_amd.io; projections as Controller.cpp:560-620)
through the device: its minimization equals the exact-mode oracle's bit for
bit (trajectory, coordinates, forces, total energy), in both coplanar modes, and
the minimized state written as a traj.h5 snapshot (Output.h:509-575 layout,
medyan_amd.traj) reads back bit for bit."""
import os

import numpy as np
import pytest

from medyan_amd import _abi, io, mechanics
from test_gpu_parity import bits_equal

pytestmark = pytest.mark.gpu
NFF = _abi.MDY_NFF
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HDF5 = os.path.exists(os.path.join(os.environ.get("MDY_HDF5_ROOT", "/opt/conda"), "include", "hdf5.h"))

# three filaments (STRAIGHT projection of two points each): the first two
# cross about 60 nm apart, joined by a linker; the third is linked to the second
TEXT = ("FILAMENT 0 400 400 400 1400 600 500\n"
        "FILAMENT 0 900 300 420 700 1300 560\n"
        "FILAMENT 0 600 900 300 760 1040 960\n")
LINKERS = [(0, 4, 0.5, 1, 1, 0.5, 8.0, 60.0), (1, 5, 0.3, 2, 3, 0.6, 8.0, 80.0)]


@pytest.fixture
def exact(oracle):
    oracle.set_exact(1)
    yield oracle
    oracle.set_exact(0)


@pytest.mark.parametrize("coplanar", [_abi.MDY_COPLANAR_REFERENCE, _abi.MDY_COPLANAR_EXACT])
def test_filament_file_network_minimizes_like_the_oracle(exact, coplanar, tmp_path):
    net = io.network_from_filament_file(TEXT, 2000.0, linkers=LINKERS)
    assert net.n_cyl == 9 + 10 + 6
    ctx = mechanics.Context(0, coplanar_mode=coplanar)
    ctx.set_network(net)
    ctx.build_neighbors()
    s = exact.system(net, coplanar_mode=coplanar)
    np.testing.assert_array_equal(ctx.pairs(), exact.nl_pairs(net))
    p = _abi.default_cg_params(net.n_beads)
    x, f, res = ctx.minimize(p, trace_cap=100000)
    tr = ctx.trace()
    xo, fo, reso, tro = s.minimize(p, trace_cap=100000)
    assert res.status == 0 and res.grad_inf_norm < 1.0
    assert (res.num_iter, res.num_energy_call, res.num_force_call) == (reso.num_iter, reso.num_energy_call,
                                                                       reso.num_force_call)
    for key in ("lambda_", "energy", "maxf", "beta"):
        assert bits_equal(tr[key], tro[key]), key
    assert bits_equal(x, xo) and bits_equal(f, fo)
    # the total is an exact sum; the per-force-field reports are plain double sums in block order
    assert bits_equal(res.energies_after[NFF:NFF + 1], reso.energies_after[NFF:NFF + 1])
    np.testing.assert_allclose(res.energies_after[:NFF], reso.energies_after[:NFF], rtol=1e-12,
                               atol=1e-12 * abs(reso.energies_after[NFF]))
    print(f"[gpu] FILAMENT network: {net.n_cyl} cylinders, {res.num_iter} iterations", flush=True)
    if HDF5:
        from medyan_amd import traj
        path = str(tmp_path / "traj.h5")
        out = traj.SnapshotOutput(path)
        out.append(traj.snapshot_of(net, x, np.array(res.energies_after[NFF:NFF + 1] + res.energies_after[:NFF])))
        out.finish()
        fils = traj.filament_beads(net)
        back = out.read_filaments(0, net.n_beads)
        want = np.concatenate([x.reshape(-1, 3)[fb] for fb in fils]).ravel()
        assert np.array_equal(back.view(np.int64), want.view(np.int64))
    ctx.close()


def test_reference_example_file_on_the_device(exact):
    """The reference's 2filaments example file (two antiparallel filaments
    30 nm apart): exactly parallel cylinder pairs, whose closed-form energy is
    not finite, go through the Simpson energy fallback; energies and forces at
    the file's coordinates equal the exact-mode oracle's bit for bit."""
    net = io.network_from_filament_file(os.path.join(GOLDEN, "filament_input_2filaments.txt"), 2000.0)
    ctx = mechanics.Context(0)
    ctx.set_network(net)
    ctx.build_neighbors()
    s = exact.system(net)
    e, per = ctx.energy(net.coord)
    eo, pero = s.energy()
    assert bits_equal([e], [eo]) and bits_equal(per, pero)
    assert bits_equal(ctx.forces(net.coord), s.forces())
    ctx.close()
