"""Post-minimization auxiliary outputs in the oracle, pinned by closed forms:
linker / motor stretch forces k (d - l0) (LinkerStretchingHarmonic.cpp:293),
branch-point force vectors (BranchingStretchingHarmonic.cpp:273-275), and
boundary load forces on filament ends (BoundaryCylinderRepulsion.cpp:124-189):
monomer i of an end pointing straight at a plane at distance r sits at
r - i*monSize and carries K exp(-(r - i*monSize)/lambda)/lambda; an end
pointing away carries none."""
import numpy as np

from medyan_amd import network


def test_stretch_forces_closed_form(oracle):
    a = np.array([[0.0, 0, 0], [108, 0, 0]]) + 400
    b = np.array([[0.0, 60, 0], [108, 60, 0]]) + 400
    c = np.array([[54.0, 10, 0], [54.0, 118, 0]]) + 400
    net = network.from_filaments([a, b, c], 1000.0, linkers=[(0, 0, 0.5, 1, 0, 0.5, 8.0, 50.0)],
                                 motors=[(0, 0, 0.25, 1, 0, 0.25, 2.5, 70.0)], branches=[(0, 0, 0.5, 2)],
                                 l_br_stretch=6.0)
    s = oracle.system(net)
    lk, mt, br = s.stretch_forces()
    assert np.isclose(lk[0], 8.0 * (60.0 - 50.0), rtol=1e-14)
    assert np.isclose(mt[0], 2.5 * (60.0 - 70.0), rtol=1e-14)
    # branch: mother midpoint (54, 0) to the daughter's first bead (54, 10): d = 10, l0 = 6
    f0 = 100.0 * (10.0 - 6.0) / 10.0
    np.testing.assert_allclose(br, [0.0, -f0 * 10.0, 0.0], rtol=1e-14, atol=1e-12)


def test_load_forces_closed_form(oracle):
    towards = np.array([[260.0, 500, 500], [152, 500, 500], [44, 500, 500]])   # plus end 19 nm from x = 25
    away = np.array([[44.0, 300, 300], [152, 300, 300], [260, 300, 300]])      # minus end 19 nm from x = 25
    net = network.from_filaments([towards, away], 1000.0)
    s = oracle.system(net)
    lp, lm = s.load_forces(net.coord, [2.7], [40], 40)
    i = np.arange(40)
    expect = 41.0 * np.exp(-(19.0 - 2.7 * i) / 2.7) / 2.7
    np.testing.assert_allclose(lp[2], expect, rtol=1e-13)
    # the second filament's minus end (bead 3) points at the same wall: same loads
    np.testing.assert_allclose(lm[3], expect, rtol=1e-13)
    # nothing else is loaded: the plus end of the second filament points away from x = 25
    # and is out of range of every other plane; the first filament's minus end is far
    mask = np.ones_like(lp, bool)
    mask[2] = False
    assert not lp[mask].any()
    mask = np.ones_like(lm, bool)
    mask[3] = False
    assert not lm[mask].any()


def test_load_forces_need_pointing_at_plane(oracle):
    """An end pointing away from a plane in range gets proj = 0."""
    f = np.array([[152.0, 500, 500], [44, 500, 500], [152, 500, 520]])        # bends back: dir away from x = 25
    net = network.from_filaments([f], 1000.0, k_bend=0.0)
    s = oracle.system(net)
    lp, lm = s.load_forces(net.coord, [2.7], [40], 40)
    d = (f[2] - f[1]) / np.linalg.norm(f[2] - f[1])
    assert -d[0] < 0 and not lp.any()
    # minus end: bead 0 at x = 152, direction bead1 -> bead0 is +x: away from x = 25
    assert not lm.any()
