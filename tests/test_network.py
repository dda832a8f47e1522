"""The seeded network generator (host code in libmdy_hip.so) behind the
benchmark workloads: determinism, counts, and the geometry rules (cylinder
length, box margins, cross-link reach, branch angle), plus the
filament-list builder."""
import math

import numpy as np
import pytest

from medyan_amd import network


def test_deterministic_per_seed(lib):
    a = network.generate(lib, "net500")
    b = network.generate(lib, "net500")
    c = network.generate(lib, "net500", seed=12345)
    np.testing.assert_array_equal(a.coord, b.coord)
    np.testing.assert_array_equal(a.ls_l0, b.ls_l0)
    assert not np.array_equal(a.coord, c.coord)


@pytest.mark.parametrize("name", ["net500", "net40k"])
def test_counts(lib, name):
    w = network.WORKLOADS[name]
    n = network.generate(lib, name)
    nf, nc = w["n_filaments"], w["cyl_per_filament"]
    assert n.n_cyl == nf * nc and n.n_beads == nf * (nc + 1)
    assert len(n.fs_k) == nf * nc and len(n.fb_k) == nf * (nc - 1)
    assert len(n.ls_k) == w["n_linkers"] and len(n.ms_k) == w["n_motors"]
    assert len(n.bs_k) == w["n_branches"] and len(n.bb_k) == w["n_branches"]
    assert len(n.plane_k) == 6


def test_geometry_rules(lib):
    n = network.generate(lib, "net500", jitter=0.0)
    x = n.coord.reshape(-1, 3)
    L = n.box
    lo, hi = 25.0 + n.bd_cutoff / 4, L - 25.0 - n.bd_cutoff / 4
    assert x.min() >= lo - 1e-9 and x.max() <= hi + 1e-9
    seg = np.linalg.norm(x[n.cyl_bead2 // 3] - x[n.cyl_bead1 // 3], axis=1)
    np.testing.assert_allclose(seg, 108.0, rtol=1e-12)
    # cross-links join different filaments at the four binding sites, within 175 nm
    for bs, p1, p2 in ((n.ls_beadset, n.ls_pos1, n.ls_pos2), (n.ms_beadset, n.ms_pos1, n.ms_pos2)):
        b = bs.reshape(-1, 4) // 3
        v1 = x[b[:, 0]] * (1 - p1[:, None]) + p1[:, None] * x[b[:, 1]]
        v2 = x[b[:, 2]] * (1 - p2[:, None]) + p2[:, None] * x[b[:, 3]]
        assert np.all(np.linalg.norm(v2 - v1, axis=1) < 175.0)
        assert set(np.unique(np.concatenate([p1, p2]))) <= {0.125, 0.375, 0.625, 0.875}


def test_branch_geometry(lib):
    n = network.generate(lib, "net500", n_filaments=40, cyl_per_filament=5, n_branches=8, jitter=0.0, seed=2)
    x = n.coord.reshape(-1, 3)
    b = n.bb_beadset.reshape(-1, 4) // 3
    m, d = x[b[:, 1]] - x[b[:, 0]], x[b[:, 3]] - x[b[:, 2]]
    ang = np.degrees(np.arccos((m * d).sum(1) / np.linalg.norm(m, axis=1) / np.linalg.norm(d, axis=1)))
    np.testing.assert_allclose(ang, 70.0, atol=1e-9)
    s = n.bs_beadset.reshape(-1, 3) // 3
    v = x[s[:, 0]] * (1 - n.bs_pos[:, None]) + n.bs_pos[:, None] * x[s[:, 1]]
    np.testing.assert_allclose(np.linalg.norm(x[s[:, 2]] - v, axis=1), 6.0, rtol=1e-12)
    assert (n.cyl_branch_child >= 0).sum() == 8


def test_from_filaments_matches_generator_rules(lib):
    g = network.generate(lib, "net500", n_filaments=3, cyl_per_filament=4, n_linkers=0, n_motors=0)
    fil = [g.coord.reshape(-1, 3)[5 * i: 5 * i + 5] for i in range(3)]
    h = network.from_filaments(fil, g.box)
    for f in ("fs_beadset", "fs_k", "fs_l0", "fb_beadset", "fb_k", "fb_theta0", "cyl_bead1", "cyl_bead2",
              "cyl_minus_end", "plane", "plane_k", "plane_screen"):
        np.testing.assert_array_equal(getattr(h, f), getattr(g, f), err_msg=f)
