"""The reference's near-coplanar excluded-volume stress input
(TestCylinderExclVolume.cpp:166-188; fixture written by
tests/golden/make_coplanar_fixture.py) as explicit tables and as a network:
its 140 cylinders form 24 chains (the second bead of one cylinder is the
first of the next), which become filaments with the fixture's equilibrium
lengths, stretching k 100 and bending k 672 as in the benchmark networks,
excluded-volume k 100 as in the reference's test, no boundary."""
import json
import os

import numpy as np

from medyan_amd import network

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    with open(os.path.join(HERE, "golden", "exvol_20210901.json")) as fh:
        d = json.load(fh)
    bs = np.array(d["beadset"], np.int32)
    eql = np.array(d["eq_length"], np.float64)
    x = np.array(d["coord"], np.float64)
    return bs, eql, x, d["k"], d["unstable_index"]


def as_network():
    bs, eql, x, k, _ = load()
    cyl = {}
    for (a, b, c, e), (l1, l2) in zip(bs.reshape(-1, 4), eql.reshape(-1, 2)):
        cyl[(int(a), int(b))] = l1
        cyl[(int(c), int(e))] = l2
    nxt = {a: b for (a, b) in cyl}
    prv = {b: a for (a, b) in cyl}
    b1, b2, fil, pos, eq = [], [], [], [], []
    for f, s in enumerate(sorted(a for (a, _) in cyl if a not in prv)):
        a, p = s, 0
        while a in nxt:
            b1.append(a); b2.append(nxt[a]); fil.append(f); pos.append(p); eq.append(cyl[(a, nxt[a])])
            a, p = nxt[a], p + 1
    nc = len(b1)
    b1, b2, eq = np.array(b1, np.int32), np.array(b2, np.int32), np.array(eq)
    fil, pos = np.array(fil, np.int32), np.array(pos, np.int32)
    fb = [(i - 1, i) for i in range(1, nc) if fil[i] == fil[i - 1]]
    nb = len(x) // 3
    return network.Network(
        coord=x.copy(), n_beads=nb, cyl_bead1=b1, cyl_bead2=b2, cyl_id=np.arange(nc, dtype=np.int64),
        cyl_filament=fil, cyl_position=pos, cyl_type=np.zeros(nc, np.int16), cyl_eq_length=eq,
        cyl_exvol_k=np.full(nc, float(k)), cyl_branch_child=np.full(nc, -1, np.int32),
        cyl_minus_end=(pos == 0).astype(np.uint8),
        fs_beadset=np.stack([b1, b2], 1).ravel(), fs_k=100.0 * (108.0 / eq), fs_l0=eq.copy(),
        fb_beadset=np.array([[b1[p], b1[c], b2[c]] for p, c in fb], np.int32).ravel(),
        fb_k=np.array([672.0 * 108.0 / eq[c] for _, c in fb]), fb_theta0=np.zeros(len(fb)),
        ls_beadset=np.zeros(0, np.int32), ls_k=np.zeros(0), ls_l0=np.zeros(0), ls_pos1=np.zeros(0),
        ls_pos2=np.zeros(0), ms_beadset=np.zeros(0, np.int32), ms_k=np.zeros(0), ms_l0=np.zeros(0),
        ms_pos1=np.zeros(0), ms_pos2=np.zeros(0), bs_beadset=np.zeros(0, np.int32), bs_k=np.zeros(0),
        bs_l0=np.zeros(0), bs_pos=np.zeros(0), bb_beadset=np.zeros(0, np.int32), bb_k=np.zeros(0),
        bb_theta0=np.zeros(0), plane=np.zeros(0), plane_k=np.zeros(0), plane_screen=np.zeros(0),
        boundary_enabled=0, box=1000.0)
