"""Writes exvol_20210901.json: the near-coplanar stress input of the
reference's excluded-volume test (src/TESTS/Mechanics/ForceField/Volume/
TestCylinderExclVolume.cpp:166-188, "the point of failure of a certain
simulation on Sep 1, 2021", disabled there because its energy/force
consistency check fails): 197 excluded-volume interactions as bead offsets
(x3), per-interaction equilibrium lengths of both cylinders, and the bead
coordinates; the interaction the reference singles out is index 68. Only the
numeric arrays are taken. Run once in a container that has the reference
checkout."""
import json
import os
import re

SRC = "/root/reference/src/TESTS/Mechanics/ForceField/Volume/TestCylinderExclVolume.cpp"
HERE = os.path.dirname(os.path.abspath(__file__))


def array(text, name):
    m = re.search(name + r"\[\]\s*\{([^}]*)\}", text)
    return [float(v) for v in m.group(1).split(",") if v.strip()]


text = open(SRC).read()
bs = [int(v) for v in array(text, "beadSet20210901")]
eql = array(text, "eqLength20210901")
coord = array(text, "coord20210901")
assert len(bs) == 4 * 197 and len(eql) == 2 * 197 and len(coord) % 3 == 0
with open(os.path.join(HERE, "exvol_20210901.json"), "w") as fh:
    json.dump({"source": "TestCylinderExclVolume.cpp:166-188", "n_interactions": 197, "unstable_index": 68,
               "k": 100.0, "beadset": bs, "eq_length": eql, "coord": coord}, fh)
print(len(bs) // 4, "interactions,", len(coord) // 3, "beads")
