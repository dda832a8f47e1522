"""MEDYAN's filament input format, as the caller of the mechanics path reads it.

* ``read_filament_file``: FilamentParser::readFilaments (Parser.cpp:3060-3094):
  lines holding '#' are skipped; a line of at least 8 tokens starting with
  ``FILAMENT`` gives the filament type and then 3-D points.
* ``filament_beads``: the bead coordinates MEDYAN creates from one entry
  (Controller.cpp:583-606 and the Filament constructor,
  Structure/Filament.cpp:102-160): two points give
  max(1, round(d / cylinderSize)) cylinders, more points one cylinder per
  point; the STRAIGHT projection (Filament.cpp:770-780) places bead i at
  p0 + dir * (i * cylinderSize), ZIGZAG alternates the direction
  (Filament.cpp:782-797), PREDEFINED keeps the points.
* ``network_from_filament_file``: a Network of those filaments
  (network.from_filaments) for the minimizer.
"""
import math

import numpy as np

from . import network


def read_filament_file(path_or_text):
    """[(type, points [n x 3])] from a FILAMENT input file (path or text)."""
    if "\n" in path_or_text or path_or_text.lstrip().startswith("FILAMENT"):
        text = path_or_text
    else:
        with open(path_or_text) as fh:
            text = fh.read()
    out = []
    for line in text.splitlines():
        if "#" in line:
            continue
        tok = line.split()
        if len(tok) >= 8 and tok[0] == "FILAMENT":
            ftype = int(tok[1])
            vals = tok[2:]
            if len(vals) % 3:
                raise ValueError("In reading filament coordinates, the number of coordinates is not divisible by 3.")
            pts = np.array([float(v) for v in vals], dtype=np.float64).reshape(-1, 3)
            out.append((ftype, pts))
    return out


def filament_beads(points, cyl_size=108.0, projection="STRAIGHT"):
    """Bead coordinates of one filament entry."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    if len(pts) < 2:
        raise ValueError("Filament data specified contains insufficient coordinates.")
    if len(pts) == 2:
        d = float(np.sqrt(((pts[1] - pts[0]) ** 2).sum()))
        # std::round: halves round away from zero (Controller.cpp:592-595)
        nseg = max(1, int(math.floor(d / cyl_size + 0.5)))
    else:
        nseg = len(pts) - 1
    nb = nseg + 1
    if projection == "PREDEFINED":
        if len(pts) != nb:
            raise ValueError("PREDEFINED projection needs one point per bead")
        return pts.copy()
    v1, v2 = pts[0], pts[1]
    dist = float(np.sqrt(((v2 - v1) ** 2).sum()))
    dirv = (v2 - v1) / dist
    if projection == "STRAIGHT":
        return np.array([v1 + dirv * (i * cyl_size) for i in range(nb)])
    if projection == "ZIGZAG":
        alt = np.array([-dirv[1], dirv[0], dirv[2]])
        return np.array([(i * cyl_size) * (dirv if i % 2 == 0 else alt) + v1 for i in range(nb)])
    raise ValueError(f"projection {projection!r} not supported (STRAIGHT, ZIGZAG, PREDEFINED)")


def network_from_filament_file(path_or_text, box, cyl_size=108.0, projection="STRAIGHT", **kw):
    """Network of the filaments in a FILAMENT input file (keyword arguments go
    to network.from_filaments: potentials, linkers, motors, branches)."""
    fils = [filament_beads(p, cyl_size, projection) for _, p in read_filament_file(path_or_text)]
    return network.from_filaments(fils, box, cyl_size=cyl_size, **kw)
