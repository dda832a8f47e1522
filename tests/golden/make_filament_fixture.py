"""Writes filament-input fixtures from the reference's 2filaments example:
its FILAMENT input file (AFB2.txt) and the bead coordinates of the first
filament in the first model of its sample PDB output (nm; the PDB stores
nm / 10). Run once in a container that has the reference checkout."""
import os

SRC = "/root/reference/examples/2filaments"
HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(SRC, "AFB2.txt")) as fh:
    text = fh.read()
with open(os.path.join(HERE, "filament_input_2filaments.txt"), "w") as fh:
    fh.write(text)
beads = []
with open(os.path.join(SRC, "sampleResults", "2filament.pdb")) as fh:
    for line in fh:
        if line.startswith("TER"):
            break
        if line.startswith("ATOM"):
            resseq = int(line[22:26])
            x, y, z = float(line[30:38]), float(line[38:46]), float(line[46:54])
            beads.append((resseq, 10 * x, 10 * y, 10 * z))
first = [b for b in beads if b[0] < 10]
with open(os.path.join(HERE, "filament_2filaments_model1_fil0.txt"), "w") as fh:
    fh.write("# bead x y z (nm): first filament, MODEL 1 of examples/2filaments/sampleResults/2filament.pdb\n")
    for b in first:
        fh.write(f"{b[0]} {b[1]:.3f} {b[2]:.3f} {b[3]:.3f}\n")
