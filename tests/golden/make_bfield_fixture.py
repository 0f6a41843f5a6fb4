"""Writes tests/golden/bfield_cathode_0p2mm.npz from the reference's
input/Bfield_data_0.2mm.csv (the magnet's field on the cathode grid, 194 x 70
cells of 0.2 mm), as InputModule::input_Bfield_data
(include/inputModule.hpp:518-565) fills gc.Bx and gc.Br: the file's cells
(columns B_x, B_r), then the ghost copies of its loops (i = 1..ni-1 for the
r ghosts, j = 1..nj-1 for the x ghosts; the corners and the ghosts next to
i = ni / j = nj keep 0).  Run in the build container, where the reference is
mounted.  The arrays are data: the reference's input field."""
import csv
import os
import sys

import numpy as np

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/input/Bfield_data_0.2mm.csv"
HERE = os.path.dirname(os.path.abspath(__file__))
NI, NJ = 194, 70

Bx = np.zeros((NI + 2, NJ + 2))
Br = np.zeros((NI + 2, NJ + 2))
with open(SRC, newline="") as f:
    rows = csv.reader(f)
    next(rows)
    for r in rows:
        i, j = int(r[0]), int(r[1])
        Bx[i, j] = float(r[4])
        Br[i, j] = float(r[5])
for a in (Bx, Br):
    for i in range(1, NI):
        a[i, 0] = a[i, 1]
        a[i, NJ + 1] = a[i, NJ]
    for j in range(1, NJ):
        a[0, j] = a[1, j]
        a[NI + 1, j] = a[NI, j]
np.savez_compressed(os.path.join(HERE, "bfield_cathode_0p2mm.npz"), Bx=Bx, Br=Br)
