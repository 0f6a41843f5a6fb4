"""Writes tests/golden/coax_impedance_er.csv from the reference's impedance-
test results (test/校正/result_deg0.csv .. result_deg17.csv, written by
OutputModule::output_phase, include/outputModule.hpp:38-180, after
EmfieldModule::solve_Microwave_impedanceTest with the setup.yaml of the same
folder).  Run in the build container, where the reference is mounted.

Kept: for deg0 (phase 0) and deg4 (phase 4 x 20 degrees), per cell i, j the
cell-centred E_r (real, imaginary) and the analytic TEM field Er_anly (real,
imaginary), exactly as printed there (six significant digits).  These are
data: the reference's own output of its coaxial-feed test.

Also writes tests/golden/result0_rows.csv: the header and the first three rows
of the reference's result file test/校正/result0.csv (OutputModule::output,
include/outputModule.hpp:185-426) and its rows inside the plasma
(jdgBnd_flc = 1) as they are, every fourth of them; the file holds the
first 48 x-columns of the cathode grid.  Data: the reference's own output."""
import csv
import os
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/校正"
HERE = os.path.dirname(os.path.abspath(__file__))
COLS = ["Er(R)", "Er(I)", "Er_anly(R)", "Er_anly(I)"]

out = []
for n in (0, 4):
    with open(os.path.join(SRC, "result_deg%d.csv" % n), newline="") as f:
        rows = list(csv.reader(f))
    h = rows[0]
    idx = [h.index(c) for c in COLS]
    for r in rows[1:]:
        out.append([str(n), r[0], r[1]] + [r[k] for k in idx])
with open(os.path.join(HERE, "coax_impedance_er.csv"), "w", newline="") as f:
    w = csv.writer(f, lineterminator="\n")
    w.writerow(["deg", "i", "j"] + COLS)
    w.writerows(out)

with open(os.path.join(SRC, "result0.csv"), newline="") as f:
    lines = f.read().split("\n")
head = lines[0].split(",")
k_flc = head.index("jdgBnd_flc")
plasma = [ln for ln in lines[4:] if ln and len(ln.split(",")) == len(head) and ln.split(",")[k_flc] == "1"]
keep = lines[:4] + plasma[::4]
with open(os.path.join(HERE, "result0_rows.csv"), "w", newline="") as f:
    f.write("\n".join(keep) + "\n")
