"""Writes the rate-table fixtures of tests/golden from the reference's input
files (run in the build container, where the reference is mounted):

- rateCoef_e_subset.csv: the header row, the threshold row and the first 161
  data rows (mean energy 0 .. 16 eV) of input/rateCoef_e.csv, every column;
- rateCoef_i.csv and coefEISEE.csv: the reference's files as they are.

They are data (BOLSIG+ rate coefficients and secondary-emission yields), kept
byte for byte so the readers see the reference's own formatting."""
import os
import shutil
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/input"
HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(SRC, "rateCoef_e.csv"), newline="") as f:
    lines = f.readlines()
with open(os.path.join(HERE, "rateCoef_e_subset.csv"), "w", newline="") as f:
    f.writelines(lines[:2 + 161])
for name in ("rateCoef_i.csv", "coefEISEE.csv"):
    shutil.copyfile(os.path.join(SRC, name), os.path.join(HERE, name))
