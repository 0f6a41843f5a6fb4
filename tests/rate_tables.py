"""Independent Python reading of the rate-table fixtures (tests/golden, from
the reference's input/rateCoef_e.csv, rateCoef_i.csv, coefEISEE.csv) into
the lookup tables of cathode_ndd_set_rate_tables, following
include/inputModule.hpp:1407-1654 and 606-659: column kinds from row 1,
thresholds [eV] from row 2, per-row sums in column order, the energy-loss
rates over the ground-state and metastable kinds, dTe = 2/3 of the energy
step in K, dEi the ion table's step in J, dTe_SEE half the yield table's step."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
E_CSV = os.path.join(GOLDEN, "rateCoef_e_subset.csv")
I_CSV = os.path.join(GOLDEN, "rateCoef_i.csv")
SEE_CSV = os.path.join(GOLDEN, "coefEISEE.csv")
E0, KB = 1.60217662e-19, 1.380649e-23
KINDS = ["Elastic", "Excitation", "Excitation_reso", "Excitation_meta", "Ionization", "Ionization_step",
         "Excitation_step"]
NAMES = ["nu_elas", "nu_exc", "nu_excReso", "nu_excMeta", "nu_ionz", "nu_ionzStep", "nu_excStep", "nu_super"]


def rows(path):
    """lines split at '\n' only and fields at ',' (a '\r' stays in the last
    field, as the reference's getline-based reader leaves it)"""
    with open(path, newline="") as f:
        text = f.read()
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    return [ln.split(",") for ln in lines]


def number(s):
    return float(s.strip())


def kind_of(name):
    return KINDS.index(name) if name in KINDS else 7


def read(path_e=E_CSV, path_i=I_CSV):
    """(tables dict for abi.rate_tables_struct, per-column data, kinds, thresholds [J])"""
    R = rows(path_e)
    head, thr_row, data = R[0], R[1], R[2:]
    nc = len(head) - 1
    kinds = [kind_of(h) for h in head[1:]]
    thr = [number(v) * E0 for v in thr_row[1:nc + 1]]
    cols = [[number(r[c + 1]) for r in data] for c in range(nc)]
    nr = len(data)
    t = {n: [0.0] * nr for n in NAMES + ["eloss_n", "eloss_m"]}
    for k in range(nr):
        for c in range(nc):
            v = cols[c][k]
            t[NAMES[kinds[c]]][k] += v
            loss = "eloss_n" if kinds[c] <= 4 else "eloss_m"
            t[loss][k] += v * thr[c]
    out = {n: np.array(v, dtype=np.float64) for n, v in t.items()}
    step = number(data[1][0]) - number(data[0][0])
    out["dTe"] = step * 2.0 / 3.0 * E0 / KB
    Ri = rows(path_i)[2:]
    out["nui_m"] = np.array([sum_in_order([number(v) for v in r[1:]]) for r in Ri], dtype=np.float64)
    out["dEi"] = (number(Ri[1][0]) - number(Ri[0][0])) * E0
    return out, np.array(cols), kinds, np.array(thr)


def sum_in_order(vals):
    s = 0.0
    for v in vals:
        s += v
    return s


def read_see(path=SEE_CSV):
    R = rows(path)[1:]
    eb = np.array([number(r[1]) for r in R])
    rd = np.array([number(r[2]) for r in R])
    ts = np.array([number(r[3]) for r in R])
    step = number(R[1][0]) - number(R[0][0])
    return eb, rd, ts, step * 0.5 * E0 / KB
