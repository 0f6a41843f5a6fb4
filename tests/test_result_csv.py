"""The result file writer (02_cathode-08_modulerev_amd/host/result_csv.hpp;
OutputModule::output, include/outputModule.hpp:185-426), without a GPU.

Pins, from the reference's own result file (tests/golden/result0_rows.csv:
the header, three rows outside the plasma and plasma rows of
test/校正/result0.csv, written by tests/golden/make_coax_fixture.py):
- the header, name for name;
- relations between columns that the reference's file satisfies to its six
  printed digits (|B|, the Hall parameters, the loss and net power, the
  direct and stepwise ionization, rho_i U_phi, |U_i|, |U_e|, the effective
  temperature, the metastable velocity, the Debye length, the metastable and
  neutral rates, the electron Mach number): the writer's output on a state
  with every field nonzero satisfies the same relations;
- outside the plasma (jdgBnd_flc = 0) every plasma column is a signed zero,
  P_ES is -0 without fluxes and phi_analytical is nan, as in the reference's
  rows; the microwave amplitudes and currents are not masked there.
"""
import csv
import math
import os
import subprocess

import numpy as np
import pytest

import host_build

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "result0_rows.csv")
E0, BOLTZ, EPS0, MASSE = 1.60217662e-19, 1.380649e-23, 8.8541878128e-12, 9.10938356e-31


def read(path):
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    return rows[0], [r for r in rows[1:] if r]


def relations(h, row):
    """(name, got, want[, scale]) of the cell-local relations, from one row's
    text; a difference of printed terms is measured against the terms' size"""
    v = {k: float(x) for k, x in zip(h, row)}
    ue = v["Ue_mag"]
    out = [
        ("Bmag", v["Bmag"], math.hypot(v["Bx"], v["Br"])),
        ("Omega", v["Omega"], E0 * v["Bmag"] / (MASSE * v["nu_m"])),
        ("Omegax", v["Omegax"], E0 * v["Bx"] / (MASSE * v["nu_m"])),
        ("Omegar", v["Omegar"], E0 * v["Br"] / (MASSE * v["nu_m"])),
        ("Loss", v["Loss"], v["rhoe"] * v["rate_eloss"]),
        ("Pnet", v["Pnet"], v["Pabs"] - v["Loss"] + v["P_ES"], abs(v["Pabs"]) + abs(v["Loss"]) + abs(v["P_ES"])),
        ("rate_ionize_direct", v["rate_ionize_direct"], v["rhoe"] * v["nu_ionz"]),
        ("rate_ionize_stepwise", v["rate_ionize_stepwise"], v["rhoe"] * v["nu_ionzStep"]),
        ("rhoUip", v["rhoUip"], v["rhoi"] * v["Uip"]),
        ("Ui_mag", v["Ui_mag"], math.hypot(v["Uix"], v["Uir"])),
        ("Ue_mag", ue, math.hypot(v["Uex"], v["Uer"])),
        ("Te_eff", v["Te_eff"], v["Te"] + MASSE * ue * ue / (3.0 * E0)),
        ("Umx", v["Umx"], v["rhoUmx"] / v["rhom"]),
        ("Umr", v["Umr"], v["rhoUmr"] / v["rhom"]),
        ("Debye", v["Debye"], math.sqrt(EPS0 * v["Te"] / ((v["rhoe"] + 1.0) * E0))),
        ("rateG_rhom", v["rateG_rhom"], v["rhoe"] * (v["nu_excMeta"] + 0.5 * v["nu_exc"])),
        ("rateL_rhom", v["rateL_rhom"], v["rhoe"] * (v["nu_ionzStep"] + v["nu_excStep"] + v["nu_super"])),
        ("rate_rhom", v["rate_rhom"], v["rateG_rhom"] - v["rateL_rhom"], abs(v["rateG_rhom"]) + abs(v["rateL_rhom"])),
        ("rate_rhon", v["rate_rhon"], (v["nu_excStep"] + v["nu_super"]) * v["rhoe"]
         - (v["nu_excMeta"] + 0.5 * v["nu_exc"] + v["nu_ionz"]) * v["rhoe"],
         v["rhoe"] * (v["nu_excStep"] + v["nu_super"] + v["nu_excMeta"] + 0.5 * v["nu_exc"] + v["nu_ionz"])),
        ("Mache", v["Mache"], ue / (math.sqrt(2.0 * BOLTZ * (v["Te"] * E0 / BOLTZ + E0 / BOLTZ) / MASSE) + 100.0)),
    ]
    return out


def check_relations(h, rows, tol):
    worst = {}
    for row in rows:
        for rel in relations(h, row):
            name, got, want = rel[:3]
            scale = max(abs(got), abs(want), rel[3] if len(rel) > 3 else 0.0, 1e-300)
            worst[name] = max(worst.get(name, 0.0), abs(got - want) / scale)
    bad = {k: e for k, e in worst.items() if not e <= tol}
    return bad


def test_reference_file_satisfies_the_relations():
    h, rows = read(FIX)
    plasma = [r for r in rows if r[h.index("jdgBnd_flc")] == "1"]
    assert len(plasma) > 50
    bad = check_relations(h, plasma, 2e-5)   # six printed digits
    assert not bad, bad


def write_state(tmp, ni, nj, seed=3):
    """A state with every field nonzero on ni x nj cells of 0.2 mm; plasma
    (jdgBnd_flc = 1) in cells i > ni / 2."""
    rng = np.random.default_rng(seed)
    shp = (ni + 2, nj + 2)
    u = lambda lo, hi: rng.uniform(lo, hi, shp)
    f = {
        "rhoe": u(1e16, 3e17), "rhoe_old": u(1e16, 3e17), "rhoi": u(1e16, 3e17), "rhoi_old": u(1e16, 3e17),
        "Uix": u(-2e3, 2e3), "Uir": u(-2e3, 2e3), "Uip": u(-50, 50), "rhoUix": u(-1e20, 1e20),
        "rhoUir": u(-1e20, 1e20), "rhoUix_wall": u(-1e18, 1e18), "rhoUir_wall": u(-1e18, 1e18),
        "rhoUex": u(-1e21, 1e21), "rhoUer": u(-1e21, 1e21), "Te": u(1e4, 6e4), "rhoeps": u(0.01, 0.2),
        "rhoeps_old": u(0.01, 0.2), "Gx": u(-20, 20), "Gr": u(-20, 20), "phi": u(0, 30), "Ex": u(-2e3, 2e3),
        "Er": u(-2e3, 2e3), "rate_ionize": u(1e19, 1e22), "nu_m": u(5e8, 2e9), "nu_m1": u(1e9, 3e9),
        "nu_ionz": u(1e3, 3e4), "nu_ionzStep": u(1e2, 2e3), "nu_excStep": u(1e3, 3e4), "nu_exc": u(1e3, 6e4),
        "nu_excMeta": u(1e3, 2e4), "nui_m": u(1e5, 4e6), "mu_para": u(100, 300), "mu_perp": u(0.5, 2),
        "mu_xx": u(10, 200), "mu_rr": u(10, 200), "mu_xr": u(-60, 60), "Ap": u(-5e-5, 5e-5),
        "Bx": u(-0.1, 0.1), "Br": u(-0.06, 0.06), "Pabs": u(10, 300), "rate_eloss": u(1e-14, 3e-13),
        "rhom": u(1e16, 7e16), "rhom_old": u(1e16, 7e16), "rhoUmx": u(-3e18, 3e18), "rhoUmr": u(-3e18, 3e18),
        "rhon": u(2e20, 3e20), "rhon_old": u(2e20, 3e20), "rhoUnx": u(-2e20, 2e20), "rhoUnr": u(-2e20, 2e20),
        "distECR": u(0, 5e-3), "scx": u(-1e-6, 1e-6), "scr": u(-1e-6, 1e-6), "nUex": u(-1e21, 1e21),
        "nUer": u(-1e21, 1e21), "nu_super": u(10, 60), "nu_ano": u(1e8, 1e9), "nu_ano_IAT": u(1e6, 3e7)}
    flc = np.zeros(shp)
    flc[ni // 2 + 1:ni + 1, 1:nj + 1] = 1.0
    f["jdgBnd_flc"] = flc
    f["jdgBnd_Ep"] = np.ones(shp)
    for k, v in f.items():
        np.ascontiguousarray(v, dtype=np.float64).tofile(os.path.join(tmp, k + ".bin"))
    for k in ("E1x", "E1r", "E1p", "J1x", "J1r", "J1p"):
        c = rng.uniform(-1e3, 1e3, shp) + 1j * rng.uniform(-1e3, 1e3, shp)
        np.ascontiguousarray(c, dtype=np.complex128).tofile(os.path.join(tmp, k + ".cbin"))
    np.linspace(1e-14, 4e-13, 50).tofile(os.path.join(tmp, "eloss_n.bin"))
    np.linspace(1e-15, 2e-14, 50).tofile(os.path.join(tmp, "eloss_m.bin"))
    h = 2e-4
    lines = [f"ni {ni}", f"nj {nj}", "xL 0", f"dx {h}", "rmin 0", f"dr {h}", "dt 1e-10", "ndt_i 1", "ndt_m 1",
             "ndt_n 1", f"masse {MASSE}", "massi 2.18017156e-25", "nu_eff 2e9", "deltaECR 1e-3",
             f"dTe {0.1 * 2 / 3 * E0 / BOLTZ}", "n_bl 6"]
    # five blocks over the grid plus a feed block (the wall flags read blocks 0..4)
    boxes = [(ni // 2 + 1, ni, 1, nj), (ni // 2 + 1, ni, 1, nj), (ni // 2 + 1, ni, 1, nj), (ni // 2 + 1, ni, 1, nj),
             (ni // 2 + 1, ni, 1, nj), (1, ni // 2, 1, nj // 2)]
    for b in boxes:
        lines += [f"box {v}" for v in b]
    with open(os.path.join(tmp, "grid.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return f


def run_writer(tmp_path, ni=12, nj=8):
    exe = host_build.build_host_check()
    write_state(str(tmp_path), ni, nj)
    out = tmp_path / "result.csv"
    r = subprocess.run([exe, "result", str(tmp_path), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "ok result" in r.stdout, r.stdout + r.stderr
    return read(str(out))


def test_header_is_the_references(tmp_path):
    h, rows = run_writer(tmp_path)
    ref_h, _ = read(FIX)
    assert h == ref_h and len(h) == 102
    assert all(len(r) == 102 for r in rows) and len(rows) == 12 * 8
    assert [(int(r[0]), int(r[1])) for r in rows[:3]] == [(1, 1), (1, 2), (1, 3)]


def test_writer_satisfies_the_reference_relations(tmp_path):
    h, rows = run_writer(tmp_path)
    plasma = [r for r in rows if r[h.index("jdgBnd_flc")] == "1"]
    assert len(plasma) == 6 * 8
    bad = check_relations(h, plasma, 2e-5)
    assert not bad, bad


def test_outside_the_plasma_zeros_and_nan(tmp_path):
    h, rows = run_writer(tmp_path)
    ref_h, ref_rows = read(FIX)
    outside = [r for r in rows if r[h.index("jdgBnd_flc")] == "0"]
    assert outside
    masked_mw = {"|Ex_mw|", "|Er_mw|", "|Ep_mw|", "Jx_exc", "Jr_exc", "Jp_exc"}
    for r in outside:
        for k, v in zip(h[4:], r[4:]):
            if k in masked_mw:
                assert float(v) != 0.0, k      # the amplitudes carry jdgBnd_Ep, the currents nothing
            elif k == "phi_analytical":
                assert v in ("nan", "-nan", "0", "-0")
            else:
                assert v in ("0", "-0"), (k, v)
    # the reference's rows outside the plasma: P_ES is -0 (no fluxes) and phi_analytical nan
    for r in ref_rows[:3]:
        d = dict(zip(ref_h, r))
        assert d["jdgBnd_flc"] == "0" and d["P_ES"] == "-0" and d["phi_analytical"] == "nan"
