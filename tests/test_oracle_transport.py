"""Transport coefficients and their tables, without a GPU.

- host/tables.hpp reads the reference's rate and yield tables (the committed
  fixtures tests/golden/rateCoef_e_subset.csv, rateCoef_i.csv, coefEISEE.csv,
  made by tests/golden/make_rate_fixture.py from the reference's input/) to
  the same bits as an independent Python reading (tests/rate_tables.py);
- that reading is pinned by the file itself: a kind with one column equals
  that column, the excitation and energy-loss tables equal exact (fsum)
  sums over the columns of their kinds (any order), and the Te step is 2/3 of
  the energy step (mean energy 3/2 k Te);
- the oracle's lookups (include/miniFuncs.hpp:35-91) are exact on linear
  tables below the last interval; from the last interval on (iR = n - 1)
  func_Te returns the last entry, func_dTe the last interval's slope, and
  func_Ei the last cross-section times the speed (the last entry at the
  table's end, proportional to sqrt(E) beyond);
- the oracle's update_transport_coef (oracle/orc_transport.c): the mobility
  tensor maps the field direction b to mu_para b and an in-plane normal to
  mu_para / (1 + W^2) (the magnetised mobility), the tables enter as
  n_n f(Te) / n_m f(Te), the ion-acoustic rate scales as sqrt(n_i) and with
  the relative drift, vanishes with it, and the Mach number is the drift of
  the face-averaged electron flux over n_e + 1e-3 max n_e.
Each planted mistake (ORC_PLANT_TR_*) fails its check.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import host_build
import oracle_lib as O
import rate_tables as RT

W = O.workloads
abi = O.abi


@pytest.fixture
def clean():
    yield
    O.set_geometry(None)
    O.lib().orc_plant(0)


def test_host_reader_matches_python_reading(tmp_path):
    exe = host_build.build_host_check()
    out = tmp_path / "tables.txt"
    r = subprocess.run([exe, "tables", RT.E_CSV, RT.I_CSV, RT.SEE_CSV, str(out)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and "ok tables" in r.stdout, r.stdout + r.stderr
    got = {}
    for line in out.read_text().splitlines():
        f = line.split()
        got[f[0]] = np.array([float(v) for v in f[2:]])
        assert int(f[1]) == len(f) - 2
    t, *_ = RT.read()
    for k in RT.NAMES + ["eloss_n", "eloss_m", "nui_m"]:
        assert np.array_equal(got[k], t[k]), k
    assert got["dTe"][0] == t["dTe"] and got["dEi"][0] == t["dEi"]
    eb, rd, ts, dte = RT.read_see()
    assert np.array_equal(got["see_eb"], eb) and np.array_equal(got["see_rd"], rd)
    assert np.array_equal(got["see_ts"], ts) and got["dTe_SEE"][0] == dte


def test_python_reading_is_pinned_by_the_file():
    t, cols, kinds, thr = RT.read()
    kinds = np.array(kinds)
    # single-column kinds are that column
    for kd, name in ((0, "nu_elas"), (3, "nu_excMeta"), (4, "nu_ionz"), (5, "nu_ionzStep"), (7, "nu_super")):
        (c,) = np.nonzero(kinds == kd)[0]
        assert np.array_equal(t[name], cols[c]), name
    # multi-column kinds and the loss rates are exact sums over their columns
    for kd, name in ((1, "nu_exc"), (2, "nu_excReso"), (6, "nu_excStep")):
        sel = np.nonzero(kinds == kd)[0]
        exact = np.array([math.fsum(cols[sel, k]) for k in range(cols.shape[1])])
        assert np.allclose(t[name], exact, rtol=1e-14, atol=0), name
    for name, sel in (("eloss_n", kinds <= 4), ("eloss_m", kinds >= 5)):
        exact = np.array([math.fsum(cols[sel, k] * thr[sel]) for k in range(cols.shape[1])])
        assert np.allclose(t[name], exact, rtol=1e-14, atol=1e-300), name
    # energy step 0.1 eV of mean energy: Te step 2/3 of it
    assert abs(t["dTe"] * RT.KB / RT.E0 - 0.1 * 2.0 / 3.0) < 1e-15
    assert len(t["nu_elas"]) == 161 and (kinds == 1).sum() == 46


def test_lookups_exact_on_linear_tables():
    L = O.lib()
    n, a, b, dTe = 50, 3.0, 0.25, 17.0
    v = np.ascontiguousarray(a + b * np.arange(n), dtype=np.float64)
    P = O.P(v)
    for Te in (0.0, 5.3, 17.0, 400.1, 47.5 * dTe):
        want = a + b * Te / dTe
        assert abs(L.orc_func_Te_tab(P, n, Te, dTe) - want) <= 1e-13 * want
        assert abs(L.orc_func_dTe_tab(P, n, Te, dTe) - b / dTe) <= 1e-13 * b / dTe
    # from the last interval on (iR = n - 1): the last entry, and the last
    # interval's slope
    assert L.orc_func_Te_tab(P, n, 48.5 * dTe, dTe) == v[-1]
    assert L.orc_func_Te_tab(P, n, 60.0 * dTe, dTe) == v[-1]
    assert abs(L.orc_func_dTe_tab(P, n, 60.0 * dTe, dTe) - b / dTe) <= 1e-13 * b / dTe
    # func_Ei: from the last interval on, the last cross-section times the
    # speed (equal to the last entry at the table's end, then ~ sqrt(E))
    dEi, m = 2.0, 3.0
    e_end = (n - 1) * dEi
    assert abs(L.orc_func_Ei_tab(P, n, 10.5 * dEi, dEi, m) - (a + b * 10.5)) <= 1e-13 * (a + b * 10.5)
    assert abs(L.orc_func_Ei_tab(P, n, e_end, dEi, m) - v[-1]) <= 1e-13 * v[-1]
    assert abs(L.orc_func_Ei_tab(P, n, 4.0 * e_end, dEi, m) - 2.0 * v[-1]) <= 1e-13 * v[-1]


def _run(h=4e-4, B_scale=1.0, **kw):
    ni, nj, g, *_ = W.reference_geometry(h)
    O.set_geometry(g)
    grid = W.default_params(ni, nj, h, h)
    a = W.transport_state(g, ni, nj, h, **kw)
    a["Bx"] *= B_scale
    a["Br"] *= B_scale
    t, *_ = RT.read()
    b = W.copy_arrays(a)
    O.transport_update(grid, W.transport_params(), t, b)
    return g, ni, nj, a, b, t


def _tensor_error():
    # fields of 30..240 uT: the Hall parameter W = e B / (m_e nu_m) around 1
    g, ni, nj, a, b, t = _run(B_scale=3e-3)
    Bx, Br = a["Bx"], a["Br"]
    Bm = np.hypot(Bx, Br)
    bx, br = Bx / Bm, Br / Bm
    M = [[b["mu_xx"], b["mu_xr"]], [b["mu_xr"], b["mu_rr"]]]
    par = (M[0][0] * bx + M[0][1] * br, M[1][0] * bx + M[1][1] * br)
    nx, nr = -br, bx
    per = (M[0][0] * nx + M[0][1] * nr, M[1][0] * nx + M[1][1] * nr)
    Wc = W.E0 * Bm / (W.MASSE * b["nu_m"])
    mu = b["mu_para"]
    e1 = max((np.abs(par[0] - mu * bx) / mu).max(), (np.abs(par[1] - mu * br) / mu).max())
    mperp = mu / (1.0 + Wc ** 2)
    e2 = max((np.abs(per[0] - mperp * nx) / mperp).max(), (np.abs(per[1] - mperp * nr) / mperp).max())
    assert Wc.min() < 1.0 < Wc.max()
    # mu_para is e / (m_e nu_m)
    e3 = np.abs(mu * W.MASSE * b["nu_m"] / W.E0 - 1.0).max()
    return max(e1, e2, e3), b, mperp


def test_mobility_tensor(clean):
    err, b, mperp = _tensor_error()
    assert err < 1e-11, err
    assert np.allclose(b["mu_perp"], mperp, rtol=1e-12, atol=0)


def test_mobility_tensor_catches_the_wrong_denominator(clean):
    O.lib().orc_plant(O.ORC_PLANT_TR_DENO)
    err, *_ = _tensor_error()
    assert err > 1e-3


def test_tables_enter_as_density_times_lookup(clean):
    g, ni, nj, a, b, t = _run()
    L = O.lib()
    rng = np.random.default_rng(3)
    for _ in range(40):
        i, j = int(rng.integers(0, ni + 2)), int(rng.integers(0, nj + 2))
        Te = a["Te"][i, j]
        for name, dens in (("nu_elas", "rhon"), ("nu_ionz", "rhon"), ("nu_ionzStep", "rhom"), ("nu_super", "rhom")):
            v = L.orc_func_Te_tab(O.P(t[name]), len(t[name]), Te, t["dTe"])
            assert b[name][i, j] == a[dens][i, j] * v, (name, i, j)
        assert b["rate_ionize"][i, j] == a["rhoe"][i, j] * (b["nu_ionz"][i, j] + b["nu_ionzStep"][i, j])


def _iat(scale_ni=1.0, scale_drift=1.0):
    h = 4e-4
    ni, nj, g, *_ = W.reference_geometry(h)
    O.set_geometry(g)
    grid = W.default_params(ni, nj, h, h)
    a = W.transport_state(g, ni, nj, h)
    a["rhoi"] = a["rhoi"] * scale_ni
    for k in ("Uix", "Uir", "rhoUex", "rhoUer"):
        a[k] = a[k] * scale_drift
    t, *_ = RT.read()
    b = W.copy_arrays(a)
    O.transport_update(grid, W.transport_params(), t, b)
    m = W.plasma_mask(g, ni, nj)
    return b["nu_ano_IAT"][m], b["Mache"][m]


def test_ion_acoustic_rate_scalings(clean):
    base, mach = _iat()
    four, _ = _iat(scale_ni=4.0)
    assert np.allclose(four, 2.0 * base, rtol=1e-14, atol=0)          # ~ omega_pi ~ sqrt(n_i)
    twice, mach2 = _iat(scale_drift=2.0)
    assert np.allclose(twice, 2.0 * base, rtol=1e-14, atol=0)         # ~ relative drift
    assert np.allclose(mach2, 2.0 * mach, rtol=1e-14, atol=0)         # Mach ~ electron drift
    zero, mach0 = _iat(scale_drift=0.0)
    assert not zero.any() and not mach0.any()


def test_ion_acoustic_rate_catches_the_wrong_density(clean):
    O.lib().orc_plant(O.ORC_PLANT_TR_IAT_DENSITY)
    base, _ = _iat()
    four, _ = _iat(scale_ni=4.0)
    assert not np.allclose(four, 2.0 * base, rtol=1e-6, atol=0)
