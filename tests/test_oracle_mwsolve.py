"""The oracle's coupled microwave solve (oracle/orc_mwsolve.c:
EmfieldModule::solve_Microwave, include/emfieldModule.hpp:2249-3321, and
solve_Microwave_impedanceTest, :5528-6141), without a GPU.

Pins:
- the reference's own impedance test (tests/golden/coax_impedance_er.csv, from
  test/校正/result_deg0.csv and result_deg4.csv): the normalised, cell-centred
  E_r of the coaxial feed at phases 0 and 80 degrees to the six printed
  digits, the excitation amplitude implied by its analytic column, and the
  amplifier's forward power of 校正後FWD_power.txt (12.57480 W).  This covers
  the E_r rows, the excitation, open and Gauss folds, the ghost rules, the
  reference-plane V and I and the power ratio;
- the plasma part of every row (with the rows of rhoe = 0 subtracted) applied
  to any field is i omega mu0 dx^2 dr^2 J, with J = sigma E computed by
  update_energy_profile's restatement (oracle/orc_mw.c, itself pinned by the
  cold-plasma momentum equation): the conductivity tensor and the averaging
  weights of the cross-component couplings;
- with B = 0 the E_phi rows decouple, and the coupled solve's E_phi equals the
  E_theta system of orc_helm (pinned by manufactured solutions) with the same
  conducting walls;
- the banded LU leaves ||b - A E|| / ||b|| at roundoff.
Each planted mistake (ORC_PLANT_MWS_*) fails its check.
"""
import csv
import os

import numpy as np
import pytest

import oracle_lib as O

W = O.workloads
abi = O.abi
HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "coax_impedance_er.csv")
FWD_POWER_AMP = 12.57480   # test/校正/校正後FWD_power.txt


@pytest.fixture
def clean():
    yield
    O.lib().orc_plant(0)


def coax_errors():
    p, lists, o, a = W.coax_impedance_test()
    rep = O.mw_solve(p, lists, o, a)
    r = p.rmin + (np.arange(p.nj + 2) - 0.5) * p.dr
    E = a["E1r"]
    err = 0.0
    with open(FIX, newline="") as f:
        rows = list(csv.reader(f))[1:]
    for row in rows:
        d, i, j = int(row[0]), int(row[1]), int(row[2])
        rL, rR = (r[j - 1] + r[j]) / 2.0, (r[j + 1] + r[j]) / 2.0
        Ec = (rR * E[i, j + 1] + rL * E[i, j]) / (2.0 * r[j]) * np.exp(1j * 2.0 * np.pi / 18.0 * d)
        ref = complex(float(row[3]), float(row[4]))
        err = max(err, abs(Ec - ref) / abs(ref))
    # analytic column: J1r_exc(new) / r * 1e3 / ln(r2 / r1) * exp(-i omega x / c0) at deg 0
    i, j = int(rows[0][1]), int(rows[0][2])
    c0 = 1.0 / np.sqrt(W.EPS0 * 2.083825280349 * W.MU0)
    x = (i - 0.5) * p.dx
    anly = complex(float(rows[0][5]), float(rows[0][6]))
    J1_fix = abs(anly) * r[j] * np.log(p.r2 / p.r1) / 1e3
    assert abs(np.angle(anly) + p.omega / c0 * x) < 1e-5
    return err, rep, J1_fix


def test_coax_impedance_fixture(clean):
    err, rep, J1_fix = coax_errors()
    assert err < 1e-5, err
    assert rep.rel_residual < 1e-12
    assert abs(rep.J1r_exc / J1_fix - 1.0) < 1e-5
    assert abs(rep.P_fwd - 8.0) < 1e-12                      # Pmw at the cathode port
    assert abs(rep.P_fwd_all / FWD_POWER_AMP - 1.0) < 5e-6   # the reference's recorded value
    assert rep.n_x == 0 and rep.n_p == 0 and rep.n_r == 194 * 6


@pytest.mark.parametrize("plant", [O.ORC_PLANT_MWS_EXCITE_SIGN, O.ORC_PLANT_MWS_GAUSS_RADII])
def test_coax_fixture_catches(clean, plant):
    O.lib().orc_plant(plant)
    err, _, _ = coax_errors()
    assert err > 1e-3, err


def sigma_error():
    p, lists, o, a = W.mw_coupled_rect(ni=20, nj=10, seed=5)
    rng = np.random.default_rng(2)
    shp = a["E1x"].shape
    E = []
    for k in ("act_x", "act_r", "act_p"):
        e = rng.uniform(-1, 1, shp) + 1j * rng.uniform(-1, 1, shp)
        E.append(np.ascontiguousarray(e * (a[k] == 1)))
    coefP = O.mw_rows(p, lists, o, a)
    vac = W.copy_arrays(a)
    vac["rhoe"] = np.zeros_like(a["rhoe"])
    coef0 = O.mw_rows(p, lists, o, vac)
    yP = O.mw_apply(p, a, o.which, coefP, *E)
    y0 = O.mw_apply(p, a, o.which, coef0, *E)
    # J = sigma E by update_energy_profile's restatement
    m = {"rhoe": a["rhoe"], "nu_m": a["nu_m"], "Bx": a["Bx"], "Br": a["Br"],
         "E1x": E[0].copy(), "E1r": E[1].copy(), "E1p": E[2].copy()}
    for k in ("J1x", "J1r", "J1p", "J1x_exc", "J1r_exc", "J1p_exc"):
        m[k] = np.zeros(shp, dtype=np.complex128)
    for k in ("Pabsx", "Pabsr", "Pabsp", "Pabs"):
        m[k] = np.zeros(shp)
    p.flag_mwRef = 0
    p.j_out, p.i_out_lo, p.i_out_hi, p.j_in_lo, p.j_in_hi = p.nj - 1, 1, p.ni, 1, p.nj
    O.mw_energy_profile(p, m)
    h4 = p.dx * p.dx * p.dr * p.dr
    iwm = 1j * p.omega * W.MU0
    worst = 0.0
    for c, (J, act) in enumerate(((m["J1x"], a["act_x"]), (m["J1r"], a["act_r"]), (m["J1p"], a["act_p"]))):
        got = (yP[c] - y0[c])[act == 1]
        want = (iwm * h4 * J)[act == 1]
        worst = max(worst, np.abs(got - want).max() / np.abs(want).max())
    return worst


def test_plasma_rows_are_the_conductivity_current(clean):
    assert sigma_error() < 1e-10   # the vacuum parts cancel to roundoff of the larger rows


def test_conductivity_check_catches_the_weight(clean):
    O.lib().orc_plant(O.ORC_PLANT_MWS_RX_WEIGHT)
    assert sigma_error() > 1e-3


def test_bzero_ephi_equals_the_etheta_system(clean):
    ni, nj, h = 24, 12, 2e-4
    p, lists, o, a = W.mw_coupled_rect(ni=ni, nj=nj, h=h, B0=0.0, seed=9)
    nu, epsr = 2e9, 1.7
    a["nu_m"][:] = nu
    a["epsr"][:] = epsr
    a["act_x"][:] = 0
    a["act_r"][:] = 0
    lists = {"x": [[], [], [], []], "r": [[], [], [], []],
             "p": [[(1, j, 0) for j in range(1, nj + 1)], [(ni, j, 0) for j in range(1, nj + 1)],
                   [(i, 1, 0) for i in range(1, ni + 1)], [(i, nj, 0) for i in range(1, ni + 1)]]}
    rep = O.mw_solve(p, lists, o, a)
    assert rep.n_x == 0 and rep.n_r == 0 and rep.n_p == ni * nj and rep.rel_residual < 1e-12
    hp = O.HelmParams(ni=ni, nj=nj, dx=h, dr=h, rmin=p.rmin, omega=p.omega, nu=nu, epsr=epsr,
                      masse=W.MASSE, pec_z=1, pec_rmax=1)
    shp = (ni + 2, nj + 2)
    co = [np.zeros(shp, dtype=np.complex128) for _ in range(6)]
    J = np.ascontiguousarray(a["J1p_exc"])
    ne = np.ascontiguousarray(a["rhoe"])
    O.lib().orc_helm_assemble(hp, O.P(ne), O.P(J.view(np.float64)), *[O.P(c.view(np.float64)) for c in co])
    x = np.zeros(shp, dtype=np.complex128)
    boxes = O.boxes_arr([1, ni, 1, nj])
    rc = O.lib().orc_helm5_direct(*[O.P(c.view(np.float64)) for c in co], boxes.ctypes.data_as(abi.c_int_p), 1,
                                  ni + 2, nj + 2, nj + 2, O.P(x.view(np.float64)))
    assert rc == 0
    inner = (slice(1, ni + 1), slice(1, nj + 1))
    err = np.abs(a["E1p"][inner] - x[inner]).max() / np.abs(x[inner]).max()
    assert err < 1e-10, err


def test_banded_lu_residual():
    p, lists, o, a = W.mw_coupled_rect(ni=30, nj=14, seed=3)
    rep = O.mw_solve(p, lists, o, a)
    assert rep.n_x == 29 * 14 and rep.n_r == 30 * 13 and rep.n_p == 30 * 14
    assert rep.rel_residual < 1e-12, rep.rel_residual
    assert np.isfinite(a["E1x"]).all() and np.isfinite(a["E1r"]).all() and np.isfinite(a["E1p"]).all()


# ---- the folds against the ghost rules and div E = 0 ----------------------
GHOST = {0: (-1, 0), 1: (1, 0), 2: (0, -1), 3: (0, 1)}   # side -> ghost offset (W, E, S, N)
NO_POST = {("x", 2, 6), ("x", 3, 5), ("x", 3, 6), ("r", 0, 5), ("r", 0, 6), ("r", 1, 5),
           ("r", 1, 1),   # convex corners (no ghost rule) and the east open E_r ghost (built from gr.Er)
           ("r", 2, 3), ("r", 3, 3), ("r", 3, 4)}   # concave E_r folds: the E_x slot one row off their
# ghost rules in the reference (:2678 vs :3185, :2700 vs :3203, :2708 vs :3209), kept as written
TARGET = {0: {O.MWS_E: (0, 1, 0), O.MWS_W: (0, -1, 0), O.MWS_N: (0, 0, 1), O.MWS_S: (0, 0, -1),
              O.MWS_NE: (1, 0, 1), O.MWS_NW: (1, -1, 1), O.MWS_SE: (1, 0, 0), O.MWS_SW: (1, -1, 0),
              O.MWS_EE: (2, 0, 0), O.MWS_WW: (2, -1, 0)},
          1: {O.MWS_E: (1, 1, 0), O.MWS_W: (1, -1, 0), O.MWS_N: (1, 0, 1), O.MWS_S: (1, 0, -1),
              O.MWS_NE: (0, 1, 0), O.MWS_SE: (0, 1, -1), O.MWS_NW: (0, 0, 0), O.MWS_SW: (0, 0, -1),
              O.MWS_NN: (2, 0, 0), O.MWS_SS: (2, 0, -1)},
          2: {O.MWS_E: (2, 1, 0), O.MWS_W: (2, -1, 0), O.MWS_N: (2, 0, 1), O.MWS_S: (2, 0, -1),
              O.MWS_EE: (0, 1, 0), O.MWS_WW: (0, 0, 0), O.MWS_NN: (1, 0, 1), O.MWS_SS: (1, 0, 0)}}


def fold_identity_error(seed=21, case=None):
    """For every face or cell in the lists whose rules all have a ghost rule:
    (P_f - P_0) E - sum_k (a_f,k - a_0,k) E_k = b_f - b_0, with the unfolded
    rows (no lists), the folded ones, and E after the solve and the ghost
    rules (the fold is the ghost rule substituted into the row)."""
    p, lists, o, a = case if case is not None else W.mw_coupled_rect(ni=20, nj=10, seed=seed)
    ni, nj = p.ni, p.nj
    empty = {"x": [[], [], [], []], "r": [[], [], [], []], "p": [[], [], [], []]}
    c0 = O.mw_rows(p, empty, o, a)
    cf = O.mw_rows(p, lists, o, a)
    g = W.copy_arrays(a)
    O.mw_solve(p, lists, o, g)
    E = [g["E1x"], g["E1r"], g["E1p"]]
    act = [a["act_x"], a["act_r"], a["act_p"]]
    entries, written = {}, {}
    for comp, c in (("x", 0), ("r", 1), ("p", 2)):
        for side in range(4):
            for (i, j, s) in lists[comp][side]:
                entries.setdefault((c, i, j), []).append((comp, side, s))
                gh = (c, i + GHOST[side][0], j + GHOST[side][1])
                written[gh] = written.get(gh, 0) + 1
    worst, n = 0.0, 0
    for (c, i, j), es in entries.items():
        if any((comp, side, s) in NO_POST for comp, side, s in es):
            continue
        # a ghost that two entries write keeps the later rule's value (a corner cell behind two walls)
        if any(written[(c, i + GHOST[side][0], j + GHOST[side][1])] > 1 for _, side, _ in es):
            continue
        ghosts = {(c, i + GHOST[side][0], j + GHOST[side][1]) for _, side, _ in es}
        lhs = (cf[c, O.MWS_P, i, j] - c0[c, O.MWS_P, i, j]) * E[c][i, j]
        ok = True
        for k, (tc, di, dj) in TARGET[c].items():
            d = cf[c, k, i, j] - c0[c, k, i, j]
            if d == 0:
                continue
            t = (tc, i + di, j + dj)
            if t not in ghosts and act[tc][t[1], t[2]] != 1:
                ok = False   # the reference drops a folded slot whose target is no unknown
                break
            lhs -= d * E[tc][t[1], t[2]]
        if not ok:
            continue
        rhs = cf[c, O.MWS_B, i, j] - c0[c, O.MWS_B, i, j]
        scale = abs(c0[c, O.MWS_P, i, j] * E[c][i, j]) + abs(rhs) + 1e-300
        worst = max(worst, abs(lhs - rhs) / scale)
        n += 1
    return worst, n


def test_folds_are_the_ghost_rules_substituted(clean):
    err, n = fold_identity_error()
    assert n > 60 and err < 1e-12, (n, err)


@pytest.mark.parametrize("plant", [O.ORC_PLANT_MWS_GAUSS_RADII, O.ORC_PLANT_MWS_CONCAVE_DENO])
def test_fold_identity_catches(clean, plant):
    O.lib().orc_plant(plant)
    err, _ = fold_identity_error()
    assert err > 1e-6, err


def test_gauss_ghosts_make_the_boundary_cell_divergence_free():
    """Type-2 ghosts (:3035-3039, :3061-3065, :3172-3176, :3193-3197): the
    discrete div E of the cell across the boundary, (E_x(i+1) - E_x(i)) / dx +
    (r_{j+1/2} E_r(j+1) - r_{j-1/2} E_r(j)) / (r_j dr), is zero."""
    p, lists, o, a = W.mw_coupled_rect(ni=20, nj=10, seed=8)
    g = W.copy_arrays(a)
    O.mw_solve(p, lists, o, g)
    Ex, Er = g["E1x"], g["E1r"]
    r = p.rmin + (np.arange(p.nj + 2) - 0.5) * p.dr
    dx, dr = p.dx, p.dr

    def div(ic, jc):
        rR, rL = (r[jc] + r[jc + 1]) / 2.0, (r[jc] + r[jc - 1]) / 2.0
        t1 = (Ex[ic + 1, jc] - Ex[ic, jc]) / dx
        t2 = (rR * Er[ic, jc + 1] - rL * Er[ic, jc]) / (r[jc] * dr)
        return abs(t1 + t2) / (abs(t1) + abs(t2) + 1e-300)

    cells = []
    for (i, j, s) in lists["x"][0]:
        if s == 2:
            cells.append((i - 1, j))
    for (i, j, s) in lists["x"][1]:
        if s == 2:
            cells.append((i, j))
    for (i, j, s) in lists["r"][2]:
        if s == 2:
            cells.append((i, j - 1))
    for (i, j, s) in lists["r"][3]:
        if s == 2:
            cells.append((i, j))
    assert len(cells) > 20
    worst = max(div(i, j) for i, j in cells)
    assert worst < 1e-12, worst


# ---- the reference's cathode: makeBoundary's lists -------------------------
def test_cathode_lists_fit_their_masks():
    """Every entry of makeBoundary's restated lists is an unknown of its
    component, its ghost is not, and no list names a face or cell twice, at
    the reference's 0.2 mm and at 0.4 mm."""
    for h in (2e-4, 4e-4):
        p, lists, o, a = W.cathode_mw_case(h=h)
        act = {"x": a["act_x"], "r": a["act_r"], "p": a["act_p"]}
        n = 0
        for comp in "xrp":
            for side in range(4):
                seen = set()
                for (i, j, s) in lists[comp][side]:
                    gi, gj = i + GHOST[side][0], j + GHOST[side][1]
                    assert (i, j) not in seen and act[comp][i, j] == 1 and act[comp][gi, gj] != 1, (comp, side, i, j)
                    seen.add((i, j))
                    n += 1
        assert n > 700


def test_cathode_folds_and_residual(clean):
    case = W.cathode_mw_case(h=4e-4)
    err, n = fold_identity_error(case=case)
    assert n > 500 and err < 1e-12, (n, err)
    p, lists, o, a = W.cathode_mw_case(h=4e-4)
    rep = O.mw_solve(p, lists, o, a)
    assert rep.rel_residual < 1e-12 and rep.n_x > 1500 and rep.n_r > 1500 and rep.n_p > 1500
