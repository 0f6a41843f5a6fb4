"""Independent pins of the oracle's microwave power profile and port
normalisation (oracle/orc_mw.c; EmfieldModule::update_energy_profile,
emfieldModule.hpp:7630-8105; mWLoss_at_boundary_xL / rR, :8488-8668).

Conductivity.  The cold electron momentum equation m (nu + i w) v =
-e (E + v x B) with J = -e n v gives, for h = e B / (m (nu + i w)) and
sigma0 = n e^2 / (m (nu + i w)) and components ordered (x, r, phi) (a
right-handed triple: x cross r = phi),

    [[1, 0, -h_r], [0, 1, h_x], [h_r, -h_x, 1]] J = sigma0 E,

solved here by numpy for uniform fields (the face averages of the reference
are then exact).  With nu = 0 the tensor is anti-Hermitian: no absorption.
With B = 0, 0.5 Re(sigma0) |E|^2 is the Drude absorption and the power sums
are that density times the exact volumes of the control volumes (annuli
pi (r_top^2 - r_bottom^2) dx).

Poynting fluxes.  H from Faraday's law, H = (i / (w mu0)) curl E.  For the
TEM wave of a coaxial line, E_r = V0 exp(-i k x) / (r ln(b/a)), the discrete
reference-plane quantities have closed forms (the r sums are exact for 1/r
fields; the x differences give sin / tan factors):
    V = V0 exp(-i k x_f) cos(k dx / 2),
    I = 4 pi V0 exp(-i k x_f) sin(k dx / 2) / (w mu0 dx ln(b/a)),
    P_in = pi |V0|^2 sin(k dx) / (w mu0 dx ln(b/a)^2) sum_j dr / r_j,
with the continuum limits V0, V0 / Z and |V0|^2 / 2Z, Z = eta ln(b/a) / 2 pi.
For smooth general fields both fluxes converge at second order to the
quadrature of the continuous flux.

Port algebra.  With b2 = (V + Z0 I) / 2 sqrt(2 Z0), a2 = (V - Z0 I) /
2 sqrt(2 Z0): P_fwd - P_ref = Re(V conj(I)) / 2, the plasma impedance is
V / I, and the input reflection of the two-port terminated by Gamma_p is
S11 + S12 S21 Gamma_p / (1 - S22 Gamma_p) (the textbook two-port formula);
after scaling the amplifier's forward power is Pmw and absorbed + boundary
loss = P_fwd - P_ref.
"""
import numpy as np
import pytest

import mw_cases as M
import oracle_lib as O

OMEGA = 2.0 * np.pi * 2.45e9
K0 = OMEGA / M.C_LIGHT


@pytest.fixture
def plant():
    yield lambda k: O.lib().orc_plant(k)
    O.lib().orc_plant(0)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ------------------------------------------------------------------ conductivity
def _sigma(n, nu, Bx, Br):
    nus = nu + 1j * OMEGA
    s0 = n * M.E0 ** 2 / (M.MASSE * nus)
    hx, hr = M.E0 * Bx / (M.MASSE * nus), M.E0 * Br / (M.MASSE * nus)
    A = np.array([[1, 0, -hr], [0, 1, hx], [hr, -hx, 1]], dtype=np.complex128)
    return np.linalg.solve(A, s0 * np.eye(3)), s0


UNI = dict(n=1e17, Bx=0.06, Br=-0.035, Ex=(1 + 2j) * 1e3, Er=(-0.5 + 1j) * 1e3, Ep=(0.7 - 0.3j) * 1e3)


def _tensor_error(nu):
    ni, nj = 6, 5
    u = UNI
    p = M.params(ni, nj, 1e-4, 1e-4, 0.0, OMEGA)
    f = M.uniform_fields(ni, nj, u["n"], nu, u["Bx"], u["Br"], u["Ex"], u["Er"], u["Ep"])
    g, _ = M.run_oracle(p, f)
    sig, s0 = _sigma(u["n"], nu, u["Bx"], u["Br"])
    want = sig @ np.array([u["Ex"], u["Er"], u["Ep"]])
    got = (g["J1x"][2:ni + 1, 1:nj + 1], g["J1r"][1:ni + 1, 2:nj + 1], g["J1p"][1:ni + 1, 1:nj + 1])
    err = max(_rel(got[c], np.full(got[c].shape, want[c])) for c in range(3))
    return err, g, s0


def test_conductivity_is_the_cold_plasma_tensor(plant):
    err, _, _ = _tensor_error(1.3e9)
    assert err <= 1e-12, err


def test_conductivity_catches_hall_sign(plant):
    plant(O.ORC_PLANT_MW_HALL_SIGN)
    err, _, _ = _tensor_error(1.3e9)
    assert err > 1e-3, err


def test_collisionless_magnetised_plasma_absorbs_nothing():
    _, g, s0 = _tensor_error(0.0)
    u = UNI
    E = np.array([u["Ex"], u["Er"], u["Ep"]])
    J = np.array([g["J1x"][3, 3], g["J1r"][3, 3], g["J1p"][3, 3]])
    # J.E* is purely imaginary; its size is |sigma0| |E|^2
    assert abs((np.conj(E) @ J).real) <= 1e-12 * abs(s0) * np.vdot(E, E).real
    # and the hall terms are not negligible here
    assert abs(J[0] - s0 * E[0]) > 0.1 * abs(s0 * E[0])


# ------------------------------------------------------------------ absorbed power
def _drude_sums():
    ni, nj, dx, dr, rmin = 7, 6, 2e-4, 1.5e-4, 0.0
    n, nu = 3e17, 2.2e9
    Ex, Er, Ep = 800 - 300j, 200 + 900j, -400 + 100j
    p = M.params(ni, nj, dx, dr, rmin, OMEGA)
    f = M.uniform_fields(ni, nj, n, nu, 0.0, 0.0, Ex, Er, Ep)
    g, rep = M.run_oracle(p, f)
    re_s0 = n * M.E0 ** 2 * nu / (M.MASSE * (nu ** 2 + OMEGA ** 2))
    pos = M.positions(ni, nj, dx, dr, rmin)
    rf, rc = pos["r_face"], pos["r_cell"]
    # x faces 2..ni over cells 1..nj; r faces 2..nj (control volumes tile
    # r_cell[1]..r_cell[nj]) over cells 1..ni; cells 1..ni x 1..nj
    want = dict(Px=0.5 * re_s0 * abs(Ex) ** 2 * (ni - 1) * np.pi * dx * (rf[nj + 1] ** 2 - rf[1] ** 2),
                Pr=0.5 * re_s0 * abs(Er) ** 2 * ni * np.pi * dx * (rc[nj] ** 2 - rc[1] ** 2),
                Pp=0.5 * re_s0 * abs(Ep) ** 2 * ni * np.pi * dx * (rf[nj + 1] ** 2 - rf[1] ** 2))
    errs = {k: abs(getattr(rep, k) - v) / v for k, v in want.items()}
    dens = 0.5 * re_s0 * (abs(Ex) ** 2 + abs(Er) ** 2 + abs(Ep) ** 2) * (p.Pmw / (rep.Px + rep.Pr + rep.Pp))
    # cells whose two x faces and two r faces are all in the boxes
    errs["Pabs"] = _rel(g["Pabs"][2:ni, 2:nj], np.full((ni - 2, nj - 2), dens))
    return errs


def test_absorbed_power_is_drude_over_exact_volumes():
    errs = _drude_sums()
    assert max(errs.values()) <= 1e-12, errs


def test_absorbed_power_catches_rface_weight(plant):
    plant(O.ORC_PLANT_MW_RFACE_WEIGHT)
    assert _drude_sums()["Pr"] > 1e-2


def test_overlapping_blocks_count_twice():
    """every loop of the reference runs over the blocks in order: a box listed
    twice contributes twice to the sums and writes the same J"""
    ni, nj = 10, 8
    f = M.random_fields(3, ni, nj)
    one = [((1, ni, 1, nj), (2, ni, 1, nj), (1, ni, 2, nj))]
    g1, r1 = M.run_oracle(M.params(ni, nj, 1e-4, 1e-4, 0.0, OMEGA, boxes=one), f)
    g2, r2 = M.run_oracle(M.params(ni, nj, 1e-4, 1e-4, 0.0, OMEGA, boxes=one * 2), f)
    for k in ("Px", "Pr", "Pp"):
        assert abs(getattr(r2, k) - 2 * getattr(r1, k)) <= 1e-13 * abs(getattr(r1, k))
    for k in ("J1x", "J1r", "J1p", "Pabsx", "Pabsr", "Pabsp"):
        assert np.array_equal(g1[k], g2[k]), k


# ------------------------------------------------------------------ TEM coaxial wave
V0 = (3 + 4j) * 10.0


def _tem_case(ni=12, nj=10, dx=1e-3, rmin=2e-3, dr=2e-4, i_ref=5, Z0=None):
    pos = M.positions(ni, nj, dx, dr, rmin)
    a, b = rmin, rmin + nj * dr
    ln = np.log(b / a)
    f = M.empty_fields(ni, nj)
    f["rhoe"][:] = 1e16
    f["nu_m"][:] = 1e9
    f["E1r"][:] = V0 * np.exp(-1j * K0 * pos["x_cell"])[:, None] / (pos["r_face"][None, :] * ln)
    kw = dict(flag_mwRef=1, i_ref=i_ref, j_in_lo=1, j_in_hi=nj, r1=a, r2=b, j_out=nj - 1, i_out_lo=1, i_out_hi=ni)
    if Z0 is not None:
        kw["Z0"] = Z0
    p = M.params(ni, nj, dx, dr, rmin, OMEGA, **kw)
    g, rep = M.run_oracle(p, f)
    xf = pos["x_face"][i_ref]
    ph = np.exp(-1j * K0 * xf)
    want = dict(V=V0 * ph * np.cos(K0 * dx / 2),
                I=4 * np.pi * V0 * ph * np.sin(K0 * dx / 2) / (OMEGA * M.MU0 * dx * ln),
                P_in=np.pi * abs(V0) ** 2 * np.sin(K0 * dx) / (OMEGA * M.MU0 * dx * ln ** 2) *
                np.sum(dr / pos["r_cell"][1:nj + 1]),
                Z_tem=M.ETA * ln / (2 * np.pi))
    return rep, want


def test_tem_reference_plane_closed_forms(plant):
    rep, w = _tem_case()
    d = M.report_dict(rep)
    assert abs(d["V_ref"] - w["V"]) <= 1e-12 * abs(w["V"])
    assert abs(d["I_ref"] - w["I"]) <= 1e-12 * abs(w["I"])
    assert abs(d["P_in"] - w["P_in"]) <= 1e-12 * w["P_in"]
    # continuum limits: |V0|^2 / 2Z and Z (k dx = 0.05, dr / r <= 0.1 here)
    assert abs(d["P_in"] - abs(V0) ** 2 / (2 * w["Z_tem"])) <= 2e-3 * d["P_in"]
    assert abs(d["Z_ref"] - w["Z_tem"]) <= 1e-3 * w["Z_tem"]


def test_tem_catches_poynting_conj(plant):
    plant(O.ORC_PLANT_MW_POYNTING_CONJ)
    rep, w = _tem_case()
    assert abs(rep.P_in - w["P_in"]) > 1e-2 * w["P_in"]


def test_tem_matched_port_reflects_nothing():
    rep0, w = _tem_case()
    rep, _ = _tem_case(Z0=M.report_dict(rep0)["Z_ref"].real)
    d = M.report_dict(rep)
    assert abs(d["Gamma_p"]) <= 1e-9
    assert d["P_ref"] <= 1e-15 * d["P_fwd"]


# ------------------------------------------------------------------ smooth fields
LX, LR, RMIN = 0.02, 0.01, 0.005


def _smooth(x, r, with_phi):
    return M.smooth(x, r, with_phi, OMEGA)


def _flux_exact(with_phi):
    t, w = np.polynomial.legendre.leggauss(64)
    # x face at LX / 2, r over the whole grid
    r = RMIN + (t + 1) / 2 * LR
    _, Er, _, Hp, _ = _smooth(LX / 2, r, False)
    p_in = np.sum(w * LR / 2 * 2 * np.pi * r * 0.5 * (Er * np.conj(Hp)).real)
    # r face at RMIN + LR / 2, x over the whole grid
    x = (t + 1) / 2 * LX
    R = RMIN + LR / 2
    Ex, _, Ep, Hp, Hx = _smooth(x, R, with_phi)
    p_out = np.sum(w * LX / 2 * 2 * np.pi * R * 0.5 * (Ep * np.conj(Hx) - Ex * np.conj(Hp)).real)
    return p_in, p_out


def _flux_discrete(n, with_phi):
    ni, nj = 2 * n, n
    dx, dr = LX / ni, LR / nj
    pos = M.positions(ni, nj, dx, dr, RMIN)
    f = M.empty_fields(ni, nj)
    f["rhoe"][:] = 1e16
    f["nu_m"][:] = 1e9
    X, R = pos["x_face"][:, None], pos["r_cell"][None, :]
    f["E1x"][:] = _smooth(X, R, with_phi)[0]
    X, R = pos["x_cell"][:, None], pos["r_face"][None, :]
    f["E1r"][:] = _smooth(X, R, with_phi)[1]
    X, R = pos["x_cell"][:, None], pos["r_cell"][None, :]
    f["E1p"][:] = _smooth(X, R, with_phi)[2]
    p = M.params(ni, nj, dx, dr, RMIN, OMEGA, flag_mwRef=1, i_ref=ni // 2 + 1, j_in_lo=1, j_in_hi=nj,
                 j_out=nj // 2, i_out_lo=1, i_out_hi=ni)
    _, rep = M.run_oracle(p, f)
    return rep.P_in, rep.P_out / rep.ratio


def _convergence(with_phi):
    ex = _flux_exact(with_phi)
    errs = []
    for n in (16, 32, 64):
        d = _flux_discrete(n, with_phi)
        errs.append([abs(d[q] - ex[q]) / abs(ex[q]) for q in range(2)])
    return np.array(errs)


@pytest.mark.parametrize("with_phi", [False, True])
def test_poynting_fluxes_converge_second_order(with_phi):
    e = _convergence(with_phi)
    cols = [1] if with_phi else [0, 1]    # the x-face flux is pinned with E_phi = 0
    for q in cols:
        assert e[2, q] <= 2e-3, e
        assert e[1, q] / e[2, q] >= 3.0, e


def test_poynting_convergence_catches_conj(plant):
    plant(O.ORC_PLANT_MW_POYNTING_CONJ)
    e = _convergence(False)
    assert e[2].min() > 1e-2, e


# ------------------------------------------------------------------ port algebra
S11, S12, S22 = 0.1 + 0.05j, 0.9 * np.exp(0.4j), -0.05 + 0.1j


def _port_case(flag, Pmw=250.0):
    ni, nj = 24, 16
    p, f = M.port_case(ni, nj, OMEGA, flag, Pmw=Pmw, S11=S11, S12=S12, S21=S12, S22=S22)
    g, rep = M.run_oracle(p, f)
    return p, f, g, M.report_dict(rep)


def _port_errors():
    p, f, g, d = _port_case(1)
    V, I, r = d["V_ref"], d["I_ref"], d["ratio"]
    Gp = d["Gamma_p"]
    e = {}
    e["fwd-ref"] = abs(d["P_fwd"] - d["P_ref"] - r * 0.5 * (V * np.conj(I)).real) / d["P_fwd"]
    e["Zp"] = abs(d["Zp"] - V / I) / abs(V / I)
    e["Gamma_p"] = abs(Gp - (V / I - p.Z0) / (V / I + p.Z0))
    e["Gamma_all"] = abs(d["Gamma_all"] - (S11 + S12 * S12 * Gp / (1 - S22 * Gp)))
    e["P_ref_all"] = abs(d["P_ref_all"] - abs(d["Gamma_all"]) ** 2 * d["P_fwd_all"]) / d["P_fwd_all"]
    e["Pmw"] = abs(d["P_fwd_all"] - p.Pmw) / p.Pmw
    e["balance"] = abs(d["power_sum"] + d["P_out"] - (d["P_fwd"] - d["P_ref"])) / d["P_fwd"]
    e["sum"] = abs(d["power_sum"] - d["ratio_eff"] * (d["Px"] + d["Pr"] + d["Pp"])) / abs(d["power_sum"])
    e["exc"] = abs(d["J1r_exc"] - p.J1r_exc * np.sqrt(r)) / d["J1r_exc"]
    s = np.sqrt(max(d["ratio_eff"], 0.0))
    for k in ("E1x", "E1r", "E1p", "J1x_exc", "J1r_exc", "J1p_exc"):
        e[k] = _rel(g[k], f[k] * s)
    return e


def test_port_identities(plant):
    e = _port_errors()
    assert max(e.values()) <= 1e-12, e


def test_port_identities_catch_swapped_waves(plant):
    plant(O.ORC_PLANT_MW_PORT_WAVES)
    e = _port_errors()
    assert e["fwd-ref"] > 1e-2, e


def test_plain_branch_scales_absorbed_power_only():
    p1, f, g1, d1 = _port_case(1)
    p0, _, g0, d0 = _port_case(0)
    ps = d0["Px"] + d0["Pr"] + d0["Pp"]
    assert d0["Px"] == d1["Px"] and d0["Pr"] == d1["Pr"] and d0["Pp"] == d1["Pp"]
    assert abs(d0["ratio"] - p0.Pmw / ps) <= 1e-15 * d0["ratio"]
    assert abs(d0["J1r_exc"] - p0.J1r_exc * np.sqrt(d0["ratio"])) <= 1e-14 * d0["J1r_exc"]
    # the same unscaled profile under the two scalings
    assert _rel(g0["Pabs"] / d0["ratio"], g1["Pabs"] / d1["ratio_eff"]) <= 1e-13
    for k in ("E1x", "E1r", "E1p", "J1x_exc", "J1r_exc", "J1p_exc"):
        assert np.array_equal(g0[k], f[k]), k
    # J and the component powers are written, not scaled, in the plain branch
    assert _rel(g0["J1x"] * np.sqrt(d1["ratio_eff"]), g1["J1x"]) <= 1e-13
    assert np.array_equal(g0["Pabsx"], g1["Pabsx"])
