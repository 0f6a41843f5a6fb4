"""Independent pins of the oracle's semi-implicit Poisson permittivity and of
its secondary-emission (SEE) wall terms, each shown to catch the mistake it
guards against by planting that mistake in the oracle (orc_plant).

Poisson (fluidModule.hpp:2627-2714 assembly, :2948-3033 correction).  The
solved system is -div(eps_f grad dphi) = e n_i - e n_e - e dt S
+ eps0 lap(phi) + e dt div(Gamma_pred) with eps_f = eps0 + e dt n_f mu on
inner faces, and the correction sets E_new = -grad(phi + dphi) and
Gamma_new = Gamma_pred + mu n_f grad dphi.  Eliminating dphi and Gamma_pred
leaves the discrete Gauss law of the corrected step in every cell,

    eps0 div(E_new) - e dt div(Gamma_new) = e (n_i - n_e) - e dt S,

(r-weighted divergence) whatever the predicted flux was.  It holds only if
the permittivity's dt, mobility and face density are the ones of the flux
correction and if the source enters with -e dt.  For a uniform plasma with
uniform Te the predicted interior flux is the pure drift mu n grad(phi), so
after one pass Gamma_new = -mu_e n E_new on every inner face: that pins the
mobility itself.

SEE (fluidModule.hpp:3250-3266 electron wall flux, :4174-4194 energy).  With
constant yields (eb, rd, ts tables flat) the change of the predicted wall
flux when emission is switched on is, per wall face,
    electrons:  -gamma_i Gamma_i - (eb + rd + ts) Gamma_th(n_e at the step start)
    energy:     -2 e T_emit (gamma_i Gamma_i + ts Gamma_th(n_e at the step start))
                - (eb + r_rd rd) q_th
where Gamma_th and q_th are the thermal particle and energy wall fluxes the
oracle gives with emission off (themselves pinned by quadrature in
test_oracle_walls.py): emitted = yield x primary.
"""
import numpy as np
import pytest

import oracle_lib as O

W = O.workloads
abi = O.abi
E0, EPS0 = W.E0, W.EPS0


@pytest.fixture
def plant():
    yield lambda k: O.lib().orc_plant(k)
    O.lib().orc_plant(0)


def _div(p, Fx, Fr):
    """r-weighted divergence of face fields on every interior cell"""
    ni, nj = p.ni, p.nj
    r = W.radii(nj, p.dr)
    qR = ((r[1:nj + 1] + r[2:nj + 2]) / 2.0 / r[1:nj + 1])[None, :]
    qL = ((r[1:nj + 1] + r[0:nj]) / 2.0 / r[1:nj + 1])[None, :]
    return ((Fx[2:ni + 2, 1:nj + 1] - Fx[1:ni + 1, 1:nj + 1]) / p.dx +
            (qR * Fr[1:ni + 1, 2:nj + 2] - qL * Fr[1:ni + 1, 1:nj + 1]) / p.dr)


def _gauss_state(seed=4, ni=32, nj=24, h=2e-4):
    """n_e != n_i, a strong ionisation source, phi = 0 with the 30 V bias:
    the one Poisson pass carries the whole potential (dphi is not small)."""
    p, st = W.config0(seed=seed, ni=ni, nj=nj, h=h)
    rng = np.random.default_rng(seed)
    shp = st["phi"].shape
    st["rhoi"] = st["rhoi"] * (1.0 + 0.2 * rng.uniform(-1, 1, shp))
    st["rate_ionize"] = rng.uniform(0.5e26, 2e26, shp)
    p.maxITR_HES_phi = 1
    return p, st


def _gauss_residual(p, st):
    """max over cells of |eps0 div E - e dt div Gamma - e(n_i - n_e) + e dt S|
    relative to the largest of the four terms"""
    ref = W.copy_state(st)
    O.ndd_step(p, ref, 1, solver=0)
    ni, nj = p.ni, p.nj
    c = (slice(1, ni + 1), slice(1, nj + 1))
    t1 = EPS0 * _div(p, ref["Ex"], ref["Er"])
    t2 = E0 * p.dt * _div(p, ref["nUex"], ref["nUer"])
    # the electrons of the Poisson routine are those at the start of the
    # electron routine, which the step keeps as rhoe_old; ions are updated first
    t3 = E0 * (ref["rhoi"][c] - ref["rhoe_old"][c])
    t4 = E0 * p.dt * st["rate_ionize"][c]
    g = t1 - t2 - t3 + t4
    scale = max(np.abs(t).max() for t in (t1, t2, t3, t4))
    return float(np.abs(g).max() / scale), ref


def test_poisson_gauss_law(plant):
    p, st = _gauss_state()
    res, ref = _gauss_residual(p, st)
    # the potential must have moved for the check to mean anything
    assert np.abs(ref["phi"][1:-1, 1:-1]).max() > 1.0
    assert res <= 1e-9, res


@pytest.mark.parametrize("mistake", ["POISSON_DT", "POISSON_MU", "POISSON_DENSITY", "POISSON_SOURCE",
                                     "POISSON_FLUX_MU"])
def test_poisson_gauss_law_catches(plant, mistake):
    plant(getattr(O, "ORC_PLANT_" + mistake))
    p, st = _gauss_state()
    res, _ = _gauss_residual(p, st)
    assert res > 1e-6, (mistake, res)


def _uniform_drift():
    p, st = W.config0(ni=24, nj=16, h=2e-4)
    shp = (p.ni + 2, p.nj + 2)
    n = np.full(shp, 1e17)
    st = W.consistent_state(p, n, n.copy(), np.full(shp, 5.0 * W.E0 / W.BOLTZ))
    p.maxITR_HES_phi = 1
    ref = W.copy_state(st)
    O.ndd_step(p, ref, 1, solver=0)
    ni, nj = p.ni, p.nj
    ex = ref["nUex"][2:ni + 1, 1:nj + 1] + p.mu_e * 1e17 * ref["Ex"][2:ni + 1, 1:nj + 1]
    er = ref["nUer"][1:ni + 1, 2:nj + 1] + p.mu_e * 1e17 * ref["Er"][1:ni + 1, 2:nj + 1]
    scale = p.mu_e * 1e17 * max(np.abs(ref["Ex"]).max(), np.abs(ref["Er"]).max())
    return max(np.abs(ex).max(), np.abs(er).max()) / scale, ref


def test_poisson_uniform_drift_flux(plant):
    res, ref = _uniform_drift()
    assert np.abs(ref["Ex"][2:-1, 1:-1]).max() > 1e3  # the bias drives a real field
    assert res <= 1e-12, res


def test_poisson_uniform_drift_catches_mobility(plant):
    plant(O.ORC_PLANT_POISSON_FLUX_MU)
    res, _ = _uniform_drift()
    assert res > 1e-3, res


# ------------------------------------------------------------------ secondary emission
EB, RD, TS, GAMMA_I, T_EMIT, R_RD = 0.03, 0.02, 0.4, 0.08, 2.5, 0.5


def _see_state(see_on):
    p, st = W.config1(ni=16, nj=12, h=2e-4)
    rng = np.random.default_rng(9)
    shp = st["phi"].shape
    # a wall ion flux into both walls, and n_e at the step start != n_e
    st["rhoUix"] = rng.uniform(1e19, 3e19, shp)
    st["rhoUix"][1, :] *= -1.0
    st["rhoe_old"] = st["rhoe"] * rng.uniform(0.7, 1.3, shp)
    # a retarding sheath on both walls (the wall 30 V below the cell), so the
    # thermal flux does not dwarf the ion-induced emission
    st["Ex"][1, :] = -3e5
    st["Ex"][p.ni + 1, :] = 3e5
    p.Te_emitSEE_eV = T_EMIT
    p.ratioEngy_EISEE_rd = R_RD
    if see_on:
        p.coefIISEE_ts = GAMMA_I
        p.n_see = 64
        p.dTe_SEE = 0.5 * W.E0 / W.BOLTZ
        for k in range(64):
            p.see_eb[k], p.see_rd[k], p.see_ts[k] = EB, RD, TS
    return p, st


def _wall_fluxes(which, see_on):
    p, st = _see_state(see_on)
    fx = O.ndd_assemble(p, st, which)[6]
    return p, st, fx


def _see_errors():
    """relative deviation of the emitted particle and energy fluxes from
    yield x primary, over both walls"""
    p, st, fe0 = _wall_fluxes(1, False)
    _, _, fe1 = _wall_fluxes(1, True)
    _, _, fq0 = _wall_fluxes(3, False)
    _, _, fq1 = _wall_fluxes(3, True)
    ni, nj = p.ni, p.nj
    errs = []
    for cell, face in ((1, 1), (ni, ni + 1)):
        j = slice(1, nj + 1)
        n, nold = st["rhoe"][cell, j], st["rhoe_old"][cell, j]
        gam_i = st["rhoUix"][face, j]
        g_th = fe0[face, j] / n * nold          # thermal particle flux of the electrons at the step start
        d_e = fe1[face, j] - fe0[face, j]
        want_e = -GAMMA_I * gam_i - (EB + RD + TS) * g_th
        d_q = fq1[face, j] - fq0[face, j]
        want_q = -2.0 * W.E0 * T_EMIT * (GAMMA_I * gam_i + TS * g_th) - (EB + R_RD * RD) * fq0[face, j]
        errs.append(np.abs(d_e - want_e).max() / np.abs(want_e).max())
        errs.append(np.abs(d_q - want_q).max() / np.abs(want_q).max())
    return max(errs)


def test_see_emitted_is_yield_times_primary(plant):
    assert _see_errors() <= 1e-12


@pytest.mark.parametrize("mistake", ["SEE_DENSITY", "SEE_ION_SIGN", "SEE_RD_RATIO", "SEE_EMIT_ENERGY"])
def test_see_check_catches(plant, mistake):
    plant(getattr(O, "ORC_PLANT_" + mistake))
    assert _see_errors() > 1e-3, mistake
