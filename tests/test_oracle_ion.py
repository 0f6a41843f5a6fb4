"""The oracle's restatement of the reference's ion block (solve_Uix_tmp,
solve_Uir_tmp, solve_Uip_tmp, solve_rhoi_constTe, correct_Ui_constTe;
include/fluidModule.hpp:123-2072; oracle/orc_ion.c), pinned without a GPU
on the five-block cathode.

Pins:
- magnetised drift: with a uniform density, an axial field E, a radial
  field Br, collisions nu and no ionization, the classical Hall drift
  Ux = (e E / (m nu)) / (1 + w^2 / nu^2), Up = (w / nu) Ux, w = e Br / m,
  is a fixed point of one update away from the boundaries (it fixes the
  signs and weights of e E, of both Lorentz terms, of nu and of dt);
- the density's discrete balance in every plasma cell,
  (n - n_old) / dt + div Gamma = S, over the corrected face fluxes, the
  wall fluxes and the open face's flux the update records (the reference's
  own balance formula, correct_Ui_constTe :1868-1871), exact with the direct
  solves; it fails when the correction's diffusion differs from the
  system's;
- the recorded wall fluxes are n times the outward flux of a Maxwellian at
  Ti drifting at the wall velocity, by quadrature;
- the wall velocities conserve energy: U_wall^2 = U_old^2 + max(2 e dphi /
  m, 0) at the adjacent face (dphi minus V_bias at the biased wall), pointing
  out of the plasma.
Each planted mistake (ORC_PLANT_ION_*) fails its check.
"""
import numpy as np
import pytest
from scipy import integrate

import oracle_lib as O

W = O.workloads
abi = O.abi
E0, KB, MI = W.E0, W.BOLTZ, W.MASSI
H = 4e-4


@pytest.fixture
def clean():
    yield
    O.set_geometry(None)
    O.lib().orc_plant(0)


def rect_geometry(ni, nj):
    """the handle's built-in rectangle in segment form"""
    g = abi.NddGeometry()
    g.n_bl = 1
    for q, v in enumerate((1, ni, 1, nj)):
        g.cell_box[q] = v
    for q, v in enumerate((2, ni, 1, nj)):
        g.xface_box[q] = v
    for q, v in enumerate((1, ni, 2, nj)):
        g.rface_box[q] = v
    segs = [(abi.SEG_X_LOW, 1, 1, nj, 0), (abi.SEG_X_HIGH, ni, 1, nj, abi.SEG_BIAS), (abi.SEG_R_OPEN, nj, 1, ni, 0),
            (abi.SEG_R_AXIS, 1, 1, ni, 0)]
    g.n_seg = len(segs)
    for k, v in enumerate(segs):
        g.seg[k] = abi.NddSegment(*v)
    return g


def cathode(h=H):
    ni, nj, g, c6, *_ = W.reference_geometry(h)
    O.set_geometry(g)
    return ni, nj, g, c6, W.default_params(ni, nj, h, h)


# ------------------------------------------------------------------ magnetised drift
E_X, NU = 1e3, 1e6
W_C = NU                        # cyclotron frequency = collision frequency
B_R = W_C * MI / E0


def _hall():
    ni, nj, g, c6, grid = cathode()
    shp = (ni + 2, nj + 2)
    ux = (E0 * E_X / (MI * NU)) / (1.0 + (W_C / NU) ** 2)
    up = W_C / NU * ux
    a = {k: np.zeros(shp) for k in abi.ION_FIELD_NAMES}
    a["rhoi"][...] = 1e16
    a["phi"][...] = 0.0
    a["Ex"][...] = E_X
    a["Br"][...] = B_R
    a["nui_m"][...] = NU
    a["Uix"][...] = ux
    a["Uip"][...] = up
    b = W.copy_arrays(a)
    O.ion_update(grid, W.ion_params(dt_i=1e-7), b)   # nu dt = 0.1: one step moves a wrong force far
    # block 4, seven cells or faces away from every boundary
    i0, i1, j0, j1 = (int(v) for v in c6[4])
    I, J = slice(i0 + 7, i1 - 6), slice(j0 + 7, j1 - 6)
    ex = np.abs(b["Uix"][I, J] - ux).max() / ux
    ep = np.abs(b["Uip"][I, J] - up).max() / up
    return ex, ep


def test_hall_drift_is_a_fixed_point(clean):
    ex, ep = _hall()
    # (the azimuthal drift's centrifugal term pushes a slow radial flow that
    # moves n by ~1e-5 and Ux through the correction by ~1e-6)
    assert ex < 1e-5 and ep < 1e-7, (ex, ep)


def test_hall_drift_catches_a_reversed_lorentz_force(clean):
    O.lib().orc_plant(O.ORC_PLANT_ION_LORENTZ_SIGN)
    ex, ep = _hall()
    assert ep > 0.05, ep


# ------------------------------------------------------------------ balance of the density
def wall_flags(g, ni, nj):
    """the reference's b*_wall flags (1 = no wall on that side; the open face counts as inner)"""
    shp = (ni + 2, nj + 2)
    bLx, bRx, bLr, bRr = (np.ones(shp) for _ in range(4))
    for s, kind, ax, fi, fj, ci, cj in W.segment_faces(g):
        if kind == abi.SEG_X_LOW:
            bLx[ci, cj] = 0
        elif kind == abi.SEG_X_HIGH:
            bRx[ci, cj] = 0
        elif kind == abi.SEG_R_LOW:
            bLr[ci, cj] = 0
        elif kind == abi.SEG_R_HIGH:
            bRr[ci, cj] = 0
    return bLx, bRx, bLr, bRr


def balance_residual(g, ni, nj, h, p, a, b):
    r = W.radii(nj, h)
    qR = ((r[1:nj + 1] + r[2:nj + 2]) / 2.0 / r[1:nj + 1])[None, :]
    qL = ((r[1:nj + 1] + r[0:nj]) / 2.0 / r[1:nj + 1])[None, :]
    I, J = slice(1, ni + 1), slice(1, nj + 1)
    bLx, bRx, bLr, bRr = (f[I, J] for f in wall_flags(g, ni, nj))
    fx, fr, fxw, frw = b["rhoUix"], b["rhoUir"], b["rhoUix_wall"], b["rhoUir_wall"]
    div = ((fx[2:ni + 2, J] * bRx - fx[I, J] * bLx) / h + (qR * fr[I, 2:nj + 2] * bRr - qL * fr[I, J] * bLr) / h +
           (fxw[2:ni + 2, J] - fxw[I, J]) / h + (qR * frw[I, 2:nj + 2] - qL * frw[I, J]) / h)
    rate = (b["rhoi"] - a["rhoi"])[I, J] / p.dt_i
    S = a["rate_ionize"][I, J]
    m = W.plasma_mask(g, ni, nj)[I, J]
    res = (rate + div - S)[m]
    return np.abs(res).max() / max(np.abs(rate[m]).max(), np.abs(S[m]).max())


def _balance():
    ni, nj, g, c6, grid = cathode()
    a = W.ion_state(g, ni, nj, H, noise=False)
    p = W.ion_params()
    b = W.copy_arrays(a)
    rep = O.ion_update(grid, p, b)
    assert rep.neg_rhoi == 0
    return balance_residual(g, ni, nj, H, p, a, b)


def test_density_balance_closes(clean):
    res = _balance()
    assert res < 1e-10, res


def test_density_balance_catches_a_mismatched_correction(clean):
    O.lib().orc_plant(O.ORC_PLANT_ION_CORR_TE)
    assert _balance() > 1e-6


# ------------------------------------------------------------------ walls
def maxwell_flux(U, sign, T=W.TI, m=MI):
    """outward flux per density of a 1-D Maxwellian at T drifting at U (sign
    +1: through a high wall, -1: through a low wall, negative), by quadrature"""
    a = m / (2.0 * KB * T)
    f = lambda v: v * np.sqrt(a / np.pi) * np.exp(-a * (v - U) ** 2)
    lim = abs(U) + 12.0 / np.sqrt(a)
    if sign > 0:
        val, _ = integrate.quad(f, 0.0, lim, epsabs=0.0, epsrel=1e-12, limit=200)
    else:
        val, _ = integrate.quad(f, -lim, 0.0, epsabs=0.0, epsrel=1e-12, limit=200)
    return val


def test_wall_fluxes_and_velocities(clean):
    ni, nj, g, c6, grid = cathode()
    a = W.ion_state(g, ni, nj, H, noise=False)
    p = W.ion_params()
    b = W.copy_arrays(a)
    O.ion_update(grid, p, b)
    phi = a["phi"]
    checked = 0
    for s, kind, ax, fi, fj, ci, cj in W.segment_faces(g):
        if kind not in (abi.SEG_X_LOW, abi.SEG_X_HIGH, abi.SEG_R_LOW, abi.SEG_R_HIGH):
            continue
        sign = -1.0 if kind in (abi.SEG_X_LOW, abi.SEG_R_LOW) else 1.0
        if ax == "x":
            U, F = b["Uix"][fi, fj], b["rhoUix_wall"][fi, fj]
            ai = fi + 1 if sign < 0 else fi - 1            # the adjacent inner face
            u_old = a["Uix"][ai, fj]
            dphi = (phi[ai, fj] + phi[ai - 1, fj]) / 2.0
            if g.seg[s].flag & abi.SEG_BIAS:
                dphi -= p.V_bias
        else:
            U, F = b["Uir"][fi, fj], b["rhoUir_wall"][fi, fj]
            aj = fj + 1 if sign < 0 else fj - 1
            u_old = a["Uir"][fi, aj]
            dphi = (phi[fi, aj] + phi[fi, aj - 1]) / 2.0
        # energy: 1/2 m U^2 = 1/2 m U_old^2 + e max(dphi, 0), out of the plasma
        want = u_old ** 2 + max(2.0 * E0 * dphi / MI, 0.0)
        assert abs(U * U - want) <= 1e-12 * want and np.sign(U) == sign, (s, fi, fj)
        # the kinetic flux of a drifting Maxwellian
        if checked % 7 == 0:
            Fq = b["rhoi"][ci, cj] * maxwell_flux(U, sign)
            assert abs(F - Fq) <= 1e-9 * abs(Fq), (s, F, Fq)
        checked += 1
    assert checked > 100
