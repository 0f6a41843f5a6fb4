"""Inputs of the microwave power profile (EmfieldModule::update_energy_profile,
emfieldModule.hpp:7630-8105) for the oracle pins and the device parity tests.

Grid conventions of the reference: cell (i, j) centred at x = (i - 1/2) dx,
r = rmin + (j - 1/2) dr; x face i at x = (i - 1) dx between cells i-1 and i;
r face j at r = rmin + (j - 1) dr between cells j-1 and j.  E1x lives on x
faces at the cell radius, E1r on r faces at the cell's x, E1p in cells.
"""
import numpy as np

import oracle_lib as O

abi = O.abi
W = O.workloads
E0, MU0, EPS0, MASSE = W.E0, W.MU0, W.EPS0, W.MASSE
ETA = np.sqrt(MU0 / EPS0)          # impedance of free space
C_LIGHT = 1.0 / np.sqrt(MU0 * EPS0)

REAL = ["rhoe", "nu_m", "Bx", "Br", "Pabsx", "Pabsr", "Pabsp", "Pabs"]
CPLX = ["E1x", "E1r", "E1p", "J1x", "J1r", "J1p", "J1x_exc", "J1r_exc", "J1p_exc"]


def positions(ni, nj, dx, dr, rmin):
    i = np.arange(ni + 2, dtype=np.float64)
    j = np.arange(nj + 2, dtype=np.float64)
    return dict(x_cell=(i - 0.5) * dx, x_face=(i - 1.0) * dx,
                r_cell=rmin + (j - 0.5) * dr, r_face=rmin + (j - 1.0) * dr)


def params(ni, nj, dx, dr, rmin, omega, boxes=None, flag_mwRef=0, Pmw=100.0, **kw):
    """boxes: list of (cell, xface, rface) inclusive boxes; default one block
    over the interior (x faces 2..ni, r faces 2..nj)"""
    p = abi.MwParams()
    p.ni, p.nj, p.dx, p.dr, p.rmin, p.omega, p.masse = ni, nj, dx, dr, rmin, omega, MASSE
    if boxes is None:
        boxes = [((1, ni, 1, nj), (2, ni, 1, nj), (1, ni, 2, nj))]
    p.n_bl = len(boxes)
    for b, (c, xf, rf) in enumerate(boxes):
        for q in range(4):
            p.cell_box[4 * b + q], p.xface_box[4 * b + q], p.rface_box[4 * b + q] = c[q], xf[q], rf[q]
    p.flag_mwRef, p.Pmw = flag_mwRef, Pmw
    p.Z0, p.r1, p.r2 = 50.0, rmin, rmin + nj * dr
    for k, v in dict(S11=(0.0, 0.0), S12=(1.0, 0.0), S21=(1.0, 0.0), S22=(0.0, 0.0)).items():
        getattr(p, k)[0], getattr(p, k)[1] = v
    p.J1r_exc = 1.0
    p.i_ref, p.j_out, p.i_out_lo, p.i_out_hi = 2, nj, 1, ni
    p.j_in_lo, p.j_in_hi = 1, nj
    for k, v in kw.items():
        if k in ("S11", "S12", "S21", "S22"):
            getattr(p, k)[0], getattr(p, k)[1] = v.real, v.imag
        else:
            setattr(p, k, v)
    return p


def empty_fields(ni, nj):
    shp = (ni + 2, nj + 2)
    f = {k: np.zeros(shp) for k in REAL}
    f.update({k: np.zeros(shp, dtype=np.complex128) for k in CPLX})
    return f


def copy_fields(f):
    return {k: (None if v is None else v.copy()) for k, v in f.items()}


def uniform_fields(ni, nj, n, nu, Bx, Br, Ex, Er, Ep):
    f = empty_fields(ni, nj)
    for k, v in dict(rhoe=n, nu_m=nu, Bx=Bx, Br=Br).items():
        f[k][:] = v
    for k, v in dict(E1x=Ex, E1r=Er, E1p=Ep).items():
        f[k][:] = v
    return f


def random_fields(seed, ni, nj, B=True):
    """a plasma bump with a collision profile, a magnetic field and random
    complex fields; every output array starts with random values so that the
    untouched (outside the boxes) and rescaled parts are both checked"""
    rng = np.random.default_rng(seed)
    shp = (ni + 2, nj + 2)
    f = empty_fields(ni, nj)
    x = np.linspace(-1, 1, ni + 2)[:, None]
    y = np.linspace(-1, 1, nj + 2)[None, :]
    f["rhoe"][:] = 1e17 * np.exp(-2 * (x * x + y * y)) * (1 + 0.1 * rng.uniform(-1, 1, shp))
    f["nu_m"][:] = 2e9 * (1 + 0.5 * rng.uniform(0, 1, shp))
    if B:
        f["Bx"][:] = 0.08 * (1 + 0.3 * np.cos(3 * x) * np.sin(2 * y)) + 0.01 * rng.uniform(-1, 1, shp)
        f["Br"][:] = 0.05 * np.sin(2 * x) * (1 + y) + 0.01 * rng.uniform(-1, 1, shp)
    for k in CPLX:
        f[k][:] = 1e3 * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
    for k in ("Pabsx", "Pabsr", "Pabsp", "Pabs"):
        f[k][:] = rng.uniform(0, 1e3, shp)
    return f


def run_oracle(p, f):
    """oracle on a copy; returns (fields after, report)"""
    g = copy_fields(f)
    rep = O.mw_energy_profile(p, g)
    return g, rep


REPORT_SCALARS = ["Px", "Pr", "Pp", "power_sum", "P_out", "P_in", "P_fwd", "P_ref", "P_fwd_all", "P_ref_all",
                  "P_err", "ratio", "ratio_eff", "J1r_exc"]
REPORT_COMPLEX = ["V_ref", "I_ref", "Z_ref", "Gamma_p", "Zp", "Gamma_all", "Zall"]


def report_dict(rep):
    d = {k: getattr(rep, k) for k in REPORT_SCALARS}
    d.update({k: complex(getattr(rep, k)[0], getattr(rep, k)[1]) for k in REPORT_COMPLEX})
    return d


# smooth analytic fields and their Faraday H = (i / (w mu0)) curl E
SMOOTH_A = ((1.0 + 0.5j) * 1e3, (0.3 - 0.8j) * 1e3, (-0.6 + 0.2j) * 1e3)
SMOOTH_K, SMOOTH_Q, SMOOTH_R0 = (120.0, 210.0, 90.0), 150.0, 0.02


def smooth(x, r, with_phi, omega):
    """E_x = A1 e^{-i k1 x} cos(q r), E_r = A2 e^{-i k2 x} sin(q r + 0.3),
    E_phi = A3 e^{-i k3 x} (1 + r / R0) (or 0); returns Ex, Er, Ep, H_phi, H_x"""
    (A1, A2, A3), (k1, k2, k3), q, R0 = SMOOTH_A, SMOOTH_K, SMOOTH_Q, SMOOTH_R0
    e1, e2, e3 = np.exp(-1j * k1 * x), np.exp(-1j * k2 * x), np.exp(-1j * k3 * x)
    Ex = A1 * e1 * np.cos(q * r)
    Er = A2 * e2 * np.sin(q * r + 0.3)
    Ep = A3 * e3 * (1 + r / R0) if with_phi else 0 * Ex
    iwm = 1j / (omega * MU0)
    Hp = iwm * (-1j * k2 * Er + q * A1 * e1 * np.sin(q * r))
    Hx = iwm * (A3 * e3 * (1 / r + 2 / R0)) if with_phi else 0 * Ex
    return Ex, Er, Ep, Hp, Hx


def port_case(ni, nj, omega, flag, seed=21, Pmw=250.0, feed=4, **kw):
    """a plasma block (cells j <= nj - feed) and a coaxial feed block (the top
    feed cells, x cells 1..4) carrying a forward TEM wave of 100 V and a 30 %
    reflected one; smooth fields of 300 V/m with 1 % noise elsewhere.  The
    reference plane is x face 3, the loss boundary the r face above cells
    (3..ni-2, nj-feed-1)."""
    dx = dr = 2e-4
    rmin = 1e-3
    pos = positions(ni, nj, dx, dr, rmin)
    f = random_fields(seed, ni, nj)
    for k in ("E1x", "E1r", "E1p"):
        f[k] *= 0.01
    f["E1x"] += 0.3 * smooth(pos["x_face"][:, None], pos["r_cell"][None, :], True, omega)[0]
    f["E1r"] += 0.3 * smooth(pos["x_cell"][:, None], pos["r_face"][None, :], True, omega)[1]
    f["E1p"] += 0.3 * smooth(pos["x_cell"][:, None], pos["r_cell"][None, :], True, omega)[2]
    j0 = nj - feed + 1
    a, b = pos["r_face"][j0], pos["r_face"][nj + 1]
    k0 = omega / C_LIGHT
    xc = pos["x_cell"]
    tem = 100.0 * (np.exp(-1j * k0 * xc) + 0.3 * np.exp(1j * k0 * xc + 0.7j))[:, None] / (
        pos["r_face"][None, :] * np.log(b / a))
    f["E1r"][:, j0:] += tem[:, j0:]
    boxes = [((1, ni, 1, j0 - 1), (2, ni, 1, j0 - 1), (1, ni, 2, j0 - 1)),
             ((1, 4, j0, nj), (2, 4, j0, nj), (1, 4, j0, nj))]
    args = dict(boxes=boxes, flag_mwRef=flag, Pmw=Pmw, i_ref=3, j_in_lo=j0, j_in_hi=nj, j_out=j0 - 2, i_out_lo=3,
                i_out_hi=ni - 2, r1=a, r2=b, Z0=50.0, J1r_exc=2.5)
    args.update(kw)
    return params(ni, nj, dx, dr, rmin, omega, **args), f
