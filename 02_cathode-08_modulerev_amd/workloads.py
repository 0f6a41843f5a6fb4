"""Synthetic NDD / microwave workloads (host numpy arrays, float64).

Each builder returns (params, state) where params is an abi.NddParams and
state a dict of (ni+2, nj+2) arrays named as in include/cathode_hip.h.  The
same arrays feed the HIP library and the CPU oracle.  Random draws use
numpy's PCG64 with the seed given; physical constants are those of the
reference (include/constants.hpp:5-8) and the material constants of its
input/setup.yaml (masse, massi, Ti).
"""
import numpy as np

from . import abi

BOLTZ = 1.380649e-23
E0 = 1.60217662e-19
EPS0 = 8.8541878128e-12
MU0 = 1.2566370614e-6
MASSE = 9.10938356e-31
MASSI = 2.18017156e-25
TI = 300.0


def default_params(ni, nj, dx, dr, dt=1e-10, V_bias=30.0, energy_on=False,
                   mode=abi.CATHODE_MODE_FP64):
    p = abi.NddParams()
    p.ni, p.nj = ni, nj
    p.dx, p.dr, p.dt = dx, dr, dt
    p.rmin = 0.0
    p.V_bias = V_bias
    p.mu_e, p.mu_i = 10.0, 0.1
    p.Ti = TI
    p.masse, p.massi = MASSE, MASSI
    # secondary emission off for the synthetic configs (no SEE in their spec)
    p.coefIISEE_ts = 0.0
    p.coefMISEE_ts = 0.0
    p.Te_emitSEE_eV = 2.5
    p.ratioEngy_EISEE_rd = 0.5
    p.dTe_SEE = 0.0
    p.n_see = 0
    p.energy_on = 1 if energy_on else 0
    # input/setup.yaml solver_setting
    p.maxITR_HES_phi = p.maxITR_HES_rhoe = p.maxITR_HES_rhoeps = p.maxITR_HES_rhoi = 2
    p.error_cnv_HES_phi = p.error_cnv_HES_rhoe = p.error_cnv_HES_rhoeps = 1e-21
    p.error_cnv_HES_rhoi = 1e-21
    p.flag_error = 3
    p.maxITR_SOR_phi = p.maxITR_SOR_rhoe = p.maxITR_SOR_rhoeps = p.maxITR_SOR_rhoi = 30
    p.error_cnv_SOR_phi = p.error_cnv_SOR_rhoe = p.error_cnv_SOR_rhoeps = p.error_cnv_SOR_rhoi = 1e-10
    p.mode = mode
    p.learned = 0
    # input/setup.yaml simulation_control: fixed dt (the report carries the CFL number)
    p.flag_adp_dt = 0
    p.CFL = 1.0
    return p


def radii(nj, dr, rmin=0.0):
    return rmin + (np.arange(nj + 2) - 0.5) * dr


def axial(ni, dx, xL=0.0):
    return xL + (np.arange(ni + 2) - 0.5) * dx


def consistent_state(p, n_e, n_i, Te_K, phi=None):
    """Fill every state array from cell densities/temperature: ghost cells,
    the potential's ghost mirrors (phi = 0 left, V_bias right, grad 0 top)
    and the face fields E = -grad phi on every face, fluxes zero."""
    ni, nj = p.ni, p.nj
    shp = (ni + 2, nj + 2)
    st = {k: np.zeros(shp) for k in abi.STATE_NAMES}

    def ghosts(a):
        a[0, :] = a[1, :]
        a[ni + 1, :] = a[ni, :]
        a[:, 0] = a[:, 1]
        a[:, nj + 1] = a[:, nj]
        return a

    st["rhoe"][...] = ghosts(n_e.copy())
    st["rhoi"][...] = ghosts(n_i.copy())
    st["Te"][...] = ghosts(Te_K.copy())
    st["rhoe_old"][...] = st["rhoe"]
    st["rhoeps"][...] = 1.5 * st["rhoe"] * BOLTZ * st["Te"]
    ph = np.zeros(shp) if phi is None else phi.copy()
    ph[0, 1:nj + 1] = -ph[1, 1:nj + 1]
    ph[ni + 1, 1:nj + 1] = 2.0 * p.V_bias - ph[ni, 1:nj + 1]
    ph[1:ni + 1, nj + 1] = ph[1:ni + 1, nj]
    st["phi"][...] = ph
    Ex = st["Ex"]
    Ex[1:ni + 2, 1:nj + 1] = -(ph[1:ni + 2, 1:nj + 1] - ph[0:ni + 1, 1:nj + 1]) / p.dx
    Er = st["Er"]
    Er[1:ni + 1, 2:nj + 1] = -(ph[1:ni + 1, 2:nj + 1] - ph[1:ni + 1, 1:nj]) / p.dr
    Er[1:ni + 1, nj + 1] = 0.0
    return st


def config0(seed=0, ni=96, nj=64, h=1e-4, V_bias=30.0, mode=abi.CATHODE_MODE_FP64):
    """Plain NDD step: Gaussian ne = ni of 1e17 peak (sigma 1 mm at z = 4.8 mm,
    r = 0) on a 1e14 floor with a shared +-1% uniform noise, Te = 5 eV."""
    p = default_params(ni, nj, h, h, V_bias=V_bias, energy_on=False, mode=mode)
    rng = np.random.Generator(np.random.PCG64(seed))
    z = axial(ni, h)[:, None]
    r = radii(nj, h)[None, :]
    zc = 0.5 * ni * h  # 4.8 mm at 96 cells of 0.1 mm
    n = 1e14 + 1e17 * np.exp(-((z - zc) ** 2 + r ** 2) / (2.0 * (1e-3) ** 2))
    u = rng.uniform(-1.0, 1.0, size=n.shape)
    n = n * (1.0 + 0.01 * u)
    Te = np.full(n.shape, 5.0 * E0 / BOLTZ)
    return p, consistent_state(p, n, n.copy(), Te)


def multi_gaussian(rng, ni, nj, h, count=4):
    z = axial(ni, h)[:, None]
    r = radii(nj, h)[None, :]
    n = np.full((ni + 2, nj + 2), 1e14)
    Lz, Lr = ni * h, nj * h
    for _ in range(count):
        amp = 10.0 ** rng.uniform(16.0, 18.0)
        sig = rng.uniform(0.5e-3, 2.0e-3)
        zc = rng.uniform(0.0, Lz)
        rc = rng.uniform(0.0, Lr)
        n = n + amp * np.exp(-((z - zc) ** 2 + (r - rc) ** 2) / (2.0 * sig * sig))
    return n


def smooth_field(rng, ni, nj, h):
    """Normalised sum of three low-order cosine modes, values in [0, 1]."""
    z = axial(ni, h)[:, None] / (ni * h)
    r = radii(nj, h)[None, :] / (nj * h)
    s = np.zeros((ni + 2, nj + 2))
    for _ in range(3):
        kz, kr = rng.integers(0, 3, size=2)
        ph = rng.uniform(0, 2 * np.pi, size=2)
        s = s + rng.uniform(0.5, 1.0) * np.cos(np.pi * kz * z + ph[0]) * np.cos(np.pi * kr * r + ph[1])
    s = s - s.min()
    return s / s.max()


def config1(seed=1, ni=512, nj=256, h=5e-5, V_bias=30.0, mode=abi.CATHODE_MODE_FP64):
    """Mixed-precision single instance: ne = ni = 1e14 + four Gaussians
    (peaks log-uniform 1e16..1e18, sigma 0.5..2 mm), Te = 2 + 8 s(x) eV with
    s a smooth seeded field, electron energy on."""
    p = default_params(ni, nj, h, h, V_bias=V_bias, energy_on=True, mode=mode)
    rng = np.random.Generator(np.random.PCG64(seed))
    n = multi_gaussian(rng, ni, nj, h)
    Te = (2.0 + 8.0 * smooth_field(rng, ni, nj, h)) * E0 / BOLTZ
    return p, consistent_state(p, n, n.copy(), Te)


def config2(seed=2, ni=1024, nj=512, h=5e-5, mode=abi.CATHODE_MODE_FP64, maxITR=100, tol=1e-10):
    """E_theta microwave solve: omega = 2.67e10 rad/s, nu = 1e9 1/s, PEC on
    every wall (axis at r = 0), n_e = 1e14 + a seeded Gaussian of peak
    5e17 m^-3 (centre z uniform in the middle half, r = 0, sigma uniform in
    [2, 5] mm), excitation: a one-cell ring current at r = 3 mm, mid z,
    carrying J dz dr = 1 A.  Returns (EthetaParams, ne, J) with J complex128."""
    p = abi.EthetaParams()
    p.ni, p.nj = ni, nj
    p.dx = p.dr = h
    p.rmin = 0.0
    p.omega = 2.67e10
    p.nu = 1e9
    p.epsr = 1.0
    p.masse = MASSE
    p.pec_z = p.pec_rmax = 1
    p.maxITR, p.tol, p.mode = maxITR, tol, mode
    rng = np.random.Generator(np.random.PCG64(seed))
    z = axial(ni, h)[:, None]
    r = radii(nj, h)[None, :]
    zc = rng.uniform(0.25, 0.75) * ni * h
    sig = rng.uniform(2e-3, 5e-3)
    ne = 1e14 + 5e17 * np.exp(-((z - zc) ** 2 + r ** 2) / (2.0 * sig * sig))
    J = np.zeros((ni + 2, nj + 2), dtype=np.complex128)
    jr = 1 + int(np.argmin(np.abs(radii(nj, h)[1:nj + 1] - 3e-3)))
    J[1 + ni // 2, jr] = 1.0 / (h * h)
    return p, np.ascontiguousarray(ne), J


SWEEP_SIZES = [(64, 32), (128, 64), (256, 128), (512, 256)]


def config3(seed=3, count=64, h=1e-4, mode=abi.CATHODE_MODE_MIXED, sizes=SWEEP_SIZES):
    """Bias sweep: `count` independent instances, each grid drawn uniformly
    from `sizes` (cells of h = 0.1 mm), bias uniform in [0, 60] V, and the
    config-0 initial state (Gaussian ne = ni of 1e17 peak, sigma 1 mm, at mid-z
    on the axis, 1e14 floor, +-1% noise, Te = 5 eV) with the instance's own
    noise seed.  Returns a list of (params, state, bias)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        ni, nj = sizes[int(rng.integers(0, len(sizes)))]
        bias = float(rng.uniform(0.0, 60.0))
        p, st = config0(seed=seed * 1000 + k, ni=ni, nj=nj, h=h, V_bias=bias, mode=mode)
        out.append((p, st, bias))
    return out


def copy_state(st):
    return {k: v.copy() for k, v in st.items()}


# ------------------------------------------------------------------ five-block cathode
# block corners of the reference's cathode (include/initialModule.hpp:222-232)
# and its domain (include/params.hpp:16-19): x1..x6, xR and r1..r5, rmax [m]
REF_XS = (0.0e-3, 5.6e-3 + 7.8e-3, 10.0e-3 + 7.8e-3, 15.0e-3 + 7.8e-3, 20.0e-3 + 7.8e-3, 0.0e-3 + 7.8e-3, 0.0388)
REF_RS = (0.6e-3, 2.0e-3, 4.0e-3, 9.0e-3, 2.0e-3, 0.014)


def reference_geometry(h, lib=None):
    """The reference's five-block cathode at spacing h in both directions:
    (ni, nj, abi.NddGeometry, six cell boxes, six x-face boxes, six r-face
    boxes) from cathode_ndd_reference_geometry (a host-only call)."""
    import ctypes
    lib = lib or abi.load()
    g = abi.NddGeometry()
    xs = np.ascontiguousarray(REF_XS, dtype=np.float64)
    rs = np.ascontiguousarray(REF_RS, dtype=np.float64)
    c6, x6, r6 = (np.zeros(24, dtype=np.int32) for _ in range(3))
    rc = lib.cathode_ndd_reference_geometry(abi.dptr(xs), abi.dptr(rs), h, h, ctypes.byref(g),
                                            c6.ctypes.data_as(abi.c_int_p), x6.ctypes.data_as(abi.c_int_p),
                                            r6.ctypes.data_as(abi.c_int_p))
    abi.check(rc, lib)
    ni = int(round(REF_XS[6] / h))
    nj = int(round(REF_RS[5] / h))
    return ni, nj, g, c6.reshape(6, 4), x6.reshape(6, 4), r6.reshape(6, 4)


def plasma_mask(g, ni, nj):
    """plasma cells of a geometry on the (ni+2, nj+2) grid"""
    m = np.zeros((ni + 2, nj + 2), dtype=bool)
    for b in range(g.n_bl):
        i0, i1, j0, j1 = (g.cell_box[4 * b + q] for q in range(4))
        m[i0:i1 + 1, j0:j1 + 1] = True
    return m


def cathode5(seed=5, h=4e-4, energy_on=True, mode=abi.CATHODE_MODE_FP64, dt=1e-10):
    """The reference's five-block cathode geometry at spacing h (the
    reference's own grid is h = 0.2 mm, 194 x 70 cells), electrode at 30 V:
    n_e = n_i = 1e15 + two seeded Gaussians (peaks log-uniform in 1e16..1e17,
    sigma 1..2 mm) in the plasma blocks and 0 outside, Te = 3..6 eV smooth,
    phi = 0, fields and fluxes 0.  Returns (params, state, geometry)."""
    ni, nj, g, _, _, _ = reference_geometry(h)
    if abs(ni * h - REF_XS[6]) > 1e-9 or abs(nj * h - REF_RS[5]) > 1e-9:
        raise ValueError("h must divide the domain 38.8 mm x 14 mm")
    p = default_params(ni, nj, h, h, dt=dt, V_bias=30.0, energy_on=energy_on, mode=mode)
    rng = np.random.Generator(np.random.PCG64(seed))
    mask = plasma_mask(g, ni, nj)
    Z = np.broadcast_to(axial(ni, h)[:, None], (ni + 2, nj + 2))
    R = np.broadcast_to(radii(nj, h)[None, :], (ni + 2, nj + 2))
    n = np.full((ni + 2, nj + 2), 1e15)
    pts = np.argwhere(mask)
    for _ in range(2):
        ci, cj = pts[rng.integers(len(pts))]
        peak = 10.0 ** rng.uniform(16.0, 17.0)
        sig = rng.uniform(1e-3, 2e-3)
        n += peak * np.exp(-((Z - Z[ci, cj]) ** 2 + (R - R[ci, cj]) ** 2) / (2 * sig * sig))
    n = np.where(mask, n, 0.0)
    Te = (3.0 + 3.0 * rng.uniform(0.0, 1.0)) * E0 / BOLTZ * np.ones((ni + 2, nj + 2))
    st = {k: np.zeros((ni + 2, nj + 2)) for k in abi.STATE_NAMES}
    st["rhoe"][...] = n
    st["rhoi"][...] = n
    st["rhoe_old"][...] = n
    st["Te"][...] = Te
    st["rhoeps"][...] = 1.5 * n * BOLTZ * Te
    return p, st, g


# ------------------------------------------------------------------ neutrals and metastables
# neutral-gas constants of the reference (include/params.hpp:35-39) and its
# input/setup.yaml (width_neutIn, Q_neutIn_mgs)
TN = 300.0
RI = 216e-12
DMN = 1.66e20
WIDTH_NEUT_IN = 0.0055
Q_NEUT_IN_MGS = 0.069


def inlet_flux(q_mgs=Q_NEUT_IN_MGS, width=WIDTH_NEUT_IN, r4=REF_RS[3], massi=MASSI):
    """fn_In = (Q [mg/s] 1e-6 / m_i) / (2 pi r4 width_neutIn) (initialModule.hpp:235-237)"""
    return q_mgs * 1e-6 / massi / (2.0 * np.pi * r4 * width)


def neutral_params(which=3, dt_n=1e-7, dt_m=1e-7, mode=abi.CATHODE_MODE_FP64, tol=1e-13, maxITR=30,
                   flag_error=3, guard_blocks=2):
    p = abi.NeutralParams()
    p.which, p.dt_n, p.dt_m = which, dt_n, dt_m
    p.Tn, p.ri, p.massi, p.DmN = TN, RI, MASSI, DMN
    p.fn_In = inlet_flux()
    p.x_in = WIDTH_NEUT_IN + REF_XS[5]
    p.xL = 0.0
    p.guard_blocks = guard_blocks
    p.maxITR, p.tol, p.mode, p.flag_error = maxITR, tol, mode, flag_error
    return p


def segment_faces(g):
    """(segment index, kind, 'x' or 'r', face i, face j, cell i, cell j) of every boundary face"""
    out = []
    for s in range(g.n_seg):
        q = g.seg[s]
        for t in range(q.lo, q.hi + 1):
            if q.kind == abi.SEG_X_LOW:
                out.append((s, q.kind, "x", q.cell, t, q.cell, t))
            elif q.kind == abi.SEG_X_HIGH:
                out.append((s, q.kind, "x", q.cell + 1, t, q.cell, t))
            elif q.kind in (abi.SEG_R_LOW, abi.SEG_R_AXIS):
                out.append((s, q.kind, "r", t, q.cell, t, q.cell))
            else:
                out.append((s, q.kind, "r", t, q.cell + 1, t, q.cell))
    return out


def neutral_state(g, ni, nj, h, seed=7, noise=True):
    """A neutral / metastable state on geometry g: n ~ 1e20 and m ~ 1e17
    (smooth, positive in every cell), n_e = 1e15 + a Gaussian of 1e17,
    rate coefficients 1e5..4e5 1/s, outward ion wall fluxes of 1e20..3e20
    and metastable wall fluxes 0.25 m vth on the wall faces, old fluxes of
    either sign below the open face.  noise: the flux arrays also carry
    random values on the faces the updates must not read or write."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shp = (ni + 2, nj + 2)
    a = {}
    a["rhon"] = 1e20 * (0.7 + 0.6 * smooth_field(rng, ni, nj, h))
    a["rhom"] = 1e17 * (0.5 + smooth_field(rng, ni, nj, h))
    Z = axial(ni, h)[:, None]
    R = radii(nj, h)[None, :]
    a["rhoe"] = 1e15 + 1e17 * np.exp(-((Z - 0.6 * ni * h) ** 2 + R ** 2) / (2 * (3e-3) ** 2))
    for k, scale in (("nu_excStep", 3e5), ("nu_super", 1e5), ("nu_excMeta", 1e5), ("nu_exc", 4e5),
                     ("nu_ionz", 2e5), ("nu_ionzStep", 2e5)):
        a[k] = scale * (0.5 + smooth_field(rng, ni, nj, h))
    names = ["rhoUix_wall", "rhoUir_wall", "rhoUnx", "rhoUnr", "rhoUnx_wall", "rhoUnr_wall", "rhoUmx", "rhoUmr",
             "rhoUmx_wall", "rhoUmr_wall"]
    for k in names:
        a[k] = rng.uniform(-1e19, 1e19, shp) if noise else np.zeros(shp)
    vth = np.sqrt(8.0 * BOLTZ * TN / (np.pi * MASSI))
    for s, kind, ax, fi, fj, ci, cj in segment_faces(g):
        if kind == abi.SEG_R_AXIS:
            continue
        if kind == abi.SEG_R_OPEN:
            # the old flux of the face below the open face, either sign
            a["rhoUnr"][ci, cj] = rng.uniform(-0.5e21, 2e21)
            a["rhoUmr"][ci, cj] = rng.uniform(-0.5e19, 2e19)
            continue
        sgn = -1.0 if kind in (abi.SEG_X_LOW, abi.SEG_R_LOW) else 1.0
        gi = "rhoUix_wall" if ax == "x" else "rhoUir_wall"
        gm = "rhoUmx_wall" if ax == "x" else "rhoUmr_wall"
        a[gi][fi, fj] = sgn * rng.uniform(1e20, 3e20)
        a[gm][fi, fj] = sgn * 0.25 * vth * a["rhom"][ci, cj]
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in a.items()}


def copy_arrays(a):
    return {k: (v.copy() if v is not None else None) for k, v in a.items()}


# ------------------------------------------------------------------ ions
TE_REP_EV = 0.1  # input/setup.yaml Te_rep_eV


def ion_params(dt_i=1e-10, mode=abi.CATHODE_MODE_FP64, V_bias=30.0, tol_Ui=1e-13, tol_rhoi=1e-13, maxITR=30):
    p = abi.IonParams()
    p.dt_i, p.Ti, p.massi, p.V_bias = dt_i, TI, MASSI, V_bias
    p.Te_rep_eV, p.rmax = TE_REP_EV, REF_RS[5]
    p.maxITR_Ui, p.tol_Ui, p.maxITR_rhoi, p.tol_rhoi = maxITR, tol_Ui, maxITR, tol_rhoi
    p.mode = mode
    return p


def ion_state(g, ni, nj, h, seed=17, V_bias=30.0, noise=True):
    """An ion state on geometry g: n_i = 1e15 + two Gaussians of 1e16..1e17
    in the plasma (0 outside), a potential rising smoothly from 0 at the
    left to V_bias at the right with E = -grad phi on the faces, velocities
    of a few hundred m/s, S = 1e22..1e23 1/(m^3 s), nui_m = 1e5..1e6 1/s,
    B = 10..50 mT.  noise: the flux arrays carry random values on the faces
    the update must not read or write."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shp = (ni + 2, nj + 2)
    m = plasma_mask(g, ni, nj)
    Z = np.broadcast_to(axial(ni, h)[:, None], shp)
    R = np.broadcast_to(radii(nj, h)[None, :], shp)
    n = np.full(shp, 1e15)
    pts = np.argwhere(m)
    for _ in range(2):
        ci, cj = pts[rng.integers(len(pts))]
        n = n + 10.0 ** rng.uniform(16.0, 17.0) * np.exp(-((Z - Z[ci, cj]) ** 2 + (R - R[ci, cj]) ** 2) / (2 * 1.5e-3 ** 2))
    a = {"rhoi": np.where(m, n, 0.0)}
    L = ni * h
    phi = V_bias * (Z / L) ** 2 + 2.0 * np.sin(np.pi * R / (nj * h)) * smooth_field(rng, ni, nj, h)
    a["phi"] = phi
    Ex = np.zeros(shp)
    Ex[1:, :] = -(phi[1:, :] - phi[:-1, :]) / h
    Er = np.zeros(shp)
    Er[:, 1:] = -(phi[:, 1:] - phi[:, :-1]) / h
    a["Ex"], a["Er"] = Ex, Er
    for k in ("Uix", "Uir", "Uip"):
        a[k] = rng.uniform(-300.0, 300.0, shp)
    a["rate_ionize"] = 1e22 * (1.0 + 9.0 * smooth_field(rng, ni, nj, h))
    a["nui_m"] = 1e5 * (1.0 + 9.0 * smooth_field(rng, ni, nj, h))
    a["Bx"] = 0.01 + 0.04 * smooth_field(rng, ni, nj, h)
    a["Br"] = 0.01 + 0.04 * smooth_field(rng, ni, nj, h)
    for k in ("rhoUix", "rhoUir", "rhoUix_wall", "rhoUir_wall"):
        a[k] = rng.uniform(-1e19, 1e19, shp) if noise else np.zeros(shp)
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in a.items()}


# ------------------------------------------------------------------ transport coefficients
def transport_params(flag_Bohm=0, flag_Sagdeev=1, nu_eff=1e7):
    p = abi.TransportParams()
    p.masse, p.massi, p.Ti, p.nu_eff = MASSE, MASSI, TI, nu_eff
    p.flag_Bohm, p.flag_Sagdeev = flag_Bohm, flag_Sagdeev
    return p


def transport_state(g, ni, nj, h, seed=41, Te_max_eV=9.0):
    """Inputs of the transport update on geometry g: n_e, n_i (Gaussians over
    1e15), n_n ~ 1e20, n_m ~ 1e17, Te = 0.5 .. Te_max_eV eV (a few cells
    beyond the fixture table's end), B of 10..80 mT, nu_ano, ion velocities
    and electron fluxes of either sign; nu_ano_IAT and Mache carry random
    values (kept where the update does not write them)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shp = (ni + 2, nj + 2)
    sm = lambda: smooth_field(rng, ni, nj, h)
    a = {}
    a["rhoe"] = 1e15 + 1e17 * sm() ** 2
    a["rhoi"] = a["rhoe"] * (1.0 + 0.1 * sm())
    a["rhon"] = 1e20 * (0.5 + sm())
    a["rhom"] = 1e17 * (0.5 + sm())
    a["Te"] = (0.5 + (Te_max_eV - 0.5) * sm()) * E0 / BOLTZ
    a["Te"][rng.integers(1, ni + 1, 5), rng.integers(1, nj + 1, 5)] = 30.0 * E0 / BOLTZ
    a["Bx"] = 0.01 + 0.07 * sm()
    a["Br"] = -0.04 + 0.08 * sm()
    a["nu_ano"] = 1e6 * sm()
    for k, sc in (("Uix", 2e3), ("Uir", 2e3), ("Uip", 5e2), ("rhoUex", 1e22), ("rhoUer", 1e22)):
        a[k] = rng.uniform(-sc, sc, shp)
    a["Uix"][rng.integers(1, ni + 1, 5), rng.integers(1, nj + 1, 5)] = 2e5   # beyond the ion table
    a["nu_ano_IAT"] = rng.uniform(0.0, 1e5, shp)
    a["Mache"] = rng.uniform(0.0, 1.0, shp)
    for k in abi.TRANSPORT_OUT:
        if k not in a:
            a[k] = np.zeros(shp)
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in a.items()}


# ---- coupled microwave solve (cathode_mw_solve) ----

def _mw_params(ni, nj, dx, dr, rmin, fmw):
    p = abi.MwParams()
    p.ni, p.nj, p.dx, p.dr, p.rmin = ni, nj, dx, dr, rmin
    p.omega = 2.0 * np.pi * fmw
    p.masse = MASSE
    return p


def _set_box(arr, b, box):
    for k in range(4):
        arr[4 * b + k] = box[k]


def _mask(ni, nj, box):
    m = np.zeros((ni + 2, nj + 2), dtype=np.int32)
    m[box[0]:box[1] + 1, box[2]:box[3] + 1] = 1
    return m


def _sparam(mag, deg):
    return mag * np.cos(deg / 180.0 * np.pi), mag * np.sin(deg / 180.0 * np.pi)


def coax_impedance_test():
    """The reference's coaxial-feed impedance test (test/校正/setup.yaml with
    flag_impTest = 1; InitialModule::makeBoundary_impedanceTest,
    include/initialModule.hpp:1061-1324): 194 x 7 cells of 0.2 mm between
    r = 0.6 and 2 mm, dielectric eps_r 2.0838, 4.25 GHz, 8 W, TEM excitation
    of 0.212687 at the west end, open east end, Gauss conditions on both
    conductors.  The stub and cable corrections of that setup are zero
    (dh_stab = dl_cable = 0), so the port S-parameters are the measured ones.
    Returns (MwParams, lists, MwSolveOpts, arrays)."""
    ni, nj, xL, xR, rmin, rmax = 194, 7, 0.0, 0.0388, 0.0006, 0.002
    dx, dr = (xR - xL) / ni, (rmax - rmin) / nj
    p = _mw_params(ni, nj, dx, dr, rmin, 4.25e9)
    p.n_bl = 1
    cell, xf, rf = (1, ni, 1, nj), (2, ni, 1, nj), (1, ni, 2, nj)
    _set_box(p.cell_box, 0, cell)
    _set_box(p.xface_box, 0, xf)
    _set_box(p.rface_box, 0, rf)
    p.flag_mwRef, p.Pmw, p.Z0 = 1, 8.0, 50.0
    p.r1, p.r2 = rmin, rmax
    p.S11[:] = _sparam(0.24403462, -552.4800212)
    p.S21[:] = _sparam(0.797613913, -619.4434408)
    p.S12[:] = _sparam(0.813217659, -619.9129232)
    p.S22[:] = _sparam(0.126026294, -850.3272909)
    p.J1r_exc = 0.212687
    p.i_ref = int(0.0 / dx + 1.5)
    lists = {
        "x": [[(xf[0], j, abi.MWBC_DIRICHLET) for j in range(xf[2], xf[3] + 1)],
              [(xf[1], j, abi.MWBC_GAUSS) for j in range(xf[2], xf[3] + 1)],
              [(i, xf[2], abi.MWBC_DIRICHLET) for i in range(xf[0], xf[1] + 1)],
              [(i, xf[3], abi.MWBC_DIRICHLET) for i in range(xf[0], xf[1] + 1)]],
        "r": [[(rf[0], j, abi.MWBC_EXCITE) for j in range(rf[2], rf[3] + 1)],
              [(rf[1], j, abi.MWBC_OPEN) for j in range(rf[2], rf[3] + 1)],
              [(i, rf[2], abi.MWBC_GAUSS) for i in range(rf[0], rf[1] + 1)],
              [(i, rf[3], abi.MWBC_GAUSS) for i in range(rf[0], rf[1] + 1)]],
        "p": [[], [], [], []],
    }
    o = abi.MwSolveOpts()
    o.which, o.j_flc0, o.nu_eff, o.max_refine, o.tol = abi.MW_IMPEDANCE, cell[2], 2.0e9, 3, 1e-13
    shp = (ni + 2, nj + 2)
    a = {"rhoe": np.full(shp, 1e17), "nu_m": np.full(shp, 5e8), "Bx": np.zeros(shp), "Br": np.zeros(shp),
         "epsr": np.ones(shp)}
    a["epsr"][cell[0]:cell[1] + 1, cell[2]:cell[3] + 1] = 2.083825280349
    a["act_p"], a["act_x"], a["act_r"] = _mask(ni, nj, cell), _mask(ni, nj, xf), _mask(ni, nj, rf)
    for k in ("J1x_exc", "J1r_exc", "J1p_exc", "E1x", "E1r", "E1p"):
        a[k] = np.zeros(shp, dtype=np.complex128)
    return p, lists, o, a


def mw_coupled_rect(ni=24, nj=12, h=2e-4, seed=11, n0=3e17, B0=0.12, fmw=4.25e9, rmin=6e-4):
    """A coupled, magnetised microwave case on one rectangular block that
    runs every fold and ghost rule of solve_Microwave: x faces 2..ni x 1..nj,
    r faces 1..ni x 2..nj, cells 1..ni x 1..nj; the entries of each list cycle
    through that list's types (plus one type no rule names); a Gaussian
    plasma (peak n0) in a field of about B0 (both components), collision
    frequency 2e9, eps_r 1 .. 2.1, excitation currents on all three
    components and a TEM excitation on the west E_r faces.  rmin > 0 (an
    inner conductor, as in the coaxial feed): the south Gauss and corner rules
    divide by the radius of the face below the first cell.  Returns
    (MwParams, lists, MwSolveOpts, arrays)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = _mw_params(ni, nj, h, h, rmin, fmw)
    p.n_bl = 1
    cell, xf, rf = (1, ni, 1, nj), (2, ni, 1, nj), (1, ni, 2, nj)
    _set_box(p.cell_box, 0, cell)
    _set_box(p.xface_box, 0, xf)
    _set_box(p.rface_box, 0, rf)
    p.flag_mwRef, p.Pmw, p.Z0 = 1, 8.0, 50.0
    p.r1, p.r2 = max(rmin, 0.5 * h), rmin + nj * h
    p.J1r_exc = 0.2
    p.i_ref = 1

    def cyc(entries, types):
        return [(i, j, types[k % len(types)]) for k, (i, j) in enumerate(entries)]

    xW = [(xf[0], j) for j in range(xf[2], xf[3] + 1)]
    xE = [(xf[1], j) for j in range(xf[2], xf[3] + 1)]
    xS = [(i, xf[2]) for i in range(xf[0], xf[1] + 1)]
    xN = [(i, xf[3]) for i in range(xf[0], xf[1] + 1)]
    rW = [(rf[0], j) for j in range(rf[2], rf[3] + 1)]
    rE = [(rf[1], j) for j in range(rf[2], rf[3] + 1)]
    rS = [(i, rf[2]) for i in range(rf[0], rf[1] + 1)]
    rN = [(i, rf[3]) for i in range(rf[0], rf[1] + 1)]
    pW = [(cell[0], j) for j in range(cell[2], cell[3] + 1)]
    pE = [(cell[1], j) for j in range(cell[2], cell[3] + 1)]
    pS = [(i, cell[2]) for i in range(cell[0], cell[1] + 1)]
    pN = [(i, cell[3]) for i in range(cell[0], cell[1] + 1)]
    lists = {
        "x": [cyc(xW, (2, 3, 4, 0, 2, 9)), cyc(xE, (2, 4, 0, 2)), cyc(xS, (0, 6, 0, 0)), cyc(xN, (0, 1, 5, 6, 1))],
        "r": [cyc(rW, (7, 7, 0, 1, 5, 6, 7)), cyc(rE, (1, 0, 5, 1)), cyc(rS, (2, 3, 2, 2)),
              cyc(rN, (2, 3, 4, 2))],
        "p": [cyc(pW, (0, 1, 1)), cyc(pE, (0,)), cyc(pS, (0,)), cyc(pN, (1, 0, 1))],
    }
    o = abi.MwSolveOpts()
    o.which, o.j_flc0, o.nu_eff, o.max_refine, o.tol = abi.MW_COUPLED, 0, 2.0e9, 4, 1e-13
    shp = (ni + 2, nj + 2)
    x = (np.arange(ni + 2) - 0.5) * h
    r = rmin + (np.arange(nj + 2) - 0.5) * h
    X, R = np.meshgrid(x, r, indexing="ij")
    xc, rc, s = x[ni // 2], r[nj // 3], 0.25 * ni * h
    a = {}
    a["rhoe"] = n0 * np.exp(-((X - xc) ** 2 + (R - rc) ** 2) / (2 * s * s)) + 1e14
    a["nu_m"] = 2e9 * (1.0 + 0.2 * rng.uniform(-1, 1, shp))
    a["Bx"] = B0 * (0.8 + 0.4 * np.cos(X / (ni * h) * np.pi))
    a["Br"] = 0.5 * B0 * np.sin(2.0 * X / (ni * h) * np.pi) * R / (nj * h)
    a["epsr"] = 1.0 + 1.1 * rng.uniform(0, 1, shp)
    a["Er_dc"] = rng.uniform(-1e3, 1e3, shp)
    a["act_p"], a["act_x"], a["act_r"] = _mask(ni, nj, cell), _mask(ni, nj, xf), _mask(ni, nj, rf)

    def cx(scale):
        return scale * (rng.uniform(-1, 1, shp) + 1j * rng.uniform(-1, 1, shp))

    a["J1x_exc"], a["J1r_exc"], a["J1p_exc"] = cx(10.0), cx(10.0), cx(10.0)
    a["E1x"], a["E1r"], a["E1p"] = cx(1.0), cx(1.0), cx(1.0)   # ghost rules read values they do not write
    return p, lists, o, {k: np.ascontiguousarray(v) for k, v in a.items()}


# the five-block cathode with its coaxial feed (InitialModule::makeBoundary,
# include/initialModule.hpp:207-1060): block corners [m]
MW_XS = {"x1": 0.0e-3, "x2": 5.6e-3 + 7.8e-3, "x3": 10.0e-3 + 7.8e-3, "x4": 15.0e-3 + 7.8e-3,
         "x5": 20.0e-3 + 7.8e-3, "x6": 0.0e-3 + 7.8e-3}
MW_RS = {"r1": 0.6e-3, "r2": 2.0e-3, "r3": 4.0e-3, "r4": 9.0e-3, "r5": 2.0e-3, "r6": 14.0e-3}
EPSR_DIELE = 2.083825280349   # input/setup.yaml epsr_diele
J1R_EXC = 0.212687            # input/setup.yaml J1r_exc


def cathode_mw_boxes(ni, nj, dx, dr, xR=0.0388, rmax=0.014):
    """The six blocks' inclusive boxes {ilo, ihi, jlo, jhi} of cells, x faces
    and r faces (:239-316; block 5 is the coaxial feed)."""
    X, R = MW_XS, MW_RS
    lo = lambda v, d: int(v / d + 0.5)
    hi = lambda v, d: int(v / d - 0.5)
    cell = [(lo(X["x6"], dx) + 1, hi(X["x2"], dx) + 1, lo(R["r3"], dr) + 1, hi(R["r4"], dr) + 1),
            (lo(X["x2"], dx) + 1, hi(X["x4"], dx) + 1, lo(R["r5"], dr) + 1, hi(R["r4"], dr) + 1),
            (lo(X["x2"], dx) + 1, hi(X["x3"], dx) + 1, lo(R["r1"], dr) + 1, hi(R["r5"], dr) + 1),
            (lo(X["x3"], dx) + 1, hi(X["x5"], dx) + 1, lo(0.0, dr) + 1, hi(R["r5"], dr) + 1),
            (lo(X["x5"], dx) + 1, hi(xR, dx) + 1, lo(0.0, dr) + 1, hi(rmax, dr) + 1),
            (lo(X["x1"], dx) + 1, hi(X["x2"], dx) + 1, lo(R["r1"], dr) + 1, hi(R["r2"], dr) + 1)]
    xf = [(lo(X["x6"], dx) + 2, hi(X["x2"], dx) + 2, lo(R["r3"], dr) + 1, hi(R["r4"], dr) + 1),
          (lo(X["x2"], dx) + 2, hi(X["x4"], dx) + 1, lo(R["r5"], dr) + 1, hi(R["r4"], dr) + 1),
          (lo(X["x2"], dx) + 2, hi(X["x3"], dx) + 2, lo(R["r1"], dr) + 1, hi(R["r5"], dr) + 1),
          (lo(X["x3"], dx) + 2, hi(X["x5"], dx) + 2, lo(0.0, dr) + 1, hi(R["r5"], dr) + 1),
          (lo(X["x5"], dx) + 2, hi(xR, dx) + 1, lo(0.0, dr) + 1, hi(rmax, dr) + 1),
          (lo(X["x1"], dx) + 2, hi(X["x2"], dx) + 2, lo(R["r1"], dr) + 1, hi(R["r2"], dr) + 1)]
    rf = [(lo(X["x6"], dx) + 1, hi(X["x2"], dx) + 1, lo(R["r3"], dr) + 2, hi(R["r4"], dr) + 1),
          (lo(X["x2"], dx) + 1, hi(X["x4"], dx) + 1, lo(R["r5"], dr) + 1, hi(R["r4"], dr) + 1),
          (lo(X["x2"], dx) + 1, hi(X["x3"], dx) + 1, lo(R["r1"], dr) + 2, hi(R["r5"], dr) + 1),
          (lo(X["x3"], dx) + 1, hi(X["x5"], dx) + 1, lo(0.0, dr) + 2, hi(R["r5"], dr) + 1),
          (lo(X["x5"], dx) + 1, hi(xR, dx) + 1, lo(0.0, dr) + 2, hi(rmax, dr) + 1),
          (lo(X["x1"], dx) + 1, hi(X["x2"], dx) + 1, lo(R["r1"], dr) + 2, hi(R["r2"], dr) + 1)]
    return cell, xf, rf


def cathode_mw_lists(c):
    """makeBoundary's twelve MicrowaveBC lists (:420-1058) from the cell boxes
    c[b] = (ilo, ihi, jlo, jhi): {"x": [W, E, S, N], "r": [...], "p": [...]},
    entries (i, j, type), in the reference's push order."""
    I0, I1, J0, J1 = 0, 1, 2, 3
    ci = lambda b, k: c[b][k]
    G, D, O_, C3, C4, V5, V6, EX = 2, 0, 1, 3, 4, 5, 6, 7
    xW = [(ci(0, I0) + 1, j, G) for j in range(ci(0, J0) + 1, ci(0, J1))]
    xW += [(ci(0, I0) + 1, ci(0, J0), C3), (ci(0, I0) + 1, ci(0, J1), C4)]
    xW += [(ci(2, I0) + 1, j, G) for j in range(ci(5, J1) + 1, ci(0, J0))]
    xW += [(ci(3, I0) + 1, j, G) for j in range(ci(3, J0), ci(2, J0))]
    xW += [(ci(4, I0) + 1, j, G) for j in range(ci(1, J0), ci(4, J1))]
    xW += [(ci(4, I0) + 1, ci(4, J1), C4)]
    xW += [(ci(5, I0) + 1, j, D) for j in range(ci(5, J0), ci(5, J1) + 1)]
    xE = [(ci(1, I1), j, G) for j in range(ci(1, J0), ci(1, J1))]
    xE += [(ci(1, I1), ci(1, J1), C4)]
    xE += [(ci(4, I1), j, G) for j in range(ci(4, J0), ci(4, J1))]
    xE += [(ci(4, I1), ci(4, J1), C4)]
    xS = [(i, ci(0, J0), D) for i in range(ci(0, I0) + 1, ci(0, I1) + 1)]
    xS += [(ci(0, I1) + 1, ci(0, J0), V6)]
    xS += [(i, ci(5, J0), D) for i in range(ci(5, I0) + 1, ci(2, I1) + 1)]
    xS += [(ci(2, I1) + 1, ci(5, J0), V6)]
    xN = [(i, ci(0, J1), D) for i in range(ci(0, I0) + 1, ci(1, I1) + 1)]
    xN += [(i, ci(5, J1), D) for i in range(ci(5, I0) + 1, ci(5, I1) + 1)]
    xN += [(ci(5, I1) + 1, ci(5, J1), V6)]
    xN += [(i, ci(3, J1), D) for i in range(ci(1, I1) + 2, ci(3, I1) + 1)]
    xN += [(ci(1, I1) + 1, ci(3, J1), V5), (ci(3, I1) + 1, ci(3, J1), V6)]
    xN += [(i, ci(4, J1), O_) for i in range(ci(4, I0) + 1, ci(4, I1) + 1)]
    rW = [(ci(0, I0), j, D) for j in range(ci(0, J0) + 1, ci(0, J1) + 1)]
    rW += [(ci(1, I0), j, D) for j in range(ci(5, J1) + 2, ci(0, J0))]
    rW += [(ci(1, I0), ci(5, J1) + 1, V5), (ci(1, I0), ci(0, J0), V6)]
    rW += [(ci(3, I0), j, D) for j in range(ci(3, J0) + 1, ci(2, J0))]
    rW += [(ci(3, I0), ci(2, J0), V6)]
    rW += [(ci(4, I0), j, D) for j in range(ci(1, J0) + 1, ci(4, J1) + 1)]
    rW += [(ci(4, I0), ci(1, J0), V5)]
    rW += [(ci(5, I0), j, EX) for j in range(ci(5, J0) + 1, ci(5, J1) + 1)]
    rE = [(ci(1, I1), j, D) for j in range(ci(1, J0) + 1, ci(1, J1) + 1)]
    rE += [(ci(1, I1), ci(1, J0), V5)]
    rE += [(ci(4, I1), j, D) for j in range(ci(4, J0) + 1, ci(4, J1) + 1)]
    rS = [(i, ci(0, J0) + 1, G) for i in range(ci(0, I0) + 1, ci(0, I1) + 1)]
    rS += [(ci(0, I0), ci(0, J0) + 1, C3)]
    rS += [(i, ci(5, J0) + 1, G) for i in range(ci(5, I0), ci(2, I1) + 1)]
    rN = [(i, ci(0, J1), G) for i in range(ci(0, I0) + 1, ci(1, I1))]
    rN += [(ci(0, I0), ci(0, J1), C3), (ci(1, I1), ci(0, J1), C4)]
    rN += [(i, ci(5, J1), G) for i in range(ci(5, I0), ci(5, I1) + 1)]
    rN += [(i, ci(3, J1), G) for i in range(ci(1, I1) + 1, ci(3, I1) + 1)]
    rN += [(i, ci(4, J1), G) for i in range(ci(4, I0) + 1, ci(4, I1))]
    rN += [(ci(4, I0), ci(4, J1), C3), (ci(4, I1), ci(4, J1), C4)]
    pW = [(ci(0, I0), j, D) for j in range(ci(0, J0), ci(0, J1) + 1)]
    pW += [(ci(2, I0), j, D) for j in range(ci(5, J1) + 1, ci(0, J0))]
    pW += [(ci(3, I0), j, D) for j in range(ci(3, J0), ci(2, J0))]
    pW += [(ci(4, I0), j, D) for j in range(ci(1, J0), ci(4, J1) + 1)]
    pW += [(ci(5, I0), j, O_) for j in range(ci(5, J0), ci(5, J1) + 1)]
    pE = [(ci(1, I1), j, D) for j in range(ci(1, J0), ci(1, J1) + 1)]
    pE += [(ci(4, I1), j, D) for j in range(ci(4, J0), ci(4, J1) + 1)]
    pS = [(i, ci(0, J0), D) for i in range(ci(0, I0), ci(0, I1) + 1)]
    pS += [(i, ci(5, J0), D) for i in range(ci(5, I0), ci(2, I1) + 1)]
    pN = [(i, ci(0, J1), D) for i in range(ci(0, I0), ci(1, I1) + 1)]
    pN += [(i, ci(5, J1), D) for i in range(ci(5, I0), ci(5, I1) + 1)]
    pN += [(i, ci(3, J1), D) for i in range(ci(1, I1) + 1, ci(3, I1) + 1)]
    pN += [(i, ci(4, J1), O_) for i in range(ci(4, I0), ci(4, I1) + 1)]
    return {"x": [xW, xE, xS, xN], "r": [rW, rE, rS, rN], "p": [pW, pE, pS, pN]}


def cathode_mw_case(h=2e-4, seed=12, Bx=None, Br=None, n0=3e17):
    """The coupled microwave solve on the reference's five-block cathode and
    coaxial feed (makeBoundary's boxes, masks and lists; input/setup.yaml's
    4.25 GHz, J1r_exc and dielectric; eps_r of the feed block from
    makeProfile, include/initialModule.hpp:1595-1600), xL = rmin = 0,
    38.8 x 14 mm at spacing h.  Bx, Br: the magnet's field on this grid (the
    reference's input/Bfield_data_0.2mm.csv at h = 0.2 mm, committed as
    tests/golden/bfield_cathode_0p2mm.npz); None gives a synthetic field of
    0.02 .. 0.4 T.  The plasma (peak n0, smooth, seeded) fills the five plasma
    blocks; the feed holds none.  Returns (MwParams, lists, MwSolveOpts, arrays)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ni, nj = int(round(0.0388 / h)), int(round(0.014 / h))
    p = _mw_params(ni, nj, h, h, 0.0, 4.25e9)
    cell, xf, rf = cathode_mw_boxes(ni, nj, h, h)
    p.n_bl = 6
    for b in range(6):
        _set_box(p.cell_box, b, cell[b])
        _set_box(p.xface_box, b, xf[b])
        _set_box(p.rface_box, b, rf[b])
    p.flag_mwRef, p.Pmw, p.Z0 = 1, 15.7024, 50.0
    p.r1, p.r2 = MW_RS["r1"], MW_RS["r2"]
    p.J1r_exc = J1R_EXC
    p.i_ref = int(0.0 / h + 0.5) + 1
    o = abi.MwSolveOpts()
    o.which, o.j_flc0, o.nu_eff, o.max_refine, o.tol = abi.MW_COUPLED, cell[0][2], 2.0e9, 3, 1e-13
    shp = (ni + 2, nj + 2)
    union = lambda boxes: np.maximum.reduce([_mask(ni, nj, b) for b in boxes]).astype(np.int32)
    a = {"act_p": union(cell), "act_x": union(xf), "act_r": union(rf)}
    a["epsr"] = np.ones(shp)
    c5 = cell[5]
    a["epsr"][c5[0]:c5[1] + 1, c5[2]:c5[3] + 1] = EPSR_DIELE
    plasma = np.maximum.reduce([_mask(ni, nj, b) for b in cell[:5]]) > 0
    a["rhoe"] = np.where(plasma, n0 * (0.05 + smooth_field(rng, ni, nj, h) ** 2), 0.0)
    a["nu_m"] = 2e9 * (0.5 + smooth_field(rng, ni, nj, h))
    if Bx is None:
        x = (np.arange(ni + 2) - 0.5) * h
        r = (np.arange(nj + 2) - 0.5) * h
        X, R = np.meshgrid(x, r, indexing="ij")
        Bx = 0.02 + 0.38 * np.exp(-((X - 0.02) ** 2) / (2 * 0.006 ** 2))
        Br = 0.1 * (X - 0.02) / 0.006 * np.exp(-((X - 0.02) ** 2) / (2 * 0.006 ** 2)) * R / 0.014
    a["Bx"], a["Br"] = np.ascontiguousarray(Bx, dtype=np.float64), np.ascontiguousarray(Br, dtype=np.float64)
    a["Er_dc"] = np.zeros(shp)
    for k in ("J1x_exc", "J1r_exc", "J1p_exc", "E1x", "E1r", "E1p"):
        a[k] = np.zeros(shp, dtype=np.complex128)
    return p, cathode_mw_lists(cell), o, {k: np.ascontiguousarray(v) for k, v in a.items()}
