"""The oracle's restatement of the neutral and metastable density updates
(FluidModule::update_rhon, include/fluidModule.hpp:4963-5411, and
update_rhom, :5416-5792; oracle/orc_neutral.c), pinned without a GPU.

Pins:
- the neutral diffusion coefficient is the mean-free-path value
  D n = lambda n vbar / 3 with lambda = 1 / (sqrt(2) n pi d^2), d = 2 ri,
  and vbar the mean of the Maxwell speed distribution by quadrature: a
  profile n0 (1 - eps ((x - xc)^2 / Lx^2 + r^2 / R^2)) fed by the source
  -D div grad n and by explicit wall and inlet fluxes equal to the profile's
  own (-D grad n), is steady to O(eps^2) (the finite-volume cylindrical
  Laplacian is exact on quadratics), with the inlet's feed fn_In entering
  the system with its sign;
- the metastables' implicit wall loss is the Knudsen flux vbar m / 4 of the
  cell (vbar by quadrature): a quadratic profile whose curvature meets that
  Robin condition on both walls is steady to roundoff;
- on the five-block cathode, each cell's discrete balance
  (m_new - m_old) / dt = S - div Gamma over the fluxes and wall fluxes the
  update writes (the reference's nabla_rhoUn / nabla_rhoUm, :5336-5339,
  :5722-5725), exact with the direct solves, for both species;
- negative values: neutrals keep the fluxes of the unreset solution and reset
  every negative cell afterwards; metastables reset only in the first
  guard_blocks blocks and take their fluxes from the reset values;
- the inlet's x <= x_in (system) against x < x_in (flux record): a cell
  centred exactly at x_in is fed by the system but its face is recorded as
  a wall.
Each planted mistake (ORC_PLANT_NEUT_*, ORC_PLANT_META_*) fails its check.
"""
import numpy as np
import pytest
from scipy import integrate

import oracle_lib as O

W = O.workloads
abi = O.abi
KB, MI = W.BOLTZ, W.MASSI


@pytest.fixture
def clean():
    yield
    O.set_geometry(None)
    O.lib().orc_plant(0)


def mean_speed(T=W.TN, m=MI):
    """mean of the Maxwell speed distribution, by quadrature"""
    a = m / (2.0 * KB * T)
    f = lambda v: 4.0 * np.pi * (a / np.pi) ** 1.5 * v ** 3 * np.exp(-a * v * v)
    vmax = 12.0 / np.sqrt(a)
    val, _ = integrate.quad(f, 0.0, vmax, epsabs=0.0, epsrel=1e-13, limit=200)
    return val


def rect_geometry(ni, nj, top, top_flag=0):
    g = abi.NddGeometry()
    g.n_bl = 1
    for q, v in enumerate((1, ni, 1, nj)):
        g.cell_box[q] = v
    for q, v in enumerate((2, ni, 1, nj)):
        g.xface_box[q] = v
    for q, v in enumerate((1, ni, 2, nj)):
        g.rface_box[q] = v
    segs = [(abi.SEG_X_LOW, 1, 1, nj, 0), (abi.SEG_X_HIGH, ni, 1, nj, abi.SEG_BIAS), (top, nj, 1, ni, top_flag),
            (abi.SEG_R_AXIS, 1, 1, ni, 0)]
    g.n_seg = len(segs)
    for k, v in enumerate(segs):
        g.seg[k] = abi.NddSegment(*v)
    return g


def grid(ni, nj, h):
    return W.default_params(ni, nj, h, h)


def zero_arrays(ni, nj):
    return {k: np.zeros((ni + 2, nj + 2)) for k in abi.NEUTRAL_FIELD_NAMES}


# ------------------------------------------------------------------ neutral D and the inlet
NI, NJ, H = 24, 12, 4e-4
N0, EPS = 1e20, 1e-6
FN_IN = 3e21


def _neutral_steady():
    ni, nj, h = NI, NJ, H
    Lx, R = ni * h, nj * h
    xc = Lx / 2.0
    vbar = mean_speed()
    sigma = np.pi * (2.0 * W.RI) ** 2
    D0 = vbar / (3.0 * np.sqrt(2.0) * sigma) / N0      # lambda vbar / 3 at n0
    g = rect_geometry(ni, nj, abi.SEG_R_HIGH, abi.SEG_INLET)
    O.set_geometry(g)
    x = W.axial(ni, h)[:, None]
    r = W.radii(nj, h)[None, :]
    prof = N0 * (1.0 - EPS * ((x - xc) ** 2 / Lx ** 2 + r ** 2 / R ** 2))
    a = zero_arrays(ni, nj)
    a["rhon"][...] = prof
    a["rhoe"][...] = 1.0
    a["nu_excStep"][...] = D0 * EPS * N0 * (2.0 / Lx ** 2 + 4.0 / R ** 2)   # -D0 lap(prof)
    # explicit wall fluxes = the profile's own flux -D0 grad n at the face
    a["rhoUix_wall"][1, 1:nj + 1] = D0 * EPS * N0 / Lx       # Gamma = -D0 eps n0 / Lx at x = 0
    a["rhoUix_wall"][ni + 1, 1:nj + 1] = -D0 * EPS * N0 / Lx
    G_top = 2.0 * D0 * EPS * N0 / R
    p = W.neutral_params(which=1, dt_n=h * h / D0)
    p.fn_In = FN_IN
    p.x_in = W.axial(ni, h)[10] + 0.3 * h                    # cells 1..10 are the inlet
    xs = W.axial(ni, h)[1:ni + 1]
    inlet = xs <= p.x_in
    a["rhoUir_wall"][1:ni + 1, nj + 1] = np.where(inlet, -G_top - FN_IN, -G_top)
    b = W.copy_arrays(a)
    rep = O.neutral_update(grid(ni, nj, h), p, b)
    dev = np.abs(b["rhon"] - prof)[1:ni + 1, 1:nj + 1].max() / (EPS * N0)
    return dev, rep, b, a, inlet


def test_neutral_diffusion_is_kinetic_theory_and_the_inlet_feeds(clean):
    dev, rep, b, a, inlet = _neutral_steady()
    assert rep.solve_n.rel_residual < 1e-13
    assert dev < 1e-4, dev
    # the flux record: inlet faces carry the feed and keep their wall copy,
    # the other top faces record the wall flux
    nj = NJ
    top = b["rhoUnr"][1:NI + 1, nj + 1]
    want = np.where(inlet, -a["rhoUir_wall"][1:NI + 1, nj + 1] - FN_IN, -a["rhoUir_wall"][1:NI + 1, nj + 1])
    assert np.array_equal(top, want)
    assert np.array_equal(b["rhoUnr_wall"][1:NI + 1, nj + 1], np.where(inlet, 0.0, want))


@pytest.mark.parametrize("plant", [O.ORC_PLANT_NEUT_SIGMA, O.ORC_PLANT_NEUT_INLET_SIGN])
def test_neutral_steady_profile_catches(clean, plant):
    O.lib().orc_plant(plant)
    dev, *_ = _neutral_steady()
    assert dev > 1e-2, dev


# ------------------------------------------------------------------ metastable Knudsen walls
def _meta_steady():
    ni, nj, h = NI, NJ, H
    Lx = ni * h
    xc = Lx / 2.0
    vbar = mean_speed()
    D = W.DMN / N0
    c = (Lx - h) / 2.0                                      # wall cell centre to xc
    acoef = -0.25 * vbar / (D * Lx + 0.25 * vbar * c * c)   # -D m'(wall) = vbar m(cell) / 4
    O.set_geometry(None)                                    # walls in x, open top, axis
    x = W.axial(ni, h)[:, None] + 0.0 * W.radii(nj, h)[None, :]
    m0 = 1e17
    prof = m0 * (1.0 + acoef * (x - xc) ** 2)
    a = zero_arrays(ni, nj)
    a["rhon"][...] = N0
    a["rhom"][...] = prof
    a["rhoe"][...] = 1.0
    a["nu_excMeta"][...] = -2.0 * acoef * D * m0            # -D lap(prof)
    p = W.neutral_params(which=2, dt_m=1e3 * h * h / D)  # the profile is steady at any dt
    b = W.copy_arrays(a)
    rep = O.neutral_update(grid(ni, nj, h), p, b)
    dev = np.abs(b["rhom"] - prof)[1:ni + 1, 1:nj + 1].max() / m0
    return dev, rep, b, vbar


def test_metastable_wall_loss_is_the_knudsen_flux(clean):
    dev, rep, b, vbar = _meta_steady()
    assert dev < 1e-12, dev
    # the recorded wall flux is vbar m / 4 of the wall cell
    m = b["rhom"]
    assert np.allclose(b["rhoUmx_wall"][NI + 1, 1:NJ + 1], 0.25 * vbar * m[NI, 1:NJ + 1], rtol=1e-12, atol=0)
    assert np.allclose(b["rhoUmx_wall"][1, 1:NJ + 1], -0.25 * vbar * m[1, 1:NJ + 1], rtol=1e-12, atol=0)


def test_metastable_steady_profile_catches_the_wrong_mean_speed(clean):
    O.lib().orc_plant(O.ORC_PLANT_META_VTH)
    dev, *_ = _meta_steady()
    assert dev > 1e-3, dev


# ------------------------------------------------------------------ balance on the five-block cathode
def wall_flags(g, ni, nj, h, p, sp):
    """the reference's b*_wall flags of every cell (1 = no wall on that side)"""
    shp = (ni + 2, nj + 2)
    bLx, bRx, bLr, bRr = (np.ones(shp) for _ in range(4))
    xs = W.axial(ni, h, p.xL)
    for s, kind, ax, fi, fj, ci, cj in W.segment_faces(g):
        if kind == abi.SEG_X_LOW:
            bLx[ci, cj] = 0
        elif kind == abi.SEG_X_HIGH:
            bRx[ci, cj] = 0
        elif kind == abi.SEG_R_LOW:
            bLr[ci, cj] = 0
        elif kind == abi.SEG_R_HIGH:
            inlet = sp == 0 and (g.seg[s].flag & abi.SEG_INLET) and xs[ci] <= p.x_in
            if not inlet:
                bRr[ci, cj] = 0
    return bLx, bRx, bLr, bRr


def balance_residual(g, ni, nj, h, p, a, b, sp):
    """max over plasma cells of |(n - n_old)/dt - S + div| / max|(n - n_old)/dt|"""
    r = W.radii(nj, h)
    qR = ((r[1:nj + 1] + r[2:nj + 2]) / 2.0 / r[1:nj + 1])[None, :]
    qL = ((r[1:nj + 1] + r[0:nj]) / 2.0 / r[1:nj + 1])[None, :]
    key = "n" if sp == 0 else "m"
    dens = "rho" + key
    fx, fr = b["rhoU%sx" % key], b["rhoU%sr" % key]
    fxw, frw = b["rhoU%sx_wall" % key], b["rhoU%sr_wall" % key]
    bLx, bRx, bLr, bRr = (f[1:ni + 1, 1:nj + 1] for f in wall_flags(g, ni, nj, h, p, sp))
    I = slice(1, ni + 1)
    J = slice(1, nj + 1)
    div = ((fx[2:ni + 2, J] * bRx - fx[I, J] * bLx) / h + (qR * fr[I, 2:nj + 2] * bRr - qL * fr[I, J] * bLr) / h +
           (fxw[2:ni + 2, J] - fxw[I, J]) / h + (qR * frw[I, 2:nj + 2] - qL * frw[I, J]) / h)
    ne = a["rhoe"][I, J]
    if sp == 0:
        S = (a["nu_excStep"] + a["nu_super"])[I, J] * ne - (a["nu_excMeta"] + 0.5 * a["nu_exc"] + a["nu_ionz"])[I, J] * ne
        dt = p.dt_n
    else:
        S = ne * (a["nu_excMeta"] + 0.5 * a["nu_exc"])[I, J] - ne * (a["nu_ionzStep"] + a["nu_excStep"] + a["nu_super"])[I, J]
        dt = p.dt_m
    rate = (b[dens] - a[dens])[I, J] / dt
    m = W.plasma_mask(g, ni, nj)[I, J]
    res = (rate - S + div)[m]
    return np.abs(res).max() / np.abs(rate[m]).max(), (rate - S + div), m


def _cathode_balance(sp, seed=7):
    h = 4e-4
    ni, nj, g, *_ = W.reference_geometry(h)
    O.set_geometry(g)
    a = W.neutral_state(g, ni, nj, h, seed=seed, noise=False)
    p = W.neutral_params(which=3)
    b = W.copy_arrays(a)
    rep = O.neutral_update(grid(ni, nj, h), p, b)
    assert rep.neg_rhon == 0 and rep.neg_rhom == 0
    # the metastable balance reads the metastable wall fluxes the update wrote;
    # the neutral one the inputs (neutrals run first)
    return balance_residual(g, ni, nj, h, p, a, b, sp)[0]


@pytest.mark.parametrize("sp", [0, 1])
def test_cathode5_cell_balance(clean, sp):
    res = _cathode_balance(sp)
    assert res < 1e-10, res


@pytest.mark.parametrize("sp,plant", [(0, O.ORC_PLANT_NEUT_INLET_SIGN), (1, O.ORC_PLANT_META_RWALL_Q)])
def test_cathode5_balance_catches(clean, sp, plant):
    O.lib().orc_plant(plant)
    res = _cathode_balance(sp)
    assert res > 1e-4, res


def test_inlet_cell_at_x_in_is_fed_but_recorded_as_wall(clean):
    """x_in exactly at a cell centre: the system feeds that cell (x <= x_in,
    fluidModule.hpp:5042-5043), the flux record writes a wall there (x < x_in,
    :5232), so the reference's balance formula, which counts that face as
    inlet and adds its wall record, is off there by the feed plus the wall
    flux; every other cell closes"""
    h = 4e-4
    ni, nj, g, c6, *_ = W.reference_geometry(h)
    O.set_geometry(g)
    a = W.neutral_state(g, ni, nj, h, seed=3, noise=False)
    p = W.neutral_params(which=1)
    i_edge = int(c6[0][0]) + 5
    p.x_in = p.xL + (i_edge - 0.5) * h                # the same expression as the cell centre
    b = W.copy_arrays(a)
    O.neutral_update(grid(ni, nj, h), p, b)
    _, res, m = balance_residual(g, ni, nj, h, p, a, b, 0)
    j_top = int(c6[0][3])
    r = W.radii(nj, h)
    qR = (r[j_top] + r[j_top + 1]) / 2.0 / r[j_top]
    Gw = -a["rhoUir_wall"][i_edge, j_top + 1] - a["rhoUmr_wall"][i_edge, j_top + 1]
    want = qR * (p.fn_In + Gw) / h
    got = res[i_edge - 1, j_top - 1]
    assert abs(got - want) <= 1e-9 * want, (got, want)
    res[i_edge - 1, j_top - 1] = 0.0
    rate_scale = np.abs(((b["rhon"] - a["rhon"]) / p.dt_n)[1:ni + 1, 1:nj + 1][m]).max()
    assert np.abs(res[m]).max() < 1e-10 * rate_scale


# ------------------------------------------------------------------ negative values
def _sink_state(h=4e-4):
    ni, nj, g, c6, *_ = W.reference_geometry(h)
    a = W.neutral_state(g, ni, nj, h, seed=11, noise=False)
    # a strong loss in the columns of blocks 0 and 2 (block 2 is not guarded)
    cols = slice(int(c6[0][0]), int(c6[2][1]) + 1)
    a["nu_ionzStep"][cols, :] *= 1e5
    a["nu_ionz"][cols, :] *= 1e7
    return ni, nj, g, c6, a


def test_negative_values(clean):
    h = 4e-4
    ni, nj, g, c6, a = _sink_state(h)
    O.set_geometry(g)
    m = W.plasma_mask(g, ni, nj)
    blk = np.full(m.shape, -1)
    for bl in range(g.n_bl):
        i0, i1, j0, j1 = (g.cell_box[4 * bl + q] for q in range(4))
        blk[i0:i1 + 1, j0:j1 + 1] = bl
    # metastables: unguarded run for the raw solution, then the guarded one
    raw = W.copy_arrays(a)
    p0 = W.neutral_params(which=2, guard_blocks=0)
    r0 = O.neutral_update(grid(ni, nj, h), p0, raw)
    neg = (raw["rhom"] < 0) & m
    assert r0.neg_rhom == 0 and neg[(blk == 0) | (blk == 1)].any() and neg[blk == 2].any()
    gd = W.copy_arrays(a)
    r2 = O.neutral_update(grid(ni, nj, h), W.neutral_params(which=2), gd)
    guard = neg & ((blk == 0) | (blk == 1))
    assert r2.neg_rhom == guard.sum()
    assert np.array_equal(gd["rhom"][guard], np.maximum(a["rhom"][guard], 0.0))
    assert np.array_equal(gd["rhom"][m & ~guard], raw["rhom"][m & ~guard])
    # fluxes from the reset values: they change exactly on faces that touch a reset cell
    touch_x = np.zeros(m.shape, dtype=bool)
    touch_x[1:, :] = guard[1:, :] | guard[:-1, :]
    diff_x = gd["rhoUmx"] != raw["rhoUmx"]
    assert diff_x.any() and not (diff_x & ~touch_x).any()
    # neutrals: every negative cell is reset, after the fluxes were taken
    n_raw = W.copy_arrays(a)
    rn = O.neutral_update(grid(ni, nj, h), W.neutral_params(which=1), n_raw)
    assert rn.neg_rhon > 0
    reset = (n_raw["rhon"] == np.maximum(a["rhon"], 0.0)) & m
    _, res, mm = balance_residual(g, ni, nj, h, W.neutral_params(which=1), a, n_raw, 0)
    # the balance of a reset cell gives back its (negative) solution value
    p = W.neutral_params(which=1)
    implied = n_raw["rhon"][1:-1, 1:-1] - p.dt_n * res     # the value the fluxes were taken from
    cells = reset[1:-1, 1:-1] & mm
    assert cells.sum() == rn.neg_rhon
    assert (implied[cells] < 0).all()
    scale = np.abs((n_raw["rhon"] - a["rhon"])[1:-1, 1:-1][mm & ~cells]).max() / p.dt_n
    assert np.abs(res[mm & ~cells]).max() < 1e-10 * scale
