"""check_CFL (fluidModule.hpp:57-115) in the oracle, pinned by closed forms:
dt_min = eps0 / (e max(n_e) max(mu)) with the maximum taken over every cell,
ghosts included; the electron speed of uniform fluxes |Gamma| / n plus the
thermal speed sqrt(8 k Te_max / (pi m_e)); and the adaptive step
(flag_adp_dt = 1) equal to a fixed step at dt = CFL dt_min."""
import numpy as np

import oracle_lib as O

W = O.workloads


def _uniform(n=3e16, Te_eV=4.0, gx=2e20, gr=-5e19):
    p, _ = W.config0(ni=20, nj=12, h=2e-4)
    shp = (p.ni + 2, p.nj + 2)
    st = W.consistent_state(p, np.full(shp, n), np.full(shp, n), np.full(shp, Te_eV * W.E0 / W.BOLTZ))
    st["rhoUex"][...] = gx
    st["rhoUer"][...] = gr
    return p, st


def test_dt_min_and_speed_closed_forms():
    n, Te, gx, gr = 3e16, 4.0, 2e20, -5e19
    p, st = _uniform(n, Te, gx, gr)
    dt_min, nmax, mumax, temax, ue = O.check_cfl(p, st)
    assert abs(dt_min / (W.EPS0 / (W.E0 * n * p.mu_e)) - 1) < 1e-15
    assert nmax == n and mumax == p.mu_e and temax == Te * W.E0 / W.BOLTZ
    vth = np.sqrt(8 * W.BOLTZ * temax / (np.pi * W.MASSE))
    assert abs(ue / (np.hypot(gx, gr) / n + vth) - 1) < 1e-13


def test_maxima_include_ghost_cells():
    p, st = _uniform()
    st["rhoe"][0, 0] = 7e16        # a ghost corner
    st["Te"][p.ni + 1, 3] = 9e5    # a ghost on the right wall
    dt_min, nmax, _, temax, _ = O.check_cfl(p, st)
    assert nmax == 7e16 and temax == 9e5
    assert abs(dt_min / (W.EPS0 / (W.E0 * 7e16 * p.mu_e)) - 1) < 1e-15


def test_adaptive_step_is_the_fixed_step_at_cfl_dt_min():
    p, st = W.config0(ni=32, nj=24, h=2e-4)
    dt_min = O.check_cfl(p, st)[0]
    pa, _ = W.config0(ni=32, nj=24, h=2e-4)
    pa.flag_adp_dt, pa.CFL = 1, 0.4
    a = W.copy_state(st)
    ra = O.ndd_step(pa, a, 1, solver=0)
    assert abs(ra.dt / (0.4 * dt_min) - 1) < 1e-15 and ra.CFL == 0.4
    pf, _ = W.config0(ni=32, nj=24, h=2e-4)
    pf.dt = 0.4 * dt_min
    b = W.copy_state(st)
    rf = O.ndd_step(pf, b, 1, solver=0)
    assert abs(rf.CFL / 0.4 - 1) < 1e-14
    for k in ("phi", "rhoe", "rhoi", "Ex", "rhoUex"):
        np.testing.assert_array_equal(a[k], b[k])
