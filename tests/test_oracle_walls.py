"""Pins of the oracle's wall fluxes and of the energy equation's J.E source
(no GPU).

The wall fluxes are moments of a Maxwellian, so a numerical quadrature of the
distribution fixes them independently of the closed forms the oracle codes:
- ions (fluidModule.hpp:1502-1526): the flux of a Maxwellian at the ion
  temperature drifting at mu_i E, through the wall face, on both walls;
- electrons (:3250-3367): the flux of a Maxwellian at rest over a retarding
  barrier e dphi, dphi = phi_cell - phi_wall when positive: with E = -dphi/dx
  that is max(-E h/2, 0) on a wall below the cell and max(+E h/2, 0) on a
  wall above it (the reference's :3255, :3335), so electrons are held back
  where the wall sits below the plasma potential;
- electron energy (:4174-4316): the energy those electrons carry, the
  parallel energy left after the barrier plus kT of the two other directions.
Secondary emission is off (no coefficients, no EISEE table), as in the
configurations of the workload.  A dropped factor, a sign or a swapped wall
fails these; the GPU path is held to the oracle by the GPU parity tests.
"""
import numpy as np
import pytest
from scipy import integrate

import oracle_lib as O

W = O.workloads
KB, E0 = 1.380649e-23, 1.60217662e-19


def moment(m, T, U, lo, hi, g=None):
    """integral over lo < v < hi of v g(v) f(v) dv, f the 1-D Maxwellian of
    mass m and temperature T drifting at U (g = 1 by default)"""
    s = np.sqrt(KB * T / m)
    wlo = max((lo - U) / s, -40.0)
    whi = min((hi - U) / s, 40.0)
    if whi <= wlo:
        return 0.0
    gg = g if g is not None else (lambda v: 1.0)
    fn = lambda w: (U + s * w) * gg(U + s * w) * np.exp(-0.5 * w * w) / np.sqrt(2.0 * np.pi)
    pts = [0.0] if wlo < 0.0 < whi else None
    val, _ = integrate.quad(fn, wlo, whi, points=pts, epsabs=0.0, epsrel=1e-13, limit=400)
    return val


def test_ion_wall_flux_is_drifting_maxwellian_half_range():
    p, st = W.config0(ni=8, nj=4, h=1e-4)
    Exs = np.array([-8e3, -1.5e3, 0.0, 4e3])  # |mu_i E / u_thermal| up to ~2.3
    st["Ex"][1, 1:5] = Exs
    st["Ex"][p.ni + 1, 1:5] = -Exs[::-1]
    fx = O.ndd_assemble(p, st, 2)[6]
    for j in range(1, 5):
        U = p.mu_i * st["Ex"][1, j]
        expect = st["rhoi"][1, j] * moment(p.massi, p.Ti, U, -np.inf, 0.0)  # toward -z, negative
        assert fx[1, j] == pytest.approx(expect, rel=1e-9)
        U = p.mu_i * st["Ex"][p.ni + 1, j]
        expect = st["rhoi"][p.ni, j] * moment(p.massi, p.Ti, U, 0.0, np.inf)
        assert fx[p.ni + 1, j] == pytest.approx(expect, rel=1e-9)


def test_electron_wall_flux_is_maxwellian_over_the_sheath_barrier():
    p, st = W.config0(ni=8, nj=4, h=1e-4)
    Exs = np.array([-3e4, -1e2, 0.0, 2e4])  # retarding on the left wall when E < 0
    st["Ex"][1, 1:5] = Exs
    st["Ex"][p.ni + 1, 1:5] = Exs
    fx = O.ndd_assemble(p, st, 1)[6]
    for j in range(1, 5):
        for face, cell, sgn in ((1, 1, -1.0), (p.ni + 1, p.ni, 1.0)):
            Te = st["Te"][cell, j]
            dphi = max(sgn * st["Ex"][face, j] * p.dx / 2.0, 0.0)
            vmin = np.sqrt(2.0 * E0 * dphi / p.masse)
            expect = sgn * st["rhoe"][cell, j] * moment(p.masse, Te, 0.0, vmin, np.inf)
            assert fx[face, j] == pytest.approx(expect, rel=1e-9), (face, j)


def test_electron_wall_energy_flux_is_carried_energy_over_the_barrier():
    ni, nj, h = 8, 4, 1e-4
    p = W.default_params(ni, nj, h, h, energy_on=True)
    rng = np.random.default_rng(4)
    n = 1e16 * (1.0 + 0.3 * rng.uniform(size=(ni + 2, nj + 2)))
    Te = (3.0 + 4.0 * rng.uniform(size=(ni + 2, nj + 2))) * E0 / KB
    st = W.consistent_state(p, n, n.copy(), Te)
    Exs = np.array([-3e4, 0.0, 5e3, 2e4])
    st["Ex"][1, 1:5] = Exs
    st["Ex"][ni + 1, 1:5] = -Exs
    fx = O.ndd_assemble(p, st, 3)[6]
    for j in range(1, 5):
        for face, cell, sgn in ((1, 1, -1.0), (ni + 1, ni, 1.0)):
            T = st["Te"][cell, j]
            dphi = max(sgn * st["Ex"][face, j] * p.dx / 2.0, 0.0)
            vmin = np.sqrt(2.0 * E0 * dphi / p.masse)
            # parallel kinetic energy after the barrier, plus kT of the two other directions
            energy = lambda v: 0.5 * p.masse * (v * v - vmin * vmin) + KB * T
            ne = st["rhoeps"][cell, j] / (1.5 * KB * T)
            expect = sgn * ne * moment(p.masse, T, 0.0, vmin, np.inf, energy)
            assert fx[face, j] == pytest.approx(expect, rel=1e-9), (face, j)


def test_energy_source_is_minus_e_flux_dot_field():
    """Uniform electron density, temperature and axial field, a uniform axial
    electron flux and none in r: the interior energy right-hand side is the
    J.E source alone, 3/5 (-e Gamma E) dx dr (fluidModule.hpp:4619-4637); the
    r-weighted divergence of the radial drift floor (EPS = 1e-3 m/s) is below
    1e-8 of it at this field and flux."""
    ni, nj, h = 10, 6, 1e-4
    p = W.default_params(ni, nj, h, h, energy_on=True)
    n = np.full((ni + 2, nj + 2), 2e16)
    Te = np.full((ni + 2, nj + 2), 4.0 * E0 / KB)
    st = W.consistent_state(p, n, n.copy(), Te)
    Ez, G = 1.5e4, 3e22
    st["Ex"][...] = Ez
    st["Er"][...] = 0.0
    st["rhoUex"][...] = G
    st["rhoUer"][...] = 0.0
    b = O.ndd_assemble(p, st, 3)[5]
    expect = 0.6 * (-E0 * G * Ez) * p.dx * p.dr
    inner = b[3:ni - 1, 1:nj + 1]
    np.testing.assert_allclose(inner, expect, rtol=1e-6)


def test_energy_source_absorbed_power_minus_loss():
    """Uniform state, no field and no flux: the interior energy right-hand side
    is 3/5 (Pabs - n_e eloss) dx dr (fluidModule.hpp:4619-4637)."""
    ni, nj, h = 10, 6, 1e-4
    p = W.default_params(ni, nj, h, h, energy_on=True)
    n = np.full((ni + 2, nj + 2), 2e16)
    Te = np.full((ni + 2, nj + 2), 4.0 * E0 / KB)
    st = W.consistent_state(p, n, n.copy(), Te)
    st["Ex"][...] = 0.0
    st["Er"][...] = 0.0
    rng = np.random.default_rng(9)
    st["Pabs"] = 1e8 * (1.0 + rng.uniform(size=n.shape))
    st["rate_eloss"] = 2e-9 * (1.0 + rng.uniform(size=n.shape))
    b = O.ndd_assemble(p, st, 3)[5]
    sl = (slice(3, ni - 1), slice(1, nj + 1))
    expect = 0.6 * (st["Pabs"][sl] - n[sl] * st["rate_eloss"][sl]) * p.dx * p.dr
    np.testing.assert_allclose(b[sl], expect, rtol=1e-7)


@pytest.mark.parametrize("Te_eV", [300.0, 2000.0])
def test_te_update_floor_and_cap(Te_eV):
    """After a step that barely moves a uniform state (dt = 1e-14 s, no bias):
    Te = n_eps / (3/2 k (n_e + 1e-3 max n_e)) = Te0 / 1.001, capped at 1000 eV
    (fluidModule.hpp:4897-4905)."""
    ni, nj, h = 12, 6, 1e-4
    p = W.default_params(ni, nj, h, h, dt=1e-14, V_bias=0.0, energy_on=True)
    n = np.full((ni + 2, nj + 2), 1e16)
    Te = np.full((ni + 2, nj + 2), Te_eV * E0 / KB)
    st = W.consistent_state(p, n, n.copy(), Te)
    O.ndd_step(p, st, 1, solver=0)
    sl = (slice(4, ni - 2), slice(1, nj + 1))
    if Te_eV < 1000.0:
        np.testing.assert_allclose(st["Te"][sl] / Te[sl], 1.0 / 1.001, rtol=1e-9)
    else:
        assert np.all(st["Te"][1:-1, 1:-1] == 1000.0 * E0 / KB)
