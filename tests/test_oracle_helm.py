"""Pins of the oracle's E_theta microwave system (oracle/orc_helm.c) by
mathematics the reference fixes, independent of the oracle's own formulas.

emfieldModule.hpp:2414-2454 divides the r-weighted TE equation
    d2E/dz2 + (1/r) d/dr(r dE/dr) - E/r^2 + (k0^2 eps_r - i w mu0 sigma) E = i w mu0 J
by -dz^2 dr^2.  A manufactured solution E = sin(pi z/L) sin(pi r/R) vanishes
on the PEC walls (z = 0, L and r = R) and like r on the axis, so the discrete
solution must converge to it at second order; a dropped 1/r^2 term, swapped
qL/qR, or a wrong PEC fold breaks that.  The first-order open boundary
(:2721-2726, :2759-2770) is the centred outgoing-wave condition; its
reflection factor has the factored closed forms checked below.
"""
import numpy as np
import pytest

import oracle_lib as O

W = O.workloads


def _params(ni, nj, L, R, ne0, pec_z=1, pec_rmax=1, rmin=0.0):
    p = W.abi.EthetaParams()
    p.ni, p.nj = ni, nj
    p.dx, p.dr = L / ni, (R - rmin) / nj
    p.rmin = rmin
    p.omega = 2.67e10
    p.nu = 1e9
    p.epsr = 1.0
    p.masse = W.MASSE
    p.pec_z, p.pec_rmax = pec_z, pec_rmax
    return p


def _sigma(p, ne):
    return ne * W.E0 ** 2 / (p.masse * (p.nu + 1j * p.omega))


def _manufactured_error(n, L=5e-3, R=5e-3, ne0=1e17):
    p = _params(n, n, L, R, ne0)
    z = W.axial(n, p.dx)[:, None]
    r = W.radii(n, p.dr)[None, :]
    a, b = np.pi / L, np.pi / R
    rr = np.where(r > 0, r, 1.0)
    E = np.sin(a * z) * np.sin(b * r)
    ne = np.full((n + 2, n + 2), ne0)
    k2 = p.omega ** 2 * W.EPS0 * W.MU0 * p.epsr
    LE = np.sin(a * z) * (-a * a * np.sin(b * r) + b * np.cos(b * r) / rr - b * b * np.sin(b * r)
                          - np.sin(b * r) / rr ** 2) + (k2 - 1j * p.omega * W.MU0 * _sigma(p, ne)) * E
    J = np.ascontiguousarray(LE / (1j * p.omega * W.MU0))
    J[0, :] = J[-1, :] = 0
    J[:, 0] = J[:, -1] = 0
    Eh, _, _ = O.helm_solve_direct(p, ne, J)
    s = (slice(1, n + 1), slice(1, n + 1))
    return float(np.abs(Eh[s] - E[s]).max() / np.abs(E[s]).max())


def test_etheta_manufactured_second_order():
    errs = [_manufactured_error(n) for n in (16, 32, 64)]
    ratios = [errs[k] / errs[k + 1] for k in range(2)]
    assert errs[-1] < 2e-3, errs
    for q in ratios:
        assert 3.3 < q < 4.7, (errs, ratios)


def test_etheta_assembly_structure():
    """Interior rows: aE = aW = dr^2, aN + aS = 2 dx^2 with aN/aS = r_{j+1/2}/r_{j-1/2};
    the row sum leaves dx^2 dr^2 (1/r^2 - k0^2 + i w mu0 sigma)."""
    n = 12
    p = _params(n, n, 3e-3, 3e-3, 1e17)
    ne = np.full((n + 2, n + 2), 3e16)
    J = np.zeros((n + 2, n + 2), dtype=np.complex128)
    aP, aE, aW, aN, aS, b = O.helm_assemble(p, ne, J)
    r = W.radii(n, p.dr)
    dx2, dr2 = p.dx ** 2, p.dr ** 2
    s = (slice(2, n), slice(2, n))  # rows without folds
    np.testing.assert_allclose(aE[s], dr2, rtol=1e-15)
    np.testing.assert_allclose(aW[s], dr2, rtol=1e-15)
    np.testing.assert_allclose((aN + aS)[s].real, 2 * dx2, rtol=1e-13)
    jj = np.arange(2, n)
    np.testing.assert_allclose((aN[2, 2:n] / aS[2, 2:n]).real, (r[jj] + r[jj + 1]) / (r[jj] + r[jj - 1]), rtol=1e-13)
    k2 = p.omega ** 2 * W.EPS0 * W.MU0
    excess = (aP - aE - aW - aN - aS)[s]
    want = dx2 * dr2 * (1.0 / r[None, 2:n] ** 2 - k2 + 1j * p.omega * W.MU0 * _sigma(p, ne[s]))
    np.testing.assert_allclose(excess, want, rtol=1e-9)
    # axis: no coupling below the first ring
    np.testing.assert_array_equal(aS[1:n + 1, 1], 0)


def test_open_boundary_reflection_closed_form():
    n = 10
    c0 = 1.0 / np.sqrt(W.EPS0 * W.MU0)
    p = _params(n, n, 2e-3, 2e-3, 0.0, pec_z=0, pec_rmax=0)
    ne = np.zeros((n + 2, n + 2))
    J = np.zeros((n + 2, n + 2), dtype=np.complex128)
    open_ = O.helm_assemble(p, ne, J)
    p2 = _params(n, n, 2e-3, 2e-3, 0.0, pec_z=1, pec_rmax=1)
    pec = O.helm_assemble(p2, ne, J)
    # left wall: aP_open = aP_free - aW R, R = (1 - i k dz / 2) / (1 + i k dz / 2)
    kh = p.omega * p.dx / c0
    R = (1 - 0.5j * kh) / (1 + 0.5j * kh)
    aW_free = p.dr ** 2
    free_left = pec[0][1, 2:n] - aW_free  # PEC added aW once
    np.testing.assert_allclose(open_[0][1, 2:n], free_left - aW_free * R, rtol=1e-13)
    assert abs(abs(R) - 1) < 1e-14
    # outer r wall: R_r = (2c sqrt(qL) - i w dr qL) / (2c sqrt(qR) + i w dr qR),
    # q = r_j / r_face, r_face = (r_j + r_{j+1})/2
    r = W.radii(n, p.dr)
    rt = 0.5 * (r[n] + r[n + 1])
    qL, qR = r[n] / rt, r[n + 1] / rt
    Rr = (2 * c0 * np.sqrt(qL) - 1j * p.omega * p.dr * qL) / (2 * c0 * np.sqrt(qR) + 1j * p.omega * p.dr * qR)
    aN_free = 0.5 * (r[n] + r[n + 1]) / r[n] * p.dx ** 2
    free_top = pec[0][2:n, n] - r[n] / r[n + 1] * aN_free
    np.testing.assert_allclose(open_[0][2:n, n], free_top - aN_free * Rr, rtol=1e-12)
    assert open_[3][2:n, n].max() == 0 and open_[2][1, 2:n].max() == 0


@pytest.mark.parametrize("rmin", [0.0, 1e-3])
def test_etheta_post_ghosts_and_power(rmin):
    """Ghosts satisfy the wall conditions of the folds; Pabs = 0.5 Re(sigma) |E|^2
    (the cold-plasma form of 0.5 Re(J E*)); P = sum Pabs 2 pi r dz dr."""
    n = 8
    p = _params(n, n, 2e-3, 2e-3 + rmin, 0.0, rmin=rmin)
    rng = np.random.default_rng(3)
    E = np.zeros((n + 2, n + 2), dtype=np.complex128)
    E[1:-1, 1:-1] = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    ne = np.abs(rng.standard_normal((n + 2, n + 2))) * 1e17
    Pabs, Pt = O.helm_post(p, ne, E)
    r = W.radii(n, p.dr, rmin)
    np.testing.assert_allclose(E[0, 1:-1], -E[1, 1:-1], rtol=0, atol=0)
    np.testing.assert_allclose(E[-1, 1:-1], -E[-2, 1:-1], rtol=0, atol=0)
    np.testing.assert_allclose(r[n + 1] * E[1:-1, -1], -r[n] * E[1:-1, -2], rtol=1e-14)
    if rmin > 0:
        np.testing.assert_allclose(r[0] * E[1:-1, 0], -r[1] * E[1:-1, 1], rtol=1e-14)
    want = 0.5 * _sigma(p, ne).real * np.abs(E) ** 2
    np.testing.assert_allclose(Pabs[1:-1, 1:-1], want[1:-1, 1:-1], rtol=1e-12)
    tot = (want[1:-1, 1:-1] * 2 * np.pi * r[None, 1:-1] * p.dx * p.dr).sum()
    assert abs(Pt - tot) <= 1e-12 * abs(tot)
