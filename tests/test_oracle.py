"""Pins of the CPU oracle (no GPU).

Each check fixes a part of the restatement by something other than the
restatement itself: a library routine (numpy.linalg.solve), a closed form, an
algebraic identity, or a hand-evaluated sweep -- chosen so that a dropped
term, a wrong sign or a shifted index fails.
"""
import numpy as np
import pytest

import oracle_lib as O

W = O.workloads
L = O.lib


# ---------------------------------------------------------------- Bernoulli
def B(z):
    return z / np.expm1(z) if z != 0 else 1.0


@pytest.mark.parametrize("z", [-80.0, -3.0, -1e-3, -1e-9, 1e-12, 1e-7, 0.5, 7.0, 60.0])
def test_bernoulli_weights(z):
    D_h = 2.5e4  # D / h
    vd = z * D_h
    bp = L().orc_bern_bp(vd, z, D_h, 0)
    bm = L().orc_bern_bm(vd, z, D_h, 0)
    # bm = (D/h) B(z), bp = (D/h) B(-z), and B(-z) - B(z) = z
    assert bm == pytest.approx(D_h * B(z), rel=1e-13)
    assert bp == pytest.approx(D_h * B(-z), rel=1e-13)
    assert bp - bm == pytest.approx(vd, rel=1e-12, abs=1e-12 * D_h)


def test_bernoulli_limits():
    D_h = 3.0
    assert L().orc_bern_bm(0.0, 0.0, D_h, 0) == D_h  # B(0) = 1
    z = 1e-6
    slope = (L().orc_bern_bm(z * D_h, z, D_h, 0) - D_h) / z
    assert slope == pytest.approx(-0.5 * D_h, rel=1e-5)  # B'(0) = -1/2
    # the clamp at lim_exp = 100 (fluidModule.hpp:3098)
    assert L().orc_bern_bm(150 * D_h, 150.0, D_h, 0) == pytest.approx(150 * D_h / np.expm1(100.0))
    # canonical and reference-literal forms agree away from z ~ 0 ...
    for z in (-5.0, 0.3, 9.0):
        a = L().orc_bern_bm(z * D_h, z, D_h, 0)
        b = L().orc_bern_bm(z * D_h, z, D_h, 1)
        assert a == pytest.approx(b, rel=1e-14)
    # ... and the literal exp()-1 form loses digits near z = 0
    z = 1e-9
    lit = L().orc_bern_bm(z * D_h, z, D_h, 1)
    assert abs(lit / (D_h * B(z)) - 1) > 1e-10


def test_func_Te_lerp():
    tab = np.array([0.0, 1.0, 4.0, 9.0, 16.0])
    dT = 2.0
    f = lambda T: L().orc_func_Te(O.P(tab), len(tab), T, dT)
    assert f(0.0) == 0.0
    assert f(1.0) == pytest.approx(0.5)       # between entries 0 and 1
    assert f(5.0) == pytest.approx(6.5)       # between 4 and 9
    assert f(6.5) == 16.0                     # beyond n-2: last entry (miniFuncs.hpp:42-44)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_calc_res_modes(mode):
    rng = np.random.default_rng(mode)
    a = rng.standard_normal((7, 6))
    b = a + 1e-3 * rng.standard_normal((7, 6))
    e = L().orc_calc_res(O.P(a), O.P(b), 6, 1, 5, 2, 4, 1e-100, mode)
    d = (a - b)[1:6, 2:5]
    o = b[1:6, 2:5]
    ref = [np.abs(d).max(), np.sqrt(np.mean(d * d / (o * o + 1e-100))), np.abs(d).max() / np.abs(o).max(),
           np.sqrt((d * d).sum() / (o * o).sum())][mode]
    assert e == pytest.approx(ref, rel=1e-13)


# ---------------------------------------------------------------- solves
def dense_from(aP, aE, aW, aN, aS, b, act):
    idx = -np.ones(aP.shape, int)
    cells = np.argwhere(act)
    for q, (i, j) in enumerate(cells):
        idx[i, j] = q
    n = len(cells)
    A = np.zeros((n, n), dtype=aP.dtype)
    rhs = np.zeros(n, dtype=b.dtype)
    for q, (i, j) in enumerate(cells):
        A[q, q] = aP[i, j]
        rhs[q] = b[i, j]
        for (di, dj, a) in ((1, 0, aE), (-1, 0, aW), (0, 1, aN), (0, -1, aS)):
            t = idx[i + di, j + dj]
            if t >= 0:
                A[q, t] -= a[i, j]
    return A, rhs, cells


def union(boxes, shape):
    act = np.zeros(shape, bool)
    for b in boxes:
        act[b[0]:b[1] + 1, b[2]:b[3] + 1] = True
    return act


@pytest.mark.parametrize("boxes", [[[1, 6, 1, 5]], [[1, 3, 1, 5], [4, 6, 2, 3]], [[1, 2, 1, 2], [5, 6, 4, 5]]])
def test_struct5_direct_vs_numpy(boxes):
    """HYPRE box semantics: couplings leaving the union are dropped
    (solverFuncs.hpp:40-102), x written only in the union (:143-152)."""
    rng = np.random.default_rng(1)
    shp = (8, 7)
    aE, aW, aN, aS = [rng.uniform(0.1, 1.0, shp) for _ in range(4)]
    aP = aE + aW + aN + aS + rng.uniform(0.0, 0.5, shp)
    b = rng.standard_normal(shp)
    act = union(boxes, shp)
    x = O.struct5_direct(aP, aE, aW, aN, aS, b, boxes, x=np.full(shp, 7.0))
    A, rhs, cells = dense_from(aP, aE, aW, aN, aS, b, act)
    xs = np.linalg.solve(A, rhs)
    for q, (i, j) in enumerate(cells):
        assert x[i, j] == pytest.approx(xs[q], rel=1e-12, abs=1e-14)
    assert np.all(x[~act] == 7.0)


def test_helm5_direct_vs_numpy():
    rng = np.random.default_rng(2)
    shp = (7, 6)
    c = lambda: rng.uniform(0.2, 1, shp) + 1j * rng.uniform(-0.3, 0.3, shp)
    aE, aW, aN, aS = c(), c(), c(), c()
    aP = (1 - 0.8j) * (aE + aW + aN + aS) - 1.5   # indefinite real part
    b = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    boxes = [[1, 5, 1, 4]]
    x = O.helm5_direct(aP, aE, aW, aN, aS, b, boxes)
    A, rhs, cells = dense_from(aP, aE, aW, aN, aS, b, union(boxes, shp))
    xs = np.linalg.solve(A, rhs)
    for q, (i, j) in enumerate(cells):
        assert abs(x[i, j] - xs[q]) <= 1e-11 * np.abs(xs).max()


def test_lu9_direct_vs_numpy():
    rng = np.random.default_rng(3)
    shp = (7, 7)
    co = [rng.uniform(0.1, 1, shp) for _ in range(8)]
    aP = sum(co) + 0.3
    b = rng.standard_normal(shp)
    boxes = [[1, 5, 1, 5]]
    act = union(boxes, shp)
    x = O.lu9_direct([aP] + co, b, boxes)
    idx = -np.ones(shp, int)
    cells = np.argwhere(act)
    for q, (i, j) in enumerate(cells):
        idx[i, j] = q
    A = np.zeros((len(cells),) * 2)
    offs = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1)]
    for q, (i, j) in enumerate(cells):
        A[q, q] = aP[i, j]
        for (di, dj), a in zip(offs, co):
            t = idx[i + di, j + dj]
            if t >= 0:
                A[q, t] -= a[i, j]
    xs = np.linalg.solve(A, b[act])
    np.testing.assert_allclose(x[act], xs, rtol=1e-11, atol=1e-13)


def sor_python(aP, aE, aW, aN, aS, b, boxes, alpha, maxITR, tol, x):
    """solverFuncs.hpp:169-216 evaluated with Python floats, cell by cell."""
    n = 0
    while True:
        n += 1
        err = 0.0
        for bx in boxes:
            for i in range(bx[0], bx[1] + 1):
                for j in range(bx[2], bx[3] + 1):
                    t = (aE[i, j] * x[i + 1, j] + aW[i, j] * x[i - 1, j] + aN[i, j] * x[i, j + 1] +
                         aS[i, j] * x[i, j - 1] + b[i, j]) / (aP[i, j] + 1e-100)
                    xn = (1.0 - alpha) * x[i, j] + alpha * t
                    e = abs((xn - x[i, j]) / (xn + x[i, j] + 1e-6)) * 2.0
                    err = max(err, e)
                    x[i, j] = xn
        if n > maxITR or not (err > tol):
            return n, err


def test_sor_matches_sequential_python():
    rng = np.random.default_rng(4)
    shp = (7, 6)
    aE, aW, aN, aS = [rng.uniform(0.1, 1, shp) for _ in range(4)]
    aP = aE + aW + aN + aS + 0.1
    b = rng.standard_normal(shp)
    boxes = [[1, 3, 1, 4], [4, 5, 2, 4]]
    x0 = rng.standard_normal(shp)
    xo = x0.copy()
    n, err = O.sor(aP, aE, aW, aN, aS, b, boxes, 1.9, 6, 1e-10, xo)
    xp = x0.copy()
    n2, err2 = sor_python(aP, aE, aW, aN, aS, b, boxes, 1.9, 6, 1e-10, xp)
    assert n == n2 == 7
    np.testing.assert_array_equal(xo, xp)
    assert err == err2


def test_cpu_mg_matches_direct():
    p, st = W.config0()
    box = [1, p.ni, 1, p.nj]
    for which in (0, 1, 2):
        aP, aE, aW, aN, aS, b = O.ndd_assemble(p, st, which)[:6]
        xd = O.struct5_direct(aP, aE, aW, aN, aS, b, box)
        xm, it, rr = O.mg_solve(aP, aE, aW, aN, aS, b, box)
        assert rr <= 1e-10 and it <= 8
        assert O.rel_l2(xm[1:-1, 1:-1], xd[1:-1, 1:-1]) < 1e-9


# ---------------------------------------------------------------- physics pins
def test_poisson_vacuum_is_linear_in_z():
    """Vacuum limit (no charge, negligible conduction): phi = 0 at z = 0 and
    V at z = L with grad phi = 0 on r walls has the exactly linear discrete
    solution phi_i = V (i - 1/2) / ni for any r-weighting."""
    p, st = W.config0(ni=24, nj=16)
    for k in ("rhoe", "rhoi", "rhoe_old", "rhoeps"):
        st[k][...] = 1e-30
    rep = O.ndd_step(p, st, 1, solver=0)
    i = np.arange(1, p.ni + 1)[:, None]
    exact = p.V_bias * (i - 0.5) / p.ni * np.ones((1, p.nj))
    np.testing.assert_allclose(st["phi"][1:-1, 1:-1], exact, rtol=1e-10, atol=1e-10 * p.V_bias)
    np.testing.assert_allclose(st["Ex"][1:-1, 1:-1], -p.V_bias / (p.ni * p.dx), rtol=1e-9)
    assert np.abs(st["Er"][1:-1, 1:-1]).max() < 1e-6
    assert rep.solves == 6


def test_sg_flux_vanishes_on_boltzmann_profile():
    """1-D SG exactness: with uniform E and Te the profile
    n_i ~ exp(vd x_i / D) carries zero flux through every interior x-face."""
    p, st = W.config0(ni=20, nj=6)
    E = -2.0e3
    st["Ex"][...] = E
    st["Er"][...] = 0.0
    Te = st["Te"][1, 1]
    mu = p.mu_e
    vd = mu * (-E) + 1e-3
    D = 1.380649e-23 * Te / 1.60217662e-19 * mu
    x = W.axial(p.ni, p.dx)[:, None] * np.ones((1, p.nj + 2))
    st["rhoe"][...] = 1e15 * np.exp(vd * x / D)
    st["rhoUex"][...] = 0.0
    st["rhoUer"][...] = 0.0
    out = O.ndd_assemble(p, st, 1)
    fx = out[6]
    scale = vd * st["rhoe"].max()
    assert np.abs(fx[2:p.ni + 1, 1:-1]).max() <= 1e-9 * scale
    # and it is not trivially zero: a uniform density gives vd n
    st["rhoe"][...] = 1e15
    fx2 = O.ndd_assemble(p, st, 1)[6]
    np.testing.assert_allclose(fx2[2:p.ni + 1, 1:-1], vd * 1e15, rtol=1e-10)


@pytest.mark.parametrize("which", [1, 2, 3])
def test_sg_assembly_mmatrix_and_conservation(which):
    """Off-diagonals >= 0, positive diagonal, and the r-weighted sum of the
    divergence telescopes to the boundary fluxes (fluidModule.hpp:3654-3657)."""
    p, st = W.config1(ni=32, nj=16, h=8e-4)
    if which == 3:
        # no electron flux -> no J.E heating: the energy RHS is a pure divergence
        st["rhoUex"][...] = 0.0
        st["rhoUer"][...] = 0.0
    aP, aE, aW, aN, aS, b, fx, fr = O.ndd_assemble(p, st, which)
    inner = (slice(1, p.ni + 1), slice(1, p.nj + 1))
    for a in (aE, aW, aN, aS):
        assert np.all(a[inner] >= 0.0)
    assert np.all(aP[inner] > 0.0)
    r = W.radii(p.nj, p.dr)
    # first PC iteration: n = n_old, S = 0  ->  b = -div F dx dr (x 3/5 for energy)
    fac = 0.6 if which == 3 else 1.0
    lhs = (r[None, 1:-1] * b[inner]).sum() / fac
    rf = 0.5 * (r[:-1] + r[1:])  # face radii r_{j-1/2}, j = 1..nj+1
    rhs = -(p.dr * (r[1:-1] * (fx[p.ni + 1, 1:-1] - fx[1, 1:-1])).sum() +
            p.dx * (rf[p.nj] * fr[1:-1, p.nj + 1] - rf[0] * fr[1:-1, 1]).sum())
    assert lhs == pytest.approx(rhs, rel=1e-9)


def test_abcd_round_trip():
    rng = np.random.default_rng(6)
    S = rng.uniform(-0.7, 0.7, 8)
    T = np.zeros(8)
    S2 = np.zeros(8)
    L().orc_s_to_abcd(O.P(S), 50.0, O.P(T))
    L().orc_abcd_to_s(O.P(T), 50.0, O.P(S2))
    np.testing.assert_allclose(S2, S, atol=1e-13)
    # reciprocal lossless thru line of length 0: S = [[0,1],[1,0]] -> ABCD = identity
    S = np.array([0, 0, 1, 0, 1, 0, 0, 0.0])
    L().orc_s_to_abcd(O.P(S), 50.0, O.P(T))
    np.testing.assert_allclose(T, [1, 0, 0, 0, 0, 0, 1, 0], atol=1e-15)
