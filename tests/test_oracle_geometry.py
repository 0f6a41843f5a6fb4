"""The block geometry of the NDD step: the reference's five-block cathode
(include/initialModule.hpp:239-343 boxes; the boundary loops of the PC
routines, fluidModule.hpp:2152-2249) and the oracle's restatement of the
routines over blocks and boundary segments (no GPU).

Pins:
- the block boxes of the reference's own parameters (x1..x6, r1..r5 of
  initialModule.hpp:222-232, xR, rmax of params.hpp:17-19, dx = dr = 0.2 mm)
  against values worked out by hand from its int(x/dx +- 0.5) + 1 formulas;
- topology, checked here independently of the library's validation: the
  face boxes hold exactly the faces between two plasma cells, and every
  face between a plasma cell and a non-plasma cell lies in exactly one
  boundary segment of the matching side;
- the rectangle given as an explicit geometry reproduces the built-in one
  bit for bit;
- on the five-block cathode, the discrete Gauss law of the corrected
  Poisson pass in every plasma cell (the same identity as in
  test_oracle_pins.py: eps0 div E - e dt div Gamma = e (n_i - n_e) - e dt S,
  now with walls on every side, internal corners and the open face), and the
  electrons' discrete balance n_new = n_old + dt (S - div Gamma_new) in
  every plasma cell (wall fluxes explicit, inner faces corrected with the
  weights of the coefficients), exact with the banded-LU solves.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

W = O.workloads
abi = O.abi
E0, EPS0 = W.E0, W.EPS0

HAND = {  # block: cell box {i0, i1, j0, j1}
    0: (40, 67, 21, 45),    # x6 = 7.8 mm .. x2 = 13.4 mm, r3 = 4 .. r4 = 9 mm
    1: (68, 114, 11, 45),   # x2 .. x4 = 22.8 mm, r5 = 2 .. r4
    2: (68, 89, 4, 10),     # x2 .. x3 = 17.8 mm, r1 = 0.6 .. r5
    3: (90, 139, 1, 10),    # x3 .. x5 = 27.8 mm, 0 .. r5
    4: (140, 194, 1, 70),   # x5 .. xR = 38.8 mm, 0 .. rmax = 14 mm
    5: (1, 67, 4, 10),      # the dielectric: x1 = 0 .. x2, r1 .. r2 = 2 mm
}


@pytest.fixture
def rect_geo():
    yield
    O.set_geometry(None)
    O.lib().orc_plant(0)


def test_reference_block_boxes_by_hand():
    ni, nj, g, c6, x6, r6 = W.reference_geometry(2e-4)
    assert (ni, nj) == (194, 70)
    for b, box in HAND.items():
        assert tuple(c6[b]) == box, b
    # face boxes: x faces from the block's second column (or the first, where
    # the block to the left is plasma), r faces likewise
    assert tuple(x6[0]) == (41, 68, 21, 45) and tuple(r6[0]) == (40, 67, 22, 45)
    assert tuple(x6[1]) == (69, 114, 11, 45) and tuple(r6[1]) == (68, 114, 11, 45)
    assert g.n_bl == 5 and g.n_seg == 12 and g.n_ne == 5 and g.n_te == 5


def _segment_faces(q):
    """(axis, face i, face j, plasma cell, neighbour cell) of each face of a segment"""
    out = []
    for t in range(q.lo, q.hi + 1):
        if q.kind == abi.SEG_X_LOW:
            out.append(("x", q.cell, t, (q.cell, t), (q.cell - 1, t)))
        elif q.kind == abi.SEG_X_HIGH:
            out.append(("x", q.cell + 1, t, (q.cell, t), (q.cell + 1, t)))
        elif q.kind in (abi.SEG_R_LOW, abi.SEG_R_AXIS):
            out.append(("r", t, q.cell, (t, q.cell), (t, q.cell - 1)))
        else:
            out.append(("r", t, q.cell + 1, (t, q.cell), (t, q.cell + 1)))
    return out


def _check_topology(g, ni, nj):
    m = W.plasma_mask(g, ni, nj)
    inner_x = np.zeros_like(m)
    inner_r = np.zeros_like(m)
    for b in range(g.n_bl):
        i0, i1, j0, j1 = (g.xface_box[4 * b + q] for q in range(4))
        inner_x[i0:i1 + 1, j0:j1 + 1] = True
        i0, i1, j0, j1 = (g.rface_box[4 * b + q] for q in range(4))
        inner_r[i0:i1 + 1, j0:j1 + 1] = True
    # inner faces: exactly the faces between two plasma cells
    both_x = m[1:, :] & m[:-1, :]      # x face i between cells i-1 and i -> index i
    both_r = m[:, 1:] & m[:, :-1]
    assert np.array_equal(inner_x[1:, :], both_x)
    assert np.array_equal(inner_r[:, 1:], both_r)
    # boundary faces: plasma on one side only, each in exactly one segment
    seen = {}
    for s in range(g.n_seg):
        for ax, fi, fj, c, nb in _segment_faces(g.seg[s]):
            assert m[c] and not m[nb], (s, c, nb)
            assert (ax, fi, fj) not in seen, (s, ax, fi, fj)
            seen[(ax, fi, fj)] = s
    one_x = m[1:, :] ^ m[:-1, :]
    one_r = m[:, 1:] ^ m[:, :-1]
    want = {("x", i + 1, j) for i, j in zip(*np.nonzero(one_x))}
    want |= {("r", i, j + 1) for i, j in zip(*np.nonzero(one_r))}
    assert set(seen) == want


def test_reference_geometry_topology():
    ni, nj, g, _, _, _ = W.reference_geometry(2e-4)
    _check_topology(g, ni, nj)
    ni, nj, g, _, _, _ = W.reference_geometry(4e-4)
    _check_topology(g, ni, nj)


def _rect(ni, nj):
    g = abi.NddGeometry()
    g.n_bl = 1
    for q, v in enumerate((1, ni, 1, nj)):
        g.cell_box[q] = v
    for q, v in enumerate((2, ni, 1, nj)):
        g.xface_box[q] = v
    for q, v in enumerate((1, ni, 2, nj)):
        g.rface_box[q] = v
    segs = [(abi.SEG_X_LOW, 1, 1, nj, 0), (abi.SEG_X_HIGH, ni, 1, nj, 1), (abi.SEG_R_OPEN, nj, 1, ni, 0),
            (abi.SEG_R_AXIS, 1, 1, ni, 0)]
    tes = [(abi.SEG_X_LOW, 1, 1, nj, 1), (abi.SEG_X_HIGH, ni, 1, nj, 1), (abi.SEG_R_LOW, 1, 1, ni, 1),
           (abi.SEG_R_HIGH, nj, 1, ni, 1)]
    g.n_seg, g.n_te, g.n_ne = len(segs), len(tes), 0
    for k, v in enumerate(segs):
        g.seg[k] = abi.NddSegment(*v)
    for k, v in enumerate(tes):
        g.te[k] = abi.NddSegment(*v)
    return g


def test_rect_topology_and_explicit_rectangle_is_the_default(rect_geo):
    p, st = W.config1(ni=20, nj=12, h=4e-4)
    _check_topology(_rect(p.ni, p.nj), p.ni, p.nj)
    a = W.copy_state(st)
    O.ndd_step(p, a, 2, solver=0)
    O.set_geometry(_rect(p.ni, p.nj))
    b = W.copy_state(st)
    O.ndd_step(p, b, 2, solver=0)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def _div(p, Fx, Fr):
    """r-weighted divergence of face fields on every cell 1..ni x 1..nj"""
    ni, nj = p.ni, p.nj
    r = W.radii(nj, p.dr)
    qR = ((r[1:nj + 1] + r[2:nj + 2]) / 2.0 / r[1:nj + 1])[None, :]
    qL = ((r[1:nj + 1] + r[0:nj]) / 2.0 / r[1:nj + 1])[None, :]
    return ((Fx[2:ni + 2, 1:nj + 1] - Fx[1:ni + 1, 1:nj + 1]) / p.dx +
            (qR * Fr[1:ni + 1, 2:nj + 2] - qL * Fr[1:ni + 1, 1:nj + 1]) / p.dr)


def _cathode_step(seed=5, h=4e-4, maxpc=1):
    p, st, g = W.cathode5(seed=seed, h=h)
    rng = np.random.default_rng(seed)
    m = W.plasma_mask(g, p.ni, p.nj)
    st["rhoi"] = st["rhoi"] * np.where(m, 1.0 + 0.2 * rng.uniform(-1, 1, m.shape), 1.0)
    st["rate_ionize"] = np.where(m, rng.uniform(0.5e24, 2e24, m.shape), 0.0)
    p.maxITR_HES_phi = maxpc
    O.set_geometry(g)
    ref = W.copy_state(st)
    O.ndd_step(p, ref, 1, solver=0)
    return p, st, ref, m[1:-1, 1:-1]


def _gauss_residual():
    p, st, ref, m = _cathode_step()
    assert np.abs(ref["phi"][1:-1, 1:-1][m]).max() > 1.0  # the bias drives the potential
    t1 = EPS0 * _div(p, ref["Ex"], ref["Er"])
    t2 = E0 * p.dt * _div(p, ref["nUex"], ref["nUer"])
    t3 = E0 * (ref["rhoi"][1:-1, 1:-1] - ref["rhoe_old"][1:-1, 1:-1])
    t4 = E0 * p.dt * st["rate_ionize"][1:-1, 1:-1]
    res = (t1 - t2 - t3 + t4)[m]
    scale = max(np.abs(t[m]).max() for t in (t1, t2, t3, t4))
    return np.abs(res).max() / scale


def test_cathode5_poisson_gauss_law(rect_geo):
    res = _gauss_residual()
    assert res <= 1e-9, res


def test_cathode5_gauss_law_catches_a_missing_fold(rect_geo):
    O.lib().orc_plant(O.ORC_PLANT_GEO_FOLD)
    assert _gauss_residual() > 1e-6


def _balance_residual():
    p, st, ref, m = _cathode_step(maxpc=2)
    n_new, n_old = ref["rhoe"][1:-1, 1:-1], ref["rhoe_old"][1:-1, 1:-1]
    lhs = (n_new - n_old) / p.dt
    rhs = st["rate_ionize"][1:-1, 1:-1] - _div(p, ref["rhoUex"], ref["rhoUer"])
    scale = np.abs(lhs[m]).max()
    assert scale > 0
    # walls carry flux: the boundary terms are not all zero
    assert np.abs(ref["rhoUex"]).max() > 0 and np.abs(ref["rhoUer"]).max() > 0
    return np.abs((lhs - rhs)[m]).max() / scale


def test_cathode5_electron_balance(rect_geo):
    res = _balance_residual()
    assert res <= 1e-9, res


def test_cathode5_balance_catches_an_implicit_wall(rect_geo):
    O.lib().orc_plant(O.ORC_PLANT_GEO_WALL_COEF)
    assert _balance_residual() > 1e-6
