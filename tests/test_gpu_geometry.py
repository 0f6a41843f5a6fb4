"""GPU parity of the implicit NDD step on a block geometry
(cathode_ndd_set_geometry): the reference's five-block cathode
(include/initialModule.hpp:239-399, the boundary loops of fluidModule.hpp)
against the oracle's restatement over blocks and segments.

Bars: every field within 1e-9 relative L2 of the oracle after 1 and 3 steps
(exact banded-LU solves in the oracle, residual 1e-10 on the device), the
electrode current within 1e-9; the rectangle given as an explicit geometry
reproduces the built-in rectangle bit for bit on the device; a geometry with
a plasma face that is neither inner nor a boundary segment is refused.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

abi = O.abi
W = O.workloads
pytestmark = pytest.mark.gpu

FIELDS = ["phi", "rhoe", "rhoi", "Ex", "Er", "rhoUex", "rhoUer", "rhoUix", "rhoUir", "nUex", "nUer", "rhoeps", "Te",
          "Gx", "Gr"]


@pytest.fixture
def rect_oracle():
    yield
    O.set_geometry(None)


def gpu_steps(lib, p, st, nsteps, g):
    h = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), lib)
    try:
        abi.check(lib.cathode_ndd_set_geometry(h, ctypes.byref(g) if g is not None else None), lib)
        f = abi.fields_struct(st)
        abi.check(lib.cathode_ndd_set_state(h, ctypes.byref(f)), lib)
        rep = abi.NddReport()
        abi.check(lib.cathode_ndd_step(h, nsteps, ctypes.byref(rep)), lib)
        abi.check(lib.cathode_ndd_get_state(h, ctypes.byref(f)), lib)
    finally:
        lib.cathode_ndd_destroy(h)
    return rep


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("nsteps", [1, 3])
def test_cathode5_matches_oracle(hiplib, rect_oracle, mode, nsteps):
    p, st, g = W.cathode5(seed=5, h=4e-4, mode=mode)
    rng = np.random.default_rng(5)
    m = W.plasma_mask(g, p.ni, p.nj)
    st["rate_ionize"] = np.where(m, rng.uniform(0.5e24, 2e24, m.shape), 0.0)
    print(f"ndd cathode5 {p.ni}x{p.nj} mode={mode} steps={nsteps}", flush=True)
    O.set_geometry(g)
    ref = W.copy_state(st)
    rr = O.ndd_step(p, ref, nsteps, solver=0)
    rep = gpu_steps(hiplib, p, st, nsteps, g)
    assert rep.nonconverged == 0 and rep.max_rel_residual <= 1e-10
    worst = {k: O.rel_l2(st[k], ref[k]) for k in FIELDS}
    bad = {k: v for k, v in worst.items() if not v <= 1e-9}
    assert not bad, (bad, worst)
    assert abs(rep.I_anode - rr.I_anode) <= 1e-9 * abs(rr.I_anode)
    assert np.abs(st["rhoUex"]).max() > 0


def test_explicit_rectangle_is_the_builtin_one(hiplib):
    p, st = W.config1(ni=64, nj=32, h=4e-4)
    g = abi.NddGeometry()
    g.n_bl = 1
    for q, v in enumerate((1, p.ni, 1, p.nj)):
        g.cell_box[q] = v
    for q, v in enumerate((2, p.ni, 1, p.nj)):
        g.xface_box[q] = v
    for q, v in enumerate((1, p.ni, 2, p.nj)):
        g.rface_box[q] = v
    segs = [(abi.SEG_X_LOW, 1, 1, p.nj, 0), (abi.SEG_X_HIGH, p.ni, 1, p.nj, 1), (abi.SEG_R_OPEN, p.nj, 1, p.ni, 0),
            (abi.SEG_R_AXIS, 1, 1, p.ni, 0)]
    tes = [(abi.SEG_X_LOW, 1, 1, p.nj, 1), (abi.SEG_X_HIGH, p.ni, 1, p.nj, 1), (abi.SEG_R_LOW, 1, 1, p.ni, 1),
           (abi.SEG_R_HIGH, p.nj, 1, p.ni, 1)]
    g.n_seg, g.n_te = len(segs), len(tes)
    for k, v in enumerate(segs):
        g.seg[k] = abi.NddSegment(*v)
    for k, v in enumerate(tes):
        g.te[k] = abi.NddSegment(*v)
    print(f"ndd rectangle 64x32 explicit geometry", flush=True)
    a = W.copy_state(st)
    b = W.copy_state(st)
    gpu_steps(hiplib, p, a, 2, None)
    gpu_steps(hiplib, p, b, 2, g)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_geometry_with_an_open_plasma_face_is_refused(hiplib):
    p, st, g = W.cathode5(seed=5, h=4e-4)
    g.n_seg = g.n_seg - 2  # drop x5 (open) and x6 (axis)
    h = ctypes.c_void_p()
    abi.check(hiplib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), hiplib)
    try:
        rc = hiplib.cathode_ndd_set_geometry(h, ctypes.byref(g))
        assert rc == abi.CATHODE_ERR_ARG
        assert b"neither inner nor a boundary" in hiplib.cathode_last_error()
    finally:
        hiplib.cathode_ndd_destroy(h)
