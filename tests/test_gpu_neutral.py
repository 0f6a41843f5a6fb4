"""GPU parity of the neutral and metastable density updates
(cathode_ndd_neutral_update; FluidModule::update_rhon / update_rhom,
include/fluidModule.hpp:4963-5792) against the oracle's restatement.

Cases: the reference's five-block cathode (inlet on the x1 wall, open face,
axis, every wall kind) in FP64 and mixed precision, the built-in rectangle,
a strong sink that drives both species negative (the reset rules and their
counts), the two updates called one at a time against one call, and at the
reference's own grid (194 x 70 at 0.2 mm) the cells' discrete balance on the
device's output.  Bars: every array within 1e-9 relative L2 of the oracle
(the oracle solves exactly, the device to a relative residual of 1e-13), the
reset counts equal, calcRes within 1e-8 relative; the balance within 1e-9
of the largest rate.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import test_oracle_neutral as TN

pytestmark = pytest.mark.gpu

W = O.workloads
abi = O.abi


@pytest.fixture
def rect_oracle():
    yield
    O.set_geometry(None)


def device(lib, grid, geo, params, arrays):
    h = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(grid), ctypes.byref(h)), lib)
    out = W.copy_arrays(arrays)
    try:
        abi.check(lib.cathode_ndd_set_geometry(h, ctypes.byref(geo) if geo is not None else None), lib)
        f = abi.neutral_fields_struct(out)
        rep = abi.NeutralReport()
        abi.check(lib.cathode_ndd_neutral_update(h, ctypes.byref(params), ctypes.byref(f), ctypes.byref(rep)), lib)
    finally:
        lib.cathode_ndd_destroy(h)
    return out, rep


def compare(lib, grid, geo, params, arrays, tag):
    print(f"neutral {tag} {grid.ni}x{grid.nj} which={params.which} mode={params.mode}", flush=True)
    O.set_geometry(geo)
    ref = W.copy_arrays(arrays)
    rr = O.neutral_update(grid, params, ref, solver=0)
    dev, rd = device(lib, grid, geo, params, arrays)
    bad = {}
    for k in abi.NEUTRAL_FIELD_NAMES:
        e = O.rel_l2(dev[k], ref[k])
        if not e <= 1e-9:
            bad[k] = e
    assert not bad, bad
    for sp, (sd, sr) in enumerate(((rd.solve_n, rr.solve_n), (rd.solve_m, rr.solve_m))):
        if params.which & (1 << sp):
            assert sd.converged and sd.rel_residual <= params.tol, (sp, sd.rel_residual)
    assert (rd.neg_rhon, rd.neg_rhom) == (rr.neg_rhon, rr.neg_rhom)
    for a, b in ((rd.error_rhon, rr.error_rhon), (rd.error_rhom, rr.error_rhom)):
        assert abs(a - b) <= 1e-8 * abs(b) + 1e-300, (a, b)
    return dev, rd, ref


@pytest.mark.parametrize("mode", [abi.CATHODE_MODE_FP64, abi.CATHODE_MODE_MIXED])
def test_cathode5_matches_oracle(hiplib, rect_oracle, mode):
    h = 4e-4
    ni, nj, g, *_ = W.reference_geometry(h)
    grid = W.default_params(ni, nj, h, h)
    a = W.neutral_state(g, ni, nj, h, seed=7, noise=True)
    dev, rd, _ = compare(hiplib, grid, g, W.neutral_params(which=3, mode=mode), a, "cathode5")
    m = W.plasma_mask(g, ni, nj)
    assert np.abs(dev["rhon"] - a["rhon"])[m].max() > 1e-3 * a["rhon"][m].max()   # the update moved things


def test_rectangle_matches_oracle(hiplib, rect_oracle):
    ni, nj, h = 64, 32, 4e-4
    grid = W.default_params(ni, nj, h, h)
    g = TN.rect_geometry(ni, nj, abi.SEG_R_OPEN)
    a = W.neutral_state(g, ni, nj, h, seed=9, noise=True)
    compare(hiplib, grid, None, W.neutral_params(which=3), a, "rectangle")


def test_negative_values_match_oracle(hiplib, rect_oracle):
    h = 4e-4
    ni, nj, g, c6, a = TN._sink_state(h)
    grid = W.default_params(ni, nj, h, h)
    _, rd, _ = compare(hiplib, grid, g, W.neutral_params(which=3), a, "sink")
    assert rd.neg_rhon > 0 and rd.neg_rhom > 0


def test_one_call_equals_two(hiplib):
    """update_rhon then update_rhom in one call, or in two: the same bits"""
    h = 4e-4
    ni, nj, g, *_ = W.reference_geometry(h)
    grid = W.default_params(ni, nj, h, h)
    a = W.neutral_state(g, ni, nj, h, seed=5, noise=True)
    print("neutral one call vs two", flush=True)
    one, _ = device(hiplib, grid, g, W.neutral_params(which=3), a)
    two, _ = device(hiplib, grid, g, W.neutral_params(which=1), a)
    two, _ = device(hiplib, grid, g, W.neutral_params(which=2), two)
    for k in abi.NEUTRAL_FIELD_NAMES:
        assert np.array_equal(one[k], two[k]), k


@pytest.mark.parametrize("sp", [0, 1])
def test_reference_grid_cell_balance(hiplib, rect_oracle, sp):
    """at the reference's grid (0.2 mm, 194 x 70) the device's output closes
    every plasma cell's balance (the oracle's formula of that identity)"""
    h = 2e-4
    ni, nj, g, *_ = W.reference_geometry(h)
    grid = W.default_params(ni, nj, h, h)
    a = W.neutral_state(g, ni, nj, h, seed=13, noise=False)
    p = W.neutral_params(which=3)
    print(f"neutral balance 194x70 species={sp}", flush=True)
    dev, rd = device(hiplib, grid, g, p, a)
    assert rd.neg_rhon == 0 and rd.neg_rhom == 0
    res = TN.balance_residual(g, ni, nj, h, p, a, dev, sp)[0]
    assert res < 1e-9, res


def test_bad_arguments_are_refused(hiplib):
    ni, nj, h = 16, 8, 4e-4
    grid = W.default_params(ni, nj, h, h)
    hd = ctypes.c_void_p()
    abi.check(hiplib.cathode_ndd_create(ctypes.byref(grid), ctypes.byref(hd)), hiplib)
    try:
        g = TN.rect_geometry(ni, nj, abi.SEG_R_OPEN)
        g.seg[0].flag = abi.SEG_INLET                    # an inlet on an x wall
        assert hiplib.cathode_ndd_set_geometry(hd, ctypes.byref(g)) == abi.CATHODE_ERR_ARG
        assert b"CATHODE_SEG_INLET" in hiplib.cathode_last_error()
        a = W.neutral_state(TN.rect_geometry(ni, nj, abi.SEG_R_OPEN), ni, nj, h)
        a["rhoUmx"] = None                               # metastables need their fluxes
        f = abi.neutral_fields_struct(a)
        rep = abi.NeutralReport()
        assert hiplib.cathode_ndd_neutral_update(hd, ctypes.byref(W.neutral_params(which=3)), ctypes.byref(f),
                                                 ctypes.byref(rep)) == abi.CATHODE_ERR_ARG
        p = W.neutral_params(which=4)
        assert hiplib.cathode_ndd_neutral_update(hd, ctypes.byref(p), ctypes.byref(f),
                                                 ctypes.byref(rep)) == abi.CATHODE_ERR_ARG
    finally:
        hiplib.cathode_ndd_destroy(hd)


def test_device_arrays_match_host_arrays(hiplib):
    """cathode_ndd_neutral_update_dev on device copies gives the host entry's bits"""
    h = 4e-4
    ni, nj, g, *_ = W.reference_geometry(h)
    grid = W.default_params(ni, nj, h, h)
    a = W.neutral_state(g, ni, nj, h, seed=29, noise=True)
    p = W.neutral_params(which=3)
    print("neutral device arrays vs host arrays", flush=True)
    host, _ = device(hiplib, grid, g, p, a)
    hd = ctypes.c_void_p()
    abi.check(hiplib.cathode_ndd_create(ctypes.byref(grid), ctypes.byref(hd)), hiplib)
    try:
        abi.check(hiplib.cathode_ndd_set_geometry(hd, ctypes.byref(g)), hiplib)
        f, D, names = abi.device_fields(hiplib, a, abi.NEUTRAL_FIELD_NAMES, abi.NeutralFields)
        rep = abi.NeutralReport()
        abi.check(hiplib.cathode_ndd_neutral_update_dev(hd, ctypes.byref(p), ctypes.byref(f), ctypes.byref(rep)),
                  hiplib)
        out = abi.download_fields(D, names)
        D.free()
    finally:
        hiplib.cathode_ndd_destroy(hd)
    for k in abi.NEUTRAL_FIELD_NAMES:
        assert np.array_equal(out[k], host[k]), k
