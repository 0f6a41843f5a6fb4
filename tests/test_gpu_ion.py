"""GPU parity of the ion block (cathode_ndd_ion_update: FluidModule::
solve_Uix_tmp, solve_Uir_tmp, solve_Uip_tmp, solve_rhoi_constTe,
correct_Ui_constTe; include/fluidModule.hpp:123-2072) against the oracle,
which restates the reference's loops literally over the five block boxes
while the device derives its boundary operations from the segments.

Cases: the five-block cathode at 0.4 mm and at the reference's own grid
(0.2 mm, 194 x 70) in FP64, mixed precision at 0.4 mm, a negative patch
that the update resets (the reset and its count), and the rectangle (no
oracle: finite, and the cell balance of the density closes).  Bars: every
array within 1e-9 relative L2 of the oracle (exact solves in the oracle,
relative residual 1e-13 on the device), equal reset counts.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import test_oracle_ion as TI

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
        f = abi.ion_fields_struct(out)
        rep = abi.IonReport()
        abi.check(lib.cathode_ndd_ion_update(h, ctypes.byref(params), ctypes.byref(f), ctypes.byref(rep)), lib)
    finally:
        lib.cathode_ndd_destroy(h)
    return out, rep


def compare(lib, h, params, arrays_fn, tag):
    ni, nj, g, *_ = W.reference_geometry(h)
    grid = W.default_params(ni, nj, h, h)
    a = arrays_fn(g, ni, nj, h)
    print(f"ion {tag} {ni}x{nj} mode={params.mode}", flush=True)
    O.set_geometry(g)
    ref = W.copy_arrays(a)
    rr = O.ion_update(grid, params, ref, solver=0)
    dev, rd = device(lib, grid, g, params, a)
    bad = {}
    for k in abi.ION_FIELD_NAMES:
        e = O.rel_l2(dev[k], ref[k])
        if not e <= 1e-9:
            bad[k] = e
    assert not bad, bad
    for s in (rd.solve_Uix, rd.solve_Uir, rd.solve_Uip, rd.solve_rhoi):
        assert s.converged, s.rel_residual
    assert rd.neg_rhoi == rr.neg_rhoi
    return dev, rd, a


@pytest.mark.parametrize("h,mode", [(4e-4, abi.CATHODE_MODE_FP64), (4e-4, abi.CATHODE_MODE_MIXED),
                                    (2e-4, abi.CATHODE_MODE_FP64)])
def test_cathode5_matches_oracle(hiplib, rect_oracle, h, mode):
    dev, _, a = compare(hiplib, h, W.ion_params(mode=mode), lambda g, ni, nj, hh: W.ion_state(g, ni, nj, hh), "cathode5")
    assert np.abs(dev["Uix"] - a["Uix"]).max() > 10.0   # the update moved things


def test_negative_density_matches_oracle(hiplib, rect_oracle):
    def sink(g, ni, nj, h):
        # a patch that starts slightly negative and has no source stays
        # negative after the solve (the continuity is positivity preserving
        # otherwise) and is put back to max(old, 0) = 0
        a = W.ion_state(g, ni, nj, h, seed=23)
        m = W.plasma_mask(g, ni, nj)
        i0, i1, j0, j1 = (g.cell_box[4 * 4 + q] for q in range(4))
        ic, jc = (i0 + i1) // 2, (j0 + j1) // 2
        a["rate_ionize"][ic - 4:ic + 5, jc - 4:jc + 5] = 0.0
        a["rhoi"][ic - 2:ic + 3, jc - 2:jc + 3] = -1e12
        assert m[ic - 2:ic + 3, jc - 2:jc + 3].all()
        return a
    _, rd, _ = compare(hiplib, 4e-4, W.ion_params(), sink, "sink")
    assert rd.neg_rhoi > 0


def test_rectangle_runs_and_closes_the_balance(hiplib):
    ni, nj, h = 64, 32, 4e-4
    grid = W.default_params(ni, nj, h, h)
    g = TI.rect_geometry(ni, nj)
    a = W.ion_state(g, ni, nj, h, noise=False)
    p = W.ion_params()
    print("ion rectangle 64x32", flush=True)
    dev, rd = device(hiplib, grid, None, p, a)
    assert rd.neg_rhoi == 0
    for k in abi.ION_FIELD_NAMES:
        assert np.isfinite(dev[k]).all(), k
    res = TI.balance_residual(g, ni, nj, h, p, a, dev)
    assert res < 1e-9, res


def test_device_arrays_match_host_arrays(hiplib):
    """cathode_ndd_ion_update_dev on device copies gives the host entry's bits"""
    h = 4e-4
    ni, nj, g, *_ = W.reference_geometry(h)
    grid = W.default_params(ni, nj, h, h)
    a = W.ion_state(g, ni, nj, h, seed=31)
    p = W.ion_params()
    print("ion device arrays vs host arrays", flush=True)
    host, _ = device(hiplib, grid, g, p, a)
    hd = ctypes.c_void_p()
    abi.check(hiplib.cathode_ndd_create(ctypes.byref(grid), ctypes.byref(hd)), hiplib)
    try:
        abi.check(hiplib.cathode_ndd_set_geometry(hd, ctypes.byref(g)), hiplib)
        f, D, names = abi.device_fields(hiplib, a, abi.ION_FIELD_NAMES, abi.IonFields)
        rep = abi.IonReport()
        abi.check(hiplib.cathode_ndd_ion_update_dev(hd, ctypes.byref(p), ctypes.byref(f), ctypes.byref(rep)), hiplib)
        out = abi.download_fields(D, names)
        D.free()
    finally:
        hiplib.cathode_ndd_destroy(hd)
    for k in abi.ION_FIELD_NAMES:
        assert np.array_equal(out[k], host[k]), k
