"""GPU parity of the transport coefficients (cathode_ndd_set_rate_tables,
cathode_ndd_transport_update[_dev]; FluidModule::update_transport_coef,
include/fluidModule.hpp:5799-5976) against the oracle, with the tables read
from the reference's own rate files (tests/golden fixtures).

Cases: the five-block cathode and the rectangle (Te up to 9 eV with a few
cells beyond the fixture table, ion speeds beyond the ion table); the
device-array entry bit-identical to the host one; a call before the tables
are set refused.  Bar: every output within 1e-13 of the oracle relative to
the array's largest value (the same FP64 expressions without contraction).
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import rate_tables as RT
import test_oracle_ion as TI

pytestmark = pytest.mark.gpu

W = O.workloads
abi = O.abi
OUTS = abi.TRANSPORT_OUT + ["nu_ano_IAT"]


@pytest.fixture
def rect_oracle():
    yield
    O.set_geometry(None)


def device(lib, grid, geo, tabs, arrays, dev=False):
    h = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(grid), ctypes.byref(h)), lib)
    out = W.copy_arrays(arrays)
    try:
        abi.check(lib.cathode_ndd_set_geometry(h, ctypes.byref(geo) if geo is not None else None), lib)
        t = abi.rate_tables_struct(tabs)
        abi.check(lib.cathode_ndd_set_rate_tables(h, ctypes.byref(t)), lib)
        p = W.transport_params()
        if dev:
            f, D, names = abi.device_fields(lib, out, abi.TRANSPORT_FIELD_NAMES, abi.TransportFields)
            abi.check(lib.cathode_ndd_transport_update_dev(h, ctypes.byref(p), ctypes.byref(f)), lib)
            out = abi.download_fields(D, names)
            D.free()
        else:
            f = abi.transport_fields_struct(out)
            abi.check(lib.cathode_ndd_transport_update(h, ctypes.byref(p), ctypes.byref(f)), lib)
    finally:
        lib.cathode_ndd_destroy(h)
    return out


@pytest.mark.parametrize("where", ["cathode5", "rectangle"])
def test_transport_matches_oracle(hiplib, rect_oracle, where):
    h = 4e-4
    if where == "cathode5":
        ni, nj, g, *_ = W.reference_geometry(h)
    else:
        ni, nj, g = 64, 32, None
    grid = W.default_params(ni, nj, h, h)
    a = W.transport_state(g if g is not None else TI.rect_geometry(ni, nj), ni, nj, h)
    tabs, *_ = RT.read()
    print(f"transport {where} {ni}x{nj}", flush=True)
    O.set_geometry(g)
    ref = W.copy_arrays(a)
    O.transport_update(grid, W.transport_params(), tabs, ref)
    dev = device(hiplib, grid, g, tabs, a)
    bad = {}
    for k in OUTS:
        scale = max(np.abs(ref[k]).max(), 1e-300)
        e = np.abs(dev[k] - ref[k]).max() / scale
        if not e <= 1e-13:
            bad[k] = e
    assert not bad, bad
    assert np.abs(dev["nu_m"] - a["nu_m"]).max() > 0


def test_device_arrays_match_host_arrays(hiplib):
    h = 4e-4
    ni, nj, g, *_ = W.reference_geometry(h)
    grid = W.default_params(ni, nj, h, h)
    a = W.transport_state(g, ni, nj, h, seed=43)
    tabs, *_ = RT.read()
    print("transport device arrays vs host arrays", flush=True)
    host = device(hiplib, grid, g, tabs, a)
    dev = device(hiplib, grid, g, tabs, a, dev=True)
    for k in OUTS:
        assert np.array_equal(dev[k], host[k]), k


def test_refused_without_tables(hiplib):
    ni, nj, h = 16, 8, 4e-4
    grid = W.default_params(ni, nj, h, h)
    hd = ctypes.c_void_p()
    abi.check(hiplib.cathode_ndd_create(ctypes.byref(grid), ctypes.byref(hd)), hiplib)
    try:
        a = W.transport_state(TI.rect_geometry(ni, nj), ni, nj, h)
        f = abi.transport_fields_struct(a)
        rc = hiplib.cathode_ndd_transport_update(hd, ctypes.byref(W.transport_params()), ctypes.byref(f))
        assert rc == abi.CATHODE_ERR_ARG and b"cathode_ndd_set_rate_tables" in hiplib.cathode_last_error()
    finally:
        hiplib.cathode_ndd_destroy(hd)
