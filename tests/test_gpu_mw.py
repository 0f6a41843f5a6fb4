"""GPU parity of the microwave power profile and port normalisation
(cathode_mw_energy_profile; EmfieldModule::update_energy_profile,
emfieldModule.hpp:7630-8105) against the oracle.

Both branches of flag_mwRef, a magnetised plasma, overlapping block boxes
(counted once per block as in the reference's block loops), outputs that keep
their values outside the boxes, NULL excitation arrays, and shape changes
between calls.  Bars: every array within 1e-12 of the oracle relative to the
array's largest value (the same FP64 formulas, contraction off; the last bits
differ only through the complex division's evaluation), the report's sums and
port quantities within 1e-11 relative (the device sums per workgroup, the
oracle in loop order).
"""
import ctypes

import numpy as np
import pytest

import mw_cases as M
import oracle_lib as O

pytestmark = pytest.mark.gpu

abi = O.abi
OMEGA = 2.0 * np.pi * 2.45e9
S11, S12, S22 = 0.1 + 0.05j, 0.9 * np.exp(0.4j), -0.05 + 0.1j


def _device(lib, p, f):
    g = M.copy_fields(f)
    fs = abi.mw_fields_struct(g)
    rep = abi.MwReport()
    abi.check(lib.cathode_mw_energy_profile(ctypes.byref(p), ctypes.byref(fs), ctypes.byref(rep)), lib)
    return g, M.report_dict(rep)


def _compare(lib, p, f, tol_arr=1e-12, tol_rep=1e-11):
    print(f"mw case ni={p.ni} nj={p.nj} n_bl={p.n_bl} flag_mwRef={p.flag_mwRef}", flush=True)
    g_ref, rep = M.run_oracle(p, f)
    r_ref = M.report_dict(rep)
    g_dev, r_dev = _device(lib, p, f)
    for k in M.REAL + M.CPLX:
        if f[k] is None:
            continue
        scale = max(np.abs(g_ref[k]).max(), 1e-300)
        err = np.abs(g_dev[k] - g_ref[k]).max() / scale
        assert err <= tol_arr, (k, err)
    for k, v in r_ref.items():
        err = abs(r_dev[k] - v) / max(abs(v), 1e-300)
        assert err <= tol_rep or abs(r_dev[k] - v) <= 1e-300, (k, r_dev[k], v, err)
    return g_dev, r_dev


@pytest.mark.parametrize("flag", [1, 0])
@pytest.mark.parametrize("shape", [(24, 16), (128, 64), (512, 256)])
def test_mw_port_case_matches_oracle(hiplib, flag, shape):
    ni, nj = shape
    p, f = M.port_case(ni, nj, OMEGA, flag, S11=S11, S12=S12, S21=S12, S22=S22)
    _, d = _compare(hiplib, p, f)
    if flag == 1:
        assert d["ratio_eff"] > 0 and abs(d["P_fwd_all"] - p.Pmw) <= 1e-12 * p.Pmw


def test_mw_overlapping_blocks_and_null_excitation(hiplib):
    ni, nj = 40, 30
    f = M.random_fields(5, ni, nj)
    for k in ("J1x_exc", "J1r_exc", "J1p_exc"):
        f[k] = None
    boxes = [((1, ni, 1, nj), (2, ni, 1, nj), (1, ni, 2, nj)),
             ((5, 20, 3, 12), (5, 21, 3, 12), (5, 20, 3, 13)),
             ((1, ni, 1, nj), (1, ni + 1, 1, nj), (1, ni, 1, nj + 1))]
    p = M.params(ni, nj, 1e-4, 1.2e-4, 2e-3, OMEGA, boxes=boxes, Pmw=40.0)
    _compare(hiplib, p, f)


def test_mw_unmagnetised_and_shape_changes(hiplib):
    """B = 0, then the buffers of a larger and a smaller grid in turn"""
    for seed, (ni, nj) in enumerate([(64, 32), (200, 120), (33, 17)]):
        f = M.random_fields(seed, ni, nj, B=False)
        _compare(hiplib, M.params(ni, nj, 1e-4, 1e-4, 0.0, OMEGA), f)


def test_mw_rejects_boxes_outside_the_grid(hiplib):
    ni, nj = 16, 12
    f = M.random_fields(1, ni, nj)
    p = M.params(ni, nj, 1e-4, 1e-4, 0.0, OMEGA, boxes=[((1, ni, 1, nj), (2, ni + 2, 1, nj), (1, ni, 2, nj))])
    fs = abi.mw_fields_struct(M.copy_fields(f))
    rep = abi.MwReport()
    rc = hiplib.cathode_mw_energy_profile(ctypes.byref(p), ctypes.byref(fs), ctypes.byref(rep))
    assert rc == abi.CATHODE_ERR_ARG
    assert b"block boxes" in hiplib.cathode_last_error()
