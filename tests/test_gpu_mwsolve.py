"""GPU parity of the coupled microwave solve (cathode_mw_solve[_dev];
EmfieldModule::solve_Microwave, include/emfieldModule.hpp:2249-3321, and
solve_Microwave_impedanceTest, :5528-6141) against the oracle and the
reference's own impedance-test output.

Cases: the reference's coaxial-feed impedance test (E_r alone, normalised to
8 W) against tests/golden/coax_impedance_er.csv and the oracle; a coupled,
magnetised rectangle whose lists run every fold and ghost rule, at two sizes;
the device-array entry bit-identical to the host-array one; argument errors.
Bars: the direct solves differ in elimination order only (block cyclic
reduction with refinement here, banded LU there), so every field agrees with
the oracle within 1e-9 of the field's largest value and the report's port
quantities within 1e-9 relative; the coax E_r within 1e-5 of the six printed
digits of the reference's file.
"""
import csv
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

W = O.workloads
abi = O.abi
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "coax_impedance_er.csv")
FIELDS = ("E1x", "E1r", "E1p")


def device(lib, p, lists, o, a, dev=False):
    g = W.copy_arrays(a)
    bc, keep = abi.mw_bc_struct(lists)
    rep = abi.MwSolveReport()
    if not dev:
        f = abi.mw_solve_fields_struct(g)
        abi.check(lib.cathode_mw_solve(ctypes.byref(p), ctypes.byref(bc), ctypes.byref(o), ctypes.byref(f),
                                       ctypes.byref(rep)), lib)
        return g, rep
    D = abi.MwSolveDevice(lib, g)
    try:
        abi.check(lib.cathode_mw_solve_dev(ctypes.byref(p), ctypes.byref(bc), ctypes.byref(o), ctypes.byref(D.f),
                                           ctypes.byref(rep)), lib)
        g = D.download()
    finally:
        D.free()
    return g, rep


def compare(lib, p, lists, o, a, tol=1e-9):
    ref = W.copy_arrays(a)
    rrep = O.mw_solve(p, lists, o, ref)
    got, drep = device(lib, p, lists, o, a)
    for k in FIELDS:
        scale = max(np.abs(ref[k]).max(), 1e-300)
        err = np.abs(got[k] - ref[k]).max() / scale
        assert err <= tol, (k, err)
    assert (drep.n_x, drep.n_r, drep.n_p) == (rrep.n_x, rrep.n_r, rrep.n_p)
    assert drep.rel_residual <= max(o.tol, 1e-12), drep.rel_residual
    return got, drep, ref, rrep


def test_coax_impedance_matches_reference_and_oracle(hiplib):
    p, lists, o, a = W.coax_impedance_test()
    print("mw_solve coax impedance test 194x7", flush=True)
    got, drep, ref, rrep = compare(hiplib, p, lists, o, a)
    for k in ("ratio", "P_fwd", "P_fwd_all", "J1r_exc"):
        assert abs(getattr(drep, k) / getattr(rrep, k) - 1.0) < 1e-9, k
    # the reflected powers are differences of nearly equal waves: bounded against the forward power
    for k in ("P_ref", "P_ref_all"):
        assert abs(getattr(drep, k) - getattr(rrep, k)) < 1e-9 * rrep.P_fwd, k
    assert abs(drep.P_fwd_all / 12.57480 - 1.0) < 5e-6   # test/校正/校正後FWD_power.txt
    # the reference's printed, cell-centred E_r
    r = p.rmin + (np.arange(p.nj + 2) - 0.5) * p.dr
    E = got["E1r"]
    with open(FIX, newline="") as f:
        rows = list(csv.reader(f))[1:]
    worst = 0.0
    for row in rows:
        d, i, j = int(row[0]), int(row[1]), int(row[2])
        rL, rR = (r[j - 1] + r[j]) / 2.0, (r[j + 1] + r[j]) / 2.0
        Ec = (rR * E[i, j + 1] + rL * E[i, j]) / (2.0 * r[j]) * np.exp(1j * 2.0 * np.pi / 18.0 * d)
        refv = complex(float(row[3]), float(row[4]))
        worst = max(worst, abs(Ec - refv) / abs(refv))
    assert worst < 1e-5, worst
    # the reference overwrites n_e and nu in the caller's arrays
    assert not got["rhoe"].any() and (got["nu_m"] == o.nu_eff).all()


@pytest.mark.parametrize("shape", [(24, 12), (96, 40)])
def test_coupled_magnetised_rect_matches_oracle(hiplib, shape):
    ni, nj = shape
    p, lists, o, a = W.mw_coupled_rect(ni=ni, nj=nj, seed=ni)
    print(f"mw_solve coupled rect {ni}x{nj}", flush=True)
    got, drep, ref, rrep = compare(hiplib, p, lists, o, a)
    assert drep.block == 3 * (nj + 2)
    # a corner no list names keeps its value
    for k in FIELDS:
        assert got[k][0, 0] == a[k][0, 0]


def test_device_arrays_match_host_arrays(hiplib):
    p, lists, o, a = W.mw_coupled_rect(ni=40, nj=16, seed=4)
    print("mw_solve device arrays vs host arrays", flush=True)
    host, _ = device(hiplib, p, lists, o, a)
    dev, _ = device(hiplib, p, lists, o, a, dev=True)
    for k in FIELDS:
        assert np.array_equal(host[k], dev[k]), k


def test_argument_errors(hiplib):
    p, lists, o, a = W.mw_coupled_rect(ni=12, nj=6, seed=1)
    bad = {k: [list(s) for s in v] for k, v in lists.items()}
    bad["x"][0].append(bad["x"][0][0])          # the same face twice in one list
    g = W.copy_arrays(a)
    bc, keep = abi.mw_bc_struct(bad)
    f = abi.mw_solve_fields_struct(g)
    rep = abi.MwSolveReport()
    rc = hiplib.cathode_mw_solve(ctypes.byref(p), ctypes.byref(bc), ctypes.byref(o), ctypes.byref(f),
                                 ctypes.byref(rep))
    assert rc == abi.CATHODE_ERR_ARG and b"twice" in hiplib.cathode_last_error()
    bad = {k: [list(s) for s in v] for k, v in lists.items()}
    bad["r"][2].append((3, 1, 2))               # an E_r entry below j = 2
    bc, keep = abi.mw_bc_struct(bad)
    rc = hiplib.cathode_mw_solve(ctypes.byref(p), ctypes.byref(bc), ctypes.byref(o), ctypes.byref(f),
                                 ctypes.byref(rep))
    assert rc == abi.CATHODE_ERR_ARG


BFIELD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bfield_cathode_0p2mm.npz")


@pytest.mark.parametrize("h", [4e-4, 2e-4])
def test_reference_cathode_matches_oracle(hiplib, h):
    """makeBoundary's geometry and lists; at the reference's 0.2 mm grid with
    the magnet's field of input/Bfield_data_0.2mm.csv (an ECR surface inside
    the plasma), at 0.4 mm with a synthetic field."""
    B = np.load(BFIELD) if h == 2e-4 else None
    p, lists, o, a = W.cathode_mw_case(h=h, Bx=None if B is None else B["Bx"], Br=None if B is None else B["Br"])
    print(f"mw_solve reference cathode h={h}", flush=True)
    got, drep, ref, rrep = compare(hiplib, p, lists, o, a)
    assert drep.n_x > 1500 and drep.n_r > 1500 and drep.n_p > 1500
