"""Pre-reduced global sums in the fused solver kernels (cathode_set_presum_min).

Every workgroup of the fused V-cycle, x-update, restart, setup and Te-update
launches needs a global sum (or maximum) of the previous launch's partials.
Above a threshold of partials one workgroup reduces them first and the launch
reads one value; below it each workgroup sums them itself.  Both forms must
give the same solution to rounding: the threshold is forced to 0 (always
pre-reduce) on a small grid and compared with the default path and with the
FP64 oracle (fluidModule.hpp:2080-4957 restated, its own multigrid to the same
1e-10 relative residual); and the default threshold is exercised where it
applies, on a 2048 x 1024 grid, against the oracle.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

abi = O.abi
W = O.workloads
pytestmark = pytest.mark.gpu

FIELDS = ["phi", "rhoe", "rhoi", "rhoeps", "Te", "Ex", "Er", "nUex", "nUer", "Gx", "Gr"]


def run_steps(lib, p, st, steps):
    """steps single NDD steps on a fresh handle; the state after each"""
    s = W.copy_state(st)
    h = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), lib)
    out, reps = [], []
    try:
        f = abi.fields_struct(s)
        abi.check(lib.cathode_ndd_set_state(h, ctypes.byref(f)), lib)
        for _ in range(steps):
            rep = abi.NddReport()
            abi.check(lib.cathode_ndd_step(h, 1, ctypes.byref(rep)), lib)
            abi.check(lib.cathode_ndd_get_state(h, ctypes.byref(f)), lib)
            out.append(W.copy_state(s))
            reps.append(rep)
    finally:
        lib.cathode_ndd_destroy(h)
    return out, reps


def paths(ni, nj):
    h = 5e-5 * 512 / ni
    res = {}
    for name, mode, learned in (("fp64", abi.CATHODE_MODE_FP64, 0), ("mixed", abi.CATHODE_MODE_MIXED, 0),
                                ("warm", abi.CATHODE_MODE_MIXED, 1)):
        p, st = W.config1(ni=ni, nj=nj, h=h, mode=mode)
        p.learned = learned
        res[name] = (p, st)
    return res


def test_forced_presum_matches_default_and_oracle(hiplib):
    steps = 3
    cases = paths(256, 128)
    p64, st0 = cases["fp64"]
    ref = W.copy_state(st0)
    O.lib().orc_mg_set_threads(8)
    refs = []
    for _ in range(steps):
        assert O.ndd_step(p64, ref, 1, solver=1).nonconverged == 0
        refs.append(W.copy_state(ref))
    try:
        for name, (p, st) in cases.items():
            hiplib.cathode_set_presum_min(-1)
            base, _ = run_steps(hiplib, p, st, steps)
            hiplib.cathode_set_presum_min(0)
            pre, reps = run_steps(hiplib, p, st, steps)
            for s in range(steps):
                assert reps[s].nonconverged == 0 and reps[s].max_rel_residual <= 1e-10, (name, s)
                e_b = max(O.rel_l2(pre[s][k], base[s][k]) for k in FIELDS)
                e_o = max(O.rel_l2(pre[s][k], refs[s][k]) for k in FIELDS)
                print(f"{name} step {s + 1}: always pre-reduced vs default {e_b:.2e}, vs oracle {e_o:.2e}",
                      flush=True)
                assert e_b <= 1e-9 and e_o <= 1e-9, (name, s, e_b, e_o)
    finally:
        hiplib.cathode_set_presum_min(-1)


def test_forced_presum_struct5(hiplib):
    """the generic box-union solve (setup, restart, fused iterations) with
    every sum pre-reduced: converged, and the same solution to rounding"""
    p, st = W.config0(ni=96, nj=64, h=1e-4)
    aP, aE, aW, aN, aS, b = O.ndd_assemble(p, st, 0)[:6]
    box = np.array([1, p.ni, 1, p.nj], dtype=np.int32)
    xs = []
    try:
        for m in (-1, 0):
            hiplib.cathode_set_presum_min(m)
            x = np.zeros_like(b)
            s = abi.Status()
            abi.check(hiplib.cathode_struct5_solve(*[abi.dptr(a) for a in (aP, aE, aW, aN, aS, b)],
                                                   box.ctypes.data_as(abi.c_int_p), 1, p.ni + 2, p.nj + 2,
                                                   p.nj + 2, 50, 1e-10, abi.CATHODE_MODE_MIXED, abi.dptr(x),
                                                   ctypes.byref(s)), hiplib)
            assert s.converged == 1
            assert O.rel_residual(aP, aE, aW, aN, aS, b, box, x) <= 1e-10
            xs.append(x)
    finally:
        hiplib.cathode_set_presum_min(-1)
    assert O.rel_l2(xs[1], xs[0]) <= 1e-9


def test_large_grid_default_threshold_vs_oracle(hiplib):
    """2048 x 1024 (8192 partials per 64 x 4 block grid: the default
    threshold pre-reduces): two steps of the FP64, mixed and warm-start
    paths against the oracle"""
    steps = 2
    cases = paths(2048, 1024)
    p64, st0 = cases["fp64"]
    ref = W.copy_state(st0)
    O.lib().orc_mg_set_threads(8)
    refs = []
    for _ in range(steps):
        assert O.ndd_step(p64, ref, 1, solver=1).nonconverged == 0
        refs.append(W.copy_state(ref))
    hiplib.cathode_set_presum_min(-1)
    for name, (p, st) in cases.items():
        out, reps = run_steps(hiplib, p, st, steps)
        for s in range(steps):
            assert reps[s].nonconverged == 0 and reps[s].max_rel_residual <= 1e-10, (name, s)
            e = {k: O.rel_l2(out[s][k], refs[s][k]) for k in FIELDS}
            print(f"2048x1024 {name} step {s + 1}: vs oracle max {max(e.values()):.2e}", flush=True)
            bad = {k: v for k, v in e.items() if not v <= 1e-9}
            assert not bad, (name, s, bad)
