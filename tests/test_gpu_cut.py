"""V-cycle cuts of mass-dominated operators in the NDD step.

A solve whose operator is mass-dominated (the row excess fraction
(aP - sum a_nb) / aP is large on a level) stops its V-cycles at that level:
level 0 when the fraction there is >= 0.9 (the ion continuity rows, whose
time term h^2/dt dwarfs the ion diffusion), else the last level of the first
fine launch unit below level 0 when its fraction is >= 0.35 (the electron and
energy rows, where the 4^l growth of the time term over the coarse levels
makes them diagonally dominant).  The cut only changes the preconditioner, so
every solve still reaches the 1e-10 relative residual and the fields agree
with the uncut run to the solver tolerance.

The metric's grid (512 x 256 cells of 0.05 mm, seed 1) is stepped with the
cut on and off (CATHODE_CUT=0, read when a handle is created), in FP64 and in
mixed precision, and each step is compared with the other and with the
oracle's multigrid step.
"""
import ctypes
import os

import pytest

import oracle_lib as O

abi = O.abi
W = O.workloads
pytestmark = pytest.mark.gpu

FIELDS = ["phi", "rhoe", "rhoi", "rhoeps", "Te", "Gx", "Gr"]
TOL = 1e-9
STEPS = 4
# report slots [2 eq + pass]: eq 0 electrons, 1 ions, 2 energy, 3 Poisson
ELECTRONS, IONS, ENERGY, POISSON = (0, 1), (2, 3), (4, 5), (6, 7)


def _handle(lib, p, st, cut):
    old = os.environ.get("CATHODE_CUT")
    os.environ["CATHODE_CUT"] = "1" if cut else "0"
    try:
        h = ctypes.c_void_p()
        abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), lib)
    finally:
        if old is None:
            del os.environ["CATHODE_CUT"]
        else:
            os.environ["CATHODE_CUT"] = old
    f = abi.fields_struct(st)
    abi.check(lib.cathode_ndd_set_state(h, ctypes.byref(f)), lib)
    return h, f


@pytest.mark.parametrize("mode", [abi.CATHODE_MODE_FP64, abi.CATHODE_MODE_MIXED])
def test_cut_matches_uncut_and_oracle(hiplib, mode):
    p, st0 = W.config1(mode=mode)
    st_cut, st_full, st_ref = W.copy_state(st0), W.copy_state(st0), W.copy_state(st0)
    hc, fc = _handle(hiplib, p, st_cut, True)
    hf, ff = _handle(hiplib, p, st_full, False)
    try:
        for step in range(STEPS):
            rc, rf = abi.NddReport(), abi.NddReport()
            abi.check(hiplib.cathode_ndd_step(hc, 1, ctypes.byref(rc)), hiplib)
            abi.check(hiplib.cathode_ndd_step(hf, 1, ctypes.byref(rf)), hiplib)
            abi.check(hiplib.cathode_ndd_get_state(hc, ctypes.byref(fc)), hiplib)
            abi.check(hiplib.cathode_ndd_get_state(hf, ctypes.byref(ff)), hiplib)
            O.ndd_step(p, st_ref, 1, solver=1)
            cuts = list(rc.cut_by_solve)
            print(f"step {step}: cuts {cuts}, iterations cut {list(rc.iterations_by_solve)} "
                  f"uncut {list(rf.iterations_by_solve)}")
            for r in (rc, rf):
                assert r.nonconverged == 0 and r.max_rel_residual <= 1e-10
            assert all(c == -1 for c in rf.cut_by_solve)
            if step == 0:
                assert all(c == -1 for c in cuts)  # no earlier step to judge from
            else:
                assert all(cuts[s] == 0 for s in IONS), cuts
                assert all(cuts[s] == 2 for s in ELECTRONS), cuts
                assert all(cuts[s] == -1 for s in POISSON), cuts
                assert all(cuts[s] in (-1, 2) for s in ENERGY), cuts
            for k in FIELDS:
                e_full = O.rel_l2(st_cut[k][1:-1, 1:-1], st_full[k][1:-1, 1:-1])
                e_ref = O.rel_l2(st_cut[k][1:-1, 1:-1], st_ref[k][1:-1, 1:-1])
                assert e_full <= TOL and e_ref <= TOL, (step, k, e_full, e_ref)
    finally:
        hiplib.cathode_ndd_destroy(hc)
        hiplib.cathode_ndd_destroy(hf)


def test_cut_in_batched_sweep_is_bitwise_sequential(hiplib):
    """Cut decisions are per handle and per solve: a batch of instances
    stepped together equals the same instances stepped one at a time."""
    import gpu_api as G

    states, params = [], []
    for q, (ni, nj) in enumerate([(128, 64), (256, 128), (128, 64)]):
        p, st = W.config0(seed=q, ni=ni, nj=nj, h=2e-4, V_bias=10.0 + 15 * q,
                          mode=abi.CATHODE_MODE_MIXED)
        params.append(p)
        states.append(st)
    os.environ.pop("CATHODE_CUT", None)
    seq = []
    for p, st in zip(params, states):
        s = W.copy_state(st)
        h = G.ndd_handle(hiplib, p, s)
        rep = abi.NddReport()
        abi.check(hiplib.cathode_ndd_step(h, 3, ctypes.byref(rep)), hiplib)
        seq.append(G.ndd_state(hiplib, h, s))
        hiplib.cathode_ndd_destroy(h)
    hs = [G.ndd_handle(hiplib, p, W.copy_state(st)) for p, st in zip(params, states)]
    try:
        reps = G.ndd_batch(hiplib, hs, 3, 3)
        for q, h in enumerate(hs):
            got = G.ndd_state(hiplib, h, states[q])
            for k in ("phi", "rhoe", "rhoi"):
                assert (got[k] == seq[q][k]).all(), (q, k)
            assert any(c >= 0 for c in reps[q].cut_by_solve), list(reps[q].cut_by_solve)
    finally:
        for h in hs:
            hiplib.cathode_ndd_destroy(h)
