"""Graph replay of the fused BiCGStab iterations: every iteration after a
restart is captured once (per mode, iteration parity and buffer order) and
replayed with hipGraphLaunch.  The replay launches the same kernels with the
same arguments as the eager path, so fields and iteration counts must be
identical, step after step, with the solver's own eager launches (restart,
setup, assembly) and host synchronisations between the replays.
CATHODE_GRAPHS=1 switches the capture on (it is off by default)."""
import ctypes
import os

import numpy as np
import pytest

import gpu_api as G
from oracle_lib import abi, workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("phi", "rhoe", "rhoi", "Te", "rhoeps")


def _run(lib, p, st, nsteps, graphs):
    """fields after each step and the per-step iteration totals"""
    old = os.environ.get("CATHODE_GRAPHS")
    os.environ["CATHODE_GRAPHS"] = "1" if graphs else "0"
    try:
        h = G.ndd_handle(lib, p, st)
        out, its = [], []
        try:
            for _ in range(nsteps):
                rep = abi.NddReport()
                abi.check(lib.cathode_ndd_step(h, 1, ctypes.byref(rep)), lib)
                its.append((rep.total_iterations, rep.solves, rep.nonconverged))
                out.append(G.ndd_state(lib, h, st))
        finally:
            lib.cathode_ndd_destroy(h)
    finally:
        if old is None:
            del os.environ["CATHODE_GRAPHS"]
        else:
            os.environ["CATHODE_GRAPHS"] = old
    return out, its


@pytest.mark.parametrize("mode", [abi.CATHODE_MODE_FP64, abi.CATHODE_MODE_MIXED])
@pytest.mark.parametrize("shape", [(256, 128), (512, 256)])
def test_graph_replay_identical(hiplib, mode, shape):
    ni, nj = shape
    p, st = W.config1(seed=11, ni=ni, nj=nj, mode=mode)
    print(f"graphs vs eager: config1 {ni}x{nj} mode={mode}, 4 steps", flush=True)
    eager, its_e = _run(hiplib, p, st, 4, graphs=False)
    p, st = W.config1(seed=11, ni=ni, nj=nj, mode=mode)
    graph, its_g = _run(hiplib, p, st, 4, graphs=True)
    assert its_g == its_e
    assert sum(t for t, _, _ in its_g) > 2 * sum(s for _, s, _ in its_g)  # replays did happen
    for k, (a, b) in enumerate(zip(eager, graph)):
        for name in FIELDS:
            assert np.array_equal(a[name], b[name]), (k, name)


def test_graph_replay_batched(hiplib):
    """instances stepped side by side on worker threads, each capturing on
    its own stream, match the eager batch exactly"""
    inst = W.config3(seed=5, count=5, sizes=[(128, 64), (256, 128)])
    res = {}
    for graphs in (False, True):
        os.environ["CATHODE_GRAPHS"] = "1" if graphs else "0"
        try:
            hs = [G.ndd_handle(hiplib, p, st) for p, st, _ in inst]
            try:
                reps = G.ndd_batch(hiplib, hs, 3, 3)
                res[graphs] = ([G.ndd_state(hiplib, h, st) for h, (_, st, _) in zip(hs, inst)],
                               [(r.total_iterations, r.nonconverged) for r in reps])
            finally:
                for h in hs:
                    hiplib.cathode_ndd_destroy(h)
        finally:
            del os.environ["CATHODE_GRAPHS"]
    assert res[True][1] == res[False][1]
    for k, (a, b) in enumerate(zip(res[False][0], res[True][0])):
        for name in FIELDS:
            d = np.linalg.norm(a[name] - b[name]) / max(np.linalg.norm(a[name]), 1e-300)
            assert np.array_equal(a[name], b[name]), (k, name, d)
