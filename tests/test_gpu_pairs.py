"""Two fine levels per launch (k_vpair_down / k_vpair_up): a workgroup runs a
tile of the second level and the tile of the first level it aggregates, so
the first level's halo grows to cover the second level's.  Every value the
pair kernels write is computed by the same operations in the same order as
the single-level kernels, so fields and iteration counts must be identical
with CATHODE_PAIR=1 (default) and CATHODE_PAIR=0, step after step, for every
pair tile configuration (CATHODE_PAIR_TILE)."""
import ctypes
import os

import numpy as np
import pytest

import gpu_api as G
from oracle_lib import abi, workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("phi", "rhoe", "rhoi", "Te", "rhoeps")


def _run(lib, p, st, nsteps, pair, tile=0):
    old = {k: os.environ.get(k) for k in ("CATHODE_PAIR", "CATHODE_PAIR_TILE")}
    os.environ["CATHODE_PAIR"] = "1" if pair else "0"
    os.environ["CATHODE_PAIR_TILE"] = str(tile)
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
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    return out, its


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])  # the pair tile configurations (VP_CFGS)
@pytest.mark.parametrize("mode", [abi.CATHODE_MODE_FP64, abi.CATHODE_MODE_MIXED])
@pytest.mark.parametrize("shape", [(256, 128), (512, 256), (250, 130)])
def test_pair_kernels_bitwise(hiplib, mode, shape, tile):
    ni, nj = shape
    print(f"pairs (tile {tile}) vs single levels: config1 {ni}x{nj} mode={mode}, 3 steps", flush=True)
    p, st = W.config1(seed=3, ni=ni, nj=nj, mode=mode)
    single, its_s = _run(hiplib, p, st, 3, pair=False)
    p, st = W.config1(seed=3, ni=ni, nj=nj, mode=mode)
    paired, its_p = _run(hiplib, p, st, 3, pair=True, tile=tile)
    assert its_p == its_s
    for k, (a, b) in enumerate(zip(single, paired)):
        for name in FIELDS:
            assert np.array_equal(a[name], b[name]), (k, name)
