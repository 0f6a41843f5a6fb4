"""cathode_struct5_solve_dev_noalloc (the struct5 device solve that allocates
nothing: its multigrid lives in a static device array) against the entry
point it mirrors, cathode_struct5_solve_dev, through the C ABI.

Two calls in a row, the second on a smaller grid, so the second runs in a
layout of the same scratch that the first left filled.  Both entry points run
the same kernels in the same order when the tile shape is fixed
(CATHODE_TILE=0, the no-allocation path's default), so x must agree bit for
bit; independently, x must satisfy the system to the relative residual 1e-10
the solve stops at (recomputed by the oracle in FP64).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O
from oracle_lib import abi
from test_gpu_solvers import cathode_blocks, random_mmatrix

W = O.workloads
pytestmark = pytest.mark.gpu


class DevArrays:
    """Device copies of host float64 arrays (cathode_dev_alloc / _copy)."""

    def __init__(self, lib):
        self.lib = lib
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        abi.check(self.lib.cathode_dev_alloc(a.nbytes, ctypes.byref(p)), self.lib)
        self.ptrs.append(p)
        abi.check(self.lib.cathode_dev_copy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 0), self.lib)
        return p

    def get(self, p, like):
        out = np.empty_like(like, dtype=np.float64)
        abi.check(self.lib.cathode_dev_copy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 1), self.lib)
        return out

    def free(self):
        for p in self.ptrs:
            self.lib.cathode_dev_free(p)
        self.ptrs = []


def _solve(lib, fn, system, boxes, mode, x0):
    """x (host) after fn on device copies of the system; x0 fills x first
    (off-box cells must keep it)."""
    aP = system[0]
    ni2, ld = aP.shape
    d = DevArrays(lib)
    try:
        ptrs = [ctypes.cast(d.put(a), abi.c_double_p) for a in system]
        xp = d.put(x0)
        bx = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1))
        st = abi.Status()
        rc = fn(*ptrs, bx.ctypes.data_as(abi.c_int_p), len(bx) // 4, ni2, ld, ld, 30, 1e-10, mode,
                ctypes.cast(xp, abi.c_double_p), None, ctypes.byref(st))
        abi.check(rc, lib)
        return d.get(xp, x0), st
    finally:
        d.free()


@pytest.fixture
def fixed_tiles():
    old = os.environ.get("CATHODE_TILE")
    os.environ["CATHODE_TILE"] = "0"
    yield
    if old is None:
        os.environ.pop("CATHODE_TILE", None)
    else:
        os.environ["CATHODE_TILE"] = old


def _cases():
    rng = np.random.default_rng(11)
    # 1) a 96 x 64 Poisson correction system of the config-0 state, one box
    p, st = W.config0()
    sys1 = tuple(O.ndd_assemble(p, st, 0)[:6])
    # 2) a smaller random M-matrix on the five-box cathode union (60 x 30)
    aP, aE, aW, aN, aS = random_mmatrix(rng, 60, 30)
    sys2 = (aP, aE, aW, aN, aS, rng.standard_normal(aP.shape))
    return [(sys1, [[1, p.ni, 1, p.nj]]), (sys2, cathode_blocks(60, 30))]


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_matches_dev_solve_twice(hiplib, fixed_tiles, mode):
    for k, (system, boxes) in enumerate(_cases()):
        x0 = np.full(system[0].shape, 123.0)
        print(f"noalloc call {k + 1}: {system[0].shape[0] - 2}x{system[0].shape[1] - 2}, "
              f"{len(boxes)} boxes, mode {mode}", flush=True)
        xa, sa = _solve(hiplib, hiplib.cathode_struct5_solve_dev_noalloc, system, boxes, mode, x0)
        xr, sr = _solve(hiplib, hiplib.cathode_struct5_solve_dev, system, boxes, mode, x0)
        assert sa.converged == 1 and sr.converged == 1
        assert sa.iterations == sr.iterations
        np.testing.assert_array_equal(xa, xr)
        assert O.rel_residual(*system, boxes, xa) <= 1e-10
        np.testing.assert_array_equal(xa == 123.0, xr == 123.0)


def test_noalloc_rejects_grid_beyond_scratch(hiplib):
    ni, nj = 512, 256  # needs about 31 MiB, the scratch holds 16 MiB
    shp = (ni + 2, nj + 2)
    rng = np.random.default_rng(3)
    aP, aE, aW, aN, aS = random_mmatrix(rng, ni, nj)
    system = (aP, aE, aW, aN, aS, rng.standard_normal(shp))
    with pytest.raises(RuntimeError, match="scratch"):
        _solve(hiplib, hiplib.cathode_struct5_solve_dev_noalloc, system, [[1, ni, 1, nj]], 0, np.zeros(shp))
