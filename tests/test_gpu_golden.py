"""GPU path against the committed golden vectors (no oracle call at run
time): one implicit NDD step from the stored state, the stored Poisson
system and the stored E_theta case, in FP64 and mixed mode, within 1e-9
relative L2 (fields) / 1e-8 (E_theta, cond(A) * 1e-10)."""
import ctypes
import os

import numpy as np
import pytest

import gpu_api as G
from oracle_lib import abi, rel_l2, workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as g:
        return {k: g[k] for k in g.files}


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name,builder,kw", [
    ("ndd_cfg0_32x24", W.config0, dict(seed=0, ni=32, nj=24, h=3e-4)),
    ("ndd_cfg1_32x16", W.config1, dict(seed=1, ni=32, nj=16, h=2e-4))])
def test_ndd_step_matches_golden(hiplib, mode, name, builder, kw):
    g = _load(name)
    p, _ = builder(mode=mode, **kw)
    st = {k[3:]: np.ascontiguousarray(v) for k, v in g.items() if k.startswith("in_")}
    print(f"golden {name} mode={mode}", flush=True)
    h = ctypes.c_void_p()
    abi.check(hiplib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), hiplib)
    try:
        abi.check(hiplib.cathode_ndd_set_state(h, ctypes.byref(abi.fields_struct(st))), hiplib)
        rep = abi.NddReport()
        abi.check(hiplib.cathode_ndd_step(h, 1, ctypes.byref(rep)), hiplib)
        out = {k: np.zeros_like(v) for k, v in st.items()}
        abi.check(hiplib.cathode_ndd_get_state(h, ctypes.byref(abi.fields_struct(out))), hiplib)
    finally:
        hiplib.cathode_ndd_destroy(h)
    assert rep.nonconverged == 0
    for k in ("phi", "rhoe", "rhoi", "Te", "rhoeps", "rhoUex", "rhoUer", "rhoUix", "rhoUir", "Ex", "Er"):
        want = g["out_" + k]
        if not np.any(want):
            continue
        err = rel_l2(out[k], want)
        assert err <= 1e-9, f"{k}: {err:.3e}"
    assert abs(rep.I_anode - float(g["I_anode"])) <= 1e-9 * abs(float(g["I_anode"])) + 1e-30


@pytest.mark.parametrize("mode", [0, 1])
def test_poisson_system_matches_golden(hiplib, mode):
    g = _load("struct5_poisson_32x24")
    ni, nj = g["b"].shape[0] - 2, g["b"].shape[1] - 2
    x, st = G.struct5(hiplib, g["aP"], g["aE"], g["aW"], g["aN"], g["aS"], g["b"], [1, ni, 1, nj], mode=mode)
    assert st.converged == 1
    assert rel_l2(x[1:-1, 1:-1], g["x"][1:-1, 1:-1]) < 5e-9


@pytest.mark.parametrize("mode", [0, 1])
def test_etheta_matches_golden(hiplib, mode):
    g = _load("etheta_24x16")
    p, _, _ = W.config2(seed=3, ni=24, nj=16, h=2e-4, mode=mode)
    E, Pabs, Pt, st = G.etheta(hiplib, p, g["ne"], g["J"])
    assert st.converged == 1
    s = (slice(1, -1), slice(1, -1))
    assert rel_l2(E[s], g["E"][s]) < 1e-8
    assert rel_l2(Pabs[s], g["Pabs"][s]) < 1e-8
    assert abs(Pt - float(g["P_total"])) <= 1e-8 * abs(float(g["P_total"]))
