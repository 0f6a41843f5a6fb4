"""Writes the golden vectors of tests/golden/*.npz from the oracle.

The reference ships no fixture for this path (its tests pin neither
solver_SMG nor the NDD assembly), so the vectors are the oracle's answers on
small seeded inputs, frozen so that later changes to the oracle, the
workload builders or the HIP path are caught against a fixed record:

  ndd_cfg0_32x24.npz      config-0 state (ni=32, nj=24, h=3e-4, seed 0) and the
                          state after one implicit step (exact banded-LU solves)
  ndd_cfg1_32x16.npz      config-1 state with electron energy (ni=32, nj=16,
                          h=2e-4, seed 1) and after one step
  struct5_poisson_32x24.npz  the first Poisson correction system of the
                          config-0 state and its exact solution
  etheta_24x16.npz        E_theta inputs (config-2 physics, ni=24, nj=16,
                          h=2e-4, seed 3), E with ghosts, Pabs, P_total

Run from the repository root:  python3 tests/golden/make_golden.py
Arrays are plain float64 / complex128 (numpy .npz, no pickles).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

W = O.workloads
CFG0 = dict(seed=0, ni=32, nj=24, h=3e-4)
CFG1 = dict(seed=1, ni=32, nj=16, h=2e-4)
ETH = dict(seed=3, ni=24, nj=16, h=2e-4)


def ndd_case(builder, kw):
    p, st = builder(**kw)
    out = W.copy_state(st)
    rep = O.ndd_step(p, out, 1, solver=0)
    d = {"in_" + k: v for k, v in st.items()}
    d.update({"out_" + k: v for k, v in out.items()})
    d["I_anode"] = np.array(rep.I_anode)
    return d


def etheta_case():
    p, ne, J = W.config2(seed=ETH["seed"], ni=ETH["ni"], nj=ETH["nj"], h=ETH["h"])
    E, Pabs, Pt = O.helm_solve_direct(p, ne, J)
    return {"ne": ne, "J": J, "E": E, "Pabs": Pabs, "P_total": np.array(Pt)}


def build_all():
    p, st = W.config0(**CFG0)
    aP, aE, aW, aN, aS, b = O.ndd_assemble(p, st, 0)[:6]
    x = O.struct5_direct(aP, aE, aW, aN, aS, b, [1, p.ni, 1, p.nj])
    return {
        "ndd_cfg0_32x24": ndd_case(W.config0, CFG0),
        "ndd_cfg1_32x16": ndd_case(W.config1, CFG1),
        "struct5_poisson_32x24": dict(aP=aP, aE=aE, aW=aW, aN=aN, aS=aS, b=b, x=x),
        "etheta_24x16": etheta_case(),
    }


if __name__ == "__main__":
    for name, arrays in build_all().items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(name, sum(a.nbytes for a in arrays.values()), "bytes")
