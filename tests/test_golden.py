"""The committed golden vectors (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle) are reproduced by today's oracle
and workload builders; a drift in either shows up here on the CPU."""
import os

import numpy as np
import pytest

import oracle_lib as O
from golden import make_golden as MG

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fresh():
    return MG.build_all()


@pytest.mark.parametrize("name", ["ndd_cfg0_32x24", "ndd_cfg1_32x16", "struct5_poisson_32x24", "etheta_24x16"])
def test_golden_reproduced(fresh, name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as g:
        want = {k: g[k] for k in g.files}
    got = fresh[name]
    assert sorted(want) == sorted(got)
    for k in want:
        if k.startswith("in_") or k in ("ne", "J") or name.startswith("struct5") and k != "x":
            np.testing.assert_array_equal(got[k], want[k], err_msg=f"{name}:{k} (inputs are exact)")
        else:
            err = O.rel_l2(np.asarray(got[k]), np.asarray(want[k]))
            assert err <= 1e-12, f"{name}:{k} rel L2 {err:.3e}"
