"""V-cycle cuts in the oracle's multigrid (the host copy of the device's).

A mass-dominated operator stops its V-cycles at level 0 (row excess fraction
delta_0 = min (aP - sum a_nb) / aP >= 0.9) or at the device's second
candidate level (delta >= 0.35 there; level 1 at 256 x 128, where the
one-workgroup tail starts at level 2).  The cut changes the preconditioner
only: the solve still reaches its tolerance, and the answer is the direct
solve's to cond(A) * tol.  An operator with no mass (Poisson) is never cut,
so its solve is bitwise the uncut one.
"""
import numpy as np
import pytest

import oracle_lib as O


def _operator(ni, nj, mass, seed=0):
    """5-point diffusion with unit couplings (Dirichlet walls folded into aP)
    plus mass * 4 on the diagonal: delta_0 = mass / (1 + mass)."""
    rng = np.random.default_rng(seed)
    shape = (ni + 2, nj + 2)
    aE, aW, aN, aS = (np.ones(shape) for _ in range(4))
    aP = 4.0 + 4.0 * mass + np.zeros(shape)
    aW[1, :] = 0.0
    aP[1, :] += 1.0  # phi = 0 ghost mirror on the low wall
    aE[ni, :] = 0.0
    aS[:, 1] = 0.0
    aN[:, nj] = 0.0
    b = np.zeros(shape)
    b[1:-1, 1:-1] = rng.uniform(-1.0, 1.0, size=(ni, nj))
    return aP, aE, aW, aN, aS, b


def _solve(coefs, box, cut):
    O.lib().orc_mg_set_cut(1 if cut else 0)
    try:
        x, it, rel = O.mg_solve(*coefs, box)
        return x, it, rel, O.lib().orc_mg_last_cut()
    finally:
        O.lib().orc_mg_set_cut(1)


@pytest.mark.parametrize("ni,nj,mass,want", [
    (96, 64, 20.0, 0),     # delta_0 = 20/21 >= 0.9: level 0
    (256, 128, 1.0, 1),    # delta_0 = 0.5, delta_1 = 4/5 >= 0.35: level 1
    (256, 128, 0.05, -1),  # delta_1 = 0.2 / 1.2 < 0.35: every level
])
def test_cut_level_and_answer(ni, nj, mass, want):
    coefs = _operator(ni, nj, mass)
    box = [[1, ni, 1, nj]]
    x_cut, it_cut, rel_cut, cut = _solve(coefs, box, True)
    x_full, it_full, rel_full, none = _solve(coefs, box, False)
    assert cut == want and none == -1
    assert rel_cut <= 1e-10 and rel_full <= 1e-10
    assert O.rel_residual(*coefs, box, x_cut) <= 1e-10
    x_ref = O.struct5_direct(*coefs, box)
    err = np.linalg.norm(x_cut - x_ref) / np.linalg.norm(x_ref)
    assert err <= 1e-9, err
    # smoothing alone converges on these operators about as fast as the
    # whole V-cycle
    assert it_cut <= it_full + 1, (it_cut, it_full)


def test_no_mass_is_never_cut():
    coefs = _operator(128, 64, 0.0)
    box = [[1, 128, 1, 64]]
    x_cut, it_cut, _, cut = _solve(coefs, box, True)
    x_full, it_full, _, _ = _solve(coefs, box, False)
    assert cut == -1
    assert it_cut == it_full and np.array_equal(x_cut, x_full)

