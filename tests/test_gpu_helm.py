"""GPU parity of the complex E_theta path against the oracle.

cathode_helm5_solve is the complex 5-point solve (complex128 multigrid, or
complex64 V-cycles with complex128 refinement); cathode_etheta_solve adds the
device assembly (emfieldModule.hpp:2414-2454, folds :2713-2773), ghosts
(:3213-3264) and Pabs (:7733-7743).  Bars: relative residual <= tol, E and
Pabs within 1e-8 relative L2 of the oracle's exact banded-LU solution on
grids the oracle factors in seconds (cond(A) * 1e-10 with cond(A) of order
1e2 here), and at the full 1024 x 512 workload: the oracle-assembled
operator's residual of the device E <= 1e-10, the oracle's post-processing of
the device E reproduces its Pabs, and complex64+refinement agrees with
complex128 within 1e-8 relative L2 in Pabs.
"""
import numpy as np
import pytest

import gpu_api as G
import oracle_lib as O

W = O.workloads
pytestmark = pytest.mark.gpu


def _random_helm(rng, ni, nj, shift):
    shp = (ni + 2, nj + 2)
    aE, aW, aN, aS = [rng.uniform(0.5, 2.0, shp) + 0j for _ in range(4)]
    aP = aE + aW + aN + aS + shift
    aP[1, :] += aW[1, :]
    aW[1, :] = 0
    b = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    return aP, aE, aW, aN, aS, b


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("shift", [0.05 + 0.3j, -0.02 + 0.1j])
def test_helm5_random_matches_direct(hiplib, mode, shift):
    rng = np.random.default_rng(11)
    ni, nj = 48, 40
    aP, aE, aW, aN, aS, b = _random_helm(rng, ni, nj, shift)
    box = [1, ni, 1, nj]
    print(f"helm5 random {ni}x{nj} shift={shift} mode={mode}", flush=True)
    x, st = G.helm5(hiplib, aP, aE, aW, aN, aS, b, box, mode=mode)
    xd = O.helm5_direct(aP, aE, aW, aN, aS, b, box)
    rr = O.rel_residual_box(aP, aE, aW, aN, aS, b, box, x)
    print(f"  its={st.iterations} rel={st.rel_residual:.3e} recomputed={rr:.3e}", flush=True)
    assert st.converged == 1 and rr <= 1.5e-10
    assert O.rel_l2(x[1:-1, 1:-1], xd[1:-1, 1:-1]) < 1e-8


@pytest.mark.parametrize("mode", [0, 1])
def test_helm5_box_union(hiplib, mode):
    from test_gpu_solvers import cathode_blocks
    rng = np.random.default_rng(5)
    ni, nj = 60, 30
    aP, aE, aW, aN, aS, b = _random_helm(rng, ni, nj, 0.1 + 0.2j)
    boxes = cathode_blocks(ni, nj)
    x0 = np.full(aP.shape, 7.0 + 7.0j)
    x, st = G.helm5(hiplib, aP, aE, aW, aN, aS, b, boxes, mode=mode, x=x0.copy())
    xd = O.helm5_direct(aP, aE, aW, aN, aS, b, boxes, x=x0.copy())
    assert st.converged == 1
    np.testing.assert_array_equal(x == x0, xd == x0)
    m = xd != x0
    assert O.rel_l2(x[m], xd[m]) < 1e-8


def _small_etheta(pec_z=1, pec_rmax=1, rmin=0.0, mode=0):
    p, ne, J = W.config2(seed=4, ni=128, nj=96, mode=mode)
    p.pec_z, p.pec_rmax, p.rmin = pec_z, pec_rmax, rmin
    return p, ne, J


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("bc", [(1, 1, 0.0), (0, 0, 0.0), (1, 0, 1e-3)])
def test_etheta_small_matches_oracle(hiplib, mode, bc):
    p, ne, J = _small_etheta(*bc, mode=mode)
    print(f"etheta 128x96 bc={bc} mode={mode}", flush=True)
    E, Pabs, Pt, st = G.etheta(hiplib, p, ne, J)
    Ed, Pd, Ptd = O.helm_solve_direct(p, ne, J)
    print(f"  its={st.iterations} rel={st.rel_residual:.3e} P={Pt:.6e} oracle {Ptd:.6e}", flush=True)
    assert st.converged == 1
    s = (slice(1, -1), slice(1, -1))
    assert O.rel_l2(E[s], Ed[s]) < 1e-8
    # ghost rows / columns written by the device post step
    g = np.zeros(E.shape, dtype=bool)
    g[0, 1:-1] = g[-1, 1:-1] = True
    g[1:-1, -1] = True
    if p.rmin > 0:
        g[1:-1, 0] = True
    assert O.rel_l2(E[g], Ed[g]) < 1e-8
    assert O.rel_l2(Pabs[s], Pd[s]) < 1e-8
    assert abs(Pt - Ptd) <= 1e-8 * abs(Ptd)


@pytest.mark.parametrize("mode", [0, 1])
def test_etheta_config2_full(hiplib, mode):
    p, ne, J = W.config2(mode=mode)
    print(f"etheta config2 1024x512 mode={mode}", flush=True)
    E, Pabs, Pt, st = G.etheta(hiplib, p, ne, J)
    print(f"  its={st.iterations} rel={st.rel_residual:.3e} P={Pt:.6e} t={st.seconds:.3f}s", flush=True)
    assert st.converged == 1 and st.rel_residual <= p.tol
    aP, aE, aW, aN, aS, b = O.helm_assemble(p, ne, J)
    box = [1, p.ni, 1, p.nj]
    rr = O.rel_residual_box(aP, aE, aW, aN, aS, b, box, E)
    print(f"  oracle-operator residual {rr:.3e}", flush=True)
    assert rr <= 1.5e-10
    E2 = E.copy()
    P2, Pt2 = O.helm_post(p, ne, E2)
    s = (slice(1, -1), slice(1, -1))
    assert O.rel_l2(Pabs[s], P2[s]) < 1e-12
    assert abs(Pt - Pt2) <= 1e-12 * abs(Pt2)
    assert Pt > 0


def test_etheta_mixed_matches_fp64_full(hiplib):
    p, ne, J = W.config2(mode=0)
    _, P64, Pt64, _ = G.etheta(hiplib, p, ne, J)
    p.mode = 1
    _, P32, Pt32, st = G.etheta(hiplib, p, ne, J)
    s = (slice(1, -1), slice(1, -1))
    err = O.rel_l2(P32[s], P64[s])
    print(f"mixed vs fp64 Pabs rel L2 {err:.3e}, its={st.iterations}", flush=True)
    assert err < 1e-8
