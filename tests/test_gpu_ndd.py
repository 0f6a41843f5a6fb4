"""GPU parity of the device-resident implicit NDD step against the oracle.

Bar: after each implicit step solved to relative residual
1e-10, every field agrees with the FP64 oracle to <= 1e-9 relative L2.  The
oracle solves its systems exactly (banded LU) on the small grids and with its
own multigrid copy to the same tolerance on the full 512x256 grid.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

abi = O.abi
W = O.workloads
pytestmark = pytest.mark.gpu

FIELDS = ["phi", "rhoe", "rhoi", "Ex", "Er", "rhoUex", "rhoUer", "rhoUix", "rhoUir", "nUex", "nUer"]
TOL = 1e-9


def gpu_steps(lib, p, st, nsteps):
    h = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), lib)
    try:
        f = abi.fields_struct(st)
        abi.check(lib.cathode_ndd_set_state(h, ctypes.byref(f)), lib)
        rep = abi.NddReport()
        abi.check(lib.cathode_ndd_step(h, nsteps, ctypes.byref(rep)), lib)
        abi.check(lib.cathode_ndd_get_state(h, ctypes.byref(f)), lib)
    finally:
        lib.cathode_ndd_destroy(h)
    return rep


def compare(gpu, ref, fields, tol=TOL):
    worst = {}
    for k in fields:
        e = O.rel_l2(gpu[k], ref[k])
        worst[k] = e
    bad = {k: v for k, v in worst.items() if not (v <= tol)}
    assert not bad, f"fields beyond {tol}: {bad} (all: {worst})"
    return worst


@pytest.mark.parametrize("mode,learned", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("nsteps", [1, 3])
def test_config0_steps(hiplib, mode, learned, nsteps):
    """learned = 1 without a network: every solve starts from the least-squares
    combination of its last solutions and runs its expected iterations
    without checks; the converged fields must not change."""
    p, st = W.config0(mode=mode)
    p.learned = learned
    print(f"ndd config0 96x64 mode={mode} learned={learned} steps={nsteps}", flush=True)
    ref = W.copy_state(st)
    rr = O.ndd_step(p, ref, nsteps, solver=0)
    rep = gpu_steps(hiplib, p, st, nsteps)
    assert rep.steps == nsteps and rep.solves == rr.solves == 6 * nsteps
    assert rep.nonconverged == 0 and rep.max_rel_residual <= 1e-10
    compare(st, ref, FIELDS)
    assert abs(rep.I_anode - rr.I_anode) <= 1e-9 * abs(rr.I_anode)


@pytest.mark.parametrize("mode,learned", [(0, 0), (1, 0), (1, 1)])
def test_config1_small_energy(hiplib, mode, learned):
    p, st = W.config1(ni=128, nj=64, h=2e-4, mode=mode)
    p.learned = learned
    nsteps = 4 if learned else 2  # warm starts need a history
    print(f"ndd config1 128x64 energy mode={mode} learned={learned} steps={nsteps}", flush=True)
    ref = W.copy_state(st)
    O.ndd_step(p, ref, nsteps, solver=0)
    rep = gpu_steps(hiplib, p, st, nsteps)
    assert rep.nonconverged == 0
    compare(st, ref, FIELDS + ["rhoeps", "Te", "Gx", "Gr"])


@pytest.mark.parametrize("shape", [(75, 53), (33, 130), (129, 67)])
@pytest.mark.parametrize("mode,learned", [(0, 0), (1, 0), (1, 1)])
def test_odd_grids_match_oracle(hiplib, shape, mode, learned):
    """Odd interior sizes (and a grid taller than it is long): partial tiles
    on every level, levels that coarsen into a half-filled last row or
    column, and a hierarchy whose directions stop coarsening at different
    levels; three steps of the full equation set against the oracle's exact
    solves."""
    ni, nj = shape
    p, st = W.config1(ni=ni, nj=nj, h=2.56e-2 / max(ni, 2 * nj), mode=mode, seed=11)
    p.learned = learned
    nsteps = 3
    print(f"ndd odd grid {ni}x{nj} mode={mode} learned={learned} steps={nsteps}", flush=True)
    ref = W.copy_state(st)
    O.ndd_step(p, ref, nsteps, solver=0)
    rep = gpu_steps(hiplib, p, st, nsteps)
    assert rep.nonconverged == 0 and rep.max_rel_residual <= 1e-10
    compare(st, ref, FIELDS + ["rhoeps", "Te", "Gx", "Gr"])


def test_sources_and_see(hiplib):
    """rate_ionize / Pabs / rate_eloss and an EISEE table exercise every RHS term."""
    p, st = W.config1(ni=64, nj=32, h=4e-4)
    rng = np.random.default_rng(5)
    shp = st["phi"].shape
    st["rate_ionize"] = rng.uniform(0, 1e24, shp)
    st["Pabs"] = rng.uniform(0, 1e5, shp)
    st["rate_eloss"] = rng.uniform(0, 1e-15, shp)
    p.coefIISEE_ts = 0.08
    p.n_see = 201
    p.dTe_SEE = 1160.45  # 0.1 eV in K
    x = np.arange(201) * 0.1
    for k in range(201):
        p.see_eb[k] = 0.02 * x[k] / (1 + 0.05 * x[k])
        p.see_rd[k] = 0.01 * np.sqrt(x[k])
        p.see_ts[k] = 0.5 * (1 - np.exp(-x[k] / 20))
    print("ndd sources+see 64x32", flush=True)
    ref = W.copy_state(st)
    O.ndd_step(p, ref, 2, solver=0)
    gpu_steps(hiplib, p, st, 2)
    compare(st, ref, FIELDS + ["rhoeps", "Te", "Gx", "Gr"])


@pytest.mark.parametrize("mode", [0, 1])
def test_config1_full_size(hiplib, mode):
    """512x256 (the metric grid), two steps; the oracle uses its multigrid."""
    p, st = W.config1(mode=mode)
    print(f"ndd config1 512x256 mode={mode}", flush=True)
    ref = W.copy_state(st)
    O.lib().orc_mg_set_threads(8)
    rr = O.ndd_step(p, ref, 2, solver=1)
    rep = gpu_steps(hiplib, p, st, 2)
    assert rep.nonconverged == 0 and rr.nonconverged == 0
    compare(st, ref, FIELDS + ["rhoeps", "Te", "Gx", "Gr"])


@pytest.mark.parametrize("mode,learned", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("cnv", [1e-21, 1e-9, 1e-5])
def test_pc_loop_depth_and_early_stop(hiplib, mode, learned, cnv):
    """Four PC passes per routine with a stopping threshold: the never-met
    threshold runs every pass (the start-of-pass copy sets alternate three
    times); the others stop the loop after a pass whose calcRes meets it
    (the speculative solve aborts and the predicted fluxes are restored).
    learned = 1 (warm starts, expected iterations unchecked) over 4 steps, so
    the aborted solves also start from a history."""
    p, st = W.config0(mode=mode)
    p.learned = learned
    p.maxITR_HES_phi = p.maxITR_HES_rhoe = p.maxITR_HES_rhoeps = p.maxITR_HES_rhoi = 4
    p.error_cnv_HES_phi = p.error_cnv_HES_rhoe = p.error_cnv_HES_rhoeps = p.error_cnv_HES_rhoi = cnv
    nsteps = 4 if learned else 2
    print(f"ndd config0 96x64 pc depth 4 cnv={cnv} mode={mode} learned={learned} steps={nsteps}", flush=True)
    ref = W.copy_state(st)
    O.ndd_step(p, ref, nsteps, solver=0)
    rep = gpu_steps(hiplib, p, st, nsteps)
    assert rep.nonconverged == 0
    compare(st, ref, FIELDS)


def test_cfl_report_matches_oracle(hiplib):
    """check_CFL before the last step (fixed dt): CFL number, dt_min and the
    maxima it is built from, against the oracle's (fluidModule.hpp:57-115)."""
    p, st = W.config0()
    print("ndd config0 96x64 check_CFL report, 3 steps", flush=True)
    ref = W.copy_state(st)
    rr = O.ndd_step(p, ref, 3, solver=0)
    rep = gpu_steps(hiplib, p, st, 3)
    assert rep.dt == p.dt and rr.dt == p.dt
    for k in ("CFL", "dt_min", "rhoe_max", "Te_max", "ue_max"):
        g, o = getattr(rep, k), getattr(rr, k)
        assert abs(g - o) <= 1e-9 * abs(o), (k, g, o)
    assert rep.ue_max > 1e5  # the electron fluxes of the earlier steps enter


@pytest.mark.parametrize("mode", [0, 1])
def test_adaptive_dt_steps(hiplib, mode):
    """flag_adp_dt = 1: every step takes dt = CFL eps0 / (e max n_e mu_e)."""
    p, st = W.config0(mode=mode)
    p.flag_adp_dt, p.CFL = 1, 0.3
    print(f"ndd config0 96x64 adaptive dt CFL 0.3 mode={mode}, 3 steps", flush=True)
    ref = W.copy_state(st)
    rr = O.ndd_step(p, ref, 3, solver=0)
    rep = gpu_steps(hiplib, p, st, 3)
    assert rep.nonconverged == 0
    assert abs(rep.dt - rr.dt) <= 1e-12 * rr.dt and rep.CFL == 0.3
    assert abs(rr.dt / (0.3 * W.EPS0 / (W.E0 * rr.rhoe_max * p.mu_e)) - 1) < 1e-14
    compare(st, ref, FIELDS)
