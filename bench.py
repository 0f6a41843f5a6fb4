#!/usr/bin/env python3
"""Benchmark: implicit NDD steps/s at 512x256 (every linear solve to relative
residual 1e-10) on one MI355X, mixed precision vs plain FP64.

Workload (one "step" = one implicit time step of the whole equation set):
512(z) x 256(r) cells of 0.05 mm, ions + Poisson + electrons + electron
energy, each a two-iteration predictor-corrector loop with one 5-point solve
per iteration (8 solves per step), densities = 1e14 floor + four seeded
Gaussians, Te = 2..10 eV smooth seeded field, bias 30 V, dt = 1e-10 s.
Synthetic state generated on the host (seed 1) and resident on the GPU before
the timed region.

Prints one JSON line.  `value` is the steps/s of the learned + mixed path
(the learned initial guess of the first Poisson correction solve, FP32
V-cycles with FP64 refinement); the plain-FP64 and mixed GPU paths are
reported beside it.
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = "02_cathode-08_modulerev_amd"
abi = importlib.import_module(PKG + ".abi")
W = importlib.import_module(PKG + ".workloads")
L = importlib.import_module(PKG + ".learned")

METRIC = "NDD steps/s at 512x256 (resid 1e-10): learned+mixed vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0


def rel_l2(a, b):
    import numpy as np
    d = float(np.linalg.norm((a - b).ravel()))
    n = float(np.linalg.norm(b.ravel()))
    return d / n if n > 0 else d


REPORT_FIELDS = {"phi": "phi", "n_e": "rhoe", "n_i": "rhoi", "n_eps": "rhoeps", "Te": "Te"}


def sig(v):
    """floats to 6 significant digits (small residuals and errors stay readable)"""
    return float(f"{v:.6g}") if isinstance(v, float) else v


def run_path(lib, mode, steps, warmup, seed, net=None, warm=False, grid=(512, 256)):
    ni, nj = grid
    # the same 25.6 mm x 12.8 mm domain at a finer or coarser spacing
    p, st = W.config1(seed=seed, ni=ni, nj=nj, h=5e-5 * 512 / ni, mode=mode)
    # learned mode: previous-solution warm starts, plus the network's guess
    # for the first Poisson correction solve when a network is attached
    p.learned = 1 if (net is not None or warm) else 0
    h = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), lib)
    if net is not None:
        abi.check(lib.cathode_ndd_set_learned(h, net), lib)
    f = abi.fields_struct(st)
    abi.check(lib.cathode_ndd_set_state(h, ctypes.byref(f)), lib)
    rep = abi.NddReport()
    if warmup > 0:
        abi.check(lib.cathode_ndd_step(h, warmup, ctypes.byref(rep)), lib)
    # cathode_ndd_step returns after synchronising the library stream
    t0 = time.perf_counter()
    abi.check(lib.cathode_ndd_step(h, steps, ctypes.byref(rep)), lib)
    t1 = time.perf_counter()
    wall = t1 - t0
    its = rep.total_iterations
    names = ["electrons", "electrons_pc", "ions", "ions_pc", "energy", "energy_pc", "poisson", "poisson_pc"]
    by_solve = {nm: rep.iterations_by_solve[q] / steps for q, nm in enumerate(names)}
    solves = rep.solves
    nonconv = rep.nonconverged
    maxres = rep.max_rel_residual
    # roofline pass: HIP events around every launch of the finest-level
    # fused down kernel (k_vtile ..., DOWN, TOP), on the step's stream
    nprof = min(5, steps)
    abi.check(lib.cathode_ndd_profile(h, 1), lib)
    abi.check(lib.cathode_ndd_step(h, nprof, ctypes.byref(rep)), lib)
    launches = ctypes.c_int()
    tot_ms = ctypes.c_double()
    bpl = ctypes.c_double()
    abi.check(lib.cathode_ndd_profile_read(h, ctypes.byref(launches), ctypes.byref(tot_ms), ctypes.byref(bpl)), lib)
    abi.check(lib.cathode_ndd_profile(h, 0), lib)
    # the finest level's tile shape the autotuner picked for this run
    ti, tj, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    abi.check(lib.cathode_ndd_tile(h, 0 if mode == abi.CATHODE_MODE_FP64 else 1, 0, ctypes.byref(ti),
                                   ctypes.byref(tj), ctypes.byref(nt)), lib)
    # the state after warmup + steps + profiled steps, for the per-field
    # comparison with the FP64 path (same seed, same number of steps)
    abi.check(lib.cathode_ndd_get_state(h, ctypes.byref(f)), lib)
    lib.cathode_ndd_destroy(h)
    avg_ms = tot_ms.value / max(launches.value, 1)
    return {
        "steps_per_s": steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "iterations_per_solve": its / max(solves, 1),
        "solves": solves,
        "nonconverged": nonconv,
        "max_rel_residual": maxres,
        "smooth_launches": launches.value,
        "smooth_avg_us": 1e3 * avg_ms,
        "smooth_bytes": bpl.value,
        "smooth_GBps": bpl.value / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0,
        "smooth_tile": f"{ti.value}x{tj.value}/{nt.value}",
        "steps_total": warmup + steps + nprof,
        "iterations_per_step_by_solve": {k: round(v, 3) for k, v in by_solve.items()},
    }, st


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_vtile_down_l0_512x256.json")


def traffic_per_launch(path, tile):
    """HBM bytes per launch of the finest-level down kernel with this tile
    shape, from the committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    measurement (calibrated against a copy kernel of known size), or None
    when that path and shape have not been measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["paths"][path][tile]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def calibrate_copy(mib=256, reps=5):
    """A device-to-device copy of known size for the FETCH_SIZE / WRITE_SIZE
    calibration: reps copies of `mib` MiB (read mib, write mib each)."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": mib * 1024 * 1024, "copy_bytes_written": mib * 1024 * 1024,
                      "reps": reps}))


def run_sweep(lib, steps, workers):
    """Bias sweep (64 instances of 64x32..512x256 cells, bias 0-60 V, 50 steps
    each by default): all instances stepped together by cathode_ndd_batch_step,
    then one at a time, on fresh handles from the same state."""
    inst = W.config3(mode=abi.CATHODE_MODE_MIXED)

    def handles():
        hs = []
        for p, st, _ in inst:
            h = ctypes.c_void_p()
            abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(h)), lib)
            abi.check(lib.cathode_ndd_set_state(h, ctypes.byref(abi.fields_struct(st))), lib)
            hs.append(h)
        return hs

    def run(w):
        hs = handles()
        arr = (ctypes.c_void_p * len(hs))(*[h.value for h in hs])
        reps = (abi.NddReport * len(hs))()
        t0 = time.perf_counter()
        abi.check(lib.cathode_ndd_batch_step(arr, len(hs), steps, w, reps), lib)
        dt = time.perf_counter() - t0
        for h in hs:
            lib.cathode_ndd_destroy(h)
        return dt, [r.I_anode for r in reps], sum(r.nonconverged for r in reps)

    t_batch, I_batch, nc = run(workers)
    t_seq, I_seq, _ = run(1)
    dev = max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(I_batch, I_seq))
    n = len(inst)
    return {"instances": n, "steps": steps, "workers": workers,
            "sizes": [[p.ni, p.nj] for p, _, _ in inst], "bias_V": [round(b, 4) for _, _, b in inst],
            "instance_steps_per_s": round(n * steps / t_batch, 2),
            "instance_steps_per_s_one_at_a_time": round(n * steps / t_seq, 2),
            "collector_current_A": I_batch, "max_rel_dev_vs_one_at_a_time": dev, "nonconverged": nc}


ROWS_H = 2e-4  # the reference's own grid: the five-block cathode at 0.2 mm, 194 x 70 cells


def rows_workload(lib):
    ni, nj, g, *_ = W.reference_geometry(ROWS_H, lib)
    grid = W.default_params(ni, nj, ROWS_H, ROWS_H)
    ion = W.ion_state(g, ni, nj, ROWS_H, noise=False)
    neu = W.neutral_state(g, ni, nj, ROWS_H, noise=False)
    return ni, nj, g, grid, ion, neu


def run_rows(lib, reps=20):
    """The ion block (solve_Uix/Uir/Uip_tmp, solve_rhoi_constTe,
    correct_Ui_constTe: 4 solves) and the neutral + metastable update
    (update_rhon, update_rhom: 2 solves) on the reference's own cathode grid
    with device-resident arrays (cathode_ndd_*_update_dev): wall time per
    call after 2 untimed calls, each solve to relative residual 1e-10."""
    ni, nj, g, grid, ion, neu = rows_workload(lib)
    hd = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(grid), ctypes.byref(hd)), lib)
    out = {}
    try:
        abi.check(lib.cathode_ndd_set_geometry(hd, ctypes.byref(g)), lib)
        cases = (("ion_block", ion, abi.ION_FIELD_NAMES, abi.IonFields, W.ion_params(tol_Ui=1e-10, tol_rhoi=1e-10),
                  lib.cathode_ndd_ion_update_dev, abi.IonReport,
                  ("solve_Uix", "solve_Uir", "solve_Uip", "solve_rhoi")),
                 ("neutral_update", neu, abi.NEUTRAL_FIELD_NAMES, abi.NeutralFields, W.neutral_params(tol=1e-10),
                  lib.cathode_ndd_neutral_update_dev, abi.NeutralReport, ("solve_n", "solve_m")))
        for name, state, names, cls, prm, fn, rcls, solves in cases:
            f, D, present = abi.device_fields(lib, state, names, cls)
            rep = rcls()
            for _ in range(2):
                abi.check(fn(hd, ctypes.byref(prm), ctypes.byref(f), ctypes.byref(rep)), lib)
            t0 = time.perf_counter()
            for _ in range(reps):
                abi.check(fn(hd, ctypes.byref(prm), ctypes.byref(f), ctypes.byref(rep)), lib)
            t1 = time.perf_counter()
            D.free()
            out[name] = {"ms_per_call": sig(1e3 * (t1 - t0) / reps), "calls": reps,
                         "iterations": [getattr(rep, k).iterations for k in solves],
                         "max_rel_residual": max(getattr(rep, k).rel_residual for k in solves)}
    finally:
        lib.cathode_ndd_destroy(hd)
    out["workload"] = (f"five-block cathode at 0.2 mm ({ni}x{nj} cells, the reference's grid), "
                       "device-resident arrays, FP64 solves to 1e-10")
    out["transport_update"] = run_transport(lib)
    out["mw_solve"] = run_mwsolve(lib)
    return out


BFIELD_0P2MM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                            "bfield_cathode_0p2mm.npz")


def mw_cathode_case():
    """The reference cathode's microwave configuration at its 0.2 mm grid with
    the magnet's field of input/Bfield_data_0.2mm.csv (committed fixture)."""
    import numpy as np
    B = np.load(BFIELD_0P2MM)
    return W.cathode_mw_case(h=2e-4, Bx=B["Bx"], Br=B["Br"])


def _time_mwsolve(lib, case, reps):
    p, lists, o, a = case
    bc, keep = abi.mw_bc_struct(lists)
    D = abi.MwSolveDevice(lib, a)
    rep = abi.MwSolveReport()
    try:
        abi.check(lib.cathode_mw_solve_dev(ctypes.byref(p), ctypes.byref(bc), ctypes.byref(o), ctypes.byref(D.f),
                                           ctypes.byref(rep)), lib)
        t0 = time.perf_counter()
        for _ in range(reps):
            abi.check(lib.cathode_mw_solve_dev(ctypes.byref(p), ctypes.byref(bc), ctypes.byref(o),
                                               ctypes.byref(D.f), ctypes.byref(rep)), lib)
        t1 = time.perf_counter()
    finally:
        D.free()
    return {"ms_per_call": sig(1e3 * (t1 - t0) / reps), "calls": reps, "unknowns": rep.n_x + rep.n_r + rep.n_p,
            "block": rep.block, "columns": p.ni + 2, "refine_steps": rep.refine, "rel_residual": rep.rel_residual}


def run_mwsolve(lib, reps=5):
    """The coupled (E_x, E_r, E_phi) magnetised microwave solve
    (solve_Microwave) with device-resident arrays (cathode_mw_solve_dev): wall
    time per call after one untimed call (assembly, folds, block cyclic
    reduction, refinement to 1e-13, ghosts).  Main line: the reference
    cathode's configuration (makeBoundary's five blocks and coaxial feed,
    194 x 70 cells of 0.2 mm, the magnet's field); also a 194 x 70 rectangle
    whose lists run every rule, and the reference's coaxial impedance test
    (194 x 7, E_r alone)."""
    out = _time_mwsolve(lib, mw_cathode_case(), reps)
    out["workload"] = ("coupled magnetised microwave solve on the reference cathode (makeBoundary's blocks and lists, "
                       "194x70 cells of 0.2 mm, input/Bfield_data_0.2mm.csv), device-resident arrays")
    out["rectangle"] = _time_mwsolve(lib, W.mw_coupled_rect(ni=194, nj=70, h=2e-4, seed=7), reps)
    out["rectangle"]["workload"] = "194x70 rectangle, every list rule, synthetic magnetised plasma"
    p, lists, o, a = W.coax_impedance_test()
    bc, keep = abi.mw_bc_struct(lists)
    D = abi.MwSolveDevice(lib, a)
    rep = abi.MwSolveReport()
    try:
        t0 = time.perf_counter()
        for _ in range(reps):
            abi.check(lib.cathode_mw_solve_dev(ctypes.byref(p), ctypes.byref(bc), ctypes.byref(o),
                                               ctypes.byref(D.f), ctypes.byref(rep)), lib)
        t1 = time.perf_counter()
    finally:
        D.free()
    out["impedance_test"] = {"ms_per_call": sig(1e3 * (t1 - t0) / reps), "P_fwd_all_W": sig(rep.P_fwd_all),
                             "workload": "the reference's coaxial impedance test, 194x7, E_r alone"}
    return out


TRANSPORT_BYTES_PER_CELL = 288  # 14 planes read (13 inputs + nu_ano_IAT), 22 written (21 outputs + nu_ano_IAT)


def run_transport(lib, reps=50, grid=(512, 256)):
    """update_transport_coef on the bench's 512x256 rectangle (every cell
    plasma), device-resident arrays, synthetic tables of the reference's size
    (1800 electron entries, 39 ion entries): wall time per call (a max
    reduction of n_e and the pointwise kernel)."""
    import numpy as np
    ni, nj = grid
    h = 0.05e-3 * 512 / ni
    p = W.default_params(ni, nj, h, h)
    rng = np.random.Generator(np.random.PCG64(7))
    tabs = {n: np.ascontiguousarray(1e-14 * (1.0 + rng.uniform(0, 1, 1800))) for n in abi.RATE_TABLE_NAMES}
    tabs["nui_m"] = np.ascontiguousarray(1e-15 * (1.0 + rng.uniform(0, 1, 39)))
    tabs["dTe"], tabs["dEi"] = 0.1 * 2.0 / 3.0 * W.E0 / W.BOLTZ, 0.1 * W.E0
    a = W.transport_state(W_rect(ni, nj), ni, nj, h)
    hd = ctypes.c_void_p()
    abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(hd)), lib)
    try:
        t = abi.rate_tables_struct(tabs)
        abi.check(lib.cathode_ndd_set_rate_tables(hd, ctypes.byref(t)), lib)
        f, D, names = abi.device_fields(lib, a, abi.TRANSPORT_FIELD_NAMES, abi.TransportFields)
        prm = W.transport_params()
        for _ in range(3):
            abi.check(lib.cathode_ndd_transport_update_dev(hd, ctypes.byref(prm), ctypes.byref(f)), lib)
        t0 = time.perf_counter()
        for _ in range(reps):
            abi.check(lib.cathode_ndd_transport_update_dev(hd, ctypes.byref(prm), ctypes.byref(f)), lib)
        t1 = time.perf_counter()
        D.free()
    finally:
        lib.cathode_ndd_destroy(hd)
    cells = (ni + 2) * (nj + 2)
    return {"ms_per_call": sig(1e3 * (t1 - t0) / reps), "calls": reps, "cells": cells,
            "bytes_per_call": cells * TRANSPORT_BYTES_PER_CELL,
            "workload": f"{ni}x{nj} rectangle, device-resident arrays, synthetic tables of 1800 / 39 entries"}


def W_rect(ni, nj):
    """the built-in rectangle in segment form (for the workload builders)"""
    g = abi.NddGeometry()
    g.n_bl = 1
    for q, v in enumerate((1, ni, 1, nj)):
        g.cell_box[q] = v
    return g


def cpu_model():
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seed, budget_s=12.0):
    """The oracle's host copy of the same algorithm (FP64 multigrid-BiCGStab
    to the same 1e-10 relative residual), built for speed (oracle/Makefile:
    liboracle_fast.so, gcc -O3 -march=x86-64-v3 -fopenmp), on a bounded
    number of steps of the same workload: one thread (the reference runs one
    rank, one thread) and every thread of this process's CPU share (the
    multigrid's sweeps, restrictions and products in OpenMP)."""
    # the oracle (test infrastructure) is loaded here only to be timed
    path = os.path.join(ROOT, "oracle", "liboracle_fast.so")
    lib = ctypes.CDLL(path)
    lib.orc_ndd_step.restype = ctypes.c_int
    lib.orc_ndd_step.argtypes = [ctypes.POINTER(abi.NddParams), ctypes.POINTER(abi.NddFields), ctypes.c_int,
                                 ctypes.c_int, ctypes.POINTER(abi.NddReport)]
    lib.orc_mg_set_threads.argtypes = [ctypes.c_int]
    nproc = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    threads_all = max(1, min(share, nproc))

    def timed(threads):
        lib.orc_mg_set_threads(threads)
        p, st = W.config1(seed=seed)
        f = abi.fields_struct(st)
        rep = abi.NddReport()
        t0 = time.perf_counter()
        assert lib.orc_ndd_step(ctypes.byref(p), ctypes.byref(f), 1, 1, ctypes.byref(rep)) == 0
        t1 = time.perf_counter()
        n = max(1, min(30, int(budget_s / max(t1 - t0, 1e-3))))
        t0 = time.perf_counter()
        assert lib.orc_ndd_step(ctypes.byref(p), ctypes.byref(f), n, 1, ctypes.byref(rep)) == 0
        t1 = time.perf_counter()
        return n / (t1 - t0), n

    v1, n1 = timed(1)
    va, na = timed(threads_all)
    # the ion block and the neutral update on the same cathode workload as
    # the device rows (multigrid solves to 1e-10), one thread
    lib.orc_mg_set_threads(1)
    lib.orc_ndd_set_geometry.argtypes = [ctypes.POINTER(abi.NddGeometry)]
    lib.orc_ion_update.argtypes = [ctypes.POINTER(abi.NddParams), ctypes.POINTER(abi.IonParams),
                                   ctypes.POINTER(abi.IonFields), ctypes.POINTER(abi.IonReport), ctypes.c_int]
    lib.orc_neutral_update.argtypes = [ctypes.POINTER(abi.NddParams), ctypes.POINTER(abi.NeutralParams),
                                       ctypes.POINTER(abi.NeutralFields), ctypes.POINTER(abi.NeutralReport),
                                       ctypes.c_int]
    ni, nj, g, grid, ion, neu = rows_workload(abi.load())
    lib.orc_ndd_set_geometry(ctypes.byref(g))
    rows = {}
    for name, state, fn, prm, fs, rcls in (
            ("ion_block", ion, lib.orc_ion_update, W.ion_params(tol_Ui=1e-10, tol_rhoi=1e-10), abi.ion_fields_struct,
             abi.IonReport),
            ("neutral_update", neu, lib.orc_neutral_update, W.neutral_params(tol=1e-10), abi.neutral_fields_struct,
             abi.NeutralReport)):
        st = W.copy_arrays(state)
        f = fs(st)
        rep = rcls()
        assert fn(ctypes.byref(grid), ctypes.byref(prm), ctypes.byref(f), ctypes.byref(rep), 1) == 0
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            assert fn(ctypes.byref(grid), ctypes.byref(prm), ctypes.byref(f), ctypes.byref(rep), 1) == 0
        rows[name] = {"ms_per_call": 1e3 * (time.perf_counter() - t0) / reps, "calls": reps, "cores": 1}
    lib.orc_ndd_set_geometry(None)
    # the coupled microwave solve of the device row (the reference cathode at
    # 0.2 mm): the oracle's banded LU with partial pivoting, one call, one thread
    lib.orc_mw_solve.argtypes = [ctypes.POINTER(abi.MwParams), ctypes.POINTER(abi.MwBc),
                                 ctypes.POINTER(abi.MwSolveOpts), ctypes.POINTER(abi.MwSolveFields),
                                 ctypes.POINTER(abi.MwSolveReport)]
    p, lists, o, a = mw_cathode_case()
    bc, keep = abi.mw_bc_struct(lists)
    f = abi.mw_solve_fields_struct(a)
    rep = abi.MwSolveReport()
    t0 = time.perf_counter()
    assert lib.orc_mw_solve(ctypes.byref(p), ctypes.byref(bc), ctypes.byref(o), ctypes.byref(f),
                            ctypes.byref(rep)) == 0
    rows["mw_solve"] = {"ms_per_call": 1e3 * (time.perf_counter() - t0), "calls": 1, "cores": 1,
                        "algorithm": "banded LU with partial pivoting over the same rows (the reference's Eigen "
                                     "SparseLU is not installed)", "rel_residual": rep.rel_residual}
    return {"value": va, "unit": "steps/s", "cores": threads_all, "kind": "port", "rows": rows,
            "sample": f"{na} steps of the same 512x256 workload after 1 untimed step, {threads_all} threads",
            "one_core": {"value": v1, "cores": 1, "sample": f"{n1} steps after 1 untimed step"},
            "build": "oracle/orc_*.c, gcc -O3 -march=x86-64-v3 -fopenmp (oracle/liboracle_fast.so)",
            "nproc": nproc, "cpu_model": cpu_model()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--paths", default="fp64,mixed,learned",
                    help="fp64, mixed (FP32 V-cycles + FP64 refinement), learned (mixed + learned initial guess), "
                         "warm (mixed + previous-solution warm starts, no network: the learned path's ablation)")
    ap.add_argument("--grid", default="512x256",
                    help="NIxNJ: the same domain on another grid (the metric is quoted at 512x256; larger grids "
                         "show the stencil kernels away from the latency-bound regime)")
    ap.add_argument("--no-sweep", dest="sweep", action="store_false",
                    help="skip the 64-instance bias sweep (batched vs one at a time), run by default")
    ap.add_argument("--sweep-steps", type=int, default=50)
    ap.add_argument("--sweep-workers", type=int, default=16)
    ap.add_argument("--rows-only", action="store_true",
                    help="only time the ion block, the neutral update and the transport update (one JSON line)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the copy kernel used to calibrate the HBM traffic counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the single-instance implicit step does not shard across GPUs here", file=sys.stderr)
        sys.exit(2)
    grid = tuple(int(v) for v in args.grid.lower().split("x"))
    if len(grid) != 2 or grid[0] < 8 or grid[1] < 8 or grid[0] * 256 != grid[1] * 512:
        print("bench.py: --grid must be NIxNJ with NI = 2 NJ (the 512x256 domain's aspect)", file=sys.stderr)
        sys.exit(2)
    lib = abi.load()
    abi.check(lib.cathode_init(0), lib)
    if args.rows_only:
        print(json.dumps({"rows": run_rows(lib)}))
        lib.cathode_finalize()
        return
    paths = [s.strip() for s in args.paths.split(",") if s.strip()]
    res = {}
    states = {}
    net = None
    for name in paths:
        mode = abi.CATHODE_MODE_FP64 if name == "fp64" else abi.CATHODE_MODE_MIXED
        if name == "learned" and net is None:
            net = L.load_net(lib)  # weights trained by learned.py, committed with the package
        res[name], states[name] = run_path(lib, mode, args.steps, args.warmup, args.seed,
                                           net if name == "learned" else None, warm=name == "warm", grid=grid)
        print(f"# {name}: {res[name]['steps_per_s']:.2f} steps/s, "
              f"{res[name]['iterations_per_solve']:.2f} BiCGStab it/solve, "
              f"smooth {res[name]['smooth_avg_us']:.2f} us ({res[name]['smooth_GBps']:.0f} GB/s)",
              file=sys.stderr, flush=True)
    # per-field relative L2 against the FP64 path after the same steps from
    # the same state (every solve of every path to relative residual 1e-10)
    if "fp64" in states:
        for name in res:
            if name != "fp64":
                res[name]["field_rel_l2_vs_fp64"] = {
                    lab: sig(rel_l2(states[name][k][1:-1, 1:-1], states["fp64"][k][1:-1, 1:-1]))
                    for lab, k in REPORT_FIELDS.items()}
    # the metric is the learned + mixed path against plain FP64
    best = "learned" if "learned" in res else ("mixed" if "mixed" in res else paths[0])
    b = res[best]
    roof = {"bound": "hbm", "achieved": round(b["smooth_GBps"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(b["smooth_GBps"] / HBM_PEAK_GBS, 4),
            "traffic": (traffic_per_launch("mixed" if best == "learned" else best, b["smooth_tile"])
                        if grid == (512, 256) else None),
            "kernel": "k_vtile finest-level down kernel (BiCGStab p/s update + 2 red-black sweeps + residual + "
                      "restriction), "
                      + ("FP64" if best == "fp64" else "FP32") + ", tile " + b["smooth_tile"],
            "bytes_model": "per cell: aP aE aW aN aS + activity byte read, x + b + 1/4 coarse b written at "
                           "level precision; r, v (+ old p on every other launch) read and new p or s written "
                           "in FP64 (fused BiCGStab update): 58 B (FP32 levels), 87 B (FP64)",
            "bytes_per_launch": b["smooth_bytes"], "avg_us": round(b["smooth_avg_us"], 3)}
    out = {
        "metric": METRIC,
        "value": round(b["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32+f64", "learned": "f32+f64 (bf16 network)", "warm": "f32+f64"}[best],
        "data": "synthetic",
        "config": {"workload": f"implicit NDD step, {grid[0]}x{grid[1]} cells of {0.05 * 512 / grid[0]:.4g} mm, "
                               "ions+Poisson+electrons+energy, each solve to relative residual 1e-10",
                   "grid": list(grid), "seed": args.seed,
                   "path": {"fp64": "plain FP64", "mixed": "FP32 V-cycles + FP64 refinement",
                            "learned": "learned initial guess + FP32 V-cycles + FP64 refinement",
                            "warm": "previous-solution warm starts + FP32 V-cycles + FP64 refinement"}[best]},
        "roofline": roof,
        "paths": {k: {kk: sig(vv) for kk, vv in v.items()} for k, v in res.items()},
    }
    if "fp64" in res and best != "fp64":
        out["speedup_vs_fp64"] = round(b["steps_per_s"] / res["fp64"]["steps_per_s"], 3)
    if args.sweep:
        out["sweep"] = run_sweep(lib, args.sweep_steps, args.sweep_workers)
        s = out["sweep"]
        print(f"# sweep: {s['instance_steps_per_s']} instance-steps/s batched ({s['workers']} workers), "
              f"{s['instance_steps_per_s_one_at_a_time']} one at a time, max dev {s['max_rel_dev_vs_one_at_a_time']:.2e}",
              file=sys.stderr, flush=True)
    out["rows"] = run_rows(lib)
    print(f"# rows: ion block {out['rows']['ion_block']['ms_per_call']} ms/call, neutral update "
          f"{out['rows']['neutral_update']['ms_per_call']} ms/call, transport update "
          f"{out['rows']['transport_update']['ms_per_call']} ms/call", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.seed)
    if net is not None:
        lib.cathode_cnn_destroy(net)
    lib.cathode_finalize()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
