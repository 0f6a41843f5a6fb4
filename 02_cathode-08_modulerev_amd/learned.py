"""Training and evaluation of the learned initial guess (csrc/learned.hip).

Pairs come from the solver itself: NDD trajectories of seeded multi-Gaussian
states (the config-1 recipe) with biases drawn uniformly in [0, 60] V; each
step's first Poisson correction system is captured on the device
(cathode_ndd_capture) as three feature planes and the normalised solution.
Training runs the library's own forward / backward / Adam kernels
(cathode_cnn_train_step_dev) on device buffers owned through the library.

    python3 02_cathode-08_modulerev_amd/learned.py --out weights.npy
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np

if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
PKG = __name__.rsplit(".", 1)[0] if "." in __name__ else "02_cathode-08_modulerev_amd"
abi = importlib.import_module(PKG + ".abi")
W = importlib.import_module(PKG + ".workloads")

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "learned_poisson_512x256.npy")


def sweep_state(seed, ni, nj, h, mode):
    """config-1 recipe with a bias drawn from the seed"""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    bias = float(rng.uniform(0.0, 60.0))
    p, st = W.config1(seed=seed, ni=ni, nj=nj, h=h, V_bias=bias, mode=mode)
    return p, st


def generate_pairs(lib, seeds, steps, ni=512, nj=256, h=5e-5, mode=abi.CATHODE_MODE_MIXED):
    """(features, targets) as abi.DeviceArray of shapes (S, 3, ni+2, nj+2) and
    (S, ni+2, nj+2), S = len(seeds) * steps"""
    S = len(seeds) * steps
    feat = abi.DeviceArray(lib, (S, 3, ni + 2, nj + 2))
    target = abi.DeviceArray(lib, (S, ni + 2, nj + 2))
    k = 0
    for s in seeds:
        p, st = sweep_state(s, ni, nj, h, mode)
        hd = ctypes.c_void_p()
        abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(hd)), lib)
        abi.check(lib.cathode_ndd_set_state(hd, ctypes.byref(abi.fields_struct(st))), lib)
        for _ in range(steps):
            abi.check(lib.cathode_ndd_capture(hd, feat.at(k), target.at(k)), lib)
            rep = abi.NddReport()
            abi.check(lib.cathode_ndd_step(hd, 1, ctypes.byref(rep)), lib)
            k += 1
        lib.cathode_ndd_destroy(hd)
    return feat, target


def train(lib, net, feat, target, epochs=4, batch=8, lr=1e-3, seed=0, log=None):
    """Adam on shuffled mini-batches (assembled on the device), learning
    rate halved every epoch; returns the mean loss of each epoch"""
    S = feat.shape[0]
    ni, nj = feat.shape[2] - 2, feat.shape[3] - 2
    rng = np.random.default_rng(seed)
    fb = abi.DeviceArray(lib, (batch,) + feat.shape[1:])
    tb = abi.DeviceArray(lib, (batch,) + target.shape[1:])
    losses = []
    try:
        for ep in range(epochs):
            order = rng.permutation(S)
            lr_ep = lr * (0.5 ** ep)
            ep_loss = []
            for b0 in range(0, S - batch + 1, batch):
                for q in range(batch):
                    fb.copy_from(q, feat, int(order[b0 + q]))
                    tb.copy_from(q, target, int(order[b0 + q]))
                loss = ctypes.c_double()
                abi.check(lib.cathode_cnn_train_step_dev(net, ni, nj, batch, fb.ptr, tb.ptr, ctypes.c_float(lr_ep),
                                                         ctypes.byref(loss)), lib)
                ep_loss.append(loss.value)
            losses.append(float(np.mean(ep_loss)))
            if log:
                log(f"epoch {ep}: mean loss {losses[-1]:.4f} (lr {lr_ep:.1e})")
    finally:
        fb.free()
        tb.free()
    return losses


def train_streaming(lib, net, pairs, chunk, steps_per_seed, ni, nj, h, batch=8, lr=1e-3, seed0=100, log=None):
    """One pass over `pairs` fresh pairs, generated `chunk` at a time (the
    trajectories of chunk / steps_per_seed seeds, then trained on, shuffled,
    and freed: the whole set need not fit in memory); the learning rate
    falls linearly to a tenth over the pass.  Returns the mean loss of each
    chunk and the seconds spent generating and training."""
    per_chunk_seeds = max(1, chunk // steps_per_seed)
    nchunks = max(1, pairs // (per_chunk_seeds * steps_per_seed))
    rng = np.random.default_rng(seed0)
    losses, t_gen, t_train = [], 0.0, 0.0
    fb = abi.DeviceArray(lib, (batch, 3, ni + 2, nj + 2))
    tb = abi.DeviceArray(lib, (batch, ni + 2, nj + 2))
    try:
        for c in range(nchunks):
            t0 = time.perf_counter()
            seeds = list(range(seed0 + c * per_chunk_seeds, seed0 + (c + 1) * per_chunk_seeds))
            feat, target = generate_pairs(lib, seeds, steps_per_seed, ni, nj, h)
            t1 = time.perf_counter()
            S = feat.shape[0]
            order = rng.permutation(S)
            lr_c = lr * (1.0 - 0.9 * c / max(1, nchunks - 1))
            cl = []
            for b0 in range(0, S - batch + 1, batch):
                for q in range(batch):
                    fb.copy_from(q, feat, int(order[b0 + q]))
                    tb.copy_from(q, target, int(order[b0 + q]))
                loss = ctypes.c_double()
                abi.check(lib.cathode_cnn_train_step_dev(net, ni, nj, batch, fb.ptr, tb.ptr, ctypes.c_float(lr_c),
                                                         ctypes.byref(loss)), lib)
                cl.append(loss.value)
            feat.free()
            target.free()
            t2 = time.perf_counter()
            t_gen += t1 - t0
            t_train += t2 - t1
            losses.append(float(np.mean(cl)))
            if log:
                log(f"chunk {c + 1}/{nchunks}: {S} pairs (seeds {seeds[0]}..{seeds[-1]}), mean loss {losses[-1]:.4f} "
                    f"(lr {lr_c:.1e}); generate {t1 - t0:.1f} s, train {t2 - t1:.1f} s")
    finally:
        fb.free()
        tb.free()
    return losses, t_gen, t_train


def get_params(lib, net):
    n = lib.cathode_cnn_num_params()
    w = np.zeros(n, dtype=np.float32)
    abi.check(lib.cathode_cnn_get_params(net, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), lib)
    return w


def load_net(lib, path=WEIGHTS):
    """network with the committed weights (raises if the file is absent)"""
    w = np.load(path, allow_pickle=False).astype(np.float32)
    net = ctypes.c_void_p()
    abi.check(lib.cathode_cnn_create(0, ctypes.byref(net)), lib)
    if w.size != lib.cathode_cnn_num_params():
        raise ValueError(f"{path}: {w.size} parameters, network has {lib.cathode_cnn_num_params()}")
    abi.check(lib.cathode_cnn_set_params(net, np.ascontiguousarray(w).ctypes.data_as(
        ctypes.POINTER(ctypes.c_float))), lib)
    return net


VARIANTS = ("plain", "warm", "learned")
FP64_VARIANT = "plain_fp64"  # every solve from x = 0 in FP64: the reference's method on the device


def evaluate(lib, net, seeds, steps, ni=512, nj=256, h=5e-5, mode=abi.CATHODE_MODE_MIXED, fp64=False):
    """Same held-out trajectories three ways: plain (every solve from x = 0,
    as the reference), warm (learned mode without a network: every solve
    from its previous-step solution, least-squares scaled) and learned (the
    network's guess combined with that previous solution for the first
    Poisson correction solve, warm starts for the others).  BiCGStab
    iterations over all solves, wall time, and the largest field difference
    to the plain run (relative L2)."""
    variants = VARIANTS + ((FP64_VARIANT,) if fp64 else ())
    out = {v: {"iterations": 0, "seconds": 0.0, "steps": 0} for v in variants}
    out["max_field_rel_l2"] = 0.0
    for s in seeds:
        finals = {}
        for name in variants:
            p, st = sweep_state(s, ni, nj, h, abi.CATHODE_MODE_FP64 if name == FP64_VARIANT else mode)
            p.learned = 0 if name in ("plain", FP64_VARIANT) else 1
            hd = ctypes.c_void_p()
            abi.check(lib.cathode_ndd_create(ctypes.byref(p), ctypes.byref(hd)), lib)
            if name == "learned":
                abi.check(lib.cathode_ndd_set_learned(hd, net), lib)
            abi.check(lib.cathode_ndd_set_state(hd, ctypes.byref(abi.fields_struct(st))), lib)
            rep = abi.NddReport()
            abi.check(lib.cathode_ndd_step(hd, 1, ctypes.byref(rep)), lib)  # warm (autotune)
            t0 = time.perf_counter()
            abi.check(lib.cathode_ndd_step(hd, steps, ctypes.byref(rep)), lib)
            out[name]["seconds"] += time.perf_counter() - t0
            out[name]["iterations"] += rep.total_iterations
            out[name]["steps"] += steps
            fin = {k: np.zeros_like(v) for k, v in st.items()}
            abi.check(lib.cathode_ndd_get_state(hd, ctypes.byref(abi.fields_struct(fin))), lib)
            finals[name] = fin
            lib.cathode_ndd_destroy(hd)
        for v in [x for x in variants if x != "plain"]:
            for k in ("phi", "rhoe", "rhoi", "rhoeps"):
                a, b = finals[v][k], finals["plain"][k]
                d = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
                out["max_field_rel_l2"] = max(out["max_field_rel_l2"], float(d))
    for name in variants:
        o = out[name]
        o["steps_per_s"] = o["steps"] / o["seconds"] if o["seconds"] > 0 else 0.0
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--train-seeds", type=int, default=96)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-seeds", type=int, default=4)
    ap.add_argument("--evaluate-only", action="store_true",
                    help="evaluate the committed weights on held-out seeds (seeds 900..) without training")
    ap.add_argument("--grid", default="512x256",
                    help="NIxNJ of the evaluation (the same domain; 1024x512 is the learned-initialisation "
                         "workload of 1024 x 512 cells)")
    ap.add_argument("--pairs", type=int, default=0,
                    help="stream this many fresh pairs at --grid (chunks of --chunk, one pass) instead of "
                         "training epochs on --train-seeds x --steps pairs at 512x256")
    ap.add_argument("--chunk", type=int, default=500)
    ap.add_argument("--steps-per-seed", type=int, default=20)
    args = ap.parse_args()
    lib = abi.load()
    abi.check(lib.cathode_init(0), lib)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.evaluate_only:
        ni, nj = (int(v) for v in args.grid.lower().split("x"))
        net = load_net(lib)
        ev = evaluate(lib, net, list(range(900, 900 + args.eval_seeds)), args.steps, ni=ni, nj=nj,
                      h=5e-5 * 512 / ni, fp64=True)
        print(json.dumps({"grid": [ni, nj], "seeds": args.eval_seeds, "steps_per_seed": args.steps,
                          "weights": os.path.basename(WEIGHTS), "eval": ev}))
        lib.cathode_cnn_destroy(net)
        lib.cathode_finalize()
        return
    if args.pairs > 0:
        ni, nj = (int(v) for v in args.grid.lower().split("x"))
        h = 5e-5 * 512 / ni
        net = ctypes.c_void_p()
        abi.check(lib.cathode_cnn_create(1, ctypes.byref(net)), lib)
        losses, t_gen, t_train = train_streaming(lib, net, args.pairs, args.chunk, args.steps_per_seed, ni, nj, h,
                                                 args.batch, args.lr, log=log)
        np.save(args.out, get_params(lib, net), allow_pickle=False)
        ev = evaluate(lib, net, list(range(900, 900 + args.eval_seeds)), args.steps, ni=ni, nj=nj, h=h, fp64=True)
        print(json.dumps({"grid": [ni, nj], "pairs": len(losses) * (args.chunk // args.steps_per_seed)
                          * args.steps_per_seed, "chunk": args.chunk, "generate_seconds": t_gen,
                          "train_seconds": t_train, "loss_per_chunk": losses, "eval_seeds": args.eval_seeds,
                          "eval_steps_per_seed": args.steps, "eval": ev}))
        lib.cathode_cnn_destroy(net)
        lib.cathode_finalize()
        return
    t0 = time.perf_counter()
    feat, target = generate_pairs(lib, list(range(100, 100 + args.train_seeds)), args.steps)
    log(f"{feat.shape[0]} pairs in {time.perf_counter() - t0:.1f} s")
    net = ctypes.c_void_p()
    abi.check(lib.cathode_cnn_create(1, ctypes.byref(net)), lib)
    t0 = time.perf_counter()
    losses = train(lib, net, feat, target, args.epochs, args.batch, args.lr, log=log)
    t_train = time.perf_counter() - t0
    w = get_params(lib, net)
    np.save(args.out, w, allow_pickle=False)
    feat.free()
    target.free()
    ev = evaluate(lib, net, list(range(900, 900 + args.eval_seeds)), args.steps)
    print(json.dumps({"pairs": args.train_seeds * args.steps, "epochs": args.epochs, "train_seconds": t_train,
                      "loss_per_epoch": losses, "eval": ev}))
    lib.cathode_cnn_destroy(net)
    lib.cathode_finalize()


if __name__ == "__main__":
    main()
