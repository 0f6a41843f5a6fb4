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
