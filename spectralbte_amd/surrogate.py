"""Training data, training loop and evaluation of the learned collision surrogate.

The network and its kernels are in the library (include/spbte.h, spbte_net_*);
this module generates (f, Q(f)) pairs with the exact operator and drives the
training.  Cells come from Mach 1.05-1.6 normal-shock profiles on the 1D grid of
the operator (Mott-Smith blends of the Rankine-Hugoniot states, several layer
widths), optionally advanced a few whole steps of the 1D solver on the GPU so
the set also holds states the solver produces.

    python -m spectralbte_amd.surrogate --cells 2048 --steps 2000
"""
import argparse
import json
import sys
import time

import numpy as np


def shock_cells(N, Lv, n, seed, nX=64):
    """n cells drawn from shock profiles with seeded Mach numbers and layer widths."""
    from spectralbte_amd import workloads
    rng = np.random.default_rng(seed)
    out = []
    while sum(len(x) for x in out) < n:
        mach = rng.uniform(1.05, 1.6)
        width = rng.uniform(6.0, 40.0)
        prof = workloads.shock_profile(N, Lv, nX, mach=mach, blend_cells=width)
        out.append(prof)
    f = np.concatenate(out)
    return f[rng.permutation(len(f))[:n]]


def significant(Q, frac=1e-2):
    """Cells whose |Q| is at least `frac` of the largest (the rest are equilibrium
    cells, whose Q is the discretisation residual of the operator)."""
    nrm = np.linalg.norm(Q, axis=1)
    return nrm >= frac * nrm.max()


def judged_errors(spb, op, net, f, floor_rel=0.1, maxpreserve=False):
    """Per cell: relative L2 error of the surrogate Q (projected) against the
    exact FP64 Q (projected), and the same with the denominator floored at
    floor_rel times the largest |Q| of the set."""
    Qs = net.predict(f, conserve=True)
    Qe = op.compute_q(f, maxpreserve=maxpreserve, conserve=True)
    d = np.linalg.norm(Qs - Qe, axis=1)
    q = np.linalg.norm(Qe, axis=1)
    return d / q, d / np.maximum(q, floor_rel * q.max()), q / q.max()


def forward_flops(M, C, r, L):
    """Multiply-add flops of one cell's forward pass (mixing and channel GEMMs)."""
    return L * (2 * 2 * M * r * C + 2 * 2 * M * C * 2 * C)


def time_predict(spb, op, net, f, reps=3):
    """Seconds per surrogate evaluation of the cells f (inputs resident in HBM)."""
    d_f = spb.DeviceArray.from_host(f)
    d_Q = spb.DeviceArray(f.shape)
    net.predict_dev(d_f.ptr, d_Q.ptr, len(f))
    op.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.predict_dev(d_f.ptr, d_Q.ptr, len(f))
    op.sync()
    return (time.perf_counter() - t0) / reps


def build_dataset(spb, op, net, f, chunk=256):
    """Append (features, exact targets) of the cells f to the network's dataset."""
    n = len(f)
    net.reserve(n)
    d_Q = spb.DeviceArray((chunk, op.N ** 3))
    for c0 in range(0, n, chunk):
        part = f[c0:c0 + chunk]
        d_f = spb.DeviceArray.from_host(part)
        net.add(d_f.ptr, d_Q.ptr, len(part))
        d_f.free()
    d_Q.free()


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--N", type=int, default=32)
    p.add_argument("--Lv", type=float, default=8.0)
    p.add_argument("--Kb", type=int, default=24)
    p.add_argument("--width", type=int, default=64)
    p.add_argument("--rank", type=int, default=128)
    p.add_argument("--layers", type=int, default=4)
    p.add_argument("--cells", type=int, default=2048)
    p.add_argument("--test-cells", type=int, default=256)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--batch", type=int, default=32)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--seed", type=int, default=1)
    a = p.parse_args()
    import spectralbte_amd as spb
    from spectralbte_amd import workloads
    v, eta = workloads.grid_1d(a.N, a.Lv)
    op = spb.Operator(a.N, a.Lv, v, eta)
    op.generate_weights_iso(1.0)
    ftr = shock_cells(a.N, a.Lv, a.cells, a.seed)
    fte = shock_cells(a.N, a.Lv, a.test_cells, a.seed + 1000)
    net = spb.Net(op, a.Kb, a.width, a.rank, a.layers, maxpreserve=False, seed=a.seed)
    d_f = spb.DeviceArray.from_host(ftr[:256])
    d_Q = spb.DeviceArray((256, a.N ** 3))
    net.calibrate(d_f.ptr, d_Q.ptr, 256)
    t0 = time.time()
    build_dataset(spb, op, net, ftr)  # training cells only: the trainer samples the whole dataset
    t_data = time.time() - t0
    print(f"# dataset {a.cells} cells in {t_data:.1f}s; {net.npar} parameters", file=sys.stderr, flush=True)
    log = []
    t_train = 0.0
    done = 0
    while done < a.steps:
        k = min(500, a.steps - done)
        t0 = time.time()
        losses = net.train(k, a.batch, a.lr * (0.3 if done > 0.7 * a.steps else 1.0), seed=a.seed + done)
        t_train += time.time() - t0
        done += k
        rel, flo, qn = judged_errors(spb, op, net, fte)  # held-out cells
        sig = qn >= 1e-2
        log.append({"step": done, "train_loss": float(np.mean(losses[-50:])),
                    "heldout_rel_err_significant_median": float(np.median(rel[sig])),
                    "heldout_floor01_err_median": float(np.median(flo))})
        print("# " + json.dumps(log[-1]), file=sys.stderr, flush=True)
    rel, flo, qn = judged_errors(spb, op, net, fte)
    sig = qn >= 1e-2
    summary = {"train_s": round(t_train, 1), "data_s": round(t_data, 1), "log": log,
               "judged_rel_err_significant_median": float(np.median(rel[sig])),
               "judged_rel_err_significant_p90": float(np.percentile(rel[sig], 90)),
               "judged_floor01_err_median": float(np.median(flo)), "judged_floor01_err_max": float(flo.max()),
               "cells_significant": int(sig.sum()), "cells": int(len(rel))}
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
