#!/usr/bin/env python3
"""Benchmark of the MI355X collision operator (one GPU).

Workload: the collision step of the 1D-3V Mach-1.2 normal shock -- 256 spatial
cells sharing one N=32, L_v=8 hard-sphere (lambda=1) weight table, each cell
evaluated with ComputeQ (src/collisions.c:212-221) followed by the projection
(src/conserve.c:207-264).  A "step" is one batched evaluation of Q for all 256
cells; the inputs are resident in HBM before timing starts.

Printed: one JSON line.  `value` = Q evaluations per second of the batched
FP32-streamed form (W streamed as FP32, FP64 arithmetic).  Alongside: the plain
FP64 form (one cell per pass over W, the reference method), the batched FP64
form, the FP32 form's relative L2 error against FP64, the roofline of the
weighted-sum kernel and the CPU oracle timed on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_VALU_TFLOPS = 78.6    # MI355X FP64 vector


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--ncell", type=int, default=256)
    p.add_argument("--form", default="fp32", choices=["fp64", "fp32", "fp16x2", "fp32a"])
    p.add_argument("--tile", default=None, help="tr,lines,tb (default: autotune)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-ref", action="store_true", help="skip the FP64 comparisons")
    p.add_argument("--cpu-cells", type=int, default=0, help="cells in the CPU sample (0: about 15 s of work)")
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (cells would shard across "
              "GPUs with a replicated weight table; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    import spectralbte_amd as spb
    from spectralbte_amd import workloads

    N, Lv, lam = 32, 8.0, 1.0
    ncell = a.ncell
    n3 = N ** 3
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    t0 = time.time()
    op.generate_weights_iso(lam)
    op.sync()
    t_gen = time.time() - t0
    f = workloads.shock_profile(N, Lv, ncell)
    d_f = spb.DeviceArray.from_host(f)
    d_Q = spb.DeviceArray((ncell, n3))
    flags = spb.CONSERVE | spb.SELF

    if a.tile:
        tile = tuple(int(x) for x in a.tile.split(","))
        op.set_tile(a.form, False, *tile)
    else:
        tile = op.autotune(ncell, a.form, False)
    print(f"# form={a.form} tile(tr,lines,tb)={tile} weights generated in {t_gen:.2f}s", file=sys.stderr, flush=True)

    for _ in range(a.warmup):
        op.compute_q_dev(d_f.ptr, None, d_Q.ptr, ncell, a.form, flags)
    op.sync()
    ms_step, ms_wsum = op.time_compute_q(d_f.ptr, None, d_Q.ptr, ncell, a.form, flags, iters=a.steps)
    op.sync()
    value = ncell / (ms_step * 1e-3)
    print(f"# {a.form}: {ms_step:.3f} ms/step, weighted-sum kernel {ms_wsum:.3f} ms", file=sys.stderr, flush=True)
    Q = d_Q.download()

    # ---- roofline of the weighted-sum kernel
    tb = tile[2]
    sw = {"fp64": 8, "fp32": 4, "fp16x2": 4, "fp32a": 4}[a.form]
    w_bytes = -(-ncell // tb) * sw * float(n3) * n3          # one pass over W per cell group
    cell_bytes = ncell * n3 * (16 * 2 + 16)                    # F, G read, Qhat written
    alg_bytes = w_bytes + cell_bytes
    achieved = alg_bytes / (ms_wsum * 1e-3) / 1e9
    flops = 10.0 * float(n3) * n3 * ncell
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "wsum_traffic.json")
    if os.path.exists(tpath):
        try:
            tj = json.load(open(tpath))
            key = f"{a.form}:{ncell}:{tile[0]},{tile[1]},{tile[2]}"
            if key in tj:
                traffic = tj[key]
        except (OSError, ValueError):
            traffic = None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_wsum", "kernel_ms": round(ms_wsum, 4), "alg_bytes_per_launch": alg_bytes,
            "valu_tflops": round(flops / (ms_wsum * 1e-3) / 1e12, 2),
            "valu_frac_fp64": round(flops / (ms_wsum * 1e-3) / 1e12 / FP64_VALU_TFLOPS, 4)}

    extra = {}
    if not a.no_ref:
        # FP64 batched on the same inputs: error of the measured form
        tile64 = op.autotune(ncell, "fp64", False)
        ms64, k64 = op.time_compute_q(d_f.ptr, None, d_Q.ptr, ncell, "fp64", flags, iters=2)
        Q64 = d_Q.download()
        # relative L2 over the batch, and per cell over the cells whose |Q| is not
        # round-off (equilibrium cells far upstream/downstream have Q ~ 0)
        nrm = np.linalg.norm(Q64, axis=1)
        sig = nrm >= 1e-3 * nrm.max()
        err = [float(np.linalg.norm(Q[b] - Q64[b]) / nrm[b]) for b in range(ncell) if sig[b]]
        err_batch = float(np.linalg.norm(Q - Q64) / np.linalg.norm(Q64))
        # plain FP64: the reference method, one cell per pass over W
        op.set_tile("fp64", False, 4, 1, 1)
        npl = min(ncell, 64)
        msp, kp = op.time_compute_q(d_f.ptr, None, d_Q.ptr, npl, "fp64", flags, iters=1)
        op.set_tile("fp64", False, 0, 0, 0)
        plain = npl / (msp * 1e-3)
        w64 = -(-npl // 1) * 8.0 * n3 * n3
        extra = {"fp64_batched_q_per_s": round(ncell / (ms64 * 1e-3), 1), "fp64_batched_tile": list(tile64),
                 "fp64_plain_q_per_s": round(plain, 1), "fp64_plain_hbm_gbs": round(w64 / (kp * 1e-3) / 1e9, 1),
                 "speedup_vs_plain_fp64": round(value / plain, 2),
                 "rel_l2_err_vs_fp64_batch": err_batch, "rel_l2_err_vs_fp64_cell_max": max(err),
                 "rel_l2_err_vs_fp64_cell_median": float(np.median(err)), "cells_in_err_stats": len(err)}
        # conservation of the measured form's output
        w = np.ones(N); w[0] = w[-1] = 0.5
        I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        p = (w[I] * w[J] * w[K] * (v[1] - v[0]) ** 3).ravel()
        C = np.stack([p, p * v[I].ravel(), p * v[J].ravel(), p * v[K].ravel(),
                      p * 0.5 * (v[I] ** 2 + v[J] ** 2 + v[K] ** 2).ravel()])
        extra["conservation_max_rel"] = float(np.max(np.abs(Q @ C.T) / (np.abs(Q) @ np.abs(C.T))))

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(op, f, N, v, eta, a.cpu_cells)

    out = {"metric": "Q(f) evals/s at N=32 (batched FP32 vs plain FP64) at <=1e-6 rel err; HBM GB/s",
           "value": round(value, 1), "unit": "Q evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32" if a.form == "fp32a" else "f64", "weights_dtype": {"fp64": "f64", "fp32": "f32",
                                                                                "fp16x2": "f16x2", "fp32a": "f32"}[a.form],
           "data": "synthetic",
           "config": {"workload": "1D-3V Mach-1.2 shock collision step: 256 cells, N=32, L_v=8, hard spheres "
                                  "(lambda=1), ComputeQ + projection per cell",
                      "ncell": ncell, "N": N, "L_v": Lv, "form": a.form, "tile_tr_lines_tb": list(tile)},
           "roofline": roof, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out))


def cpu_baseline(op, f, N, v, eta, ncells):
    """The CPU oracle (C + OpenMP restatement of the reference loop nest) on a
    bounded sample of the same workload: ComputeQ + projection of a few cells."""
    import oracle
    oracle.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    threads = max(1, min(threads, 16))
    W = op.get_weights()
    cons = oracle.Conserver(N, v)
    t0 = time.time()
    q = cons(oracle.compute_q(N, v, eta, W, f[0], f[0], nthreads=threads).ravel())
    t1 = time.time() - t0
    n = ncells or max(1, min(16, int(15.0 / max(t1, 1e-3))))
    t0 = time.time()
    for b in range(1, n + 1):
        q = cons(oracle.compute_q(N, v, eta, W, f[b % len(f)], f[b % len(f)], nthreads=threads).ravel())
    dt = time.time() - t0
    del W
    return {"value": round(n / dt, 4), "unit": "Q evals/s", "cores": threads, "kind": "port",
            "sample": f"{n} cells of the same workload (ComputeQ + projection, N=32, FP64), oracle/spbte_oracle.c "
                      f"with {threads} OpenMP threads over output rows, {dt:.1f} s"}


if __name__ == "__main__":
    main()
