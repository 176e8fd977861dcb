#!/usr/bin/env python3
"""Benchmark of the MI355X collision operator (one GPU).

Headline workload: the N=32 half of the 0D-3V ensemble -- 512 independent
problems at N=32, L_v=6, hard spheres (lambda=1), each a seeded sum of two
Maxwellians (weights U(0.2,0.8), means U(-1.5,1.5)^3, temperatures U(0.5,2),
unit density).  A step is one batched evaluation of the 0D operator of the
reference, ComputeQ_maxPreserve (src/collisions.c:178-210) followed by the
projection (src/conserve.c:207-264), for all 512 problems; the inputs are
resident in HBM before timing starts.

`value` = Q evaluations per second with the weight table streamed as FP32 and
the complex products in FP32 (FP32 partial sums of 32 terms folded into FP64
accumulators).  Alongside, each with its precision form and its error against
FP64 stated: the same batch in FP64 and in the plain FP64 form (one cell per
pass over the table, the reference method); the BKW relaxation at N=16; the
256-cell N=32 Mach-1.2 shock collision step (ComputeQ + projection) in every
form; whole-step runs (shock, mixed-N ensemble, heat transfer on the
glancing-collision table); the anisotropic table generators; the learned
surrogate; the roofline of the weighted-sum kernel against the published peak
and against the issue ceiling measured in the same run; and the reference
method's CPU restatement timed on this host.  Progress goes to stderr, the
JSON line to stdout.
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
FP64_VALU_TFLOPS = 78.6    # MI355X FP64 vector peak (FP64-arithmetic forms)
FP32_VALU_TFLOPS = 157.3   # MI355X FP32 vector peak as published (it counts packed FP32, which this build does not issue)


def log(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def guarded(name, fn, *args):
    """fn(*args), or {"error": ...} with the failure logged."""
    try:
        return fn(*args)
    except Exception as e:  # noqa: BLE001 -- any failure of a secondary measurement
        log(f"{name} failed: {type(e).__name__}: {e}")
        return {"error": f"{type(e).__name__}: {e}"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--ncell", type=int, default=512, help="problems in the ensemble batch")
    p.add_argument("--form", default="fp32a", choices=["fp64", "fp32", "fp16x2", "fp32a"])
    p.add_argument("--tile", default=None, help="kind,tr,lines,tb (default: autotune)")
    p.add_argument("--workload", default="ensemble", choices=["ensemble", "shock"],
                   help="timed workload: the ensemble (default) or the shock collision step")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-ref", action="store_true", help="skip the FP64 comparisons")
    p.add_argument("--no-extra", action="store_true", help="skip every row beside the headline")
    p.add_argument("--no-runs", action="store_true", help="skip the whole-step runs (shock, mixed ensemble)")
    p.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate measurement")
    p.add_argument("--full-runs", action="store_true",
                   help="also run configurations 2-4 at their full lengths (about 4 minutes)")
    p.add_argument("--rows", default=None,
                   help="comma-separated secondary rows to run, e.g. c1_bkw16 (default: every row the flags above keep)")
    return p.parse_args()


def cmat(N, v):
    w = np.ones(N); w[0] = w[-1] = 0.5
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    p = (w[I] * w[J] * w[K] * (v[1] - v[0]) ** 3).ravel()
    return np.stack([p, p * v[I].ravel(), p * v[J].ravel(), p * v[K].ravel(),
                     p * 0.5 * (v[I] ** 2 + v[J] ** 2 + v[K] ** 2).ravel()])


class Workload:
    """One batch on one weight table: inputs in HBM, timed evaluation, FP64 checks."""

    def __init__(self, spb, kind, ncell):
        from spectralbte_amd import workloads
        self.spb, self.kind, self.ncell = spb, kind, ncell
        self.N = 32
        if kind == "ensemble":
            self.Lv = 6.0
            self.v, self.eta = workloads.grid_0d(self.N, self.Lv)
            self.f = workloads.two_maxwellian_ensemble(self.N, self.Lv, ncell, seed=0x5EED0002)
            self.mp = True
            self.desc = (f"0D-3V ensemble, N=32 half: {ncell} seeded two-Maxwellian problems, N=32, L_v=6, hard "
                         "spheres (lambda=1), ComputeQ_maxPreserve + projection per problem")
        else:
            self.Lv = 8.0
            self.v, self.eta = workloads.grid_1d(self.N, self.Lv)
            self.f = workloads.shock_profile(self.N, self.Lv, ncell)
            self.mp = False
            self.desc = (f"1D-3V Mach-1.2 shock collision step: {ncell} cells, N=32, L_v=8, hard spheres "
                         "(lambda=1), ComputeQ + projection per cell")
        self.op = spb.Operator(self.N, self.Lv, self.v, self.eta)
        t0 = time.time()
        self.op.generate_weights_iso(1.0)
        self.op.sync()
        log(f"{kind}: weights generated on the GPU in {time.time() - t0:.2f}s")
        self.d_f = spb.DeviceArray.from_host(self.f)
        self.d_Q = spb.DeviceArray((ncell, self.N ** 3))
        self.flags = spb.CONSERVE | spb.SELF | (spb.MAXPRESERVE if self.mp else 0)

    def run(self, form, steps, warmup, tile=None, ncell=None):
        n = ncell or self.ncell
        if tile:
            self.op.set_tile(form, self.mp, tile[1], tile[2], tile[3], kind=tile[0])
        else:
            tile = self.op.autotune(n, form, self.mp)
        for _ in range(warmup):
            self.op.compute_q_dev(self.d_f.ptr, None, self.d_Q.ptr, n, form, self.flags)
        self.op.sync()
        ms, mk = self.op.time_compute_q(self.d_f.ptr, None, self.d_Q.ptr, n, form, self.flags, iters=steps)
        self.op.sync()
        log(f"{self.kind} {form} tile(kind,tr,lines,tb)={tile}: {ms:.3f} ms/step, weighted sum {mk:.3f} ms, "
            f"{n / ms * 1e3:.1f} Q/s")
        return ms, mk, tuple(tile), self.d_Q.download()[:n]

    def errors(self, Q, Q64):
        nrm = np.linalg.norm(Q64, axis=1)
        sig = nrm >= 1e-3 * nrm.max()   # cells whose Q is not round-off (equilibrium cells have Q ~ 0)
        err = np.array([np.linalg.norm(Q[b] - Q64[b]) / nrm[b] for b in range(len(Q)) if sig[b]])
        C = cmat(self.N, self.v)
        cons = float(np.max(np.abs(Q @ C.T) / (np.abs(Q) @ np.abs(C.T))))
        return {"rel_l2_err_vs_fp64_batch": float(np.linalg.norm(Q - Q64) / np.linalg.norm(Q64)),
                "rel_l2_err_vs_fp64_max": float(err.max()), "rel_l2_err_vs_fp64_median": float(np.median(err)),
                "cells_in_err_stats": int(len(err)), "conservation_max_rel": cons}


def flops_per_pair(mp):
    # ComputeQ: complex product (6) + weight (2) + accumulate (2); fused maxPreserve: two products
    return 18.0 if mp else 10.0


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (cells would shard across "
              "GPUs, each holding a replica of the weight table).", file=sys.stderr)
        sys.exit(2)
    import spectralbte_amd as spb

    ncell = a.ncell if a.workload == "ensemble" else min(a.ncell, 256)
    wl = Workload(spb, a.workload, ncell)
    N, n3 = wl.N, wl.N ** 3
    P = float(n3) * n3
    tile = tuple(int(x) for x in a.tile.split(",")) if a.tile else None
    ms, mk, tile, Q = wl.run(a.form, a.steps, a.warmup, tile)
    value = ncell / (ms * 1e-3)

    # ---- roofline of the weighted-sum kernel: bound by vector ALU issue (the
    # cell groups of a batch share each streamed piece of W through L2, so HBM
    # traffic per flop is small).  `peak` is the published vector peak; the
    # issue ceiling of the instructions the kernel actually uses (unpacked FP32
    # or FP64 FMA, 4 cycles per wave instruction) is measured here as well
    flops = flops_per_pair(wl.mp) * P * ncell
    sw = {"fp64": 8, "fp32": 4, "fp16x2": 4, "fp32a": 4}[a.form]
    fg = 8 if a.form == "fp32a" else 16                              # bytes per transformed complex value
    hbm_min = sw * P + ncell * n3 * (fg * (2 if not wl.mp else 4) + 16)  # W once + F/G read + Qhat written
    peak = FP32_VALU_TFLOPS if a.form == "fp32a" else FP64_VALU_TFLOPS
    traffic, tsrc = None, "not profiled for this tile: no HBM byte count"
    tpath = os.path.join(ROOT, "profiles", "wsum_traffic.json")
    tkey = f"{a.workload}:{a.form}:{ncell}:" + ",".join(map(str, tile))
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(tkey)
        except (OSError, ValueError):
            traffic = None
        if traffic is not None:
            tsrc = (f"profiles/wsum_traffic.json[{tkey}]: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this launch, "
                    "scaled by the counters' reading for a copy kernel of known size")
    achieved = flops / (mk * 1e-3) / 1e12
    ceil_tf, ceil_ghz = spb.valu_peak(fp64=a.form != "fp32a")
    # cycles per wave FMA on one SIMD (1024 SIMDs of 64 lanes, 2 flops per FMA)
    # from the FMA-chain kernel at its own clock, and the ceiling that issue
    # rate gives at the 2.4 GHz peak clock the published figures assume; per
    # (r, c) pair and problem the weighted sum issues 10 FMA/MUL instructions
    # for 18 flops (fused maxPreserve; 6 for 10 with one product), against 2
    # flops per FMA; the FP32-product maxPreserve sum at one problem per lane
    # (streaming tiles with tb = 1) takes Gauss's form of the complex products,
    # 8 instructions for the same 18 flops
    cyc = 2.0 * 64 * 1024 * ceil_ghz * 1e9 / (ceil_tf * 1e12)
    ceil_peak = 2.0 * 64 * 1024 * 2.4e9 / cyc / 1e12
    gauss = a.form == "fp32a" and wl.mp and tile[0] >= 1 and tile[3] == 1
    instr = (8.0 if gauss else 10.0) if wl.mp else 6.0
    alg_ceiling = ceil_peak * flops_per_pair(wl.mp) / (2.0 * instr)
    log(f"issue ceiling ({'FP32' if a.form == 'fp32a' else 'FP64'} FMA chains): {ceil_tf:.1f} TFLOP/s at {ceil_ghz:.2f} GHz, "
        f"{cyc:.2f} cycles per wave FMA, {ceil_peak:.1f} TFLOP/s at 2.4 GHz; this instruction mix can reach "
        f"{alg_ceiling:.1f}, the kernel reaches {achieved:.1f}")
    # one cell group per launch (the plain form) reads the table once per cell:
    # below the ridge of peak flops over HBM bandwidth the launch is HBM-bound
    hbm_bound = flops / hbm_min < peak * 1e12 / (HBM_PEAK_GBS * 1e9)
    if hbm_bound:
        main = {"bound": "hbm", "achieved": round(hbm_min / (mk * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm_min / (mk * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    else:
        main = {"bound": "valu", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4)}
    roof = {**main, "traffic": traffic, "traffic_source": tsrc, "valu_tflops": round(achieved, 2),
            "kernel": "k_wsum_stream" if tile[0] >= 1 else "k_wsum", "kernel_ms": round(mk, 4),
            "alg_flops_per_launch": flops,
            "issue_probe_tflops": round(ceil_tf, 2), "issue_probe_clock_ghz": round(ceil_ghz, 3),
            "issue_cycles_per_wave_fma": round(cyc, 3), "issue_ceiling_at_2.4ghz_tflops": round(ceil_peak, 2),
            "issue_ceiling_for_this_instruction_mix_tflops": round(alg_ceiling, 2),
            "frac_of_issue_ceiling": round(achieved / alg_ceiling, 4),
            "hbm_alg_min_bytes_per_launch": hbm_min,
            "hbm_alg_min_gbs": round(hbm_min / (mk * 1e-3) / 1e9, 1), "hbm_peak_gbs": HBM_PEAK_GBS}

    extra = {}
    if not a.no_ref:
        ms64, mk64, tile64, Q64 = wl.run("fp64", 2, 1)
        extra.update({"fp64_batched_q_per_s": round(ncell / (ms64 * 1e-3), 1), "fp64_batched_tile": list(tile64)})
        extra.update(wl.errors(Q, Q64))
        # plain FP64: the reference method -- one cell per launch, so every cell
        # streams the whole FP64 table from HBM (tile tuned for a single cell)
        msp, kp, tp, _ = wl.run("fp64", 4, 1, ncell=1)
        plain = 1.0 / (msp * 1e-3)
        extra.update({"fp64_plain_q_per_s": round(plain, 1), "fp64_plain_tile": list(tp),
                      "fp64_plain_ms_per_q": round(msp, 4), "fp64_plain_w_stream_ms": round(kp, 4),
                      "fp64_plain_w_stream_gbs": round(8.0 * P / (kp * 1e-3) / 1e9, 1),
                      "speedup_vs_plain_fp64": round(value / plain, 2)})

    # the secondary measurements report an error instead of costing the headline line
    rows = None if a.rows is None else set(a.rows.split(","))
    def row(key, label, fn):
        if rows is None or key in rows:
            extra[key] = guarded(label, fn, spb)

    if not a.no_extra:
        row("c1_bkw16", "BKW16 run", c1_bkw16)
        row("c3_shock_forms", "shock collision step per form", c3_shock_forms)
    if not a.no_extra and not a.no_runs:
        row("shock_run", "shock run", shock_run)
        row("ensemble_mixed_run", "mixed ensemble run", mixed_ensemble_run)
        row("landau_table", "Landau table", landau_table)
        row("glance_table", "glancing-collision table", glance_table)
        row("heat_glance_run", "heat-transfer run", heat_glance_run)
    if not a.no_extra and not a.no_surrogate:
        row("surrogate", "surrogate", surrogate_run)
    if a.full_runs:
        row("full_runs", "full-length runs", full_runs)

    cpu = None
    if not a.no_cpu:
        res = guarded("CPU baseline", cpu_baseline, spb, wl)
        if "error" not in res:
            cpu, extra["cpu_baseline_table"] = res

    out = {"metric": "Q(f) evals/s at N=32 (batched FP32 vs plain FP64) at <=1e-6 rel err; HBM GB/s",
           "value": round(value, 1), "unit": "Q evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32" if a.form == "fp32a" else "f64",
           "weights_dtype": {"fp64": "f64", "fp32": "f32", "fp16x2": "f16x2", "fp32a": "f32"}[a.form],
           "data": "synthetic",
           "config": {"workload": wl.desc, "ncell": ncell, "N": N, "L_v": wl.Lv, "form": a.form,
                      "tile_kind_tr_lines_tb": list(tile)},
           "roofline": roof, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out), flush=True)


def _state_change_error(f0, fa, fb):
    """Relative L2 difference of two runs' state changes from the same start:
    |(fa - f0) - (fb - f0)| / |fb - f0| (fb the FP64 run)."""
    d = fb - f0
    return float(np.linalg.norm(fa - fb) / np.linalg.norm(d))


def c1_bkw16(spb, steps=100):
    """The BKW relaxation (plain FP64 method): one 0D problem, N=16, L_v=5,
    Maxwell molecules (lambda=0), the BKW state at K=0.7 (src/initializer.c:
    123-135), Kn=1, dt=0.01, 100 Heun steps of ComputeQ_maxPreserve +
    projection (exec/boltz.c:186-250), each step one HIP graph; and one
    evaluation timed alone for the weight stream.  The FP64 table (134 MB) fits
    in the 256 MB Infinity Cache, so the stream can read faster than HBM."""
    from spectralbte_amd import workloads
    N, Lv, dt, K = 16, 5.0, 0.01, 0.7
    v, eta = workloads.grid_0d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(0.0)
    f0 = workloads.bkw_state(N, Lv, K)
    run = spb.Run.ensemble(op, f0[None], order=2, dt=dt, Kn=1.0, form="fp64")
    run.use_graph(True)
    run.step(1)
    run.set_state(f0[None])
    op.sync()
    t0 = time.perf_counter()
    run.step(steps)
    sec = time.perf_counter() - t0
    m0 = run.moments()
    C = cmat(N, v)
    fe = run.state()[0]
    drift = np.abs(C @ fe - C @ f0) / (np.abs(C) @ np.abs(f0))
    run.close()
    d_f = spb.DeviceArray.from_host(f0[None])
    d_Q = spb.DeviceArray((1, N ** 3))
    flags = spb.SELF | spb.MAXPRESERVE | spb.CONSERVE
    op.compute_q_dev(d_f.ptr, None, d_Q.ptr, 1, "fp64", flags)
    ms, mk = op.time_compute_q(d_f.ptr, None, d_Q.ptr, 1, "fp64", flags, iters=50)
    op.close()
    P = float(N) ** 6
    log(f"BKW16 run: {steps} Heun steps in {sec * 1e3:.1f} ms, {2 * steps / sec:.0f} Q/s; one Q {ms * 1e3:.1f} us, "
        f"W stream {mk * 1e3:.1f} us")
    return {"workload": "0D-3V BKW relaxation, N=16, L_v=5, lambda=0, K=0.7, Kn=1, dt=0.01, 100 Heun steps of "
                        "ComputeQ_maxPreserve + projection, one HIP graph per step",
            "form": "fp64 (plain: one problem per pass over W)", "run_ms": round(sec * 1e3, 2),
            "q_per_s": round(2 * steps / sec, 1), "single_q_us": round(ms * 1e3, 2),
            "single_q_w_stream_us": round(mk * 1e3, 2), "w_stream_gbs": round(8.0 * P / (mk * 1e-3) / 1e9, 1),
            "w_table_bytes": 8.0 * P,
            "note": "the 134 MB FP64 table fits in the 256 MB Infinity Cache: the rate is not an HBM rate",
            "final_rho_u_T": [round(float(x), 12) for x in m0[0][[0, 1, 4]]],
            "conserved_moment_drift_max_rel": float(drift.max()),
            "parity": "tests/test_gpu_configs.py::test_c1_bkw16_run_against_oracle (moments 1e-10 vs the oracle's driver)"}


def c3_shock_forms(spb, ncell=256):
    """The shock collision step (ComputeQ + projection of 256 cells, N=32,
    L_v=8, hard spheres, Mach-1.2 profile) in every precision form, each with
    its error against the FP64 form on the same cells.  ComputeQ keeps Q(M, M),
    so in near-equilibrium cells Q is a small residual of large gain and loss
    terms and FP32 rounding of W is amplified."""
    wl = Workload(spb, "shock", ncell)
    out, Qs = {"workload": wl.desc}, {}
    for form in ("fp64", "fp32", "fp16x2", "fp32a"):
        ms, mk, t, Q = wl.run(form, 3, 1)
        Qs[form] = Q
        out[form] = {"q_per_s": round(ncell / (ms * 1e-3), 1), "ms_per_step": round(ms, 3),
                     "kernel_ms": round(mk, 3), "tile": list(t)}
    for form in ("fp32", "fp16x2", "fp32a"):
        e = wl.errors(Qs[form], Qs["fp64"])
        out[form].update(e)
        out[form]["meets_1e-6"] = bool(e["rel_l2_err_vs_fp64_batch"] <= 1e-6)
    out["fp64"]["meets_1e-6"] = True
    out["fp64"]["parity"] = "tests/test_gpu_configs.py::test_shock_cells_n32_against_oracle (1e-12 vs the oracle)"
    wl.op.close()
    return out


def shock_run(spb, steps=3):
    """Whole steps of the 1D-3V Mach-1.2 shock (256 cells on a domain of length
    workloads.SHOCK_LENGTH = 10, N=32, L_v=8, lambda=1, Kn=0.1, order 2,
    dt=0.002): advectTwo, Heun collision step (2 Q per cell),
    advectTwo, replayed as one captured HIP graph per step; in FP64 (the form
    that meets 1e-6 on ComputeQ near equilibrium) and in the FP32-product form,
    with that run's state change against the FP64 run's."""
    from spectralbte_amd import workloads
    N, Lv, nX = 32, 8.0, 256
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    x, dx = _mesh(nX, workloads.SHOCK_LENGTH, 2)
    f0 = workloads.shock_profile(N, Lv, nX)
    res, states = {}, {}
    for form in ("fp64", "fp32a"):
        run = spb.Run.mesh(op, f0, x, dx, order=2, dt=0.002, Kn=0.1, bc=spb.BC_COPY, form=form)
        op.autotune(nX, form, False)
        run.use_graph(True)
        run.step(1)
        run.set_state(f0)
        op.sync()
        t0 = time.perf_counter()
        run.step(steps)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        states[form] = run.state()
        run.close()
        log(f"shock run {form}: {ms:.2f} ms/step (graph), {2 * nX / ms * 1e3:.1f} Q/s")
        res[form] = {"ms_per_step": round(ms, 3), "q_per_s": round(2 * nX / ms * 1e3, 1),
                     "projected_500_steps_s": round(500 * ms / 1e3, 1), "finite": bool(np.all(np.isfinite(states[form])))}
    res["fp32a"]["state_change_rel_err_vs_fp64"] = _state_change_error(f0, states["fp32a"], states["fp64"])
    op.close()
    return {"workload": f"1D-3V Mach-1.2 shock, 256 cells on length {workloads.SHOCK_LENGTH:g}, N=32, L_v=8, "
                        "lambda=1, Kn=0.1, order 2 (advectTwo + Heun collision + advectTwo), dt=0.002, whole step "
                        "as one HIP graph",
            "steps_timed": steps, **res}


def landau_table(spb, N=32, Lv=7.0, lam=1.0):
    """The anisotropic generator's glance-0 (Landau) table on the GPU, on the
    1D grid of the heat-transfer configuration (N=32, L_v=7, lambda=1):
    N^6 = 1.07e9 entries, each 64 x 64 Gauss-Legendre nodes (the two theta
    nodes of a symmetric pair give equal terms: 2048 J0 and sincos
    evaluations)."""
    from spectralbte_amd import workloads
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    t0 = time.perf_counter()
    op.generate_weights_landau(lam)
    sec = time.perf_counter() - t0
    op.close()
    n = float(N) ** 6
    log(f"Landau table N={N}: {sec:.2f} s, {n / sec:.3g} entries/s")
    return {"workload": f"Landau (glance 0) weight table, N={N}, L_v={Lv}, lambda={lam}, 1D grid", "seconds": round(sec, 2),
            "entries_per_s": round(n / sec, 1), "gl_nodes_per_s": round(n * 4096 / sec, 1)}


def glance_table(spb, N=32, Lv=7.0, lam=1.0, glance=0.01):
    """The anisotropic generator's glancing-collision (glance > 0) table on the
    GPU, on the same grid: N^6 = 1.07e9 entries, each a 2-D sum over at most
    112 theta x 33 psi nodes (fewer for rows of smaller |zeta|) of one
    tabulated radial function (the reference nests three adaptive quadratures
    per entry)."""
    from spectralbte_amd import workloads
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    t0 = time.perf_counter()
    op.generate_weights_glance(lam, glance)
    sec = time.perf_counter() - t0
    op.close()
    n = float(N) ** 6
    log(f"glance table N={N}: {sec:.2f} s, {n / sec:.3g} entries/s")
    return {"workload": f"glancing-collision weight table, N={N}, L_v={Lv}, lambda={lam}, glance={glance}, 1D grid",
            "seconds": round(sec, 2), "entries_per_s": round(n / sec, 1)}


def heat_glance_run(spb, steps=3):
    """Whole steps of the 1D-3V heat transfer between diffuse walls at T=1 and
    T=2 with the anisotropic glancing-collision table (grazing parameter 0.1):
    128 cells, N=32, L_v=7, lambda=1, Kn=0.05, order 1, a Maxwellian at rho=1,
    u=0, T=1.5 with 1% seeded density noise, autotuned tiles; in FP64 and in
    the FP32-product form, with the latter's error against FP64 on the
    collision operator of the initial state and on the runs' state change.
    The table carries no trapezoid or h_eta^3 factors, as the reference writes
    it, so Q is about 1e4 times the isotropic table's and the explicit step
    needs dt = workloads.HEAT_GLANCE_DT = 1e-7 (1e-6 and 3e-7 diverge within
    200 steps; 2000 steps cover 2e-4 time units)."""
    from spectralbte_amd import workloads
    N, Lv, nX = 32, 7.0, 128
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_glance(1.0, 0.1)
    x, dx = _mesh(nX, 1.0, 1)
    noise = 1.0 + 0.01 * np.random.default_rng(128).standard_normal(nX)
    V2 = (v[:, None, None] ** 2 + v[None, :, None] ** 2 + v[None, None, :] ** 2).ravel()
    f0 = np.stack([r * np.exp(-V2 / 1.5) / (1.5 * np.pi) ** 1.5 for r in noise])  # src/initializer.c:366
    res, states, qs = {}, {}, {}
    for form in ("fp64", "fp32a"):
        tile = op.autotune(nX, form, False)
        qs[form] = op.compute_q(f0, form=form, conserve=True)
        run = spb.Run.mesh(op, f0, x, dx, order=1, dt=workloads.HEAT_GLANCE_DT, Kn=0.05, bc=spb.BC_DIFFUSE, form=form)
        run.use_graph(True)
        run.step(1)
        run.set_state(f0)
        op.sync()
        t0 = time.perf_counter()
        run.step(steps)
        ms = (time.perf_counter() - t0) * 1e3 / steps
        states[form] = run.state()
        run.close()
        log(f"heat-transfer run (glance 0.1) {form}: {ms:.2f} ms/step, {nX / ms * 1e3:.1f} Q/s")
        res[form] = {"ms_per_step": round(ms, 3), "q_per_s": round(nX / ms * 1e3, 1), "tile": list(tile),
                     "projected_2000_steps_s": round(2000 * ms / 1e3, 1),
                     "finite": bool(np.all(np.isfinite(states[form])))}
    e_q = float(np.linalg.norm(qs["fp32a"] - qs["fp64"]) / np.linalg.norm(qs["fp64"]))
    res["fp32a"]["q_rel_l2_err_vs_fp64_initial_state"] = e_q
    res["fp32a"]["state_change_rel_err_vs_fp64"] = _state_change_error(f0, states["fp32a"], states["fp64"])
    res["fp32a"]["meets_1e-6"] = bool(e_q <= 1e-6)
    op.close()
    return {"workload": "1D-3V heat transfer, 128 cells, N=32, L_v=7, lambda=1, glancing-collision table "
                        f"(glance 0.1), Kn=0.05, order 1, diffuse walls T=1 and T=2, dt={workloads.HEAT_GLANCE_DT:g}, "
                        "whole step as one HIP graph",
            "steps_timed": steps, **res}


def full_runs(spb):
    """Configurations 2-4 at their full lengths, each as one timed sequence of
    graph-replayed steps: the mixed ensemble's 50 Heun steps (FP32 products and
    FP64, conservation of every problem and the FP32 run's state change
    against FP64), the shock's 500 steps in FP64 (the far-field states, which
    the Rankine-Hugoniot jump keeps, and the profile's range), and the heat
    transfer's 2000 steps in FP64 (mass between the diffuse walls)."""
    from spectralbte_amd import workloads
    out = {}
    # C2: 512 problems at N=24 and 512 at N=32, L_v=6, lambda=1, Heun dt=0.005, 50 steps
    c2 = {}
    for N, seed in ((24, 0x5EED0024), (32, 0x5EED0032)):
        v, eta = workloads.grid_0d(N, 6.0)
        op = spb.Operator(N, 6.0, v, eta)
        op.generate_weights_iso(1.0)
        f0 = workloads.two_maxwellian_ensemble(N, 6.0, 512, seed=seed)
        C = cmat(N, v)
        ends = {}
        for form in ("fp32a", "fp64"):
            op.autotune(512, form, True)
            r = spb.Run.ensemble(op, f0, order=2, dt=0.005, Kn=1.0, form=form)
            r.use_graph(True)
            r.step(1)  # captures the step's graph
            r.set_state(f0)
            op.sync()
            t0 = time.perf_counter()
            r.step(50)
            sec = time.perf_counter() - t0
            ends[form] = r.state()
            r.close()
            drift = np.abs(ends[form] @ C.T - f0 @ C.T) / (np.abs(f0) @ np.abs(C).T)
            c2[f"N{N}_{form}"] = {"seconds": round(sec, 2), "q_per_s": round(2 * 512 * 50 / sec, 1),
                                  "conserved_moment_drift_max_rel": float(drift.max()),
                                  "finite": bool(np.all(np.isfinite(ends[form])))}
            log(f"full C2 N={N} {form}: 50 steps in {sec:.2f} s, drift {drift.max():.2e}")
        c2[f"N{N}_fp32a"]["state_change_rel_err_vs_fp64"] = _state_change_error(f0, ends["fp32a"], ends["fp64"])
        op.close()
    out["c2_mixed_ensemble_50_steps"] = c2
    # C3: 256-cell Mach-1.2 shock, N=32, L_v=8, Kn=0.1, order 2, dt=0.002, 500 steps, FP64
    N, Lv, nX = 32, 8.0, 256
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    x, dx = _mesh(nX, workloads.SHOCK_LENGTH, 2)
    f0 = workloads.shock_profile(N, Lv, nX)
    op.autotune(nX, "fp64", False)
    r = spb.Run.mesh(op, f0, x, dx, order=2, dt=0.002, Kn=0.1, bc=spb.BC_COPY, form="fp64")
    r.use_graph(True)
    r.step(1)
    r.set_state(f0)
    m0 = r.moments()
    op.sync()
    t0 = time.perf_counter()
    r.step(500)
    sec = time.perf_counter() - t0
    m1 = r.moments()
    fin = bool(np.all(np.isfinite(r.state())))
    r.close()
    op.close()
    far = {"upstream_start": [round(float(q), 6) for q in m0[0][[0, 1, 4]]],
           "upstream_end": [round(float(q), 6) for q in m1[0][[0, 1, 4]]],
           "downstream_start": [round(float(q), 6) for q in m0[-1][[0, 1, 4]]],
           "downstream_end": [round(float(q), 6) for q in m1[-1][[0, 1, 4]]]}
    out["c3_shock_500_steps_fp64"] = {"seconds": round(sec, 2), "q_per_s": round(2 * nX * 500 / sec, 1),
                                      "ms_per_step": round(sec * 2.0, 3), "finite": fin,
                                      "rho_range_end": [float(m1[:, 0].min()), float(m1[:, 0].max())],
                                      "T_range_end": [float(m1[:, 4].min()), float(m1[:, 4].max())],
                                      "far_field_rho_ux_T": far}
    log(f"full C3: 500 steps in {sec:.1f} s; far field {far}")
    # C4: heat transfer between diffuse walls, glance 0.1, 128 cells, dt=1e-7, 2000 steps, FP64
    N, Lv, nX = 32, 7.0, 128
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_glance(1.0, 0.1)
    x, dx = _mesh(nX, 1.0, 1)
    noise = 1.0 + 0.01 * np.random.default_rng(128).standard_normal(nX)
    V2 = (v[:, None, None] ** 2 + v[None, :, None] ** 2 + v[None, None, :] ** 2).ravel()
    f0 = np.stack([rr * np.exp(-V2 / 1.5) / (1.5 * np.pi) ** 1.5 for rr in noise])
    op.autotune(nX, "fp64", False)
    r = spb.Run.mesh(op, f0, x, dx, order=1, dt=workloads.HEAT_GLANCE_DT, Kn=0.05, bc=spb.BC_DIFFUSE, form="fp64")
    r.use_graph(True)
    r.step(1)
    r.set_state(f0)
    m0 = r.moments()
    op.sync()
    t0 = time.perf_counter()
    r.step(2000)
    sec = time.perf_counter() - t0
    m1 = r.moments()
    fin = bool(np.all(np.isfinite(r.state())))
    r.close()
    op.close()
    d = dx[1:-1]
    mass0, mass1 = float(m0[:, 0] @ d), float(m1[:, 0] @ d)
    out["c4_heat_2000_steps_fp64"] = {"seconds": round(sec, 2), "q_per_s": round(nX * 2000 / sec, 1),
                                      "ms_per_step": round(sec / 2.0, 3), "finite": fin,
                                      "mass_rel_change": abs(mass1 - mass0) / mass0,
                                      "T_wall_cells_end": [float(m1[0, 4]), float(m1[-1, 4])]}
    log(f"full C4: 2000 steps in {sec:.1f} s, mass change {abs(mass1 - mass0) / mass0:.2e}")
    return out


def _mesh(nX, length, order):
    """Uniform cells on [0, length] with `order` ghost cells per side (src/mesh_setup.c:121-175)."""
    d = length / nX
    x = (np.arange(nX + 2 * order) - order + 0.5) * d
    return x, np.full(nX + 2 * order, d)


def mixed_ensemble_run(spb, form="fp32a", ncell=512, steps=2):
    """The whole 0D ensemble: 512 problems at N=24 and 512 at N=32 (L_v=6,
    lambda=1), Heun steps of dt=0.005 (2 Q per problem per step), timed with the
    two grids' runs queued on their own streams (overlapping) and one after the
    other; the faster schedule is reported, with the FP32-product runs' state
    change against FP64 runs of the same steps."""
    from spectralbte_amd import workloads
    runs, ops, f0s = [], [], []
    for N, seed in ((24, 0x5EED0024), (32, 0x5EED0032)):
        v, eta = workloads.grid_0d(N, 6.0)
        op = spb.Operator(N, 6.0, v, eta)
        op.generate_weights_iso(1.0)
        op.autotune(ncell, form, True)
        f0 = workloads.two_maxwellian_ensemble(N, 6.0, ncell, seed=seed)
        r = spb.Run.ensemble(op, f0, order=2, dt=0.005, Kn=1.0, form=form)
        r.use_graph(True)
        r.step(1)
        runs.append(r)
        ops.append(op)
        f0s.append(f0)
    for r, f0 in zip(runs, f0s):
        r.set_state(f0)
    t0 = time.perf_counter()
    for r in runs:
        r.step(steps, sync=False)
    for op in ops:
        op.sync()
    ms_overlap = (time.perf_counter() - t0) * 1e3 / steps
    st32 = [r.state() for r in runs]
    for r, f0 in zip(runs, f0s):
        r.set_state(f0)
    t0 = time.perf_counter()
    for r in runs:  # one grid after the other
        r.step(steps)
    ms_seq = (time.perf_counter() - t0) * 1e3 / steps
    ms = min(ms_overlap, ms_seq)
    per = {"overlapped_ms_per_step": round(ms_overlap, 2), "sequential_ms_per_step": round(ms_seq, 2)}
    for N, r in zip((24, 32), runs):
        t0 = time.perf_counter()
        r.step(1)
        per[f"N{N}_ms_per_step"] = round((time.perf_counter() - t0) * 1e3, 2)
        r.close()
    errs = {}
    for N, op, f0, st in zip((24, 32), ops, f0s, st32):
        r = spb.Run.ensemble(op, f0, order=2, dt=0.005, Kn=1.0, form="fp64")
        r.step(steps)
        errs[f"N{N}_state_change_rel_err_vs_fp64"] = _state_change_error(f0, st, r.state())
        r.close()
        op.close()
    q = 2 * 2 * ncell / ms * 1e3
    log(f"mixed ensemble {form}: {ms:.2f} ms/step for 2x{ncell} problems, {q:.1f} Q/s; {errs}")
    return {"workload": f"0D-3V ensemble, {ncell} problems at N=24 + {ncell} at N=32, L_v=6, lambda=1, Heun dt=0.005, "
                        "ComputeQ_maxPreserve + projection, the faster of overlapped and sequential grids",
            "form": form, "ms_per_step": round(ms, 2), "q_per_s": round(q, 1), **per, **errs}


def surrogate_run(spb, form="fp64", ncell=4096, train_cells=2048, steps=2500, batch=64, check_every=8):
    """Learned surrogate on the N=32, L_v=8 hard-sphere grid: trained on the GPU
    on exact Q of shock-profile cells, judged on held-out cells, then timed on a
    4096-cell batch alone and inside the checked collision step: groups of 8
    cells, one checked exactly per group (tolerance 1e-3, error floored at 0.1
    of the largest checked |Q|), a failing group evaluated exactly.  The exact
    baseline is the FP64 form (on ComputeQ near equilibrium the one that meets
    1e-6).  For comparison, the groups a zero prediction would pass under the
    same check are counted from the exact Q."""
    from spectralbte_amd import surrogate as sg, workloads
    N, Lv, Kb, C, r, L = 32, 8.0, 24, 64, 128, 4
    v, eta = workloads.grid_1d(N, Lv)
    op = spb.Operator(N, Lv, v, eta)
    op.generate_weights_iso(1.0)
    net = spb.Net(op, Kb, C, r, L, maxpreserve=False, seed=1)
    npar = net.npar
    ftr = sg.shock_cells(N, Lv, train_cells, 1)
    fte = sg.shock_cells(N, Lv, 256, 1001)
    d_f = spb.DeviceArray.from_host(ftr[:256])
    d_Q = spb.DeviceArray((256, N ** 3))
    net.calibrate(d_f.ptr, d_Q.ptr, 256)
    d_f.free(); d_Q.free()
    t0 = time.perf_counter()
    sg.build_dataset(spb, op, net, ftr)
    t_data = time.perf_counter() - t0
    log(f"surrogate: dataset of {train_cells} cells in {t_data:.1f} s; training")
    t0 = time.perf_counter()
    k1 = int(0.7 * steps)  # learning rate 1e-3, then 3e-4 for the last 30%
    losses = [net.train(k1, batch, 1e-3, seed=1)]
    log(f"surrogate: {k1} Adam steps, loss {losses[-1][-1]:.3g}")
    losses.append(net.train(steps - k1, batch, 3e-4, seed=2))
    losses = np.concatenate(losses)
    t_train = time.perf_counter() - t0
    rel, flo, qn = sg.judged_errors(spb, op, net, fte)
    sig = qn >= 1e-2
    log(f"surrogate: trained in {t_train:.1f} s; judging and timing on {ncell} cells")
    f = workloads.shock_profile(N, Lv, ncell, mach=1.2, blend_cells=20.0 * ncell / 256)
    t_pred = sg.time_predict(spb, op, net, f)
    fl = sg.forward_flops(Kb ** 3, C, r, L) * ncell
    d_f = spb.DeviceArray.from_host(f)
    d_Q = spb.DeviceArray(f.shape)
    flags = spb.SELF | spb.CONSERVE
    exact = {}
    for fm in (form, "fp32a"):  # default tiles: autotuning 4096 cells would take minutes
        op.compute_q_dev(d_f.ptr, None, d_Q.ptr, ncell, fm, flags)
        op.sync()
        t0 = time.perf_counter()
        op.compute_q_dev(d_f.ptr, None, d_Q.ptr, ncell, fm, flags)
        op.sync()
        exact[fm] = time.perf_counter() - t0
        if fm == form:
            Qe = d_Q.download()
        log(f"surrogate: exact {fm} {ncell / exact[fm]:.0f} Q/s")
    net.collide_dev(d_f.ptr, d_Q.ptr, ncell, form, check_every=check_every, phase=0, tol=1e-3, floor_rel=0.1)
    op.sync()
    t0 = time.perf_counter()
    n_exact, err = net.collide_dev(d_f.ptr, d_Q.ptr, ncell, form, check_every=check_every, phase=0, tol=1e-3,
                                   floor_rel=0.1)
    op.sync()
    t_coll = time.perf_counter() - t0
    Qc = d_Q.download()
    step_err = float(np.linalg.norm(Qc - Qe) / np.linalg.norm(Qe))
    # a zero prediction under the same check: a group passes when its checked
    # cell's |Q| is at most tol times the floor
    checked = np.arange(0, ncell, check_every)
    qn_c = np.linalg.norm(Qe[checked], axis=1)
    zero_pass = int(np.sum(qn_c <= 1e-3 * 0.1 * qn_c.max()))
    net.close()
    op.close()
    served = ncell - n_exact
    log(f"surrogate: train {steps} steps {t_train:.1f}s, loss {losses[-1]:.3g}; judged median rel err "
        f"{np.median(rel[sig]):.3g}; predict {ncell / t_pred:.0f} Q/s ({fl / t_pred / 1e12:.1f} TFLOP/s); "
        f"checked step: {served} of {ncell} cells served by the network, {n_exact} exact, max checked err {err:.3g}, "
        f"{ncell / t_coll:.0f} Q/s against {ncell / exact[form]:.0f} exact ({form}); a zero prediction passes "
        f"{zero_pass} of {len(checked)} groups")
    return {"workload": f"learned surrogate, N=32, L_v=8, lambda=1, {Kb}^3 kept modes, width {C}, rank {r}, {L} layers "
                        f"({npar} parameters), trained {steps} Adam steps of {batch} "
                        f"cells on {train_cells} shock-profile cells; timed on {ncell} cells (Mach 1.2 profile)",
            "data_s": round(t_data, 2), "train_s": round(t_train, 2), "final_train_loss": float(losses[-1]),
            "judged_rel_err_median_significant_cells": float(np.median(rel[sig])),
            "judged_err_floor0.1_median": float(np.median(flo)), "judged_err_floor0.1_max": float(flo.max()),
            "predict_q_per_s": round(ncell / t_pred, 1),
            "predict_mfma_tflops": round(fl / t_pred / 1e12, 2), "mfma_peak_bf16_tflops": 2500.0,
            "exact_q_per_s": round(ncell / exact[form], 1), "exact_form": form,
            "exact_fp32a_q_per_s": round(ncell / exact["fp32a"], 1),
            "checked_step_q_per_s": round(ncell / t_coll, 1), "checked_step_cells_exact": int(n_exact),
            "checked_step_cells_served_by_network": int(served),
            "checked_step_max_checked_err": float(err),
            "checked_step_rel_l2_err_vs_exact_batch": step_err,
            "zero_prediction_groups_passing": zero_pass, "groups": int(len(checked)),
            "check": f"groups of {check_every} cells, one checked per group, tol 1e-3, floor 0.1 x max checked |Q|; "
                     "a failing group is evaluated exactly"}


def _cpu_model():
    """lscpu's model name and socket count (/proc/cpuinfo's model name without lscpu)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        info = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        model = info.get("Model name", "").strip()
        sockets = info.get("Socket(s)", "").strip()
        if model:
            return f"{model}, {sockets} socket(s)" if sockets else model
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spb, wl):
    """The reference method's CPU restatement (oracle/spbte_oracle.c: the
    output-row-parallel FP64 loop nest of src/collisions.c:127-165 with OpenMP
    over rows, W as the reference holds it, the twiddled transform and the 5x5
    projection), built here -O3 -march=native, timed on this host with all the
    threads this process may use and with 1 thread: ComputeQ and
    ComputeQ_maxPreserve (three passes over W, as the reference) + projection
    at N=16 (the BKW grid), 24 and 32 (the ensemble grids), on the workloads'
    own states.  Returns (the headline workload's figure, the table)."""
    import tempfile

    import oracle
    from spectralbte_amd import workloads
    tmp = tempfile.mkdtemp(prefix="spbte_cpu_")
    try:
        oracle.use_library(oracle.build_native(tmp))
        build = "gcc -O3 -march=native -fopenmp (compiled on this host)"
    except Exception as e:  # noqa: BLE001 -- fall back to the portable build
        log(f"native oracle build failed ({e}); timing the portable build")
        oracle.build()
        build = "gcc -O2 -fopenmp (portable build)"
    allt = max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "0")) or 1 << 30))
    rows = []
    cases = [(16, 5.0, 0.0, "BKW K=0.7"), (24, 6.0, 1.0, "two-Maxwellian ensemble"),
             (32, 6.0, 1.0, "two-Maxwellian ensemble")]
    headline = None
    for N, Lv, lam, what in cases:
        if N == wl.N and Lv == wl.Lv:
            W, v, eta, f = wl.op.get_weights(), wl.v, wl.eta, wl.f[:3]
        else:
            v, eta = workloads.grid_0d(N, Lv)
            op = spb.Operator(N, Lv, v, eta)
            op.generate_weights_iso(lam)
            W = op.get_weights()
            op.close()
            f = (workloads.bkw_state(N, Lv, 0.7)[None].repeat(3, 0) if lam == 0.0
                 else workloads.two_maxwellian_ensemble(N, Lv, 3, seed=0x5EED0000 + N))
        cons = oracle.Conserver(N, v)
        for mp in (False, True):
            fn = oracle.compute_q_maxpreserve if mp else oracle.compute_q
            for threads in (allt, 1):
                reps = 3 if (threads > 1 or N <= 16) else 1
                t0 = time.perf_counter()
                for b in range(reps):
                    cons(fn(N, v, eta, W, f[b], f[b], nthreads=threads).ravel())
                dt = time.perf_counter() - t0
                row = {"N": N, "L_v": Lv, "lambda": lam, "state": what,
                       "operator": "ComputeQ_maxPreserve" if mp else "ComputeQ", "threads": threads,
                       "evaluations": reps, "seconds": round(dt, 3), "q_per_s": round(reps / dt, 4)}
                rows.append(row)
                log(f"CPU N={N} {row['operator']} {threads} thread(s): {reps / dt:.3f} Q/s")
                if N == wl.N and mp == wl.mp and threads == allt:
                    headline = row
        del W
    shutil_rm(tmp)
    table = {"what": "reference-method CPU restatement (oracle/spbte_oracle.c), not the reference binary",
             "build": build, "cpu_model": _cpu_model(), "logical_cpus_visible": len(os.sched_getaffinity(0)),
             "rows": rows}
    sh = [r for r in rows if r["N"] == 32 and r["operator"] == "ComputeQ" and r["threads"] == allt]
    if sh:
        table["shock_step_extrapolated_s"] = round(2 * 256 / sh[0]["q_per_s"], 1)
        table["shock_step_note"] = ("one 1D shock step (256 cells, 2 ComputeQ each) extrapolated from the N=32 "
                                    "ComputeQ rate at L_v=6; the transport stages are not timed")
    if headline is None:  # a headline grid outside the table: the N=32 maxPreserve row stands in
        headline = [r for r in rows if r["N"] == 32 and r["operator"] == "ComputeQ_maxPreserve"
                    and r["threads"] == allt][0]
    op_name = "ComputeQ_maxPreserve (three passes over W, as the reference)" if wl.mp else "ComputeQ"
    cpu = {"value": headline["q_per_s"], "unit": "Q evals/s", "cores": headline["threads"], "kind": "port",
           "sample": f"{headline['evaluations']} problems of the headline workload ({op_name} + projection, N=32, "
                     f"FP64) through oracle/spbte_oracle.c, {build}, {headline['threads']} OpenMP threads over "
                     f"output rows on {table['cpu_model']}, {headline['seconds']} s; the 1-thread and N=16/24 "
                     "figures are in cpu_baseline_table"}
    return cpu, table


def shutil_rm(path):
    import shutil
    shutil.rmtree(path, ignore_errors=True)


if __name__ == "__main__":
    main()
