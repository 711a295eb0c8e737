"""Benchmark: FP64 SpTRSV on the 64^3 7-point Laplacian (one MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]

A step is one forward triangular solve x = L^-1 b with the lower triangle L of
the 3D 7-point Laplacian on a 64^3 grid in natural ordering (n = 262144,
nnz(L) = 1036288), one right-hand side b ~ U(-1,1) (seed 0), every input
resident in HBM.  value = algorithmic GB/s = (12 nnz(L) + 20 n) bytes per
solve x K solves / the timed region's length (the reference's byte model,
test/iccg_level_merge.cpp:354-355).  The execution plan is autotuned before
the timed region (analysis time, reported separately, as the reference's
Scheduling_Time).

Further fields (not the headline): the same solve on the seed-1 random
symmetric permutation, 32 right-hand sides, IC0 and ILU0 (64^3), and ICCG
time to ||r||/||r0|| < 1e-8 on the kNN stiffness matrix (n = 36000) with
the FP64 factor and with the FP32 factor.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SpTRSV GB/s (% HBM peak); ICCG time-to-1e-8, FP32-precond vs FP64"


def bytes_sptrsv(n, nnz, k=1, vbytes=8):
    """values + colidx per nonzero, rowptr per row, b read + x written per row and RHS."""
    return (vbytes + 4) * nnz + 4 * n + 2 * vbytes * n * k


def time_events(fn, steps, warmup):
    """Time `steps` back-to-back calls on torch's current stream (the stream
    every library call of this process is issued on) with HIP events around
    each call.  Returns (whole-region ms, mean ms per call, median ms per call);
    the region runs from the first call's start event to the last call's end
    event, bracketed by device synchronisation."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in evs:
        a.record(st)
        fn()
        b.record(st)
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b in evs]
    total = evs[0][0].elapsed_time(evs[-1][1])
    return float(total), float(np.mean(ms)), float(np.median(ms))


def cpu_baseline(L, b, seconds):
    """The oracle's serial forward solve, timed on one host core."""
    import oracle as O
    O.sptrsv_lower(L, b)
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < seconds:
        O.sptrsv_lower(L, b)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    return {"value": round(bytes_sptrsv(L.n, L.nnz) / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{reps} serial FP64 forward solves of the 64^3 L (oracle/oracle.c orc_sptrsv_lower, "
                      f"ascending FMA sums), {dt * 1e3:.3f} ms per solve, ~{seconds:.0f} s"}


def traffic_from_profiles():
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/traffic.json), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None, None
    with open(p) as f:
        d = json.load(f)
    return d.get("sptrsv_lap3d64_bytes_per_launch"), d.get("plan")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only, no secondary measurements")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--plan", default="auto",
                    help="auto (autotune) or MODE:ROWS_PER_UNIT:BLOCK:DATAFLOW, MODE 1 = RANGE, 2 = SCHEDULE")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run 3 plain copy kernels of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build has no multi-GPU path (DESIGN.md explains how the "
              "solve would shard)", file=sys.stderr)
        sys.exit(2)
    import torch

    from baas_amd import _lib
    from baas_amd import matgen as g
    from baas_amd.api import PCG, Handle, Schedule

    torch.cuda.set_device(0)
    if args.calib_copy_mib:
        src = torch.ones(args.calib_copy_mib << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)  # reads and writes calib_copy_mib MiB
        torch.cuda.synchronize()
        del src, dst
    h = Handle()
    A = g.laplace3d(64)
    L = A.lower()
    b = g.rhs_uniform(L.n, 0)
    t0 = time.perf_counter()
    s = Schedule(h, L.rowptr, L.colidx, kind=_lib.KIND_TRSV)
    t_an = time.perf_counter() - t0
    lv = torch.as_tensor(L.values, device="cuda")
    bd = torch.as_tensor(b, device="cuda")
    x = torch.empty_like(bd)
    t0 = time.perf_counter()
    if args.plan == "auto":
        plan = s.autotune(lv)
    else:
        mode, rpu, blk, df = (int(v) for v in args.plan.split(":"))
        s.replan(mode, rpu, blk, dataflow=df)
        plan = {"mode": mode, "rows_per_unit": rpu, "block": blk, "dataflow": df, "us": None}
    t_tune = time.perf_counter() - t0
    print(f"[bench] analysis {t_an:.3f} s, autotune {t_tune:.2f} s, plan {plan}", file=sys.stderr, flush=True)

    total_ms, mean_ms, med_ms = time_events(lambda: s.solve(lv, bd, x), args.steps, args.warmup)
    h.check()
    by = bytes_sptrsv(L.n, L.nnz)
    traffic, traffic_plan = traffic_from_profiles()
    gbs = by * args.steps / (total_ms * 1e-3) / 1e9
    kern_gbs = by / (mean_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (3D 7-point Laplacian 64^3, b ~ U(-1,1) seed 0)",
        "config": {"workload": "FP64 forward SpTRSV x = L^-1 b, L = lower triangle of the 3D 7-point Laplacian "
                               "on a 64^3 grid, natural ordering, 1 right-hand side",
                   "n": L.n, "nnz_L": L.nnz, "bytes_per_solve": by, "plan": plan,
                   "analysis_s": round(t_an, 4), "autotune_s": round(t_tune, 3)},
        "roofline": {"bound": "hbm", "achieved": round(kern_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kern_gbs / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "traffic_plan": traffic_plan,
                     "kernel": "k_sweep_df (dataflow)" if plan["dataflow"] else "k_sweep",
                     "algorithmic_bytes_per_launch": by, "mean_launch_ms": round(mean_ms, 5),
                     "median_launch_ms": round(med_ms, 5)},
    }
    if not args.quick:
        extra = {}
        # seed-1 random symmetric permutation of the same matrix
        P, _ = g.random_symmetric_permutation(A, 1)
        LP = P.lower()
        sp = Schedule(h, LP.rowptr, LP.colidx)
        lpv = torch.as_tensor(LP.values, device="cuda")
        pplan = sp.autotune(lpv)
        _, m, _ = time_events(lambda: sp.solve(lpv, bd, x), 50, 5)
        extra["sptrsv_perm1"] = {"ms": round(m, 4), "GB/s": round(by / (m * 1e-3) / 1e9, 2), "plan": pplan}
        # 32 right-hand sides, natural ordering
        B = torch.as_tensor(g.rhs_uniform(L.n, 1, k=32), device="cuda")
        X = torch.empty_like(B)
        kplan = s.autotune(lv, nrhs=32)
        _, m, _ = time_events(lambda: s.solve(lv, B, X), 20, 3)
        byk = bytes_sptrsv(L.n, L.nnz, k=32)
        extra["sptrsv_32rhs"] = {"ms": round(m, 4), "GB/s": round(byk / (m * 1e-3) / 1e9, 2),
                                 "frac_of_peak": round(byk / (m * 1e-3) / 1e9 / HBM_PEAK_GBS, 5), "plan": kplan}
        # IC0 (64^3 Laplacian) and ILU0 (64^3 convection-diffusion)
        sa = Schedule(h, A.rowptr, A.colidx, kind=_lib.KIND_IC0)
        av = torch.as_tensor(A.values, device="cuda")
        lu = torch.empty_like(av)
        fplan = sa.autotune_factor(av)
        _, m, _ = time_events(lambda: sa.ic0(av, lu), 20, 3)
        h.check()
        nl = L.nnz
        extra["ic0_lap3d64"] = {"ms": round(m, 4), "GB/s": round((20 * nl + 8 * L.n) / (m * 1e-3) / 1e9, 2),
                                "plan": fplan}
        Cd = g.convdiff3d(64)
        sc = Schedule(h, Cd.rowptr, Cd.colidx, kind=_lib.KIND_ILU0)
        cv = torch.as_tensor(Cd.values, device="cuda")
        luc = torch.empty_like(cv)
        fplan = sc.autotune_factor(cv)
        _, m, _ = time_events(lambda: sc.ilu0(cv, luc), 20, 3)
        h.check()
        extra["ilu0_convdiff64"] = {"ms": round(m, 4),
                                    "GB/s": round((20 * Cd.nnz + 8 * Cd.n) / (m * 1e-3) / 1e9, 2), "plan": fplan}
        # ICCG time to 1e-8, FP64 factor vs FP32 factor (kNN 3x3-block stiffness, n = 36000)
        K, kb = g.knn_stiffness()
        sk = Schedule(h, K.rowptr, K.colidx, kind=_lib.KIND_IC0)
        kv = torch.as_tensor(K.values, device="cuda")
        kbd = torch.as_tensor(kb, device="cuda")
        res = {}
        for prec, name in ((0, "fp64"), (1, "fp32_factor")):
            solver = PCG(h, sk, kv, precision=prec)
            pc_plan = solver.autotune()  # plans tuned for this precision's sweeps
            solver.solve(kbd, tol=1e-8, maxiter=5000)  # warm-up: factor, capture the graph
            ts, its = [], 0
            for _ in range(5):
                xk, its, rel = solver.solve(kbd, tol=1e-8, maxiter=5000)
                ts.append(solver.times()[1])
            ic_s, _ = solver.times()
            solver.solve(kbd, tol=1e-8, maxiter=5000, refactor=True)
            res[name] = {"iterations": its, "time_to_1e-8_s": round(float(np.median(ts)), 6),
                         "ic0_s": round(solver.times()[0], 6), "relres": rel, "plan": pc_plan}
        res["speedup_fp32_factor"] = round(res["fp64"]["time_to_1e-8_s"] / res["fp32_factor"]["time_to_1e-8_s"], 3)
        extra["iccg_knn36000"] = res
        # batched ICCG: 256 2D Laplacians (sides 16..256, seed 3) x 32 RHS
        from baas_amd.api import Batch
        mats, rhs = g.batch_laplace2d()
        bt = Batch(h, mats, 32)
        Bt = torch.as_tensor(np.concatenate(rhs, axis=0), device="cuda")
        del rhs
        bt.solve(Bt, tol=1e-8, maxiter=3000)  # first call: IC0, plan tuning, graph capture
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, iters = bt.solve(Bt, tol=1e-8, maxiter=3000)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        h.check()
        extra["batch_iccg_256x32"] = {"systems": len(mats), "rows": int(sum(m.n for m in mats)), "rhs": 32,
                                      "time_to_1e-8_s": round(dt, 4), "iterations_max": int(iters.max()),
                                      "iterations_mean": round(float(iters.mean()), 1), "info": bt.info()}
        del bt, Bt
        out["extra"] = extra
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(L, b, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
