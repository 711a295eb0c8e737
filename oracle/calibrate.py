"""Time the oracle's restatements beside the compiled reference (oracle/_ref)
on the same inputs, in the container that has the reference tree.

    python oracle/calibrate.py

TEST INFRASTRUCTURE ONLY.  The reference's serial IC0 (spic0_csr_uL_serial,
merge/IC.cpp:32-89), ILU0 (ilu0csr_uplooking_ref, merge/ILU.cpp:350-386) and
SpMV (spmv_omp, blas/spmv.cpp:8-27) and the buildset() analysis
(ref_baas_schedule: the call sequence of test/SpTRSVRuntime.hpp:675-735) are
timed against oracle.c's restatements; its triangular solves (merge/SpTRSV.cpp) do not build here
(papi.h is absent), so the solve is timed on the restatement only.  Prints a
Markdown table.
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from workloads import workload  # noqa: E402


def best_of(fn, seconds=2.0):
    fn()
    t0, reps, best = time.perf_counter(), 0, 1e30
    while time.perf_counter() - t0 < seconds:
        t = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t)
        reps += 1
    return best * 1e3


def main():
    ref = O.get_ref()
    if ref is None:
        raise SystemExit("oracle/_ref is not built here")
    I32 = O._i32
    print(f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')} (the reference's spmv_omp is "
          "OpenMP-parallel, its IC0/ILU0 here are the serial kernels; the restatements are serial)")
    print("| matrix | kernel | reference (ms) | restatement (ms) | restatement / reference |")
    print("|---|---|---|---|---|")
    for name in ("C2_lap3d_64", "C3_knn", "bcsstk18", "bcsstk38"):
        _, _, A = workload(name)
        rp, ci, va, dp = I32(A.rowptr), I32(A.colidx), O._f64(A.values), I32(A.diagptr())
        x = np.random.default_rng(0).standard_normal(A.n)
        y = np.zeros(A.n)
        lu = va.copy()
        cases = (
            ("IC0 serial", lambda: (lu.__setitem__(slice(None), va), ref.ref_ic0(A.n, rp, ci, va, lu)),
             lambda: (lu.__setitem__(slice(None), va), O.lib.orc_ic0(A.n, rp, ci, dp, va, lu))),
            ("ILU0 serial", lambda: (lu.__setitem__(slice(None), va), ref.ref_ilu0(A.n, rp, ci, va, lu)),
             lambda: (lu.__setitem__(slice(None), va), O.lib.orc_ilu0(A.n, rp, ci, dp, lu))),
            ("SpMV", lambda: ref.ref_spmv(A.n, rp, ci, va, x, y), lambda: O.lib.orc_spmv(A.n, rp, ci, va, x, y)),
        )
        L, _, _ = workload(name)
        cases = cases + (
            ("analysis (buildset, SpTRSV, ncore 8)",
             lambda: O.baas_schedule(0, L.n, L.rowptr, L.colidx, ncore=8, use_ref=True),
             lambda: O.baas_schedule(0, L.n, L.rowptr, L.colidx, ncore=8)),)
        for kern, f_ref, f_orc in cases:
            r, o = best_of(f_ref), best_of(f_orc)
            print(f"| {name} | {kern} | {r:.3f} | {o:.3f} | {o / r:.2f} |", flush=True)


if __name__ == "__main__":
    main()
