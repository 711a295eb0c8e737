"""Benchmark: FP64 SpTRSV on the 64^3 7-point Laplacian (one MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
                    [--plan MODE:ROWS:BLOCK:DATAFLOW[:BAND]] [--only WORKLOAD]

A step is one forward triangular solve x = L^-1 b with the lower triangle L of
the 3D 7-point Laplacian on a 64^3 grid in natural ordering (n = 262144,
nnz(L) = 1036288), one right-hand side b ~ U(-1,1) (seed 0), every input
resident in HBM.  value = algorithmic GB/s = (12 nnz(L) + 20 n) bytes per
solve x K solves / the timed region's length (the reference's byte model,
test/iccg_level_merge.cpp:354-355).  The execution plan is autotuned before
the timed region (analysis time, reported separately, as the reference's
Scheduling_Time).

Further fields (not the headline): the same solve on the seed-1 random
symmetric permutation, both renumbered by their schedule's row order, 32
right-hand sides, IC0 and ILU0 (64^3), ICCG
time to ||r||/||r0|| < 1e-8 on the kNN stiffness matrix (n = 36000) with
the FP64 factor and with the FP32 factor, batched ICCG, and rocSPARSE's
csrsv / spsv / csric0 / csrilu0 timed on the same inputs as a vendor-library
comparator (in place of the reference's MKL baselines).  --only runs one of
them alone (under a profiler, with --plan fixing its plan so that no
autotune launches mix into the kernel statistics).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
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


def host_cpu():
    """CPU model, physical cores and the threads this process may use."""
    model, cores, sockets = "unknown", None, 1
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Core(s) per socket":
                cores = int(v)
            elif k.strip() == "Socket(s)":
                sockets = int(v)
    except (OSError, ValueError):
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, (cores * sockets if cores else None), avail


def cpu_baseline_worker(seconds):
    """Runs in a child process started with the OpenMP environment set:
    times the oracle's serial solve on one core and its OpenMP level-set and
    point-to-point restatements of the reference's solves
    (merge/SpTRSV.cpp:41-80 and :421-478) on OMP_NUM_THREADS threads, each
    for about seconds / 3, on the 64^3 lower triangle.  Prints one JSON line."""
    import oracle as O
    from baas_amd import matgen as g
    L = g.laplace3d(64).lower()
    b = g.rhs_uniform(L.n, 0)
    nt = O.omp_max_threads()
    lp, _, pto = O.levelset_rows(L.n, L.rowptr, L.colidx)
    p2p = O.P2PSolver(L, nt)
    x_ref = O.sptrsv_lower(L, b)
    runs = {"serial_1_thread": lambda: O.sptrsv_lower(L, b),
            f"levelset_omp_{nt}_threads": lambda: O.sptrsv_lower_levels_omp(L, b, lp, pto, nt),
            f"p2p_omp_{nt}_threads": lambda: p2p.solve(b)}
    res = {}
    for name, fn in runs.items():
        assert np.array_equal(fn(), x_ref), name  # every variant gives the serial bits
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < seconds / 3:
            fn()
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        res[name] = {"ms": round(dt * 1e3, 4), "GB/s": round(bytes_sptrsv(L.n, L.nnz) / dt / 1e9, 3), "reps": reps}
    print(json.dumps({"threads": nt, "runs": res}))


def cpu_baseline(seconds):
    """The reference's CPU path restated (oracle/), timed on this host: the
    OpenMP solves at the host's physical core count (capped by the cores
    this process may use) with OMP_PROC_BIND=close and OMP_PLACES=cores, as
    the reference's scripts run them (script/sub_trsv.sh:15-20), plus the
    serial solve on one core.  The fastest is the baseline value."""
    model, phys, avail = host_cpu()
    nt = max(1, min(phys or avail, avail))
    env = dict(os.environ, OMP_NUM_THREADS=str(nt), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", str(seconds)],
                       capture_output=True, text=True, env=env, timeout=max(120, 6 * seconds))
    if r.returncode != 0:
        return {"error": r.stderr[-2000:]}
    d = json.loads(r.stdout.strip().splitlines()[-1])
    best = max(d["runs"].items(), key=lambda kv: kv[1]["GB/s"])
    cores = 1 if best[0].startswith("serial") else d["threads"]
    return {"value": best[1]["GB/s"], "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": (f"FP64 forward solves of the 64^3 L, ~{seconds / 3:.0f} s per variant: oracle/oracle.c "
                       f"restatements (serial; OpenMP level-set, merge/SpTRSV.cpp:41-80; OpenMP point-to-point, "
                       f"merge/SpTRSV.cpp:421-478) at {d['threads']} threads, OMP_PROC_BIND=close, "
                       f"OMP_PLACES=cores; best: {best[0]} {best[1]['ms']} ms per solve"),
            "cpu_model": model, "physical_cores": phys, "cores_available": avail, "threads": d["threads"],
            "runs": d["runs"]}


def parse_plan(text):
    """MODE:ROWS_PER_UNIT:BLOCK:DATAFLOW[:BAND] -> the plan dict autotune returns."""
    f = [int(v) for v in text.split(":")]
    return {"mode": f[0], "rows_per_unit": f[1], "block": f[2], "dataflow": f[3],
            "band": f[4] if len(f) > 4 else 16, "us": None}


def apply_plan(s, plan):
    """Re-pack schedule s into the plan a previous autotune (or --plan) chose."""
    if plan["mode"] == 4:
        s.replan_tiles(plan["rows_per_unit"], plan["band"], plan["block"], dataflow=int(plan["dataflow"]))
    else:
        s.replan(plan["mode"], plan["rows_per_unit"], plan["block"], dataflow=int(plan["dataflow"]))


def rocsparse_comparators(L, b, x_ours, A, lu_ours, Cd, luc_ours, steps=50):
    """rocSPARSE on the same inputs, timed as a black box: csrsv and the
    generic spsv (FP64 lower solve, analysis timed apart), csric0 on A and csrilu0 on the
    convection-diffusion matrix.  Its summation order is its own, so the
    largest relative difference from this library's (bit-exact) result is
    reported, not an equality.  Factorisations overwrite their input, so each
    timed call copies A's values first; the copy's own time is subtracted."""
    import ctypes as C
    import torch
    try:
        rs = C.CDLL("librocsparse.so.1")
    except OSError:
        try:
            rs = C.CDLL("/opt/rocm/lib/librocsparse.so.1")
        except OSError as e:
            return {"error": str(e)}
    P, I = C.c_void_p, C.c_int
    NONE, REUSE, AUTO = 111, 0, 0

    def chk(st, what):
        if st != 0:
            raise RuntimeError(f"rocsparse {what}: status {st}")

    def dptr(t):
        return P(t.data_ptr())

    def reldiff(a, b_):
        a, b_ = a.double(), b_.double()
        return float(((a - b_).abs().max() / b_.abs().max().clamp_min(1e-300)).item())

    handle = P()
    chk(rs.rocsparse_create_handle(C.byref(handle)), "create_handle")
    chk(rs.rocsparse_set_stream(handle, P(torch.cuda.current_stream().cuda_stream)), "set_stream")
    out = {}
    objs = []
    try:
        # lower triangular solve
        descr, info, sz = P(), P(), C.c_size_t()
        chk(rs.rocsparse_create_mat_descr(C.byref(descr)), "descr")
        objs.append(("rocsparse_destroy_mat_descr", descr))
        chk(rs.rocsparse_set_mat_fill_mode(descr, I(0)), "fill mode")
        chk(rs.rocsparse_create_mat_info(C.byref(info)), "info")
        objs.append(("rocsparse_destroy_mat_info", info))
        rp = torch.as_tensor(L.rowptr, device="cuda")
        ci = torch.as_tensor(L.colidx, device="cuda")
        lv = torch.as_tensor(L.values, device="cuda")
        m, nnz = I(L.n), I(L.nnz)
        chk(rs.rocsparse_dcsrsv_buffer_size(handle, I(NONE), m, nnz, descr, dptr(lv), dptr(rp), dptr(ci), info,
                                            C.byref(sz)), "csrsv_buffer_size")
        buf = torch.empty(max(sz.value, 8), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        chk(rs.rocsparse_dcsrsv_analysis(handle, I(NONE), m, nnz, descr, dptr(lv), dptr(rp), dptr(ci), info,
                                         I(REUSE), I(AUTO), dptr(buf)), "csrsv_analysis")
        torch.cuda.synchronize()
        t_an = time.perf_counter() - t0
        y = torch.empty_like(b)
        alpha = C.c_double(1.0)
        solve = lambda: chk(rs.rocsparse_dcsrsv_solve(handle, I(NONE), m, nnz, C.byref(alpha), descr, dptr(lv),
                                                       dptr(rp), dptr(ci), info, dptr(b), dptr(y), I(AUTO),
                                                       dptr(buf)), "csrsv_solve")
        _, ms, _ = time_events(solve, steps, 5)
        by = bytes_sptrsv(L.n, L.nnz)
        out["csrsv_lap3d64"] = {"ms": round(ms, 4), "GB/s": round(by / (ms * 1e-3) / 1e9, 2),
                                "analysis_s": round(t_an, 4), "max_rel_diff": reldiff(y, x_ours)}
        # the generic interface (rocsparse_spsv, default algorithm) on the same solve
        I64, F64, I32T, BASE0 = C.c_int64, 152, 2, 0
        mat, vx, vy, bsz = P(), P(), P(), C.c_size_t()
        chk(rs.rocsparse_create_csr_descr(C.byref(mat), I64(L.n), I64(L.n), I64(L.nnz), dptr(rp), dptr(ci),
                                          dptr(lv), I(I32T), I(I32T), I(BASE0), I(F64)), "create_csr_descr")
        objs.append(("rocsparse_destroy_spmat_descr", mat))
        fill, diag = C.c_int(0), C.c_int(0)  # lower, non-unit
        chk(rs.rocsparse_spmat_set_attribute(mat, I(0), C.byref(fill), C.c_size_t(4)), "fill attribute")
        chk(rs.rocsparse_spmat_set_attribute(mat, I(1), C.byref(diag), C.c_size_t(4)), "diag attribute")
        y2 = torch.empty_like(b)
        chk(rs.rocsparse_create_dnvec_descr(C.byref(vx), I64(L.n), dptr(b), I(F64)), "dnvec x")
        objs.append(("rocsparse_destroy_dnvec_descr", vx))
        chk(rs.rocsparse_create_dnvec_descr(C.byref(vy), I64(L.n), dptr(y2), I(F64)), "dnvec y")
        objs.append(("rocsparse_destroy_dnvec_descr", vy))

        def spsv(stage, buf_ptr):
            chk(rs.rocsparse_spsv(handle, I(NONE), C.byref(alpha), mat, vx, vy, I(F64), I(0), I(stage),
                                  C.byref(bsz), buf_ptr), f"spsv stage {stage}")
        spsv(1, None)
        buf2 = torch.empty(max(bsz.value, 8), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        spsv(2, dptr(buf2))
        torch.cuda.synchronize()
        t_an2 = time.perf_counter() - t0
        _, ms2, _ = time_events(lambda: spsv(3, dptr(buf2)), steps, 5)
        out["spsv_generic_lap3d64"] = {"ms": round(ms2, 4), "GB/s": round(by / (ms2 * 1e-3) / 1e9, 2),
                                       "analysis_s": round(t_an2, 4), "max_rel_diff": reldiff(y2, x_ours)}

        def factor(name, M, ours, fn_pre):
            d2, inf2, sz2 = P(), P(), C.c_size_t()
            chk(rs.rocsparse_create_mat_descr(C.byref(d2)), "descr")
            objs.append(("rocsparse_destroy_mat_descr", d2))
            chk(rs.rocsparse_create_mat_info(C.byref(inf2)), "info")
            objs.append(("rocsparse_destroy_mat_info", inf2))
            arp = torch.as_tensor(M.rowptr, device="cuda")
            aci = torch.as_tensor(M.colidx, device="cuda")
            av = torch.as_tensor(M.values, device="cuda")
            work = av.clone()
            mm, nz = I(M.n), I(M.nnz)
            chk(getattr(rs, f"rocsparse_d{fn_pre}_buffer_size")(handle, mm, nz, d2, dptr(work), dptr(arp), dptr(aci),
                                                               inf2, C.byref(sz2)), f"{fn_pre}_buffer_size")
            fb = torch.empty(max(sz2.value, 8), dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            chk(getattr(rs, f"rocsparse_d{fn_pre}_analysis")(handle, mm, nz, d2, dptr(work), dptr(arp), dptr(aci),
                                                            inf2, I(REUSE), I(AUTO), dptr(fb)), f"{fn_pre}_analysis")
            torch.cuda.synchronize()
            t_a = time.perf_counter() - t0

            def run():
                work.copy_(av)
                chk(getattr(rs, f"rocsparse_d{fn_pre}")(handle, mm, nz, d2, dptr(work), dptr(arp), dptr(aci), inf2,
                                                        I(AUTO), dptr(fb)), fn_pre)
            _, ms_all, _ = time_events(run, 20, 3)
            _, ms_copy, _ = time_events(lambda: work.copy_(av), 20, 3)
            run()
            torch.cuda.synchronize()
            rows = np.repeat(np.arange(M.n), np.diff(M.rowptr))
            sel = torch.as_tensor(M.colidx <= rows if fn_pre == "csric0" else np.ones(M.nnz, bool), device="cuda")
            out[name] = {"ms": round(ms_all - ms_copy, 4), "analysis_s": round(t_a, 4),
                         "max_rel_diff": reldiff(work[sel], ours[sel])}

        factor("csric0_lap3d64", A, lu_ours, "csric0")
        factor("csrilu0_convdiff64", Cd, luc_ours, "csrilu0")
    finally:
        torch.cuda.synchronize()
        for fn, o in reversed(objs):
            getattr(rs, fn)(o)
        rs.rocsparse_destroy_handle(handle)
    return out


def traffic_from_profiles():
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/traffic.json), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None, None
    with open(p) as f:
        d = json.load(f)
    return d.get("sptrsv_lap3d64_bytes_per_launch"), d.get("plan")


class Ctx:
    """What the secondary measurements share: the handle, the 64^3 Laplacian
    A, its lower triangle L with its analysed schedule, values, b and the
    headline solve's x, and the headline plan.  `plan_override` (from --plan
    with --only) replaces a workload's autotune with that plan."""

    def __init__(self, h, A, L, s, lv, bd, x, plan, plan_override=None):
        self.h, self.A, self.L, self.s, self.lv, self.bd, self.x, self.plan = h, A, L, s, lv, bd, x, plan
        self.plan_override = plan_override
        self.sa = self.av = self.lu = self.Cd = self.sc = self.cv = self.luc = None

    def tune(self, sched, fn):
        """autotune through fn() unless a plan was forced."""
        if self.plan_override is None:
            return fn()
        apply_plan(sched, self.plan_override)
        return dict(self.plan_override)


def x_perm1(c):
    """The same solve on the seed-1 random symmetric permutation of A."""
    import torch
    from baas_amd import matgen as g
    from baas_amd.api import Schedule
    P, _ = g.random_symmetric_permutation(c.A, 1)
    LP = P.lower()
    sp = Schedule(c.h, LP.rowptr, LP.colidx)
    lpv = torch.as_tensor(LP.values, device="cuda")
    pplan = c.tune(sp, lambda: sp.autotune(lpv))
    _, m, _ = time_events(lambda: sp.solve(lpv, c.bd, c.x), 50, 5)
    by = bytes_sptrsv(c.L.n, c.L.nnz)
    return {"ms": round(m, 4), "GB/s": round(by / (m * 1e-3) / 1e9, 2), "plan": pplan}


def x_renumbered(c):
    """Locality renumbering (the GPU form of the reference's Perm variants):
    the lower triangle renumbered by its tuned schedule's row order
    (P L P^T, each work unit's rows consecutive), analysed and tuned again;
    the solve time in the renumbered system, for the natural ordering and
    for the seed-1 permutation.  b is renumbered outside the timed region,
    as in test/FusionDemo.cpp:96-112."""
    import torch
    from baas_amd import matgen as g
    from baas_amd.api import Schedule
    out = {}
    by = bytes_sptrsv(c.L.n, c.L.nnz)
    for name, M in (("lap3d64", c.A), ("lap3d64_perm1", g.random_symmetric_permutation(c.A, 1)[0])):
        L = M.lower()
        s0 = Schedule(c.h, L.rowptr, L.colidx)
        s0.autotune(torch.as_tensor(L.values, device="cuda"))
        order = s0.unit_order()
        LP = g.symmetric_permute(L, order)
        t0 = time.perf_counter()
        sp = Schedule(c.h, LP.rowptr, LP.colidx)
        t_an = time.perf_counter() - t0
        lpv = torch.as_tensor(LP.values, device="cuda")
        bp = c.bd[torch.as_tensor(order.astype(np.int64), device="cuda")].contiguous()
        xp = torch.empty_like(bp)
        plan = c.tune(sp, lambda: sp.autotune(lpv))
        _, m, _ = time_events(lambda: sp.solve(lpv, bp, xp), 50, 5)
        c.h.check()
        out[name] = {"ms": round(m, 4), "GB/s": round(by / (m * 1e-3) / 1e9, 2), "plan": plan,
                     "second_analysis_s": round(t_an, 4)}
    return out


def x_rhs32(c):
    """32 right-hand sides (row-major n x 32, U(-1,1) seed 1), natural ordering."""
    import torch
    from baas_amd import matgen as g
    B = torch.as_tensor(g.rhs_uniform(c.L.n, 1, k=32), device="cuda")
    X = torch.empty_like(B)
    kplan = c.tune(c.s, lambda: c.s.autotune(c.lv, nrhs=32))
    _, m, _ = time_events(lambda: c.s.solve(c.lv, B, X), 20, 3)
    byk = bytes_sptrsv(c.L.n, c.L.nnz, k=32)
    if c.plan_override is None and c.plan is not None:
        apply_plan(c.s, c.plan)
    return {"ms": round(m, 4), "GB/s": round(byk / (m * 1e-3) / 1e9, 2),
            "frac_of_peak": round(byk / (m * 1e-3) / 1e9 / HBM_PEAK_GBS, 5), "bytes": byk, "plan": kplan}


def x_ic0(c):
    """IC0 of A (the 64^3 Laplacian); algorithmic bytes 20 nnz(L) + 8 n."""
    import torch
    from baas_amd import _lib
    from baas_amd.api import Schedule
    c.sa = Schedule(c.h, c.A.rowptr, c.A.colidx, kind=_lib.KIND_IC0)
    c.av = torch.as_tensor(c.A.values, device="cuda")
    c.lu = torch.empty_like(c.av)
    fplan = c.tune(c.sa, lambda: c.sa.autotune_factor(c.av))
    _, m, med = time_events(lambda: c.sa.ic0(c.av, c.lu), 20, 3)
    c.h.check()
    by = 20 * c.L.nnz + 8 * c.L.n
    gbs = by / (m * 1e-3) / 1e9
    return {"ms": round(m, 4), "median_ms": round(med, 4), "GB/s": round(gbs, 2),
            "frac_of_peak": round(gbs / HBM_PEAK_GBS, 5), "bytes": by, "plan": fplan}


def x_ilu0(c):
    """ILU0 of the 64^3 convection-diffusion operator; bytes 20 nnz(A) + 8 n."""
    import torch
    from baas_amd import _lib
    from baas_amd import matgen as g
    from baas_amd.api import Schedule
    c.Cd = g.convdiff3d(64)
    c.sc = Schedule(c.h, c.Cd.rowptr, c.Cd.colidx, kind=_lib.KIND_ILU0)
    c.cv = torch.as_tensor(c.Cd.values, device="cuda")
    c.luc = torch.empty_like(c.cv)
    fplan = c.tune(c.sc, lambda: c.sc.autotune_factor(c.cv))
    _, m, med = time_events(lambda: c.sc.ilu0(c.cv, c.luc), 20, 3)
    c.h.check()
    by = 20 * c.Cd.nnz + 8 * c.Cd.n
    gbs = by / (m * 1e-3) / 1e9
    return {"ms": round(m, 4), "median_ms": round(med, 4), "GB/s": round(gbs, 2),
            "frac_of_peak": round(gbs / HBM_PEAK_GBS, 5), "bytes": by, "plan": fplan}


def x_rocsparse(c):
    """The vendor library on the same inputs, checked against this library's results."""
    if c.sa is None:
        x_ic0(c)
    if c.sc is None:
        x_ilu0(c)
    apply_plan(c.s, c.plan)
    c.s.solve(c.lv, c.bd, c.x)
    c.sa.ic0(c.av, c.lu)
    c.sc.ilu0(c.cv, c.luc)
    c.h.check()
    return rocsparse_comparators(c.L, c.bd, c.x, c.A, c.lu, c.Cd, c.luc)


def iccg(c, precisions):
    """ICCG time to ||r||/||r0|| < 1e-8 on the kNN 3x3-block stiffness matrix
    (n = 36000), per factor precision (0 FP64, 1 FP32); IC0 timed apart."""
    import torch
    from baas_amd import _lib
    from baas_amd import matgen as g
    from baas_amd.api import PCG, Schedule
    K, kb = g.knn_stiffness()
    sk = Schedule(c.h, K.rowptr, K.colidx, kind=_lib.KIND_IC0)
    kv = torch.as_tensor(K.values, device="cuda")
    kbd = torch.as_tensor(kb, device="cuda")
    res = {}
    for prec in precisions:
        name = ("fp64", "fp32_factor")[prec]
        solver = PCG(c.h, sk, kv, precision=prec)
        pc_plan = c.tune(sk, solver.autotune)  # plans tuned for this precision's sweeps
        solver.solve(kbd, tol=1e-8, maxiter=5000)  # warm-up: factor, capture the graph
        ts, its = [], 0
        for _ in range(5):
            xk, its, rel = solver.solve(kbd, tol=1e-8, maxiter=5000)
            ts.append(solver.times()[1])
        # IC0 timed on its own: the first refactorisation on a new plan
        # also builds that plan's factorisation records (host, once)
        solver.solve(kbd, tol=1e-8, maxiter=5000, refactor=True)
        ic_first = solver.times()[0]
        solver.solve(kbd, tol=1e-8, maxiter=5000, refactor=True)
        t = float(np.median(ts))
        res[name] = {"iterations": its, "time_to_1e-8_s": round(t, 6), "ms_per_iteration": round(t / its * 1e3, 5),
                     "ic0_s": round(solver.times()[0], 6), "ic0_first_s": round(ic_first, 6), "relres": rel,
                     "plan": pc_plan}
    if len(precisions) == 2:
        res["speedup_fp32_factor"] = round(res["fp64"]["time_to_1e-8_s"] / res["fp32_factor"]["time_to_1e-8_s"], 3)
    return res


def x_batch(c):
    """Batched ICCG: 256 2D Laplacians (sides 16..256, seed 3) x 32 RHS."""
    import torch
    from baas_amd import matgen as g
    from baas_amd.api import Batch
    mats, rhs = g.batch_laplace2d()
    bt = Batch(c.h, mats, 32)
    Bt = torch.as_tensor(np.concatenate(rhs, axis=0), device="cuda")
    del rhs
    bt.solve(Bt, tol=1e-8, maxiter=3000)  # first call: IC0, plan tuning, graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, iters = bt.solve(Bt, tol=1e-8, maxiter=3000)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    c.h.check()
    r = {"systems": len(mats), "rows": int(sum(m.n for m in mats)), "rhs": 32, "time_to_1e-8_s": round(dt, 4),
         "iterations_max": int(iters.max()), "iterations_mean": round(float(iters.mean()), 1), "info": bt.info()}
    del bt, Bt
    return r


EXTRAS = {"sptrsv_perm1": x_perm1, "sptrsv_renumbered": x_renumbered, "sptrsv_32rhs": x_rhs32, "ic0_lap3d64": x_ic0, "ilu0_convdiff64": x_ilu0,
          "rocsparse": x_rocsparse, "iccg_knn36000": lambda c: iccg(c, (0, 1)),
          "iccg_knn36000_fp64": lambda c: iccg(c, (0,)), "iccg_knn36000_fp32": lambda c: iccg(c, (1,)),
          "batch_iccg_256x32": x_batch}
DEFAULT_EXTRAS = ["sptrsv_perm1", "sptrsv_renumbered", "sptrsv_32rhs", "ic0_lap3d64", "ilu0_convdiff64", "rocsparse", "iccg_knn36000",
                  "batch_iccg_256x32"]


def run_extras(c, names):
    return {n: EXTRAS[n](c) for n in names}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only, no secondary measurements")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--plan", default="auto",
                    help="auto (autotune) or MODE:ROWS_PER_UNIT:BLOCK:DATAFLOW[:BAND], MODE 1 = RANGE, "
                         "2 = SCHEDULE, 3 = LEVELS, 4 = TILES (ROWS_PER_UNIT = row range, BAND = levels), "
                         "5 = BOXES, 6 = PIPE (DATAFLOW 5); with --only it applies to that workload")
    ap.add_argument("--only", choices=sorted(EXTRAS), default=None,
                    help="run one secondary measurement alone (no headline solve) and print its JSON")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run 3 plain copy kernels of this many MiB (HBM counter calibration)")
    ap.add_argument("--cpu-baseline-worker", type=float, default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_worker is not None:
        cpu_baseline_worker(args.cpu_baseline_worker)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs a solve on one GPU and has no multi-GPU path "
              "(no collectives or cross-GPU publication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    from baas_amd import _lib
    from baas_amd import matgen as g
    from baas_amd.api import Handle, Schedule

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
    bd = torch.as_tensor(b, device="cuda")
    x = torch.empty_like(bd)
    forced = None if args.plan == "auto" else parse_plan(args.plan)
    if args.only:
        # one secondary workload; the headline schedule only where it needs it
        c = Ctx(h, A, L, None, None, bd, x, None, forced)
        if args.only in ("sptrsv_32rhs", "rocsparse"):
            c.s = Schedule(h, L.rowptr, L.colidx, kind=_lib.KIND_TRSV)
            c.lv = torch.as_tensor(L.values, device="cuda")
            if args.only == "rocsparse":
                c.plan, c.plan_override = c.s.autotune(c.lv), None
        r = EXTRAS[args.only](c)
        h.check()
        print(json.dumps({"only": args.only, args.only: r}))
        return
    t0 = time.perf_counter()
    s = Schedule(h, L.rowptr, L.colidx, kind=_lib.KIND_TRSV)
    t_an = time.perf_counter() - t0
    lv = torch.as_tensor(L.values, device="cuda")
    t0 = time.perf_counter()
    if forced is None:
        plan = s.autotune(lv)
    else:
        plan = forced
        apply_plan(s, plan)
    t_tune = time.perf_counter() - t0
    an_gpu, an_host = s.times()
    # the same schedule rebuilt from its saved bytes (analysis and tuned plan)
    blob = s.save()
    t0 = time.perf_counter()
    Schedule(h, L.rowptr, L.colidx, kind=_lib.KIND_TRSV, saved=blob)
    t_load = time.perf_counter() - t0
    print(f"[bench] analysis {t_an:.3f} s, autotune {t_tune:.2f} s, load {t_load:.3f} s, plan {plan}",
          file=sys.stderr, flush=True)

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
                   "analysis_s": round(t_an, 4), "analysis_gpu_phase_s": round(an_gpu, 4),
                   "analysis_host_phase_s": round(an_host, 4), "autotune_s": round(t_tune, 3),
                   "saved_schedule_bytes": len(blob), "saved_schedule_load_s": round(t_load, 4)},
        "roofline": {"bound": "hbm", "achieved": round(kern_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kern_gbs / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "traffic_plan": traffic_plan,
                     "kernel": {0: "k_sweep", 1: "k_sweep_df (dataflow)", 2: "k_sweep_lane (dataflow, lanes)",
                                5: "k_sweep_pipe (wave-pipelined grid boxes)"}.get(
                         int(plan["dataflow"]), "k_sweep_mr") if plan["mode"] != 3 else "k_sweep_levels",
                     "algorithmic_bytes_per_launch": by, "mean_launch_ms": round(mean_ms, 5),
                     "median_launch_ms": round(med_ms, 5)},
    }
    if not args.quick:
        out["extra"] = run_extras(Ctx(h, A, L, s, lv, bd, x, plan), DEFAULT_EXTRAS)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
