"""Write the golden fixtures of tests/golden/ from the COMPILED reference.

Run in a container that has the reference tree (it builds oracle/_ref from
the reference's own spm/ merge/ blas/ sources).  The fixtures are data only:
inputs are regenerated from baas_amd.matgen (seeded), outputs are stored as
full arrays where small (config C1) and as SHA-256 digests, counts, sampled
entries and checksums where large.

    python tests/golden/make_golden.py

Reference outputs recorded (all from oracle/_ref/libbaas_ref.so):
* BAAS schedules (ref_baas_schedule: reduction, both coarsenings, group DAG,
  level set, threads per level, level merge) for every workload and kind.
* row-level level sets (DAG_levelSet_CSC).
* IC0 (spic0_csr_uL_serial) and ILU0 (ilu0csr_uplooking_ref) factors.
* SpMV / dcsrmv / waxpby on small vectors.
* Matrix Market loader output for bcsstk18 (the reference's data file).
The reference's triangular solves (merge/SpTRSV.cpp) do not build here
(papi.h is missing), so no solve outputs are recorded.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def workloads():
    from workloads import NAMES, NONSYMMETRIC, workload
    return {name: workload(name) for name in NAMES}, NONSYMMETRIC


def main():
    O.build()
    if O.ref is None:
        raise SystemExit("oracle/_ref was not built (reference tree absent)")
    manifest = {"generator": "tests/golden/make_golden.py", "schedules": {}, "levels": {},
                "ic0": {}, "ilu0": {}, "inputs": {}}
    wl, nonsym = workloads()
    arrays = {}
    for name, (L, thr, A) in wl.items():
        manifest["inputs"][name] = {"n": L.n, "nnz_lower": L.nnz,
                                    "lower_rowptr": sha(L.rowptr), "lower_colidx": sha(L.colidx)}
        lp, ntl, pto = O.levelset_rows(L.n, L.rowptr, L.colidx, use_ref=True)
        manifest["levels"][name] = {"nlevels": len(lp) - 1, "level_ptr": sha(lp),
                                    "node_to_level": sha(ntl), "perm_to_orig": sha(pto)}
        kinds = [(0, L.rowptr)] if A is None else [(0, L.rowptr), (0, thr), (1, thr), (2, thr)]
        for kind, rp in kinds:
            for ncore in (1, 2, 8, 64):
                s = O.baas_schedule(kind, L.n, L.rowptr, L.colidx, rp, ncore, use_ref=True)
                tag = f"{name}/k{kind}/{'L' if rp is L.rowptr else 'A'}/c{ncore}"
                manifest["schedules"][tag] = {
                    k: (sha(v) if isinstance(v, np.ndarray) else v) for k, v in s.items()}
                if name == "C1_lap2d_128" and ncore in (1, 8):
                    for k, v in s.items():
                        if isinstance(v, np.ndarray):
                            arrays[f"C1_sched_c{ncore}_{k}"] = v
        if name == "C1_lap2d_128":
            arrays["C1_level_ptr"], arrays["C1_node_to_level"], arrays["C1_perm_to_orig"] = lp, ntl, pto
        if A is not None and name not in nonsym:
            lu = O.ic0(A, use_ref=True)
            rng = np.random.default_rng(123)
            lower_pos = np.flatnonzero(A.colidx <= np.repeat(np.arange(A.n), np.diff(A.rowptr)))
            pick = np.sort(rng.choice(lower_pos, size=64, replace=False))
            manifest["ic0"][name] = {"sum_lower": float(lu[lower_pos].sum()),
                                     "norm_lower": float(np.linalg.norm(lu[lower_pos])),
                                     "pick": pick.tolist(), "values": lu[pick].tolist()}
        if A is not None and name in ("C2_convdiff_64", "C2_convdiff_64_perm1", "C2_lap3d_64", "bcsstk18"):
            lu = O.ilu0(A, use_ref=True)
            rng = np.random.default_rng(321)
            pick = np.sort(rng.choice(A.nnz, size=64, replace=False))
            manifest["ilu0"][name] = {"sum": float(lu.sum()), "norm": float(np.linalg.norm(lu)),
                                      "pick": pick.tolist(), "values": lu[pick].tolist()}
    # BLAS on a small system
    K = wl["bcsstk18"][2]
    rng = np.random.default_rng(77)
    x = rng.standard_normal(K.n)
    y = rng.standard_normal(K.n)
    arrays["blas_x"], arrays["blas_y"] = x, y
    arrays["blas_spmv"] = O.spmv(K, x, use_ref=True)
    for a, b in [(1.0, 0.0), (-1.0, 1.0), (0.5, -2.0), (1.25, 0.75)]:
        arrays[f"blas_dcsrmv_{a}_{b}"] = O.dcsrmv(K, x, y, a, b, use_ref=True)
    for a, b in [(1.0, 1.0), (1.0, -1.0), (1.0, 0.375), (0.625, 1.0), (-1.0, 0.0), (-1.0, 0.5),
                 (0.25, 0.0), (0.3, 0.9)]:
        arrays[f"blas_waxpby_{a}_{b}"] = O.waxpby(a, x[:1000], b, y[:1000], use_ref=True)
    # Matrix Market loaders
    for lower in (0, 1):
        m = C.c_int(); nz = C.c_int()
        path = os.path.join(HERE, "_tmp_bcsstk18.mtx")
        import gzip
        with gzip.open(os.path.join(HERE, "bcsstk18.mtx.gz"), "rb") as fi, open(path, "wb") as fo:
            fo.write(fi.read())
        O.ref.ref_load_mtx.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.c_void_p, C.c_void_p, C.c_void_p]
        O.ref.ref_load_mtx(path.encode(), lower, C.byref(m), C.byref(nz), None, None, None)
        rp = np.zeros(m.value + 1, np.int32); ci = np.zeros(nz.value, np.int32); va = np.zeros(nz.value)
        O.ref.ref_load_mtx(path.encode(), lower, C.byref(m), C.byref(nz), rp.ctypes.data,
                           ci.ctypes.data, va.ctypes.data)
        os.remove(path)
        manifest["inputs"][f"bcsstk18_loader_lower{lower}"] = {
            "n": m.value, "nnz": nz.value, "rowptr": sha(rp), "colidx": sha(ci), "values": sha(va)}
    np.savez_compressed(os.path.join(HERE, "golden_arrays.npz"), **arrays)
    with open(os.path.join(HERE, "golden_manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", len(manifest["schedules"]), "schedules,", len(arrays), "arrays")


if __name__ == "__main__":
    main()
