"""The C++ runtime's Matrix Market loaders (SPM::CSR(file) and the lower
loader CSR(file, true, 0)) against the reference loader's digests for
bcsstk18 (tests/golden, written from the compiled reference).  CPU only: the
loader makes no GPU call."""
import gzip
import hashlib
import os
import shutil
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def _tool(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        from baas_amd import build as B
        B.build_demo()
    return path


def test_cpp_loader_matches_reference_loader(golden, tmp_path):
    man, _ = golden
    src = os.path.join(ROOT, "tests", "golden", "bcsstk18.mtx.gz")
    mtx = tmp_path / "bcsstk18.mtx"
    with gzip.open(src, "rb") as fi, open(mtx, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    r = subprocess.run([_tool("mtx_dump"), str(mtx), str(tmp_path / "b18")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    for lower, tag in ((0, "full"), (1, "lower")):
        want = man["inputs"][f"bcsstk18_loader_lower{lower}"]
        arrs = {k: np.fromfile(tmp_path / f"b18.{tag}.{k}", dtype=dt)
                for k, dt in (("rowptr", np.int32), ("colidx", np.int32), ("values", np.float64))}
        assert len(arrs["rowptr"]) - 1 == want["n"] and len(arrs["colidx"]) == want["nnz"]
        for k, a in arrs.items():
            assert hashlib.sha256(a.tobytes()).hexdigest() == want[k], (tag, k)


def test_reference_shaped_mains_are_built():
    for name in ("ic0_level_merge", "ilu0_level_merge", "sptrsv_level_merge", "iccg_scheduler"):
        assert os.access(_tool(name), os.X_OK), name


def test_cpp_csr_permute(tmp_path):
    """SPM::CSR::permute (spm/Permute.cpp:215-405) against P A P^T: rows
    taken in permToOrig order, columns renamed by its inverse, entry order
    kept (sort = false) or ascending (sort = true), diagptr following."""
    import scipy.io
    import scipy.sparse as sp
    from baas_amd import matgen as g
    A = g.convdiff3d(6).to_scipy().tocsr()
    mtx = tmp_path / "cd6.mtx"
    scipy.io.mmwrite(str(mtx), A.tocoo())
    p2o = np.random.default_rng(5).permutation(A.shape[0]).astype(np.int32)
    p2o.tofile(tmp_path / "perm.int32")
    r = subprocess.run([_tool("mtx_dump"), str(mtx), str(tmp_path / "cd"), str(tmp_path / "perm.int32")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o2p = np.argsort(p2o)
    full = sp.csr_matrix(A)
    full.sort_indices()
    for tag, sort in (("perm", False), ("perm_sorted", True)):
        rp = np.fromfile(tmp_path / f"cd.{tag}.rowptr", np.int32)
        ci = np.fromfile(tmp_path / f"cd.{tag}.colidx", np.int32)
        va = np.fromfile(tmp_path / f"cd.{tag}.values", np.float64)
        dg = np.fromfile(tmp_path / f"cd.{tag}.diagptr", np.int32)
        for i in range(A.shape[0]):
            row = p2o[i]
            cols = o2p[full.indices[full.indptr[row]:full.indptr[row + 1]]]
            vals = full.data[full.indptr[row]:full.indptr[row + 1]]
            if sort:
                o = np.argsort(cols, kind="stable")
                cols, vals = cols[o], vals[o]
            np.testing.assert_array_equal(ci[rp[i]:rp[i + 1]], cols)
            np.testing.assert_array_equal(va[rp[i]:rp[i + 1]], vals)
            assert ci[dg[i]] == i
        P = sp.csr_matrix((va, ci, rp), shape=A.shape)
        assert abs(P - A[p2o][:, p2o]).max() == 0


def _ref_load(path):
    """The compiled reference's CSR(file) (a .bin name goes through loadBin)."""
    import ctypes as C
    import oracle as O
    f = O.ref.ref_load_mtx
    f.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_void_p]
    m, nz = C.c_int(), C.c_int()
    f(str(path).encode(), 0, C.byref(m), C.byref(nz), None, None, None)
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(nz.value, np.int32)
    va = np.zeros(nz.value)
    f(str(path).encode(), 0, C.byref(m), C.byref(nz), rp.ctypes.data, ci.ctypes.data, va.ctypes.data)
    return rp, ci, va


def test_cpp_csr_binary_files(tmp_path):
    """CSR::storeBin / loadBin (spm/CSR.cpp:332-435): the file layout
    {1211216, m, n, nnz, row lengths, colidx, values}; where oracle/_ref is
    built, byte-identical to the reference's storeBin and readable by its
    loadBin, and the reference's files readable here."""
    import oracle as O
    from baas_amd import matgen as g
    import scipy.io
    A = g.convdiff3d(7)
    mtx = tmp_path / "cd7.mtx"
    scipy.io.mmwrite(str(mtx), A.to_scipy().tocoo())
    r = subprocess.run([_tool("mtx_dump"), str(mtx), str(tmp_path / "cd")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    ours = (tmp_path / "cd.full.bin").read_bytes()
    want = (np.array([1211216, A.n, A.n, A.nnz], np.int32).tobytes() + np.diff(A.rowptr).astype(np.int32).tobytes()
            + A.colidx.astype(np.int32).tobytes() + A.values.astype(np.float64).tobytes())
    assert ours == want
    # read back through the runtime's .bin constructor
    r = subprocess.run([_tool("mtx_dump"), str(tmp_path / "cd.full.bin"), str(tmp_path / "rb")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for k, dt in (("rowptr", np.int32), ("colidx", np.int32), ("values", np.float64)):
        np.testing.assert_array_equal(np.fromfile(tmp_path / f"rb.full.{k}", dt), getattr(A, k).astype(dt))
    if O.ref is None:
        return
    O.ref.ref_store_bin(A.n, A.n, O._i32(A.rowptr), O._i32(A.colidx), O._f64(A.values),
                        str(tmp_path / "ref.bin").encode())
    assert (tmp_path / "ref.bin").read_bytes() == ours
    rp, ci, va = _ref_load(tmp_path / "cd.full.bin")
    np.testing.assert_array_equal(rp, A.rowptr)
    np.testing.assert_array_equal(ci, A.colidx)
    np.testing.assert_array_equal(va, A.values)
