"""Pin the CPU oracle before it is trusted as the checker.

Each oracle function is checked against (a) the golden vectors written by the
compiled reference (tests/golden/), (b) oracle/_ref live when it is built in
this container, and (c) what mathematics fixes (longest-path levels, brute
force on tiny DAGs, manufactured solutions, scipy, factor identities).
"""
import hashlib

import numpy as np
import pytest

import oracle as O
from baas_amd import matgen as g
from workloads import NAMES, workload


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---------------------------------------------------------------- schedules
@pytest.mark.parametrize("name", NAMES)
def test_schedules_match_reference_golden(golden, name):
    man, _ = golden
    L, thr, A = workload(name)
    tags = [t for t in man["schedules"] if t.startswith(name + "/")]
    assert tags
    for tag in tags:
        _, k, which, c = tag.split("/")
        rp = L.rowptr if which == "L" else thr
        s = O.baas_schedule(int(k[1:]), L.n, L.rowptr, L.colidx, rp, int(c[1:]))
        want = man["schedules"][tag]
        for key, val in s.items():
            got = sha(val) if isinstance(val, np.ndarray) else val
            assert got == want[key], f"{tag}:{key}"


def test_c1_schedule_arrays_full(golden):
    _, arr = golden
    L, _, _ = workload("C1_lap2d_128")
    for ncore in (1, 8):
        s = O.baas_schedule(0, L.n, L.rowptr, L.colidx, L.rowptr, ncore)
        for key, val in s.items():
            if isinstance(val, np.ndarray):
                np.testing.assert_array_equal(val, arr[f"C1_sched_c{ncore}_{key}"])
    lp, ntl, pto = O.levelset_rows(L.n, L.rowptr, L.colidx)
    np.testing.assert_array_equal(lp, arr["C1_level_ptr"])
    np.testing.assert_array_equal(ntl, arr["C1_node_to_level"])
    np.testing.assert_array_equal(pto, arr["C1_perm_to_orig"])
    assert len(lp) - 1 == 255


@pytest.mark.parametrize("name", NAMES)
def test_row_levels_match_golden(golden, name):
    man, _ = golden
    L, _, _ = workload(name)
    lp, ntl, pto = O.levelset_rows(L.n, L.rowptr, L.colidx)
    want = man["levels"][name]
    assert len(lp) - 1 == want["nlevels"]
    assert sha(lp) == want["level_ptr"] and sha(ntl) == want["node_to_level"]
    assert sha(pto) == want["perm_to_orig"]


def test_level_counts_closed_form():
    # 2D natural: 2*side-1 levels; 3D natural: 3*side-2 levels (longest path).
    L = g.laplace2d_lower(128)
    assert len(O.levelset_rows(L.n, L.rowptr, L.colidx)[0]) - 1 == 255
    L3 = g.laplace3d(16).lower()
    assert len(O.levelset_rows(L3.n, L3.rowptr, L3.colidx)[0]) - 1 == 46


def _random_lower(n, p, rng):
    import scipy.sparse as sp
    M = sp.random(n, n, density=p, random_state=rng, format="csr")
    M = sp.tril(M, k=-1) + sp.identity(n)
    return g.Csr.from_scipy(M)


def _longest_path_levels(L):
    lev = np.ones(L.n, dtype=np.int64)
    for i in range(L.n):
        for j in range(L.rowptr[i], L.rowptr[i + 1] - 1):
            lev[i] = max(lev[i], lev[L.colidx[j]] + 1)
    return lev


@pytest.mark.parametrize("seed", range(6))
def test_levels_brute_force(seed):
    rng = np.random.default_rng(seed)
    L = _random_lower(60, 0.08, rng)
    lp, ntl, pto = O.levelset_rows(L.n, L.rowptr, L.colidx)
    lev = _longest_path_levels(L)
    np.testing.assert_array_equal(ntl, lev)
    order = sorted(range(L.n), key=lambda v: (lev[v], v))
    np.testing.assert_array_equal(pto, order)
    np.testing.assert_array_equal(np.bincount(lev)[1:], np.diff(lp))


@pytest.mark.parametrize("seed", range(6))
def test_two_hop_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    L = _random_lower(50, 0.12, rng)
    cp, ri = O.csc_of(L.n, L.rowptr, L.colidx)
    op, os_ = O.two_hop(L.n, cp, ri)
    ch = [set(ri[cp[i] + 1: cp[i + 1]]) for i in range(L.n)]
    for i in range(L.n):
        keep = [i] + sorted(j for j in ch[i] if not any(j in ch[k] for k in ch[i]))
        assert list(os_[op[i]:op[i + 1]]) == keep
    if O.ref is not None:
        rp, rs = O.two_hop(L.n, cp, ri, use_ref=True)
        np.testing.assert_array_equal(op, rp)
        np.testing.assert_array_equal(os_, rs)


@pytest.mark.skipif(O.ref is None, reason="oracle/_ref not built here")
@pytest.mark.parametrize("seed", range(8))
def test_schedule_random_dags_vs_reference(seed):
    rng = np.random.default_rng(200 + seed)
    L = _random_lower(400, 0.006 + 0.002 * seed, rng)
    for kind in (0, 1, 2):
        for ncore in (1, 3, 8):
            a = O.baas_schedule(kind, L.n, L.rowptr, L.colidx, L.rowptr, ncore)
            b = O.baas_schedule(kind, L.n, L.rowptr, L.colidx, L.rowptr, ncore, use_ref=True)
            for k in a:
                if isinstance(a[k], np.ndarray):
                    np.testing.assert_array_equal(a[k], b[k], err_msg=f"{kind}/{ncore}/{k}")
                else:
                    assert a[k] == b[k]


def test_groups_are_a_partition_and_topological():
    L, _, _ = workload("C1_lap2d_128")
    s = O.baas_schedule(0, L.n, L.rowptr, L.colidx)
    gs, gp = s["group_set"], s["group_ptr"]
    np.testing.assert_array_equal(np.sort(gs), np.arange(L.n))
    gid = np.empty(L.n, np.int64)
    for q in range(s["ngroups"]):
        gid[gs[gp[q]:gp[q + 1]]] = q
    lvl = np.empty(s["ngroups"], np.int64)
    lvl[s["perm_to_orig"]] = np.repeat(np.arange(s["nlevels"]), np.diff(s["level_ptr"]))
    rows = np.repeat(np.arange(L.n), np.diff(L.rowptr))
    gi, gj = gid[rows], gid[L.colidx]
    assert np.all((gi == gj) | (lvl[gi] > lvl[gj]))


# ---------------------------------------------------------------- solves
@pytest.mark.parametrize("name", ["C1_lap2d_128", "C2_lap3d_64_perm1", "bcsstk18"])
def test_sptrsv_manufactured_solution(name):
    L, _, _ = workload(name)
    b = np.add.reduceat(L.values, L.rowptr[:-1])  # x = 1 (test/FusionDemo.cpp:47-77)
    x = O.sptrsv_lower(L, b)
    assert np.max(np.abs(x - 1.0)) <= 1e-10


@pytest.mark.parametrize("name", ["C1_lap2d_128", "C5_diff3d_48", "bcsstk18"])
def test_sptrsv_vs_scipy_and_transpose_identity(name):
    import scipy.sparse.linalg as spl
    L, _, _ = workload(name)
    rng = np.random.default_rng(0)
    b = rng.uniform(-1, 1, L.n)
    x = O.sptrsv_lower(L, b)
    xs = spl.spsolve_triangular(L.to_scipy(), b, lower=True)
    assert np.linalg.norm(x - xs) <= 1e-12 * np.linalg.norm(xs)
    U = g.Csr.from_scipy(L.to_scipy().T)
    gv = rng.standard_normal(L.n)
    lam = O.sptrsv_upper(U, gv)
    lam_a = O.sptrsv_upper(U, gv, descending=False)
    assert np.linalg.norm(lam - lam_a) <= 1e-13 * np.linalg.norm(lam)
    assert abs(x @ gv - b @ lam) <= 1e-11 * abs(x @ gv) + 1e-300


def test_sptrsv_levels_equals_serial_bitwise():
    L, _, _ = workload("C1_lap2d_128")
    b = g.rhs_uniform(L.n, 0)
    lp, _, pto = O.levelset_rows(L.n, L.rowptr, L.colidx)
    np.testing.assert_array_equal(O.sptrsv_lower_levels_omp(L, b, lp, pto, 4), O.sptrsv_lower(L, b))
    B = g.rhs_uniform(L.n, 1, k=3)
    X = O.sptrsv_lower_multi(L, B)
    for r in range(3):
        np.testing.assert_array_equal(X[:, r], O.sptrsv_lower(L, B[:, r].copy()))


def test_planted_mistake_is_caught():
    """A dropped term in the solve must fail the manufactured-solution check."""
    L, _, _ = workload("C1_lap2d_128")
    b = np.add.reduceat(L.values, L.rowptr[:-1])
    bad = g.Csr(L.n, L.rowptr, L.colidx, L.values.copy())
    bad.values[L.rowptr[100]] = 0.0  # as if the first off-diagonal term were skipped
    assert np.max(np.abs(O.sptrsv_lower(bad, b) - 1.0)) > 1e-6


# ---------------------------------------------------------------- factorizations
@pytest.mark.parametrize("name", ["C2_lap3d_64", "C2_lap3d_64_perm1", "C3_knn", "C5_diff3d_48",
                                  "bcsstk18", "bcsstk38"])
def test_ic0_matches_reference_golden(golden, name):
    man, _ = golden
    _, _, A = workload(name)
    lu = O.ic0(A)
    want = man["ic0"][name]
    # bit-exact: the oracle restates the reference build's self-dot blocking
    np.testing.assert_array_equal(lu[want["pick"]], want["values"])
    lower = A.colidx <= np.repeat(np.arange(A.n), np.diff(A.rowptr))
    assert abs(lu[lower].sum() - want["sum_lower"]) <= 1e-12 * abs(want["norm_lower"]) * np.sqrt(A.n)


@pytest.mark.skipif(O.ref is None, reason="oracle/_ref not built here")
@pytest.mark.parametrize("name", ["C2_lap3d_64_perm1", "C3_knn", "bcsstk38"])
def test_ic0_vs_reference_full(name):
    _, _, A = workload(name)
    a, b = O.ic0(A), O.ic0(A, use_ref=True)
    lower = A.colidx <= np.repeat(np.arange(A.n), np.diff(A.rowptr))
    np.testing.assert_array_equal(a[lower], b[lower])


@pytest.mark.parametrize("name", ["C2_lap3d_64", "C5_diff3d_48"])
def test_ic0_identity_on_pattern(name):
    _, _, A = workload(name)
    lu = O.ic0(A)
    Lf, _ = O.ic0_split(A, lu)
    Ls = Lf.to_scipy()
    M = (Ls @ Ls.T).tocsr()
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    low = A.colidx <= rows
    got = np.asarray(M[rows[low], A.colidx[low]]).ravel()
    np.testing.assert_allclose(got, A.values[low], rtol=1e-12, atol=1e-12)


def test_ic0_nonpositive_pivot_keeps_diagonal():
    # a 2x2 matrix whose Schur complement is negative: L_11 keeps A_11
    A = g.Csr(2, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
              np.array([1.0, 2.0, 2.0, 1.0]))
    lu = O.ic0(A)
    assert lu[0] == 1.0 and lu[2] == 2.0 and lu[3] == 1.0
    if O.ref is not None:
        np.testing.assert_array_equal(lu, O.ic0(A, use_ref=True))


@pytest.mark.parametrize("name", ["C2_convdiff_64", "C2_lap3d_64", "bcsstk18"])
def test_ilu0_matches_reference_golden(golden, name):
    man, _ = golden
    _, _, A = workload(name)
    lu = O.ilu0(A)
    want = man["ilu0"][name]
    np.testing.assert_array_equal(lu[want["pick"]], want["values"])


def test_ilu0_identity_on_pattern():
    import scipy.sparse as sp
    _, _, A = workload("C2_convdiff_64")
    lu = O.ilu0(A)
    F = g.Csr(A.n, A.rowptr, A.colidx, lu).to_scipy()
    Lm = sp.tril(F, k=-1) + sp.identity(A.n)
    Um = sp.triu(F)
    M = (Lm @ Um).tocsr()
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    got = np.asarray(M[rows, A.colidx]).ravel()
    np.testing.assert_allclose(got, A.values, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- BLAS / loaders
def test_blas_golden(golden):
    _, arr = golden
    _, _, K = workload("bcsstk18")
    np.testing.assert_array_equal(O.spmv(K, arr["blas_x"]), arr["blas_spmv"])
    x, y = arr["blas_x"][:1000], arr["blas_y"][:1000]
    for key in arr:
        if key.startswith("blas_waxpby_"):
            a, b = (float(t) for t in key.split("_")[2:])
            np.testing.assert_array_equal(O.waxpby(a, x, b, y), arr[key], err_msg=key)
    import scipy.sparse as sp  # noqa: F401
    np.testing.assert_allclose(O.spmv(K, arr["blas_x"]), K.to_scipy() @ arr["blas_x"], rtol=1e-12,
                               atol=1e-9 * np.abs(K.values).max())


def test_matrix_market_loader_golden(golden):
    from baas_amd import mmio
    import os
    man, _ = golden
    path = os.path.join(os.path.dirname(__file__), "golden", "bcsstk18.mtx.gz")
    for lower, A in ((0, mmio.load_matrix_market(path)), (1, mmio.load_lower_matrix_market(path))):
        want = man["inputs"][f"bcsstk18_loader_lower{lower}"]
        assert A.n == want["n"] and A.nnz == want["nnz"]
        assert sha(A.rowptr) == want["rowptr"] and sha(A.colidx) == want["colidx"]
        assert sha(A.values) == want["values"]


# ---------------------------------------------------------------- PCG
@pytest.mark.parametrize("precision", [0, 1])
def test_pcg_converges_to_true_solution(precision):
    _, _, A = workload("bcsstk18")
    rng = np.random.default_rng(9)
    b = rng.standard_normal(A.n)
    x, k, hist = O.pcg(A, b, maxiter=3000, tol=1e-8, precision=precision)
    assert k > 0 and hist[k] < 1e-8 and hist[k - 1] >= 1e-8
    S = A.to_scipy()
    assert np.linalg.norm(b - S @ x) <= 1.1e-8 * np.linalg.norm(b)


def test_pcg_mixed_close_to_fp64_iterations():
    _, _, A = workload("C3_knn")
    rng = np.random.default_rng(2)
    b = rng.standard_normal(A.n)
    _, k64, _ = O.pcg(A, b, maxiter=2000, precision=0)
    _, k32, _ = O.pcg(A, b, maxiter=2000, precision=1)
    assert k64 > 0 and k32 > 0 and abs(k64 - k32) <= max(3, k64 // 10)
