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


@pytest.mark.parametrize("name", NAMES)
def test_group_cost(name):
    """costComputation: the compiled reference's doubles on the schedule's
    groups (where oracle/_ref is built), and the invariant that the costs of
    a partition of the rows add up to nnz(L); a 3-row case by hand."""
    L, _, _ = workload(name)
    s = O.baas_schedule(0, L.n, L.rowptr, L.colidx)
    c = O.group_cost(s["group_ptr"], s["group_set"], L.rowptr)
    assert c.dtype == np.int32 and len(c) == s["ngroups"]
    assert int(c.astype(np.int64).sum()) == L.nnz
    if O.ref is not None:
        r = O.group_cost(s["group_ptr"], s["group_set"], L.rowptr, L.colidx, use_ref=True)
        np.testing.assert_array_equal(c.astype(np.float64), r)
    # rows of lengths 1, 2, 3 in groups {0, 2} and {1}
    np.testing.assert_array_equal(O.group_cost(np.array([0, 2, 3]), np.array([0, 2, 1]), np.array([0, 1, 3, 6])),
                                  [4, 2])


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


@pytest.mark.parametrize("name", ["C2_convdiff_64", "C2_convdiff_64_perm1", "C2_lap3d_64", "bcsstk18"])
def test_ilu0_matches_reference_golden(golden, name):
    man, _ = golden
    _, _, A = workload(name)
    lu = O.ilu0(A)
    want = man["ilu0"][name]
    np.testing.assert_array_equal(lu[want["pick"]], want["values"])


@pytest.mark.parametrize("name", ["C2_convdiff_64", "C2_convdiff_64_perm1"])
def test_ilu0_identity_on_pattern(name):
    import scipy.sparse as sp
    _, _, A = workload(name)
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


# ---------------------------------------------------------------- dcsrmv, IC0 split
def test_dcsrmv_golden_and_closed_forms(golden):
    """y = alpha (A x) + beta y against the compiled reference's dcsrmv
    (blas/spmv.cpp:29-45), and alpha = 1, beta = 0 against SpMV bit for bit."""
    _, arr = golden
    _, _, K = workload("bcsstk18")
    x, y = arr["blas_x"], arr["blas_y"]
    keys = [k for k in arr if k.startswith("blas_dcsrmv_")]
    assert len(keys) == 4
    for key in keys:
        a, b = (float(t) for t in key.split("_")[2:])
        np.testing.assert_array_equal(O.dcsrmv(K, x, y, a, b), arr[key], err_msg=key)
    np.testing.assert_array_equal(O.dcsrmv(K, x, y, 1.0, 0.0), O.spmv(K, x))
    # the residual form used by PCG: b - A x
    np.testing.assert_allclose(O.dcsrmv(K, x, y, -1.0, 1.0), y - K.to_scipy() @ x, rtol=0,
                               atol=1e-9 * np.abs(K.values).max() * np.abs(x).max())
    if O.ref is not None:
        for a, b in ((2.0, -0.5), (-0.75, 3.0)):
            np.testing.assert_array_equal(O.dcsrmv(K, x, y, a, b), O.dcsrmv(K, x, y, a, b, use_ref=True))


@pytest.mark.parametrize("name", ["C3_knn", "bcsstk18"])
def test_ic0_split_upper_is_transpose_of_lower(name):
    """U = L^T exactly: same pattern as scipy's transpose and the same bits
    (test/iccg_level_merge.cpp:817-891 fills U by scattering L's entries)."""
    import scipy.sparse as sp
    _, _, A = workload(name)
    Lf, Uf = O.ic0_split(A, O.ic0(A))
    T = sp.csr_matrix(Lf.to_scipy().T)
    T.sort_indices()
    np.testing.assert_array_equal(Uf.rowptr, T.indptr)
    np.testing.assert_array_equal(Uf.colidx, T.indices)
    np.testing.assert_array_equal(Uf.values, T.data)
    # lower rows end with the diagonal, upper rows start with it
    np.testing.assert_array_equal(Lf.colidx[Lf.rowptr[1:] - 1], np.arange(A.n))
    np.testing.assert_array_equal(Uf.colidx[Uf.rowptr[:-1]], np.arange(A.n))


# ---------------------------------------------------------------- FP32 sweeps
def _fma32(a, b, c):
    """Correctly rounded float32 fma(a, b, c) from float64 arithmetic.

    The product of two float32 values is exact in float64.  s = fl64(p + c)
    and its error e (TwoSum, exact) give p + c = s + e.  Rounding s to float32
    is then the correctly rounded result unless s lies exactly halfway
    between two float32 values (float32 midpoints are float64 numbers), in
    which case the sign of e decides.  No subnormals occur in these tests
    (asserted by the caller)."""
    p = a.astype(np.float64) * b.astype(np.float64)
    c = c.astype(np.float64)
    s = p + c
    bb = s - p
    e = (p - (s - bb)) + (c - bb)
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    toward = np.where(s > r64, np.float32(np.inf), np.float32(-np.inf)).astype(np.float32)
    other = np.nextafter(r, toward)
    mid = (r64 + other.astype(np.float64)) * 0.5
    at_mid = (s != r64) & (s == mid)
    up_ = at_mid & (e != 0) & ((e > 0) == (other.astype(np.float64) > r64))
    return np.where(up_, other, r).astype(np.float32)


def _levels(n, rowptr, colidx, upper):
    lev = np.zeros(n, np.int64)
    order = range(n - 1, -1, -1) if upper else range(n)
    for i in order:
        lo, hi = (rowptr[i] + 1, rowptr[i + 1]) if upper else (rowptr[i], rowptr[i + 1] - 1)
        if hi > lo:
            lev[i] = lev[colidx[lo:hi]].max() + 1
    return lev


def _sweep32_numpy(T, b, upper):
    """Independent FP32 triangular sweep: level by level, every row of a level
    at once, entry position by entry position in the summation order of the
    lower solve (ascending columns, merge/SpTRSV.cpp:24-39) or the backward
    P2P solve (descending columns, merge/SpTRSV.cpp:643-652)."""
    n, rp, ci = T.n, T.rowptr, T.colidx
    v = T.values.astype(np.float32)
    b = b.astype(np.float32)
    x = np.zeros(n, np.float32)
    lev = _levels(n, rp, ci, upper)
    diag = rp[:-1] if upper else rp[1:] - 1
    first = rp[:-1] + 1 if upper else rp[:-1]
    last = rp[1:] if upper else rp[1:] - 1  # off-diagonal range [first, last)
    for l in range(lev.max() + 1):
        rows = np.flatnonzero(lev == l)
        cnt = last[rows] - first[rows]
        s = np.zeros(rows.size, np.float32)
        for k in range(int(cnt.max()) if rows.size else 0):
            m = cnt > k
            pos = (last[rows[m]] - 1 - k) if upper else (first[rows[m]] + k)
            s[m] = _fma32(v[pos], x[ci[pos]], s[m])
        x[rows] = (b[rows] - s) / v[diag[rows]]
    assert np.all((x == 0) | (np.abs(x) >= np.finfo(np.float32).tiny))
    return x


def _fp32_cases():
    L1, _, _ = workload("C1_lap2d_128")
    _, _, A3 = workload("C3_knn")
    Lk, Uk = O.ic0_split(A3, O.ic0(A3))
    U1 = g.Csr.from_scipy(L1.to_scipy().T)
    return {"C1_lap2d_128": (L1, U1), "C3_knn_ic0": (Lk, Uk)}


@pytest.mark.parametrize("name", ["C1_lap2d_128", "C3_knn_ic0"])
def test_fp32_sweeps_equal_independent_numpy(name):
    Lt, Ut = _fp32_cases()[name]
    b = g.rhs_uniform(Lt.n, 3).astype(np.float32)
    np.testing.assert_array_equal(O.sptrsv_lower_f32(Lt, b), _sweep32_numpy(Lt, b, upper=False))
    np.testing.assert_array_equal(O.sptrsv_upper_f32(Ut, b), _sweep32_numpy(Ut, b, upper=True))


def test_fma32_emulation_against_exact_rationals():
    """The float32 fma above against exact rational arithmetic, including
    values built to land on float32 midpoints."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    a = rng.standard_normal(3000).astype(np.float32)
    b = rng.standard_normal(3000).astype(np.float32)
    c = rng.standard_normal(3000).astype(np.float32)
    # near-midpoint cases: a*b + c a hair off a float32 half-ulp
    a[:500] = np.float32(1.0) + np.float32(2.0 ** -23) * rng.integers(1, 1000, 500).astype(np.float32)
    b[:500] = np.float32(1.0) + np.float32(2.0 ** -23) * rng.integers(1, 1000, 500).astype(np.float32)
    c[:500] = np.float32(2.0 ** -24) * rng.choice([-3, -1, 1, 3], 500).astype(np.float32)
    # exact-midpoint cases: a*b = 3 (1 + m 2^-23), m odd, lies halfway between
    # two float32 numbers and a tiny c (lost in fl64(p + c)) decides the side
    a[500:1000] = np.float32(3.0)
    b[500:1000] = np.float32(1.0) + np.float32(2.0 ** -23) * (2 * rng.integers(0, 500, 500) + 1).astype(np.float32)
    c[500:1000] = np.float32(2.0 ** -60) * rng.choice([-1, 0, 1], 500).astype(np.float32)
    p = a[500:1000].astype(np.float64) * b[500:1000].astype(np.float64)
    assert np.all(p.astype(np.float32).astype(np.float64) != p)  # really halfway cases
    got = _fma32(a, b, c)

    def exact(ai, bi, ci):
        v = Fraction(float(ai)) * Fraction(float(bi)) + Fraction(float(ci))
        lo = np.float32(float(v))
        # correct rounding from the exact value: pick the nearer neighbour, ties to even
        cands = [np.nextafter(lo, np.float32(-np.inf)), lo, np.nextafter(lo, np.float32(np.inf))]
        best = min(cands, key=lambda f: (abs(Fraction(float(f)) - v),
                                         int(np.float32(f).view(np.int32)) & 1))
        return best
    want = np.array([exact(*t) for t in zip(a, b, c)], np.float32)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("name", ["C1_lap2d_128", "C3_knn_ic0"])
def test_fp32_sweeps_backward_error_bound(name):
    """Componentwise backward error of a substitution (Higham, Accuracy and
    Stability, Thm 8.5): the computed x satisfies |b - T x| <= gamma_m |T| |x|
    with m the longest row and gamma_m = m u / (1 - m u), u = 2^-24.  The
    residual is formed in float64 from the float32 data, where it is exact up
    to ~1e-16 relative."""
    Lt, Ut = _fp32_cases()[name]
    b = g.rhs_uniform(Lt.n, 4).astype(np.float32)
    u = 2.0 ** -24
    for T, x in ((Lt, O.sptrsv_lower_f32(Lt, b)), (Ut, O.sptrsv_upper_f32(Ut, b))):
        S = g.Csr(T.n, T.rowptr, T.colidx, T.values.astype(np.float32).astype(np.float64)).to_scipy()
        m = int(np.diff(T.rowptr).max())
        gam = m * u / (1 - m * u)
        x64 = x.astype(np.float64)
        r = np.abs(b.astype(np.float64) - S @ x64)
        bound = gam * (abs(S) @ np.abs(x64))
        assert np.all(r <= bound * (1 + 1e-6) + 1e-300)
    # a mistake planted in the oracle's input (one off-diagonal term dropped,
    # as a skipped index would) breaks the bound measured against the true L
    bad = g.Csr(Lt.n, Lt.rowptr, Lt.colidx, Lt.values.copy())
    i = Lt.n // 2
    assert Lt.rowptr[i + 1] - 1 > Lt.rowptr[i]
    bad.values[Lt.rowptr[i]] = 0.0
    x_bad = O.sptrsv_lower_f32(bad, b).astype(np.float64)
    S = g.Csr(Lt.n, Lt.rowptr, Lt.colidx, Lt.values.astype(np.float32).astype(np.float64)).to_scipy()
    m = int(np.diff(Lt.rowptr).max())
    r = np.abs(b.astype(np.float64) - S @ x_bad)
    assert np.any(r > m * u / (1 - m * u) * (abs(S) @ np.abs(x_bad)))


def test_fp32_sweeps_manufactured_and_vs_fp64():
    """x = 1 manufactured solution in FP32 on the 2D Laplacian (b = row sums):
    the lower factor has ||D^-1 N||_inf = 1/2, so its Skeel condition number
    is at most 3 and the error at most ~3 (m + 1) u = 9e-7; and the FP32 solve
    against the FP64 one on random b within the same bound, times |x|."""
    L, _, _ = workload("C1_lap2d_128")
    b = np.add.reduceat(L.values, L.rowptr[:-1]).astype(np.float32)
    x = O.sptrsv_lower_f32(L, b)
    assert np.max(np.abs(x.astype(np.float64) - 1.0)) <= 9e-7
    U = g.Csr.from_scipy(L.to_scipy().T)
    bu = np.add.reduceat(U.values, U.rowptr[:-1]).astype(np.float32)
    assert np.max(np.abs(O.sptrsv_upper_f32(U, bu).astype(np.float64) - 1.0)) <= 9e-7
    r = g.rhs_uniform(L.n, 8)
    x64 = O.sptrsv_lower(L, r)
    x32 = O.sptrsv_lower_f32(L, r.astype(np.float32)).astype(np.float64)
    err = np.abs(x32 - x64)
    assert np.all(err <= 9e-7 * np.abs(x64).max())
    assert err.max() > 0  # it really is a single-precision solve


# ---------------------------------------------------------------- adjoint reference
def test_iccg_adjoint_fd_reference_48():
    """The 48^3 finite-difference reference the GPU ICCG-adjoint test uses,
    against the adjoint formula dL/db = lambda, dL/dA_ij = -lambda_i x_j with
    lambda = A^-1 (x - x*) from the oracle's PCG; a transposed index
    (-lambda_j x_i) must fail the same 1e-6 bound."""
    from adjoint_fd import iccg_fd_48
    A, xs, b, rows, ents, fd_b, fd_a = iccg_fd_48()
    x, _, _ = O.pcg(A, b, maxiter=20000, tol=1e-14)
    lam, _, _ = O.pcg(A, x - xs, maxiter=20000, tol=1e-14)
    cols = A.colidx[ents]
    gb, ga = lam[rows], -(lam[rows] * x[cols])
    assert np.all(np.abs(fd_b - gb) <= 1e-6 * np.abs(gb))
    assert np.all(np.abs(fd_a - ga) <= 1e-6 * np.abs(ga))
    off = cols != rows
    wrong = -(lam[cols] * x[rows])
    assert np.any(np.abs(fd_a[off] - wrong[off]) > 1e-6 * np.abs(wrong[off]))


@pytest.mark.parametrize("name", ["C1_lap2d_128", "C3_knn", "bcsstk38"])
def test_sptrsv_p2p_omp_equals_serial_bitwise(name):
    """The point-to-point OpenMP restatement (the CPU baseline's P2P leg)
    gives the serial solve's bits at every thread count, call after call."""
    L, _, _ = workload(name)
    b = g.rhs_uniform(L.n, 2)
    want = O.sptrsv_lower(L, b)
    for nt in (1, 2, 3, 8):
        P = O.P2PSolver(L, nt)
        for _ in range(3):
            np.testing.assert_array_equal(P.solve(b), want, err_msg=f"{name} {nt}")
