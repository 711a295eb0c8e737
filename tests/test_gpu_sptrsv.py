"""GPU parity of the analysis and the triangular solves, through the C-ABI.

Schedules must equal the oracle's (itself pinned to the compiled reference)
bit for bit; solves must equal the oracle's serial solves bit for bit (same
FMA order, IEEE division), for every plan and block shape.
"""
import numpy as np
import pytest

import oracle as O
from baas_amd import matgen as g
from gpu_util import announce, down, up
from workloads import workload

pytestmark = pytest.mark.gpu

SCHED_CASES = [("C1_lap2d_128", 0), ("C2_lap3d_64", 0), ("C2_lap3d_64", 1), ("C2_lap3d_64_perm1", 1),
               ("C2_convdiff_64", 2), ("C2_convdiff_64_perm1", 2), ("C3_knn", 1), ("C5_diff3d_48", 0), ("bcsstk18", 1), ("bcsstk38", 0)]


@pytest.fixture(scope="module")
def handle():
    from baas_amd.api import Handle
    return Handle()


def _analyze(h, name, kind, ncore=8, **kw):
    from baas_amd.api import Schedule
    L, thr, A = workload(name)
    if kind == 0:
        return Schedule(h, L.rowptr, L.colidx, kind=0, ncore=ncore, **kw), L
    return Schedule(h, A.rowptr, A.colidx, kind=kind, ncore=ncore, **kw), L


@pytest.mark.parametrize("name,kind", SCHED_CASES)
def test_schedule_bit_exact(handle, name, kind):
    announce("schedule", name, kind)
    L, thr, A = workload(name)
    for ncore in (1, 8):
        s, _ = _analyze(handle, name, kind, ncore)
        got = s.schedule_dict()
        want = O.baas_schedule(kind, L.n, L.rowptr, L.colidx, L.rowptr if kind == 0 else thr, ncore)
        for k, v in want.items():
            if isinstance(v, np.ndarray):
                np.testing.assert_array_equal(got[k], v, err_msg=f"{name}/{kind}/{ncore}/{k}")
            else:
                assert got[k] == v, (name, kind, ncore, k)
        np.testing.assert_array_equal(s.export("group_cost"),
                                      O.group_cost(want["group_ptr"], want["group_set"], L.rowptr))
        lp, ntl, pto = O.levelset_rows(L.n, L.rowptr, L.colidx)
        np.testing.assert_array_equal(s.export("row_level_ptr"), lp)
        np.testing.assert_array_equal(s.export("row_level"), ntl)
        np.testing.assert_array_equal(s.export("row_perm"), pto)


def test_reduced_dag_matches_oracle(handle):
    announce("reduced dag C3")
    L, _, _ = workload("C3_knn")
    s, _ = _analyze(handle, "C3_knn", 0)
    cp, ri = O.csc_of(L.n, L.rowptr, L.colidx)
    rp, rs = O.two_hop(L.n, cp, ri)
    np.testing.assert_array_equal(s.export("reduced_ptr"), rp)
    np.testing.assert_array_equal(s.export("reduced_set"), rs)


# (plan mode, rows per unit, threads per workgroup, dataflow kernel)
PLANS = [(1, 0, 256, False), (1, 1024, 256, False), (1, 16384, 512, False), (2, 0, 256, False),
         (2, 64, 64, False), (2, 1024, 1024, False), (1, 4096, 256, True), (1, 12288, 512, True),
         (1, 2048, 192, True), (2, 256, 256, True), (2, 2048, 512, True), (1, 8192, 1024, True),
         (3, 0, 256, False), (3, 0, 1024, False), (4, 16384, 256, True), (4, 32768, 512, True),
         (4, 8192, 1024, False), (1, 1024, 1024, 2), (1, 512, 512, 2), (1, 4096, 128, 2), (2, 256, 512, 2),
         (4, 16384, 1024, 2), (3, 1, 1024, False), (3, 4, 512, False), (2, 64, 1024, 2), (2, 32, 512, 2),
         (1, 512, 1024, 3), (2, 32, 512, 4), (4, 16384, 1024, 4), (1, 256, 256, 3)]
SOLVE_CASES = ["C1_lap2d_128", "C2_lap3d_64", "C2_lap3d_64_perm1", "C5_diff3d_48", "bcsstk18", "bcsstk38"]


@pytest.mark.parametrize("name", SOLVE_CASES)
def test_forward_backward_bit_exact_all_plans(handle, name):
    import torch
    from baas_amd.api import Schedule
    L, _, _ = workload(name)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    b = g.rhs_uniform(L.n, 0)
    x_ref = O.sptrsv_lower(L, b)
    U = g.Csr.from_scipy(L.to_scipy().T)
    y_ref = O.sptrsv_upper(U, b)  # descending sum
    lv, bd = up(L.values), up(b)
    uv = s.transpose_values(lv)
    np.testing.assert_array_equal(down(uv), U.values)
    for mode, rpu, blk, df in PLANS:
        announce("solve", name, mode, rpu, blk, df)
        s.replan(mode, rpu, blk, dataflow=df)
        for rep in range(2):
            x = s.solve(lv, bd)
            y = s.solve(uv, bd, uplo=1)
            handle.check()
            np.testing.assert_array_equal(down(x), x_ref, err_msg=f"fwd {name} {mode} {rpu} {blk} {df} {rep}")
            np.testing.assert_array_equal(down(y), y_ref, err_msg=f"bwd {name} {mode} {rpu} {blk} {df} {rep}")
    torch.cuda.synchronize()


@pytest.mark.parametrize("name", ["C1_lap2d_128", "C2_lap3d_64_perm1", "bcsstk18"])
def test_manufactured_solution_and_levelsync(handle, name):
    from baas_amd.api import Schedule
    announce("manufactured", name)
    L, _, _ = workload(name)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    b = np.add.reduceat(L.values, L.rowptr[:-1])
    x = down(s.solve(up(L.values), up(b)))
    handle.check()
    assert np.max(np.abs(x - 1.0)) <= 1e-10  # test/FusionDemo.cpp:66-77
    xl = down(s.solve_levelsync(up(L.values), up(b)))
    np.testing.assert_array_equal(xl, O.sptrsv_lower(L, b))


@pytest.mark.parametrize("k", [2, 3, 32])
def test_multi_rhs_bit_exact(handle, k):
    from baas_amd.api import Schedule
    for name in ("C1_lap2d_128", "C2_lap3d_64_perm1"):
        L, _, _ = workload(name)
        s = Schedule(handle, L.rowptr, L.colidx, kind=0)
        B = g.rhs_uniform(L.n, 3, k=k)
        X_ref = O.sptrsv_lower_multi(L, B)
        for mode, rpu, blk, df in [(1, 0, 256, 0), (1, 256, 256, 0), (2, 0, 256, 0), (3, 0, 512, 0),
                                   (4, 16384, 256, 0), (3, 1, 1024, 0), (1, 32, 1024, 2), (2, 16, 512, 2),
                                   (1, 300, 128, 2), (1, 256, 1024, 3), (2, 32, 512, 4), (1, 512, 1024, 4),
                                   (1, 100, 64, 3)]:
            announce("multi", name, k, mode, rpu, blk, df)
            s.replan(mode, rpu, blk, dataflow=df)
            X = down(s.solve(up(L.values), up(B)))
            handle.check()
            np.testing.assert_array_equal(X, X_ref)
            U = g.Csr.from_scipy(L.to_scipy().T)
            Y = down(s.solve(s.transpose_values(up(L.values)), up(B), uplo=1))
            Y_ref = np.stack([O.sptrsv_upper(U, B[:, r].copy()) for r in range(k)], axis=1)
            np.testing.assert_array_equal(Y, Y_ref)


def test_f32_bit_exact(handle):
    import torch
    from baas_amd.api import Schedule
    for name in ("C2_lap3d_64", "C3_knn"):
        announce("f32", name)
        L, _, _ = workload(name)
        s = Schedule(handle, L.rowptr, L.colidx, kind=0)
        b = g.rhs_uniform(L.n, 1)
        for mode, rpu, df in ((1, 4096, False), (1, 4096, True), (3, 0, False), (4, 16384, True), (1, 1024, 2),
                              (1, 1024, 3), (2, 32, 4)):
            s.replan(mode, rpu, 256, dataflow=df)
            x = down(s.solve(up(L.values, torch.float32), up(b, torch.float32)))
            handle.check()
            np.testing.assert_array_equal(x, O.sptrsv_lower_f32(L, b))
        U = g.Csr.from_scipy(L.to_scipy().T)
        uv32 = s.transpose_values(up(L.values)).to(torch.float32)
        y = down(s.solve(uv32, up(b, torch.float32), uplo=1))
        np.testing.assert_array_equal(y, O.sptrsv_upper_f32(U, b))


@pytest.mark.parametrize("name", ["C1_lap2d_128", "C2_lap3d_64", "C3_knn", "bcsstk38"])
def test_tile_plans_bit_exact(handle, name):
    """TILES plans over row ranges and level bands, from one-level bands to
    a single row range, forward and backward, dataflow and unit kernels."""
    from baas_amd.api import Schedule
    L, _, _ = workload(name)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    b = g.rhs_uniform(L.n, 5)
    x_ref = O.sptrsv_lower(L, b)
    U = g.Csr.from_scipy(L.to_scipy().T)
    y_ref = O.sptrsv_upper(U, b)
    lv, bd = up(L.values), up(b)
    uv = s.transpose_values(lv)
    for rr, band, blk, df in [(8192, 8, 384, True), (32768, 32, 256, True), (65536, 1, 512, True),
                              (1 << 30, 16, 256, True), (4096, 64, 512, False), (16384, 16, 1024, False),
                              (8192, 16, 1024, 2)]:
        announce("tiles", name, rr, band, blk, df)
        s.replan_tiles(rr, band, blk, dataflow=df)
        pl = s.plan()
        assert pl["mode"] == 4 and pl["band"] == band and pl["block"] == blk
        x = s.solve(lv, bd)
        y = s.solve(uv, bd, uplo=1)
        handle.check()
        np.testing.assert_array_equal(down(x), x_ref, err_msg=f"fwd {name} {rr} {band} {blk} {df}")
        np.testing.assert_array_equal(down(y), y_ref, err_msg=f"bwd {name} {rr} {band} {blk} {df}")


@pytest.mark.parametrize("name", ["C1_lap2d_128", "C2_lap3d_64", "C5_diff3d_48"])
def test_box_plans_bit_exact(handle, name):
    """BOXES plans on stencil grids in natural order (recognised from the
    pattern): every kernel, forward and backward, bit for bit."""
    from baas_amd.api import Schedule
    L, _, _ = workload(name)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    b = g.rhs_uniform(L.n, 6)
    x_ref = O.sptrsv_lower(L, b)
    U = g.Csr.from_scipy(L.to_scipy().T)
    y_ref = O.sptrsv_upper(U, b)
    lv, bd = up(L.values), up(b)
    uv = s.transpose_values(lv)
    for rpu, blk, df in [(512, 1024, 2), (512, 512, 2), (1024, 384, 1), (512, 256, 0), (64, 256, 2), (100, 128, 2),
                         (900, 1024, 2), (1000, 1024, 2)]:
        announce("boxes", name, rpu, blk, df)
        s.replan(5, rpu, blk, dataflow=df)
        x = s.solve(lv, bd)
        y = s.solve(uv, bd, uplo=1)
        handle.check()
        np.testing.assert_array_equal(down(x), x_ref, err_msg=f"fwd {name} {rpu} {blk} {df}")
        np.testing.assert_array_equal(down(y), y_ref, err_msg=f"bwd {name} {rpu} {blk} {df}")


PIPE_GRIDS = [("C1_lap2d_128", None), ("C2_lap3d_64", None), ("C5_diff3d_48", None), ("grid", (13, 11, 9)),
              ("grid", (70, 5)), ("grid", (5, 17, 6)), ("grid", (20, 3, 7)), ("grid", (9, 9, 9))]


def _pipe_matrix(name, shape):
    if shape is not None:
        return g.laplace_grid(shape).lower(), f"grid{'x'.join(map(str, shape))}"
    return workload(name)[0], name


@pytest.mark.parametrize("name,shape", PIPE_GRIDS)
def test_pipe_plans_bit_exact(handle, name, shape):
    """PIPE plans (boxes swept by one wavefront as a diagonal pipeline) on
    stencil grids, partial boxes and narrow grids included: forward and
    backward bit for bit, every box depth; then four and two right-hand
    sides (two a lane: two pairs, one pair), each column against the
    oracle."""
    from baas_amd.api import Schedule
    L, tag = _pipe_matrix(name, shape)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    b = g.rhs_uniform(L.n, 7)
    x_ref = O.sptrsv_lower(L, b)
    U = g.Csr.from_scipy(L.to_scipy().T)
    y_ref = O.sptrsv_upper(U, b)
    lv, bd = up(L.values), up(b)
    uv = s.transpose_values(lv)
    B = g.rhs_uniform(L.n, 17, 4)
    X_ref = np.stack([O.sptrsv_lower(L, np.ascontiguousarray(B[:, c])) for c in range(4)], 1)
    Y_ref = np.stack([O.sptrsv_upper(U, np.ascontiguousarray(B[:, c])) for c in range(4)], 1)
    for rpu in (64, 256, 512, 1024, 2048):
        announce("pipe", tag, rpu)
        s.replan(6, rpu, 128, dataflow=5)
        for _ in range(2):  # the second sweep runs on the other publication half
            x = s.solve(lv, bd)
            y = s.solve(uv, bd, uplo=1)
            handle.check()
            np.testing.assert_array_equal(down(x), x_ref, err_msg=f"fwd {tag} {rpu}")
            np.testing.assert_array_equal(down(y), y_ref, err_msg=f"bwd {tag} {rpu}")
        for k in (4, 2):
            for _ in range(2):
                X = down(s.solve(lv, up(np.ascontiguousarray(B[:, :k]))))
                Y = down(s.solve(uv, up(np.ascontiguousarray(B[:, :k])), uplo=1))
                handle.check()
                np.testing.assert_array_equal(X, X_ref[:, :k], err_msg=f"fwd {tag} {rpu} {k} rhs")
                np.testing.assert_array_equal(Y, Y_ref[:, :k], err_msg=f"bwd {tag} {rpu} {k} rhs")


@pytest.mark.parametrize("name,shape", [("C2_lap3d_64", None), ("grid", (70, 5))])
def test_pipe_plans_f32_and_multi_rhs(handle, name, shape):
    """FP32 sweeps on PIPE plans against the oracle's FP32 solves, and
    several right-hand sides (all in one launch, each with its own box
    faces) against the FP64 oracle: two calls per count, so both
    publication halves carry them, and a change of count in between."""
    import torch
    from baas_amd.api import Schedule
    L, tag = _pipe_matrix(name, shape)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s.replan(6, 1024, 128, dataflow=5)
    U = g.Csr.from_scipy(L.to_scipy().T)
    b = g.rhs_uniform(L.n, 8)
    announce("pipe f32", tag)
    lv32 = up(L.values, torch.float32)
    uv32 = up(U.values, torch.float32)
    x = s.solve(lv32, up(b, torch.float32))
    y = s.solve(uv32, up(b, torch.float32), uplo=1)
    handle.check()
    np.testing.assert_array_equal(down(x), O.sptrsv_lower_f32(L, b))
    np.testing.assert_array_equal(down(y), O.sptrsv_upper_f32(U, b))
    announce("pipe f32 2 rhs", tag)  # two a lane
    B2 = g.rhs_uniform(L.n, 21, 2)
    X = down(s.solve(lv32, up(B2, torch.float32)))
    Y = down(s.solve(uv32, up(B2, torch.float32), uplo=1))
    handle.check()
    for c in range(2):
        np.testing.assert_array_equal(X[:, c], O.sptrsv_lower_f32(L, np.ascontiguousarray(B2[:, c])))
        np.testing.assert_array_equal(Y[:, c], O.sptrsv_upper_f32(U, np.ascontiguousarray(B2[:, c])))
    lv = up(L.values)
    uv = s.transpose_values(lv)
    for k in ((3, 32) if L.n > 100000 else (5, 2)):
        announce("pipe rhs", tag, k)
        B = g.rhs_uniform(L.n, 9 + k, k)
        want_x = [O.sptrsv_lower(L, np.ascontiguousarray(B[:, c])) for c in range(k)]
        want_y = [O.sptrsv_upper(U, np.ascontiguousarray(B[:, c])) for c in range(k)]
        for rep in range(2):
            X = down(s.solve(lv, up(B)))
            Y = down(s.solve(uv, up(B), uplo=1))
            handle.check()
            for c in range(k):
                np.testing.assert_array_equal(X[:, c], want_x[c], err_msg=f"fwd {tag} {k} {c} {rep}")
                np.testing.assert_array_equal(Y[:, c], want_y[c], err_msg=f"bwd {tag} {k} {c} {rep}")


@pytest.mark.parametrize("case", ["C2_lap3d_64_perm1", "C3_knn", "wrap", "hole"])
def test_pipe_plans_reject(handle, case):
    """PIPE plans need a full stencil grid in natural order: an unstructured
    matrix, a grid with one entry wrapping round from x = 0 to the previous
    line, and a grid missing one interior entry get UNSUPPORTED and keep
    their plan."""
    from baas_amd import _lib
    from baas_amd.api import Schedule
    announce("pipe reject", case)
    if case == "wrap":
        import scipy.sparse as sp
        M = g.laplace_grid((16, 16)).lower().to_scipy().tolil()
        M[32, 31] = -0.5  # row 32 is (x = 0, y = 2); column 31 is (15, 1)
        L = g.Csr.from_scipy(M.tocsr())
    elif case == "hole":
        M = g.laplace_grid((16, 16)).lower().to_scipy().tolil()
        M[40, 39] = 0.0  # row 40 is (8, 2): its x parent removed
        M = M.tocsr()
        M.eliminate_zeros()
        L = g.Csr.from_scipy(M)
    else:
        L = workload(case)[0]
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s.replan(1, 1024, 256)
    with pytest.raises(_lib.BaasError) as e:
        s.replan(6, 1024, 128)
    assert e.value.code == _lib.UNSUPPORTED
    assert s.plan()["mode"] == 1
    b = g.rhs_uniform(L.n, 0)
    np.testing.assert_array_equal(down(s.solve(up(L.values), up(b))), O.sptrsv_lower(L, b))


@pytest.mark.parametrize("name", ["C2_lap3d_64_perm1", "bcsstk18", "C3_knn"])
def test_box_plans_reject_unstructured(handle, name):
    """Matrices that are not stencil grids in natural order get UNSUPPORTED,
    and the schedule keeps its previous plan."""
    from baas_amd import _lib
    from baas_amd.api import Schedule
    announce("boxes reject", name)
    L, _, _ = workload(name)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s.replan(1, 1024, 256)
    with pytest.raises(_lib.BaasError) as e:
        s.replan(5, 512, 512)
    assert e.value.code == _lib.UNSUPPORTED
    assert s.plan()["mode"] == 1
    b = g.rhs_uniform(L.n, 0)
    np.testing.assert_array_equal(down(s.solve(up(L.values), up(b))), O.sptrsv_lower(L, b))


def _tiny_lower(kind):
    """Small lower triangles (diagonal last): one row, two rows, a dense 3 x 3,
    and a 1000-row bidiagonal chain (1000 levels)."""
    if kind == "one":
        return g.Csr(1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([2.5]))
    if kind == "two":
        return g.Csr(2, np.array([0, 1, 3], np.int32), np.array([0, 0, 1], np.int32), np.array([2.0, -1.0, 3.0]))
    if kind == "dense3":
        return g.Csr(3, np.array([0, 1, 3, 6], np.int32), np.array([0, 0, 1, 0, 1, 2], np.int32),
                     np.array([4.0, -1.0, 4.0, -0.5, -1.0, 4.0]))
    n = 1000
    rp = np.concatenate([[0, 1], 1 + 2 * np.arange(1, n)]).astype(np.int32)
    ci = np.concatenate([[0], np.stack([np.arange(n - 1), np.arange(1, n)], 1).ravel()]).astype(np.int32)
    va = np.concatenate([[3.0], np.tile([-1.0, 3.0], n - 1)])
    return g.Csr(n, rp, ci, va)


def _banded_lower(n, width, full, seed):
    """Lower triangle whose rows have up to `width` parents among the 2*width
    rows above them: all of the nearest `width` (full) or a random subset
    of each row's window (ragged, so chunks are partly padded)."""
    rng = np.random.default_rng(seed)
    rp, ci, va = [0], [], []
    for i in range(n):
        lo = max(0, i - 2 * width)
        if full:
            cols = list(range(max(0, i - width), i))
        else:
            pool = np.arange(lo, i)
            k = min(len(pool), int(rng.integers(0, width + 1)))
            cols = sorted(rng.choice(pool, size=k, replace=False).tolist()) if k else []
        ci += cols + [i]
        va += list(rng.uniform(-1.0, 1.0, len(cols))) + [float(width) + 1.0 + rng.uniform(0.0, 1.0)]
        rp.append(len(ci))
    return g.Csr(n, np.array(rp, np.int32), np.array(ci, np.int32), np.array(va))


@pytest.mark.parametrize("width", [1, 2, 3, 4, 5, 6, 7, 8, 9, 13])
def test_lane_chunk_widths(handle, width):
    """The lane kernel polls a row's parents in chunks sized from the longest
    row (2, 3, 4, 6 or 8 parents, two for longer rows), the level-set kernel
    gathers them in chunks of 2, 3, 4 or 8: every width, full and ragged
    rows, forward and backward, FP64 and FP32, one and eight right-hand
    sides, bit-exact."""
    import torch
    from baas_amd.api import Schedule
    for full in (True, False):
        L = _banded_lower(3000, width, full, seed=width)
        s = Schedule(handle, L.rowptr, L.colidx, kind=0)
        b = g.rhs_uniform(L.n, 4)
        U = g.Csr.from_scipy(L.to_scipy().T)
        x_ref, y_ref = O.sptrsv_lower(L, b), O.sptrsv_upper(U, b)
        lv, bd = up(L.values), up(b)
        uv = s.transpose_values(lv)
        for mode, rpu, blk in ((1, 512, 1024), (2, 32, 512), (3, 0, 256)):
            announce("chunk width", width, full, mode, rpu, blk)
            s.replan(mode, rpu, blk, dataflow=2 if mode != 3 else 0)
            x = s.solve(lv, bd)
            y = s.solve(uv, bd, uplo=1)
            handle.check()
            np.testing.assert_array_equal(down(x), x_ref, err_msg=f"fwd {width} {full} {mode}")
            np.testing.assert_array_equal(down(y), y_ref, err_msg=f"bwd {width} {full} {mode}")
        x32 = down(s.solve(up(L.values, torch.float32), up(b, torch.float32)))
        handle.check()
        np.testing.assert_array_equal(x32, O.sptrsv_lower_f32(L, b))
        B = np.random.default_rng(width).uniform(-1.0, 1.0, (L.n, 8))
        for mode, rpu, blk, df in ((3, 0, 256, 0), (2, 32, 1024, 2)):
            s.replan(mode, rpu, blk, dataflow=df)
            X = down(s.solve(lv, up(B)))
            handle.check()
            np.testing.assert_array_equal(X, O.sptrsv_lower_multi(L, B), err_msg=f"8 rhs {width} {full} {mode}")


@pytest.mark.parametrize("kind", ["one", "two", "dense3", "chain"])
def test_tiny_systems_every_plan(handle, kind):
    from baas_amd.api import Schedule
    L = _tiny_lower(kind)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    b = g.rhs_uniform(L.n, 9)
    x_ref = O.sptrsv_lower(L, b)
    U = g.Csr.from_scipy(L.to_scipy().T)
    y_ref = O.sptrsv_upper(U, b)
    lv, bd = up(L.values), up(b)
    uv = s.transpose_values(lv)
    plans = [(1, 0, 256, 0), (1, 64, 192, 1), (1, 64, 128, 2), (2, 0, 256, 0), (2, 16, 512, 2), (3, 0, 64, 0),
             (3, 1, 256, 0), (4, 64, 256, 2)] + ([(5, 100, 256, 2)] if kind == "chain" else [])
    for mode, rpu, blk, df in plans:
        announce("tiny", kind, mode, rpu, blk, df)
        if mode == 4:
            s.replan_tiles(rpu, 4, blk, dataflow=df)
        else:
            s.replan(mode, rpu, blk, dataflow=df)
        x = s.solve(lv, bd)
        y = s.solve(uv, bd, uplo=1)
        handle.check()
        np.testing.assert_array_equal(down(x), x_ref, err_msg=f"fwd {kind} {mode} {rpu} {blk} {df}")
        np.testing.assert_array_equal(down(y), y_ref, err_msg=f"bwd {kind} {mode} {rpu} {blk} {df}")


@pytest.mark.parametrize("kind", ["one", "two", "dense3", "chain"])
def test_tiny_factorisations(handle, kind):
    from baas_amd import _lib
    from baas_amd.api import Schedule
    L = _tiny_lower(kind)
    A = g.Csr.from_scipy((L.to_scipy() + L.to_scipy().T).tocsr())  # symmetric, diagonally dominant
    want_ic, want_ilu = O.ic0(A), O.ilu0(A)
    for fkind, want in ((_lib.KIND_IC0, want_ic), (_lib.KIND_ILU0, want_ilu)):
        s = Schedule(handle, A.rowptr, A.colidx, kind=fkind)
        av = up(A.values)
        for mode, rpu, blk, df in ((1, 0, 256, 0), (2, 16, 512, 2), (1, 64, 1024, 2), (3, 0, 64, 0)):
            announce("tiny factor", kind, fkind, mode, rpu, blk, df)
            s.replan(mode, rpu, blk, dataflow=df)
            lu = s.ic0(av) if fkind == _lib.KIND_IC0 else s.ilu0(av)
            st = handle.status()
            assert st in (_lib.OK, _lib.NONPOSITIVE_PIVOT)
            np.testing.assert_array_equal(down(lu), want)


def test_empty_system_rejected(handle):
    from baas_amd import _lib
    from baas_amd.api import Schedule
    announce("empty")
    with pytest.raises(_lib.BaasError) as e:
        Schedule(handle, np.zeros(1, np.int32), np.zeros(0, np.int32), kind=0)
    assert e.value.code == _lib.INVALID_ARG


def test_plan_saved_and_reapplied(handle):
    """An autotuned plan, saved as its five numbers and applied to a fresh
    schedule of the same matrix, is the same plan and gives the same bits."""
    import json
    from baas_amd.api import Schedule
    announce("plan save/apply")
    L, _, _ = workload("C2_lap3d_64")
    lv = up(L.values)
    b = up(g.rhs_uniform(L.n, 2))
    s1 = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s1.autotune(lv)
    saved = json.dumps(s1.plan())
    x1 = down(s1.solve(lv, b))
    s2 = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s2.apply_plan(json.loads(saved))
    assert s2.plan() == json.loads(saved)
    x2 = down(s2.solve(lv, b))
    handle.check()
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(x2, O.sptrsv_lower(L, down(b)))


def test_division_and_sqrt_are_ieee(handle):
    """The kernels' f64 division matches the CPU's correctly rounded one."""
    import torch
    from baas_amd.api import Schedule
    announce("ieee div")
    rng = np.random.default_rng(11)
    n = 4096
    # a diagonal-only L: x_i = b_i / d_i, one division per row
    L = g.Csr(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32),
              np.exp(rng.uniform(-30, 30, n)) * rng.choice([-1, 1], n))
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    b = rng.standard_normal(n) * np.exp(rng.uniform(-30, 30, n))
    x = down(s.solve(up(L.values), up(b)))
    handle.check()
    np.testing.assert_array_equal(x, b / L.values)


def test_watchdog_reports_timeout_without_hanging(handle):
    """A one-poll spin limit forces every kind of wait (per-row flags,
    dataflow exchange and LDS polls, level barriers) to give up: the kernel
    still exits and the error is reported (the values are then meaningless),
    and once the limit is restored the next solves are exact again."""
    from baas_amd import _lib
    from baas_amd.api import Schedule
    L, _, _ = workload("C2_lap3d_64_perm1")
    s = Schedule(handle, L.rowptr, L.colidx, kind=0, plan_mode=2)
    b = up(g.rhs_uniform(L.n, 0))
    want = O.sptrsv_lower(L, down(b))
    lv = up(L.values)
    for mode, rpu, blk, df in ((2, 0, 256, False), (1, 4096, 256, True), (3, 0, 256, False), (1, 1024, 512, 2)):
        announce("watchdog", mode, rpu, blk, df)
        s.replan(mode, rpu, blk, dataflow=df)
        handle.set_spin_limit(1)
        try:
            for _ in range(4):
                s.solve(lv, b)
            st = handle.status()
        finally:
            handle.set_spin_limit(1 << 22)
        assert st in (_lib.OK, _lib.TIMEOUT)
        for _ in range(3):  # both publication halves of the dataflow plan
            x = down(s.solve(lv, b))
            handle.check()
            np.testing.assert_array_equal(x, want)


def test_invalid_pattern_rejected(handle):
    from baas_amd import _lib
    from baas_amd.api import Schedule
    announce("invalid pattern")
    # row 1 misses its diagonal
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 0], np.int32)
    with pytest.raises(_lib.BaasError) as e:
        Schedule(handle, rp, ci, kind=0)
    assert e.value.code == _lib.INVALID_ARG


@pytest.mark.parametrize("name", ["C2_lap3d_64", "C2_lap3d_64_perm1"])
def test_autotune_keeps_bits(handle, name):
    from baas_amd.api import Schedule
    announce("autotune", name)
    L, _, _ = workload(name)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    lv = up(L.values)
    for uplo in (0, 1):
        vals = lv if uplo == 0 else s.transpose_values(lv)
        chosen = s.autotune(vals, uplo=uplo)
        print("[autotune]", name, uplo, chosen)
        assert chosen["mode"] in (1, 2, 3, 4, 5, 6) and chosen["us"] > 0
        b = g.rhs_uniform(L.n, 0)
        got = down(s.solve(vals, up(b), uplo=uplo))
        handle.check()
        if uplo == 0:
            np.testing.assert_array_equal(got, O.sptrsv_lower(L, b))
        else:
            U = g.Csr.from_scipy(L.to_scipy().T)
            np.testing.assert_array_equal(got, O.sptrsv_upper(U, b))


@pytest.mark.parametrize("name", ["C2_lap3d_64", "C2_lap3d_64_perm1", "C3_knn", "bcsstk18"])
def test_renumbered_by_unit_order(handle, name):
    """Locality renumbering: the plan's row order is a topological
    permutation; L renumbered by it (P L P^T) is lower triangular, its solve
    equals the oracle's on the renumbered system bit for bit, and mapped
    back it agrees with the solve in the original numbering to rounding."""
    from baas_amd.api import Schedule
    announce("renumber", name)
    L, _, _ = workload(name)
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s.autotune(up(L.values))
    order = s.unit_order()
    assert np.array_equal(np.sort(order), np.arange(L.n))
    pos = np.empty(L.n, np.int64)
    pos[order] = np.arange(L.n)
    rows = np.repeat(np.arange(L.n), np.diff(L.rowptr))
    assert np.all(pos[L.colidx] <= pos[rows])  # parents first
    LP = g.symmetric_permute(L, order)
    assert np.all(LP.colidx <= np.repeat(np.arange(L.n), np.diff(LP.rowptr)))
    sp_ = Schedule(handle, LP.rowptr, LP.colidx, kind=0)
    sp_.autotune(up(LP.values))
    b = g.rhs_uniform(L.n, 0)
    xp = down(sp_.solve(up(LP.values), up(b[order])))
    handle.check()
    np.testing.assert_array_equal(xp, O.sptrsv_lower(LP, b[order]))
    x = O.sptrsv_lower(L, b)
    back = np.empty_like(xp)
    back[order] = xp
    np.testing.assert_allclose(back, x, rtol=1e-12, atol=1e-12 * np.abs(x).max())


@pytest.mark.parametrize("name,kind,plan", [("C2_lap3d_64", 0, None), ("C3_knn", 1, None),
                                             ("C1_lap2d_128", 0, (4, 2048, 512, 2, 8)),
                                             ("C2_convdiff_64", 2, None)])
def test_schedule_save_load(handle, name, kind, plan):
    """A saved schedule rebuilds the same analysis, plan and bits without
    re-analysing; bytes for another pattern or kind, or damaged ones, are
    refused."""
    import time
    from baas_amd import _lib as L
    from baas_amd.api import Schedule
    announce("save/load", name, kind, plan)
    Lw, thr, A = workload(name)
    M = Lw if kind == 0 else A
    s = Schedule(handle, M.rowptr, M.colidx, kind=kind)
    if plan is None:
        if kind == 0:
            s.autotune(up(Lw.values))
        else:
            s.autotune_factor(up(A.values))
    else:
        s.replan_tiles(plan[1], plan[4], plan[2], plan[3])
    blob = s.save()
    t0 = time.perf_counter()
    r = Schedule(handle, M.rowptr, M.colidx, kind=kind, saved=blob)
    print("[save/load]", name, len(blob), "bytes, load", round(time.perf_counter() - t0, 4), "s", r.plan())
    assert r.plan() == s.plan() and r.info() == s.info()
    for k in list(L.EXPORTS):
        np.testing.assert_array_equal(r.export(k), s.export(k), err_msg=k)
    b = g.rhs_uniform(Lw.n, 0)
    if kind == 0:
        lv = up(Lw.values)
        np.testing.assert_array_equal(down(r.solve(lv, up(b))), down(s.solve(lv, up(b))))
    else:
        av = up(A.values)
        f = "ic0" if kind == 1 else "ilu0"
        np.testing.assert_array_equal(down(getattr(r, f)(av)), down(getattr(s, f)(av)))
    handle.check()
    with pytest.raises(L.BaasError):  # another kind
        Schedule(handle, M.rowptr, M.colidx, kind=(kind + 1) % 3, saved=blob)
    # another valid pattern of the same size (a symmetric renumbering)
    import scipy.sparse as sps
    perm = np.random.default_rng(9).permutation(Lw.n)
    S = Lw.to_scipy()
    full = A if A is not None else g.Csr.from_scipy(S + S.T - sps.diags(S.diagonal()))
    other = g.symmetric_permute(full, perm)
    other = other.lower() if kind == 0 else other
    assert other.n == M.n and len(other.colidx) == len(M.colidx)
    with pytest.raises(L.BaasError):
        Schedule(handle, other.rowptr, other.colidx, kind=kind, saved=blob)
    with pytest.raises(L.BaasError):  # another size
        small = g.laplace2d_lower(32)
        Schedule(handle, small.rowptr, small.colidx, kind=kind, saved=blob)
    bad = bytearray(blob)
    bad[0] ^= 1
    with pytest.raises(L.BaasError):
        Schedule(handle, M.rowptr, M.colidx, kind=kind, saved=bytes(bad))
    with pytest.raises(L.BaasError):
        Schedule(handle, M.rowptr, M.colidx, kind=kind, saved=blob[: len(blob) // 2])
