"""The allocation-free solve (baas_sptrsv_f64_noalloc) through the C-ABI.

It must give the bits of baas_sptrsv_f64 on every plan kind, forward and
backward, on two calls in a row whose second input is smaller, reject
solves beyond its fixed scratch, and leave free device memory unchanged
(baas_sptrsv_f64 allocates a plan's publication area on its first dataflow
call; this entry point never does).
"""
import ctypes as C

import numpy as np
import pytest

from baas_amd import _lib
from baas_amd import matgen as g
from gpu_util import announce, down, up

pytestmark = pytest.mark.gpu

# (plan mode, rows per unit, threads per workgroup, dataflow kernel)
PLANS = [(5, 900, 1024, 2), (6, 1024, 128, 5), (2, 64, 512, 2), (1, 4096, 256, 1), (1, 1024, 256, 0), (3, 0, 256, 0)]


@pytest.fixture(scope="module")
def handle():
    from baas_amd.api import Handle
    return Handle()


def _noalloc(h, s, uplo, vals, b, x, nrhs):
    lib = _lib.load()
    return lib.baas_sptrsv_f64_noalloc(h.ptr, s.ptr, int(uplo), C.c_void_p(vals.data_ptr()),
                                       C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), int(nrhs))


def _case(h, L, plan, nrhs, seed):
    import torch
    from baas_amd.api import Schedule
    s = Schedule(h, L.rowptr, L.colidx, kind=0)
    mode, rpu, blk, df = plan
    if mode == 5 and L.n != 64 ** 3:
        mode, rpu = 1, 512  # boxes need the 3D grid; ranges elsewhere (PIPE plans take 2D grids too)
    s.replan(mode, rpu, blk, dataflow=df)
    lv = up(L.values)
    uv = s.transpose_values(lv)
    b = up(g.rhs_uniform(L.n, seed, k=nrhs) if nrhs > 1 else g.rhs_uniform(L.n, seed))
    out = []
    for uplo, vals in ((0, lv), (1, uv)):
        want = s.solve(vals, b, uplo=uplo)
        got = torch.full_like(b, float("nan"))
        st = _noalloc(h, s, uplo, vals, b, got, nrhs)
        assert st == _lib.OK, _lib.BaasError(st, "baas_sptrsv_f64_noalloc")
        h.check()
        out.append((down(got), down(want)))
    return out


@pytest.mark.parametrize("plan", PLANS)
def test_noalloc_equals_sptrsv_two_calls_smaller_second(handle, plan):
    big = g.laplace3d(64).lower()
    small = g.laplace2d_lower(128)
    for tag, L, nrhs in (("64^3", big, 1), ("128^2", small, 1), ("128^2 x3", small, 3), ("128^2 x4", small, 4)):
        announce("noalloc", plan, tag)
        for got, want in _case(handle, L, plan, nrhs, seed=len(tag)):
            np.testing.assert_array_equal(got, want, err_msg=f"{plan} {tag}")


def test_noalloc_rejects_beyond_scratch(handle):
    import torch
    from baas_amd.api import Schedule
    announce("noalloc too large")
    L = g.laplace3d(64).lower()
    nrhs = 16  # 2 x 262144 x 16 values > the fixed 4194304
    assert 2 * L.n * nrhs > _lib.NOALLOC_MAX_VALUES
    s = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s.replan(5, 900, 1024, dataflow=2)
    b = torch.zeros(L.n, nrhs, dtype=torch.float64, device="cuda")
    x = torch.full_like(b, 7.0)
    assert _noalloc(handle, s, 0, up(L.values), b, x, nrhs) == _lib.TOO_LARGE
    handle.check()
    assert bool(torch.all(x == 7.0))  # nothing was launched


def test_noalloc_leaves_free_memory_unchanged(handle):
    import torch
    from baas_amd.api import Schedule
    announce("noalloc free memory")
    L = g.laplace3d(64).lower()
    # warm both entry points' kernels on one schedule first
    s0 = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s0.replan(5, 900, 1024, dataflow=2)
    lv, b = up(L.values), up(g.rhs_uniform(L.n, 1))
    x = torch.empty_like(b)
    s0.solve(lv, b, x)
    assert _noalloc(handle, s0, 0, lv, b, x, 1) == _lib.OK
    handle.check()
    # a fresh plan: baas_sptrsv_f64 would allocate its publication area here
    s1 = Schedule(handle, L.rowptr, L.colidx, kind=0)
    s1.replan(5, 900, 1024, dataflow=2)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        assert _noalloc(handle, s1, 0, lv, b, x, 1) == _lib.OK
    handle.check()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
    np.testing.assert_array_equal(down(x), down(s0.solve(lv, b)))
