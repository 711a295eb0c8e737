"""Argument checks of the Python front end (no GPU needed).

Every numeric call of baas_amd.api validates dtype, size, contiguity and
device before it hands raw pointers to the C-ABI; a mismatch raises
ValueError and nothing reaches the library.  The objects here are built
without calling the library (no handle or schedule exists on a CPU box), so
a check that failed to fire would show up as an attempt to call into it.
"""
import types

import pytest
import torch

from baas_amd import _lib as L
from baas_amd import api
from baas_amd.autograd import ICCGSolve


class _Poison:
    """Stands in for the handle/schedule pointers: any use raises."""

    def __getattr__(self, name):
        raise AssertionError("the C-ABI was reached before validation")


def _sched(n=6, nnz=14, nnz_lower=10, kind=L.KIND_TRSV):
    s = object.__new__(api.Schedule)
    s.h = types.SimpleNamespace(device=0, ptr=_Poison())
    s._s = None
    s.n, s.nnz, s.nnz_lower, s.kind = n, nnz, nnz_lower, kind
    return s


def _ok_vals(s, dtype=torch.float64):
    return torch.zeros(s.nnz_lower, dtype=dtype)


def _raises(fn, match):
    with pytest.raises(ValueError, match=match):
        fn()


def test_solve_rejects_dtype_mismatch():
    s = _sched()
    b = torch.zeros(s.n, dtype=torch.float64)
    _raises(lambda: s.solve(_ok_vals(s, torch.float32), b), "vals: dtype")
    _raises(lambda: s.solve(_ok_vals(s), b.to(torch.int32)), "b: expected a float64 or float32")


def test_solve_rejects_wrong_sizes():
    s = _sched()
    _raises(lambda: s.solve(_ok_vals(s), torch.zeros(s.n + 1, dtype=torch.float64)), "b: shape")
    _raises(lambda: s.solve(torch.zeros(s.nnz_lower - 1, dtype=torch.float64),
                            torch.zeros(s.n, dtype=torch.float64)), "vals: 9 elements")
    _raises(lambda: s.solve(_ok_vals(s), torch.zeros(s.n, 0, dtype=torch.float64)), "b: shape")
    _raises(lambda: s.solve(_ok_vals(s), torch.zeros(s.n, 3, dtype=torch.float64),
                            x=torch.zeros(s.n, 2, dtype=torch.float64)), "x: shape")


def test_solve_rejects_strided_views():
    s = _sched()
    B = torch.zeros(s.n, 4, dtype=torch.float64)
    _raises(lambda: s.solve(_ok_vals(s), B[:, 1]), "b: must be contiguous")
    _raises(lambda: s.solve(_ok_vals(s), B.t().contiguous().t()), "b: must be contiguous")
    v = torch.zeros(2 * s.nnz_lower, dtype=torch.float64)[::2]
    _raises(lambda: s.solve(v, torch.zeros(s.n, dtype=torch.float64)), "vals: must be contiguous")


def test_solve_rejects_host_memory():
    s = _sched()
    _raises(lambda: s.solve(_ok_vals(s), torch.zeros(s.n, dtype=torch.float64)), "b: on cpu")


def test_factor_and_split_checks():
    s = _sched(kind=L.KIND_IC0)
    _raises(lambda: s.ic0(torch.zeros(s.nnz, dtype=torch.float32)), "a_vals: dtype")
    _raises(lambda: s.ic0(torch.zeros(s.nnz - 1, dtype=torch.float64)), "a_vals: 13 elements")
    _raises(lambda: s.ilu0(torch.zeros(s.nnz, dtype=torch.float64)), "kind")
    _raises(lambda: s.ic0(torch.zeros(s.nnz, dtype=torch.float64)), "a_vals: on cpu")
    _raises(lambda: s.ic0_split(torch.zeros(s.nnz - 2, dtype=torch.float64)), "lu: 12 elements")
    _raises(lambda: s.ic0_split(torch.zeros(s.nnz, dtype=torch.float64), dtype=torch.float16), "dtype")


def test_grad_transpose_autotune_checks():
    s = _sched()
    x = torch.zeros(s.n, dtype=torch.float64)
    _raises(lambda: s.grad_vals(torch.zeros(s.n + 1, dtype=torch.float64), x), "lam: shape")
    _raises(lambda: s.grad_vals(x, torch.zeros(s.n - 1, dtype=torch.float64)), "x: shape")
    _raises(lambda: s.transpose_values(torch.zeros(3, dtype=torch.float64)), "lvals: 3 elements")
    _raises(lambda: s.autotune(torch.zeros(s.nnz_lower, dtype=torch.int64)), "vals: expected a float64")
    _raises(lambda: s.solve_levelsync(_ok_vals(s), torch.zeros(s.n, dtype=torch.float32)), "b: dtype")


def test_blas_checks():
    h = types.SimpleNamespace(device=0, ptr=_Poison())
    rp = torch.zeros(5, dtype=torch.int32)
    ci = torch.zeros(7, dtype=torch.int32)
    _raises(lambda: api.spmv(h, rp, ci, torch.zeros(7, dtype=torch.float32), torch.zeros(4, dtype=torch.float64)),
            "vals: dtype")
    _raises(lambda: api.spmv(h, rp, ci, torch.zeros(6, dtype=torch.float64), torch.zeros(4, dtype=torch.float64)),
            "vals: 6 elements")
    _raises(lambda: api.spmv(h, rp.to(torch.int64), ci, torch.zeros(7, dtype=torch.float64),
                             torch.zeros(4, dtype=torch.float64)), "rowptr: dtype")
    _raises(lambda: api.dot(h, torch.zeros(4, dtype=torch.float64), torch.zeros(5, dtype=torch.float64)),
            "y: 5 elements")
    _raises(lambda: api.axpby(h, torch.zeros(3, dtype=torch.float64), 1.0, torch.zeros(4, dtype=torch.float64),
                              1.0, torch.zeros(4, dtype=torch.float64)), "w: 3 elements")
    _raises(lambda: api.nrm2(h, torch.zeros(4, dtype=torch.float32)), "x: dtype")


def test_pcg_and_batch_checks():
    s = _sched(kind=L.KIND_IC0)
    p = object.__new__(api.PCG)
    p.h, p.s, p._p = s.h, s, _Poison()
    _raises(lambda: p.solve(torch.zeros(s.n + 1, dtype=torch.float64)), "b: shape")
    _raises(lambda: p.solve(torch.zeros(s.n, dtype=torch.float32)), "b: dtype")
    _raises(lambda: api.PCG(s.h, _sched(kind=L.KIND_TRSV), torch.zeros(14, dtype=torch.float64)), "kind IC0")
    _raises(lambda: api.PCG(s.h, s, torch.zeros(13, dtype=torch.float64)), "a_vals: 13 elements")
    bt = object.__new__(api.Batch)
    bt.h, bt.N, bt.k, bt.sizes, bt._b = s.h, 10, 4, [4, 6], _Poison()
    _raises(lambda: bt.solve(torch.zeros(10, 3, dtype=torch.float64)), "B: shape")
    _raises(lambda: bt.solve(torch.zeros(4, 10, dtype=torch.float64).t()), "B: must be contiguous")


def test_iccg_solve_refuses_a_foreign_values_tensor():
    """The forward pass uses the buffer the PCG was created on; a fresh tensor
    (say 2*a computed from parameters) would be silently ignored."""
    s = _sched(kind=L.KIND_IC0)
    p = object.__new__(api.PCG)
    a = torch.ones(s.nnz, dtype=torch.float64)
    p.h, p.s, p.a_vals, p._p = s.h, s, a, _Poison()
    b = torch.zeros(s.n, dtype=torch.float64)
    for other in (a.clone(), 2 * a, a.to(torch.float32)):
        with pytest.raises(ValueError, match="a_vals must be the tensor"):
            ICCGSolve.apply(other, b, p, 1e-10, 100)
