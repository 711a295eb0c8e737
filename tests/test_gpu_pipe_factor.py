"""GPU parity of the pipelined IC0 / ILU0 (PIPE plans over full 7- and 5-point
stencil matrices) through the C-ABI.

On such a matrix in natural order no two pattern rows share an earlier
column, so the reference's up-looking factorisations (merge/IC.cpp:32-89,
merge/ILU.cpp:350-386) divide each lower entry by its column's diagonal and
update only the diagonal; the pipelined kernel must give the oracle's bits
on the configurations' grid matrices, on ragged 3D and 2D grids, with
nonsymmetric values (ILU0), in place and out of place, on both publication
halves (two factorisations in a row), and report the pivot status the other
kernels report.  A matrix whose lower triangle is a full stencil but whose
full pattern is not falls back to the lane kernel on the same plan, with
the same bits.
"""
import numpy as np
import pytest

import oracle as O
from baas_amd import _lib
from baas_amd import matgen as g
from gpu_util import announce, down, up
from workloads import workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def handle():
    from baas_amd.api import Handle
    return Handle()


def _nonsym(A, seed):
    """A's pattern with values scaled by 1 + U(-0.3, 0.3) (still diagonally dominant)."""
    rng = np.random.default_rng(seed)
    v = A.values * (1.0 + rng.uniform(-0.3, 0.3, A.values.size))
    return g.Csr(A.n, A.rowptr.copy(), A.colidx.copy(), v)


def _matrix(name):
    if name in ("C2_lap3d_64", "C2_convdiff_64", "C5_diff3d_48"):
        return workload(name)[2]
    kind, shape = name.split(":")
    A = g.laplace_grid(tuple(int(v) for v in shape.split("x")))
    return _nonsym(A, 7) if kind == "ns" else A


CASES = [(_lib.KIND_IC0, "C2_lap3d_64"), (_lib.KIND_IC0, "C5_diff3d_48"), (_lib.KIND_IC0, "lap:13x11x9"),
         (_lib.KIND_IC0, "lap:5x17x6"), (_lib.KIND_IC0, "lap:9x9x9"), (_lib.KIND_IC0, "lap:70x5"),
         (_lib.KIND_ILU0, "C2_convdiff_64"), (_lib.KIND_ILU0, "ns:13x11x9"), (_lib.KIND_ILU0, "ns:20x3x7"),
         (_lib.KIND_ILU0, "ns:9x9x9"), (_lib.KIND_ILU0, "ns:70x5"), (_lib.KIND_ILU0, "lap:64x64")]


@pytest.mark.parametrize("kind,name", CASES)
def test_pipe_factor_bit_exact(handle, kind, name):
    import torch
    from baas_amd.api import Schedule
    A = _matrix(name)
    want = O.ic0(A) if kind == _lib.KIND_IC0 else O.ilu0(A)
    s = Schedule(handle, A.rowptr, A.colidx, kind=kind)
    fac = s.ic0 if kind == _lib.KIND_IC0 else s.ilu0
    av = up(A.values)
    for rpu in (64, 256, 1024, 2048):
        announce("pipe factor", kind, name, rpu)
        s.replan(6, rpu, 128, dataflow=5)
        assert s.pipe_info() == (True, True), (name, rpu)
        for rep in range(2):  # both publication halves
            lu = fac(av)
            handle.check()
            np.testing.assert_array_equal(down(lu), want, err_msg=f"{name} {rpu} {rep}")
        # in place: lu starts as A's values
        lu = av.clone()
        fac(lu, lu)
        handle.check()
        np.testing.assert_array_equal(down(lu), want, err_msg=f"{name} {rpu} in place")
        assert bool(torch.equal(av, up(A.values)))  # the input is left alone


def test_pipe_ic0_nonpositive_pivot(handle):
    """A Laplacian with its diagonal scaled down: IC0 meets non-positive
    pivots, keeps those diagonals as the reference does, and reports
    NONPOSITIVE_PIVOT; the factor equals the oracle's."""
    from baas_amd.api import Schedule
    A = g.laplace_grid((9, 8, 7))
    d = A.diagptr()
    A.values[d] *= 0.35
    want = O.ic0(A)
    s = Schedule(handle, A.rowptr, A.colidx, kind=_lib.KIND_IC0)
    s.replan(6, 256, 128, dataflow=5)
    assert s.pipe_info() == (True, True)
    announce("pipe ic0 nonpositive")
    lu = s.ic0(up(A.values))
    assert handle.status() == _lib.NONPOSITIVE_PIVOT  # (read and cleared)
    np.testing.assert_array_equal(down(lu), want)


def test_pipe_factor_falls_back_off_stencil(handle):
    """A full-stencil lower triangle whose full pattern has one more upper
    entry: the PIPE plan builds (its boxes follow L), the factorisation runs
    on the lane kernel, with the oracle's bits."""
    import scipy.sparse as sp
    from baas_amd.api import Schedule
    A = g.laplace_grid((10, 9, 8))
    M = A.to_scipy().tolil()
    M[5, 40] = -0.25  # an upper entry no stencil row holds
    B = _nonsym(g.Csr.from_scipy(sp.csr_matrix(M)), 3)
    want = O.ilu0(B)
    s = Schedule(handle, B.rowptr, B.colidx, kind=_lib.KIND_ILU0)
    s.replan(6, 512, 128, dataflow=5)
    assert s.pipe_info() == (True, False)
    announce("pipe factor fallback")
    lu = s.ilu0(up(B.values))
    handle.check()
    np.testing.assert_array_equal(down(lu), want)
