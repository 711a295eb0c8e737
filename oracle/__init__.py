"""ctypes wrappers around the CPU checker libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package.

* ``lib``  -- ``oracle/liboracle.so``, the C restatement (oracle.c).
* ``ref``  -- ``oracle/_ref/libbaas_ref.so``, the reference's own sources
  compiled in place (None where it was not built).  It is loaded on first
  use only (``get_ref()`` or the attribute ``ref``), so the GPU-side tests,
  which never ask for it, do not load it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i = C.c_int
_d = C.c_double

KIND_TRSV, KIND_IC0, KIND_ILU0 = 0, 1, 2


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and oracle/_ref where the reference tree exists)."""
    out = subprocess.run(["make", "-C", HERE, "-j8"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout[-4000:] + out.stderr[-4000:])


def _load(path):
    return C.CDLL(path) if os.path.exists(path) else None


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_lib_path = os.path.join(HERE, "liboracle.so")
if not os.path.exists(_lib_path):
    build()
lib = _load(_lib_path)
_ref_lib = None
_ref_tried = False

_sig(lib, "orc_two_hop_reduce", _i, [_i, _I, _I, _I, _I])
_sig(lib, "orc_levelset_csc", _i, [_i, _I, _I, _I, _I, _I])
_sig(lib, "orc_csr_to_csc", None, [_i, _i, _I, _I, C.c_void_p, _I, _I, C.c_void_p])
_sig(lib, "orc_reverse_coarsen", _i, [_i, _I, _I, _I, _I])
_sig(lib, "orc_forward_coarsen", _i, [_i, _I, _I, _I, _I])
_sig(lib, "orc_baas_schedule", None, [_i, _i, _I, _I, _I, _i, _I, _I, _I, _I, _I, _I, _I, _I])
_sig(lib, "orc_group_cost", None, [_i, _I, _I, _I, _I])
_sig(lib, "orc_sptrsv_lower", None, [_i, _I, _I, _D, _D, _D])
_sig(lib, "orc_sptrsv_upper", None, [_i, _I, _I, _D, _D, _D])
_sig(lib, "orc_sptrsv_upper_desc", None, [_i, _I, _I, _D, _D, _D])
_sig(lib, "orc_sptrsv_lower_f32", None, [_i, _I, _I, _F, _F, _F])
_sig(lib, "orc_sptrsv_upper_desc_f32", None, [_i, _I, _I, _F, _F, _F])
_sig(lib, "orc_sptrsv_lower_levels_omp", None, [_i, _I, _I, _D, _D, _D, _i, _I, _I, _i])
_sig(lib, "orc_sptrsv_lower_multi", None, [_i, _I, _I, _D, _i, _D, _D])
_sig(lib, "orc_sptrsv_lower_p2p_omp", None, [_i, _I, _I, _D, _D, _D, _i, _I, _I, _I, _I, _i, _I, _i])
_sig(lib, "omp_max_threads", _i, [])
_sig(lib, "orc_ic0", None, [_i, _I, _I, _I, _D, _D])
_sig(lib, "orc_ilu0", None, [_i, _I, _I, _I, _D])
_sig(lib, "orc_ic0_split", None, [_i, _I, _I, _I, _D, _I, _I, _D, _I, _I, _D])
_sig(lib, "orc_spmv", None, [_i, _I, _I, _D, _D, _D])
_sig(lib, "orc_dcsrmv", None, [_i, _I, _I, _D, _D, _D, _d, _d])
_sig(lib, "orc_dot", _d, [_i, _D, _D])
_sig(lib, "orc_waxpby", None, [_i, _D, _d, _D, _d, _D])
_sig(lib, "orc_pcg", _i, [_i, _I, _I, _I, _D, _D, _D, _i, _d, _i, C.c_void_p, _i])


def get_ref():
    """The compiled reference (oracle/_ref/libbaas_ref.so), or None where it was not built."""
    global _ref_lib, _ref_tried
    if _ref_tried:
        return _ref_lib
    _ref_tried = True
    r = _load(os.path.join(HERE, "_ref", "libbaas_ref.so"))
    if r is not None:
        _sig(r, "ref_two_hop", _i, [_i, _I, _I, _I, _I])
        _sig(r, "ref_levelset_rows", _i, [_i, _I, _I, _I, _I, _I])
        _sig(r, "ref_baas_schedule", None, [_i, _i, _I, _I, _I, _i, _I, _I, _I, _I, _I, _I, _I, _I])
        _sig(r, "ref_ic0", None, [_i, _I, _I, _D, _D])
        _sig(r, "ref_ilu0", None, [_i, _I, _I, _D, _D])
        _sig(r, "ref_spmv", None, [_i, _I, _I, _D, _D, _D])
        _sig(r, "ref_dcsrmv", None, [_i, _I, _I, _D, _D, _D, _d, _d])
        _sig(r, "ref_dot", _d, [_i, _D, _D])
        _sig(r, "ref_waxpby", None, [_i, _D, _d, _D, _d, _D])
        _sig(r, "ref_group_cost", None, [_i, _I, _I, _I, _I, _D])
        _sig(r, "ref_store_bin", None, [_i, _i, _I, _I, _D, C.c_char_p])
    _ref_lib = r
    return r


def __getattr__(name):
    if name == "ref":
        return get_ref()
    raise AttributeError(name)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# structure
def csc_of(n, rowptr, colidx):
    cp = np.zeros(n + 1, np.int32)
    ri = np.zeros(max(int(rowptr[-1]), 1), np.int32)
    lib.orc_csr_to_csc(n, n, _i32(rowptr), _i32(colidx), None, cp, ri, None)
    return cp, ri[: int(rowptr[-1])]


def two_hop(n, ptr, st, use_ref=False):
    op = np.zeros(n + 1, np.int32)
    os_ = np.zeros(max(int(ptr[-1]), 1), np.int32)
    f = get_ref().ref_two_hop if use_ref else lib.orc_two_hop_reduce
    e = f(n, _i32(ptr), _i32(st), op, os_)
    return op, os_[:e]


def levelset_csc(n, ptr, st):
    lp = np.zeros(n + 1, np.int32)
    ntl = np.zeros(n, np.int32)
    pto = np.zeros(n, np.int32)
    nl = lib.orc_levelset_csc(n, _i32(ptr), _i32(st), lp, ntl, pto)
    return lp[: nl + 1], ntl, pto


def levelset_rows(n, lrowptr, lcolidx, use_ref=False):
    if use_ref:
        lp = np.zeros(n + 1, np.int32)
        ntl = np.zeros(n, np.int32)
        pto = np.zeros(n, np.int32)
        nl = get_ref().ref_levelset_rows(n, _i32(lrowptr), _i32(lcolidx), lp, ntl, pto)
        return lp[: nl + 1], ntl, pto
    cp, ri = csc_of(n, lrowptr, lcolidx)
    return levelset_csc(n, cp, ri)


def baas_schedule(kind, n, lrowptr, lcolidx, thr_rowptr=None, ncore=8, use_ref=False):
    """The BAAS buildset() analysis; returns a dict of int32 arrays."""
    if thr_rowptr is None:
        thr_rowptr = lrowptr
    gp = np.zeros(n + 1, np.int32)
    gs = np.zeros(n, np.int32)
    lp = np.zeros(n + 1, np.int32)
    pto = np.zeros(n, np.int32)
    ntl = np.zeros(n, np.int32)
    nt = np.zeros(n, np.int32)
    mp = np.zeros(n + 1, np.int32)
    cnt = np.zeros(4, np.int32)
    f = get_ref().ref_baas_schedule if use_ref else lib.orc_baas_schedule
    f(kind, n, _i32(lrowptr), _i32(lcolidx), _i32(thr_rowptr), ncore, gp, gs, lp, pto, ntl, nt, mp, cnt)
    ng, nl, nm, red = (int(v) for v in cnt)
    return dict(ngroups=ng, nlevels=nl, nmerged=nm, reduced_edges=red,
                group_ptr=gp[: ng + 1], group_set=gs, level_ptr=lp[: nl + 1],
                perm_to_orig=pto[:ng], node_to_level=ntl[:ng], nthreads=nt[:nl],
                merged_ptr=mp[: nm + 1])


def group_cost(group_ptr, group_set, lrowptr, lcolidx=None, use_ref=False):
    """costComputation (merge/Coarsen.cpp:854-891): nonzeros of each group's
    rows in the lower CSR (int32; the reference returns doubles, its
    restatement ints of the same values)."""
    ng = len(group_ptr) - 1
    if use_ref:
        out = np.zeros(ng)
        get_ref().ref_group_cost(ng, _i32(group_ptr), _i32(group_set), _i32(lrowptr), _i32(lcolidx), out)
        return out
    out = np.zeros(ng, np.int32)
    lib.orc_group_cost(ng, _i32(group_ptr), _i32(group_set), _i32(lrowptr), out)
    return out


# --------------------------------------------------------------------------
# numerics
def sptrsv_lower(L, b):
    x = np.zeros(L.n)
    lib.orc_sptrsv_lower(L.n, _i32(L.rowptr), _i32(L.colidx), _f64(L.values), _f64(b), x)
    return x


def sptrsv_upper(U, b, descending=True):
    x = np.zeros(U.n)
    f = lib.orc_sptrsv_upper_desc if descending else lib.orc_sptrsv_upper
    f(U.n, _i32(U.rowptr), _i32(U.colidx), _f64(U.values), _f64(b), x)
    return x


def sptrsv_lower_f32(L, b):
    x = np.zeros(L.n, np.float32)
    lib.orc_sptrsv_lower_f32(L.n, _i32(L.rowptr), _i32(L.colidx),
                             np.ascontiguousarray(L.values, np.float32),
                             np.ascontiguousarray(b, np.float32), x)
    return x


def sptrsv_upper_f32(U, b):
    x = np.zeros(U.n, np.float32)
    lib.orc_sptrsv_upper_desc_f32(U.n, _i32(U.rowptr), _i32(U.colidx),
                                  np.ascontiguousarray(U.values, np.float32),
                                  np.ascontiguousarray(b, np.float32), x)
    return x


class P2PSolver:
    """The point-to-point OpenMP lower solve with its ready flags kept
    between calls (one epoch per call, so nothing is reset)."""

    def __init__(self, L, nthreads):
        self.L = L
        self.lp, ntl, self.pto = levelset_rows(L.n, L.rowptr, L.colidx)
        self.lvl = _i32(ntl - 1)
        self.nthreads = max(1, int(nthreads))
        # the thread whose chunk of its level holds each row
        cnt = np.diff(self.lp)
        chunk = (cnt + self.nthreads - 1) // self.nthreads
        q = np.arange(L.n)
        lq = np.repeat(np.arange(len(cnt)), cnt)
        self.own = np.empty(L.n, np.int32)
        self.own[self.pto] = ((q - self.lp[:-1][lq]) // chunk[lq]).astype(np.int32)
        self.flags = np.zeros((len(self.lp) - 1) * self.nthreads, np.int32)
        self.epoch = 0
        self.rp, self.ci, self.va = _i32(L.rowptr), _i32(L.colidx), _f64(L.values)

    def solve(self, b, x=None):
        if x is None:
            x = np.zeros(self.L.n)
        self.epoch += 1
        lib.orc_sptrsv_lower_p2p_omp(self.L.n, self.rp, self.ci, self.va, _f64(b), x, len(self.lp) - 1, self.lp,
                                     self.pto, self.lvl, self.own, self.nthreads, self.flags, self.epoch)
        return x


def omp_max_threads():
    return lib.omp_max_threads()


def sptrsv_lower_multi(L, B):
    B = _f64(B)
    k = B.shape[1]
    X = np.zeros_like(B)
    lib.orc_sptrsv_lower_multi(L.n, _i32(L.rowptr), _i32(L.colidx), _f64(L.values), k, B, X)
    return X


def sptrsv_lower_levels_omp(L, b, level_ptr, perm_to_orig, nthreads=0):
    x = np.zeros(L.n)
    lib.orc_sptrsv_lower_levels_omp(L.n, _i32(L.rowptr), _i32(L.colidx), _f64(L.values), _f64(b), x,
                                    len(level_ptr) - 1, _i32(level_ptr), _i32(perm_to_orig), nthreads)
    return x


def ic0(A, use_ref=False):
    lu = A.values.copy()
    if use_ref:
        get_ref().ref_ic0(A.n, _i32(A.rowptr), _i32(A.colidx), _f64(A.values), lu)
    else:
        lib.orc_ic0(A.n, _i32(A.rowptr), _i32(A.colidx), _i32(A.diagptr()), _f64(A.values), lu)
    return lu


def ilu0(A, use_ref=False):
    lu = A.values.copy()
    if use_ref:
        get_ref().ref_ilu0(A.n, _i32(A.rowptr), _i32(A.colidx), _f64(A.values), lu)
    else:
        lib.orc_ilu0(A.n, _i32(A.rowptr), _i32(A.colidx), _i32(A.diagptr()), lu)
    return lu


def ic0_split(A, lu):
    from baas_amd.matgen import Csr
    n = A.n
    dp = _i32(A.diagptr())
    nl = int(np.sum(dp - A.rowptr[:-1] + 1))
    Lp = np.zeros(n + 1, np.int32); Li = np.zeros(nl, np.int32); Lx = np.zeros(nl)
    Up = np.zeros(n + 1, np.int32); Ui = np.zeros(nl, np.int32); Ux = np.zeros(nl)
    lib.orc_ic0_split(n, _i32(A.rowptr), _i32(A.colidx), dp, _f64(lu), Lp, Li, Lx, Up, Ui, Ux)
    return Csr(n, Lp, Li, Lx), Csr(n, Up, Ui, Ux)


def spmv(A, x, use_ref=False):
    y = np.zeros(A.n)
    f = get_ref().ref_spmv if use_ref else lib.orc_spmv
    f(A.n, _i32(A.rowptr), _i32(A.colidx), _f64(A.values), _f64(x), y)
    return y


def dcsrmv(A, x, y, alpha, beta, use_ref=False):
    """alpha * (A x) + beta * y (y is not modified; the result is returned)."""
    out = _f64(y).copy()
    f = get_ref().ref_dcsrmv if use_ref else lib.orc_dcsrmv
    f(A.n, _i32(A.rowptr), _i32(A.colidx), _f64(A.values), _f64(x), out, float(alpha), float(beta))
    return out


def dot(x, y):
    return lib.orc_dot(len(x), _f64(x), _f64(y))


def waxpby(alpha, x, beta, y, use_ref=False):
    w = np.zeros(len(x))
    f = get_ref().ref_waxpby if use_ref else lib.orc_waxpby
    f(len(x), w, alpha, _f64(x), beta, _f64(y))
    return w


def pcg(A, b, maxiter=1000, tol=1e-8, precision=0, stop_at_tol=True):
    """Returns (x, first_iteration_below_tol, relres history)."""
    x = np.zeros(A.n)
    hist = np.zeros(maxiter + 1)
    k = lib.orc_pcg(A.n, _i32(A.rowptr), _i32(A.colidx), _i32(A.diagptr()), _f64(A.values), _f64(b), x,
                    maxiter, tol, precision, hist.ctypes.data, 1 if stop_at_tol else 0)
    last = k if (k > 0 and stop_at_tol) else maxiter
    return x, k, hist[: last + 1]
