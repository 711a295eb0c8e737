"""Python front end over the C-ABI: handles, schedules and the numeric calls.

All arrays are torch tensors on the GPU (``cuda`` device on ROCm).  The
schedule/solve split mirrors the reference runtimes: ``Schedule(...)`` is
``buildset()`` (test/FusionDemo.cpp:79-123), the numeric calls are
``fused_code()``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _need(t, name, dtype, numel=None, shape=None):
    """Reject a tensor the C-ABI call would misread: wrong dtype (the kernel
    would read 4- or 8-byte elements past the buffer), wrong size or a
    strided view (the kernels assume dense row-major storage).  Raises
    ValueError before anything reaches the library; ``_on`` then checks the
    device of every argument."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: {t.numel()} elements, expected {numel}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    return t


def _on(device, **tensors):
    """Every tensor must live on the handle's GPU."""
    for name, t in tensors.items():
        if t.device.type != "cuda" or (t.device.index if t.device.index is not None else 0) != device:
            raise ValueError(f"{name}: on {t.device}, expected cuda:{device}")


def _float_dtype(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype not in (torch.float64, torch.float32):
        raise ValueError(f"{name}: expected a float64 or float32 tensor")
    return t.dtype


def _dev_i32(a, device="cuda"):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.int32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)


def _dev_f(a, dtype=torch.float64, device="cuda"):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), device=device).to(dtype)


class Handle:
    """One library handle bound to torch's current stream."""

    def __init__(self, stream=None):
        L.load()
        self._h = C.c_void_p()
        self.device = torch.cuda.current_device()
        s = stream if stream is not None else torch.cuda.current_stream()
        L.call("baas_create", C.byref(self._h), C.c_void_p(s.cuda_stream))

    @property
    def ptr(self):
        return self._h

    def check(self):
        st = L.load().baas_check(self._h)
        if st != L.OK:
            raise L.BaasError(st, "kernel")

    def status(self) -> int:
        return L.load().baas_check(self._h)

    def set_spin_limit(self, polls: int):
        L.call("baas_set_spin_limit", self._h, int(polls))

    def __del__(self):
        try:
            if self._h:
                L.load().baas_destroy(self._h)
        except Exception:
            pass


class Schedule:
    """BAAS analysis of an n x n pattern (lower L for TRSV, full A for IC0/ILU0)."""

    def __init__(self, h: Handle, rowptr, colidx, kind=L.KIND_TRSV, ncore=8, plan_mode=L.PLAN_AUTO,
                 rows_per_unit=0, block_threads=0, saved: bytes = None):
        """saved: bytes from save() for the same pattern and kind -- the
        schedule is rebuilt from them (analysis and plan included) instead
        of analysed."""
        self.h = h
        self.rowptr = _dev_i32(rowptr)
        self.colidx = _dev_i32(colidx)
        self.n = self.rowptr.numel() - 1
        self.kind = kind
        self._s = C.c_void_p()
        if saved is not None:
            blob = (C.c_ubyte * len(saved)).from_buffer_copy(saved)
            L.call("baas_schedule_load", h.ptr, self.n, _ptr(self.rowptr), _ptr(self.colidx), kind,
                   C.cast(blob, C.c_void_p), len(saved), C.byref(self._s))
        else:
            o = L.Opts(ncore, plan_mode, rows_per_unit, block_threads, 0)
            L.call("baas_analyze", h.ptr, self.n, _ptr(self.rowptr), _ptr(self.colidx), kind, C.byref(o),
                   C.byref(self._s))
        self.nnz = self.colidx.numel()
        self.nnz_lower = self.info()["nnz_lower"]

    def save(self) -> bytes:
        """The analysis and the plan in use (an autotuned one included) as
        bytes; Schedule(..., saved=those bytes) rebuilds this schedule."""
        n = C.c_size_t()
        L.call("baas_schedule_save", self._s, None, C.byref(n))
        buf = (C.c_ubyte * n.value)()
        L.call("baas_schedule_save", self._s, C.cast(buf, C.c_void_p), C.byref(n))
        return bytes(buf)

    @property
    def ptr(self):
        return self._s

    def info(self):
        c = (C.c_int * 8)()
        L.call("baas_schedule_info", self._s, C.cast(c, C.c_void_p))
        keys = ("ngroups", "nlevels", "nmerged", "reduced_edges", "nrowlevels", "nunits", "nnz_lower",
                "plan_mode")
        return dict(zip(keys, list(c)))

    def export(self, which) -> np.ndarray:
        code = L.EXPORTS[which] if isinstance(which, str) else which
        n = C.c_longlong()
        L.call("baas_schedule_export", self._s, code, None, C.byref(n))
        out = np.zeros(n.value, np.int32)
        L.call("baas_schedule_export", self._s, code, C.c_void_p(out.ctypes.data), C.byref(n))
        return out

    def unit_order(self) -> np.ndarray:
        """The forward execution plan's row order (work units in claim
        order, each unit's rows by sub-level): a topological order in which
        renumbering the matrix (matgen.symmetric_permute) makes every unit's
        rows consecutive -- the GPU counterpart of the reference's
        thread-contiguous renumbering (test/SpTRSVRuntime.hpp:751-772)."""
        return self.export("unit_rows")

    def schedule_dict(self):
        i = self.info()
        d = {k: self.export(k) for k in ("group_ptr", "group_set", "level_ptr", "perm_to_orig",
                                         "node_to_level", "nthreads", "merged_ptr")}
        d.update(ngroups=i["ngroups"], nlevels=i["nlevels"], nmerged=i["nmerged"],
                 reduced_edges=i["reduced_edges"])
        return d

    def times(self):
        g, h = C.c_double(), C.c_double()
        L.call("baas_schedule_times", self._s, C.byref(g), C.byref(h))
        return g.value, h.value

    def replan(self, mode, rows_per_unit=0, block_threads=0, dataflow=None):
        L.call("baas_replan", self._s, int(mode), int(rows_per_unit), int(block_threads))
        if dataflow is not None:
            L.call("baas_plan_dataflow", self._s, int(dataflow))

    def replan_tiles(self, row_range, band_levels, block_threads=0, dataflow=None):
        """TILES plan: units of row_range consecutive rows by band_levels consecutive levels."""
        L.call("baas_replan_tiles", self._s, int(row_range), int(band_levels), int(block_threads))
        if dataflow is not None:
            L.call("baas_plan_dataflow", self._s, int(dataflow))

    def plan(self):
        """The plan in use: mode, rows per unit, workgroup size, dataflow, band levels."""
        c = (C.c_int * 5)()
        L.call("baas_schedule_plan", self._s, C.cast(c, C.c_void_p))
        return {"mode": c[0], "rows_per_unit": c[1], "block": c[2], "dataflow": c[3], "band": c[4]}

    def pipe_info(self):
        """(the forward plan is a PIPE plan, IC0/ILU0 run pipelined on it)."""
        c = (C.c_int * 2)()
        L.call("baas_schedule_pipe", self._s, C.cast(c, C.c_void_p))
        return bool(c[0]), bool(c[1])

    def apply_plan(self, plan):
        """Re-pack into a plan returned by plan() or autotune() (for example one
        saved from an earlier run, so a new process skips the search)."""
        if plan["mode"] == L.PLAN_TILES:
            self.replan_tiles(plan["rows_per_unit"], plan.get("band") or 16, plan["block"], plan["dataflow"])
        else:
            self.replan(plan["mode"], plan["rows_per_unit"], plan["block"], plan["dataflow"])

    def autotune(self, vals, uplo=0, nrhs=1):
        """Time candidate plans on these values and keep the fastest."""
        c = (C.c_int * 5)()
        _need(vals, "vals", _float_dtype(vals, "vals"), numel=self.nnz_lower)
        _on(self.h.device, vals=vals)
        f = "baas_autotune_f32" if vals.dtype == torch.float32 else "baas_autotune"
        L.call(f, self.h.ptr, self._s, _ptr(vals), int(uplo), int(nrhs), C.cast(c, C.c_void_p))
        return {"mode": c[0], "rows_per_unit": c[1], "block": c[2], "dataflow": c[3], "us": c[4],
                "band": self.plan()["band"]}

    def autotune_factor(self, a_vals):
        """Time the IC0/ILU0 factorisation over plans and keep the fastest."""
        c = (C.c_int * 5)()
        _need(a_vals, "a_vals", torch.float64, numel=self.nnz)
        _on(self.h.device, a_vals=a_vals)
        L.call("baas_autotune_factor", self.h.ptr, self._s, _ptr(a_vals), C.cast(c, C.c_void_p))
        return {"mode": c[0], "rows_per_unit": c[1], "block": c[2], "dataflow": c[3], "us": c[4],
                "band": self.plan()["band"]}

    def upper_pattern(self):
        rp, ci = C.c_void_p(), C.c_void_p()
        L.call("baas_schedule_upper_pattern", self._s, C.byref(rp), C.byref(ci))
        return rp, ci

    # ---- numerics
    def solve(self, vals, b, x=None, uplo=0):
        """x = L^-1 b (uplo 0, vals in L's layout) or U^-1 b (uplo 1, vals in U = L^T layout)."""
        dt = _float_dtype(b, "b")
        if b.dim() not in (1, 2) or b.shape[0] != self.n or (b.dim() == 2 and b.shape[1] < 1):
            raise ValueError(f"b: shape {tuple(b.shape)}, expected ({self.n},) or ({self.n}, nrhs)")
        _need(b, "b", dt)
        _need(vals, "vals", dt, numel=self.nnz_lower)
        nrhs = 1 if b.dim() == 1 else b.shape[1]
        if x is None:
            x = torch.empty_like(b)
        _need(x, "x", dt, shape=b.shape)
        if x.data_ptr() == b.data_ptr():
            raise ValueError("x must not alias b")
        _on(self.h.device, b=b, vals=vals, x=x)
        f = "baas_sptrsv_f64" if b.dtype == torch.float64 else "baas_sptrsv_f32"
        L.call(f, self.h.ptr, self._s, int(uplo), _ptr(vals), _ptr(b), _ptr(x), int(nrhs))
        return x

    def solve_levelsync(self, vals, b, x=None):
        _need(b, "b", torch.float64, shape=(self.n,))
        _need(vals, "vals", torch.float64, numel=self.nnz_lower)
        if x is None:
            x = torch.zeros_like(b)
        _need(x, "x", torch.float64, shape=(self.n,))
        _on(self.h.device, b=b, vals=vals, x=x)
        L.call("baas_sptrsv_levelsync_f64", self.h.ptr, self._s, _ptr(vals), _ptr(b), _ptr(x))
        return x

    def transpose_values(self, lvals):
        _need(lvals, "lvals", torch.float64, numel=self.nnz_lower)
        _on(self.h.device, lvals=lvals)
        uv = torch.empty_like(lvals)
        L.call("baas_transpose_values_f64", self.h.ptr, self._s, _ptr(lvals), _ptr(uv))
        return uv

    def grad_vals(self, lam, x):
        if x.dim() not in (1, 2) or x.shape[0] != self.n:
            raise ValueError(f"x: shape {tuple(x.shape)}, expected ({self.n},) or ({self.n}, nrhs)")
        _need(x, "x", torch.float64)
        _need(lam, "lam", torch.float64, shape=x.shape)
        _on(self.h.device, lam=lam, x=x)
        nrhs = 1 if x.dim() == 1 else x.shape[1]
        dv = torch.empty(self.nnz_lower, dtype=torch.float64, device=x.device)
        L.call("baas_sptrsv_grad_vals_f64", self.h.ptr, self._s, _ptr(lam), _ptr(x), _ptr(dv), int(nrhs))
        return dv

    def _factor_args(self, a_vals, lu, want):
        if self.kind != want:
            raise ValueError(f"schedule analysed for kind {self.kind}, this call needs kind {want}")
        _need(a_vals, "a_vals", torch.float64, numel=self.nnz)
        if lu is None:
            lu = torch.empty_like(a_vals)
        _need(lu, "lu", torch.float64, numel=self.nnz)
        _on(self.h.device, a_vals=a_vals, lu=lu)
        return lu

    def ic0(self, a_vals, lu=None):
        lu = self._factor_args(a_vals, lu, L.KIND_IC0)
        L.call("baas_ic0_f64", self.h.ptr, self._s, _ptr(a_vals), _ptr(lu))
        return lu

    def ilu0(self, a_vals, lu=None):
        lu = self._factor_args(a_vals, lu, L.KIND_ILU0)
        L.call("baas_ilu0_f64", self.h.ptr, self._s, _ptr(a_vals), _ptr(lu))
        return lu

    def ic0_split(self, lu, dtype=torch.float64):
        if dtype not in (torch.float64, torch.float32):
            raise ValueError(f"dtype {dtype}: expected float64 or float32")
        _need(lu, "lu", torch.float64, numel=self.nnz)
        _on(self.h.device, lu=lu)
        lv = torch.empty(self.nnz_lower, dtype=dtype, device=lu.device)
        uv = torch.empty(self.nnz_lower, dtype=dtype, device=lu.device)
        f = "baas_ic0_split_f64" if dtype == torch.float64 else "baas_ic0_split_f32"
        L.call(f, self.h.ptr, self._s, _ptr(lu), _ptr(lv), _ptr(uv))
        return lv, uv

    def __del__(self):
        try:
            if self._s:
                L.load().baas_schedule_destroy(self._s)
        except Exception:
            pass


def spmv(h: Handle, rowptr, colidx, vals, x, y=None, alpha=1.0, beta=0.0, ncols=None):
    """y = alpha (A x) + beta y for an m x ncols CSR A (ncols defaults to m)."""
    m = rowptr.numel() - 1
    _need(rowptr, "rowptr", torch.int32, numel=m + 1)
    _need(colidx, "colidx", torch.int32)
    _need(vals, "vals", torch.float64, numel=colidx.numel())
    _need(x, "x", torch.float64, numel=m if ncols is None else int(ncols))
    if y is None:
        y = torch.zeros(m, dtype=torch.float64, device=x.device)
    _need(y, "y", torch.float64, numel=m)
    _on(h.device, rowptr=rowptr, colidx=colidx, vals=vals, x=x, y=y)
    L.call("baas_spmv_f64", h.ptr, m, _ptr(rowptr), _ptr(colidx), _ptr(vals), _ptr(x), _ptr(y),
           float(alpha), float(beta))
    return y


def _vec_pair(x, y):
    _need(x, "x", torch.float64)
    _need(y, "y", torch.float64, numel=x.numel())


def dot(h: Handle, x, y):
    _vec_pair(x, y)
    _on(h.device, x=x, y=y)
    r = torch.zeros(1, dtype=torch.float64, device=x.device)
    L.call("baas_dot_f64", h.ptr, x.numel(), _ptr(x), _ptr(y), _ptr(r))
    return r


def nrm2(h: Handle, x):
    _need(x, "x", torch.float64)
    _on(h.device, x=x)
    r = torch.zeros(1, dtype=torch.float64, device=x.device)
    L.call("baas_nrm2_f64", h.ptr, x.numel(), _ptr(x), _ptr(r))
    return r


def axpby(h: Handle, w, alpha, x, beta, y):
    _vec_pair(x, y)
    _need(w, "w", torch.float64, numel=x.numel())
    _on(h.device, w=w, x=x, y=y)
    L.call("baas_axpby_f64", h.ptr, w.numel(), _ptr(w), float(alpha), _ptr(x), float(beta), _ptr(y))
    return w


class PCG:
    """IC0-preconditioned CG on the GPU (pcg_opt, test/iccg_level_merge.cpp:380-533)."""

    def __init__(self, h: Handle, sched: Schedule, a_vals, precision=0):
        if sched.kind != L.KIND_IC0:
            raise ValueError("PCG needs a schedule analysed with kind IC0")
        _need(a_vals, "a_vals", torch.float64, numel=sched.nnz)
        _on(h.device, a_vals=a_vals)
        self.h, self.s, self.a_vals = h, sched, a_vals
        self._p = C.c_void_p()
        L.call("baas_pcg_create", h.ptr, sched.ptr, _ptr(a_vals), int(precision), C.byref(self._p))

    def solve(self, b, x=None, tol=1e-8, maxiter=1000, refactor=False):
        _need(b, "b", torch.float64, shape=(self.s.n,))
        if x is None:
            x = torch.zeros_like(b)
        _need(x, "x", torch.float64, shape=(self.s.n,))
        if x.data_ptr() == b.data_ptr():
            raise ValueError("x must not alias b")
        _on(self.h.device, b=b, x=x)
        it, rr = C.c_int(), C.c_double()
        L.call("baas_pcg_solve", self._p, _ptr(b), _ptr(x), float(tol), int(maxiter), int(bool(refactor)),
               C.byref(it), C.byref(rr))
        return x, it.value, rr.value

    def times(self):
        a, b = C.c_double(), C.c_double()
        L.call("baas_pcg_times", self._p, C.byref(a), C.byref(b))
        return a.value, b.value

    def autotune(self):
        """Tune the schedule's plans for this solver's triangular sweeps."""
        c = (C.c_int * 5)()
        L.call("baas_pcg_autotune", self._p, C.cast(c, C.c_void_p))
        return {"mode": c[0], "rows_per_unit": c[1], "block": c[2], "dataflow": c[3], "us": c[4]}

    def __del__(self):
        try:
            if self._p:
                L.load().baas_pcg_destroy(self._p)
        except Exception:
            pass


class Batch:
    """Many SPD systems x many right-hand sides: IC0 + PCG (FP32 factor, FP64 CG)."""

    def __init__(self, h: Handle, mats, nrhs, block_threads=0):
        import numpy as _np
        self.h = h
        self.sizes = [m.n for m in mats]
        self.sys_ptr = _np.zeros(len(mats) + 1, _np.int32)
        _np.cumsum(self.sizes, out=self.sys_ptr[1:])
        N = int(self.sys_ptr[-1])
        rp, ci, va = [_np.zeros(1, _np.int64)], [], []
        off, nz = 0, 0
        for m in mats:
            rp.append(m.rowptr[1:].astype(_np.int64) + nz)
            ci.append(m.colidx.astype(_np.int64) + off)
            va.append(m.values)
            off += m.n
            nz += m.nnz
        self.rowptr = _dev_i32(_np.concatenate(rp))
        self.colidx = _dev_i32(_np.concatenate(ci))
        self.vals = _dev_f(_np.concatenate(va))
        self.N, self.k = N, int(nrhs)
        o = L.Opts(8, L.PLAN_RANGE, 0, block_threads, 0)
        self._b = C.c_void_p()
        L.call("baas_batch_create", h.ptr, len(mats), C.c_void_p(self.sys_ptr.ctypes.data), N, _ptr(self.rowptr),
               _ptr(self.colidx), _ptr(self.vals), self.k, C.byref(o), C.byref(self._b))

    def solve(self, B, X=None, tol=1e-8, maxiter=1000):
        """B, X: (N, nrhs) float64 device tensors.  Returns X and per-(system, rhs) iterations."""
        import numpy as _np
        _need(B, "B", torch.float64, shape=(self.N, self.k))
        if X is None:
            X = torch.zeros_like(B)
        _need(X, "X", torch.float64, shape=(self.N, self.k))
        if X.data_ptr() == B.data_ptr():
            raise ValueError("X must not alias B")
        _on(self.h.device, B=B, X=X)
        it = _np.zeros(len(self.sizes) * self.k, _np.int32)
        st = L.load().baas_batch_pcg(self._b, _ptr(B), _ptr(X), float(tol), int(maxiter),
                                      C.c_void_p(it.ctypes.data))
        if st not in (L.OK, L.NOT_CONVERGED):
            raise L.BaasError(st, "baas_batch_pcg")
        return X, it.reshape(len(self.sizes), self.k)

    def info(self):
        t = (C.c_double * 3)()
        r = (C.c_int * 3)()
        L.call("baas_batch_info", self._b, C.cast(t, C.c_void_p), C.cast(r, C.c_void_p))
        return {"ic0_setup_s": t[0], "cg_s": t[1], "tune_s": t[2], "rows_per_unit": list(r)}

    def __del__(self):
        try:
            if self._b:
                L.load().baas_batch_destroy(self._b)
        except Exception:
            pass
