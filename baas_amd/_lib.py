"""ctypes binding of libbaas_hip.so (the C-ABI declared in include/baas_hip.h).

Device pointers come from torch tensors (``tensor.data_ptr()``); torch is only
the allocator and stream provider here.  Loading fails loudly when the HIP
library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbaas_hip.so")

OK, INVALID_ARG, ZERO_PIVOT, NONPOSITIVE_PIVOT, HIP_ERROR, TIMEOUT, NOT_CONVERGED, ALLOC, UNSUPPORTED, TOO_LARGE = range(10)
NOALLOC_MAX_VALUES = 4194304
KIND_TRSV, KIND_IC0, KIND_ILU0 = 0, 1, 2
PLAN_AUTO, PLAN_RANGE, PLAN_SCHEDULE, PLAN_LEVELS, PLAN_TILES, PLAN_BOXES, PLAN_PIPE = 0, 1, 2, 3, 4, 5, 6
EXPORTS = {
    "group_ptr": 0, "group_set": 1, "level_ptr": 2, "perm_to_orig": 3, "node_to_level": 4,
    "nthreads": 5, "merged_ptr": 6, "row_level_ptr": 7, "row_perm": 8, "row_level": 9,
    "reduced_ptr": 10, "reduced_set": 11, "unit_ptr": 12, "unit_rows": 13, "group_cost": 14,
}


class BaasError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        msg = _lib.baas_status_string(code).decode() if _lib is not None else str(code)
        super().__init__(f"{where}: {msg} (status {code})")


class Opts(C.Structure):
    _fields_ = [("ncore", C.c_int), ("plan_mode", C.c_int), ("rows_per_unit", C.c_int),
                ("block_threads", C.c_int), ("merge_threshold", C.c_int)]


_lib = None
_P = C.c_void_p
_i = C.c_int
_d = C.c_double

_SIGS = {
    "baas_create": (_i, [C.POINTER(_P), _P]),
    "baas_destroy": (_i, [_P]),
    "baas_set_stream": (_i, [_P, _P]),
    "baas_check": (_i, [_P]),
    "baas_status_string": (C.c_char_p, [_i]),
    "baas_set_spin_limit": (_i, [_P, C.c_longlong]),
    "baas_analyze": (_i, [_P, _i, _P, _P, _i, C.POINTER(Opts), C.POINTER(_P)]),
    "baas_schedule_destroy": (_i, [_P]),
    "baas_schedule_info": (_i, [_P, _P]),
    "baas_schedule_export": (_i, [_P, _i, _P, C.POINTER(C.c_longlong)]),
    "baas_schedule_times": (_i, [_P, C.POINTER(_d), C.POINTER(_d)]),
    "baas_schedule_upper_pattern": (_i, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "baas_schedule_save": (_i, [_P, _P, C.POINTER(C.c_size_t)]),
    "baas_schedule_load": (_i, [_P, _i, _P, _P, _i, _P, C.c_size_t, C.POINTER(_P)]),
    "baas_replan": (_i, [_P, _i, _i, _i]),
    "baas_replan_tiles": (_i, [_P, _i, _i, _i]),
    "baas_schedule_plan": (_i, [_P, _P]),
    "baas_schedule_pipe": (_i, [_P, _P]),
    "baas_plan_dataflow": (_i, [_P, _i]),
    "baas_autotune": (_i, [_P, _P, _P, _i, _i, _P]),
    "baas_autotune_f32": (_i, [_P, _P, _P, _i, _i, _P]),
    "baas_autotune_factor": (_i, [_P, _P, _P, _P]),
    "baas_pcg_autotune": (_i, [_P, _P]),
    "baas_sptrsv_f64": (_i, [_P, _P, _i, _P, _P, _P, _i]),
    "baas_sptrsv_f32": (_i, [_P, _P, _i, _P, _P, _P, _i]),
    "baas_sptrsv_f64_noalloc": (_i, [_P, _P, _i, _P, _P, _P, _i]),
    "baas_sptrsv_levelsync_f64": (_i, [_P, _P, _P, _P, _P]),
    "baas_transpose_values_f64": (_i, [_P, _P, _P, _P]),
    "baas_cast_f64_f32": (_i, [_P, _i, _P, _P]),
    "baas_cast_f32_f64": (_i, [_P, _i, _P, _P]),
    "baas_sptrsv_grad_vals_f64": (_i, [_P, _P, _P, _P, _P, _i]),
    "baas_csr_outer_grad_f64": (_i, [_P, _i, _P, _P, _P, _P, _P, _i]),
    "baas_ic0_f64": (_i, [_P, _P, _P, _P]),
    "baas_ilu0_f64": (_i, [_P, _P, _P, _P]),
    "baas_ic0_split_f64": (_i, [_P, _P, _P, _P, _P]),
    "baas_ic0_split_f32": (_i, [_P, _P, _P, _P, _P]),
    "baas_spmv_f64": (_i, [_P, _i, _P, _P, _P, _P, _P, _d, _d]),
    "baas_dot_f64": (_i, [_P, _i, _P, _P, _P]),
    "baas_nrm2_f64": (_i, [_P, _i, _P, _P]),
    "baas_axpby_f64": (_i, [_P, _i, _P, _d, _P, _d, _P]),
    "baas_pcg_create": (_i, [_P, _P, _P, _i, C.POINTER(_P)]),
    "baas_pcg_destroy": (_i, [_P]),
    "baas_pcg_solve": (_i, [_P, _P, _P, _d, _i, _i, C.POINTER(_i), C.POINTER(_d)]),
    "baas_pcg_times": (_i, [_P, C.POINTER(_d), C.POINTER(_d)]),
    "baas_batch_create": (_i, [_P, _i, _P, _i, _P, _P, _P, _i, C.POINTER(Opts), C.POINTER(_P)]),
    "baas_batch_destroy": (_i, [_P]),
    "baas_batch_pcg": (_i, [_P, _P, _P, _d, _i, _P]),
    "baas_batch_info": (_i, [_P, _P, _P]),
}


def declared_symbols():
    """Names of every entry point include/baas_hip.h declares."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "baas_hip.h")
    with open(hdr) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(baas_[a-z0-9_]+)\s*\(", txt)))


def load(path: str = LIB_PATH):
    """Load the HIP library; raise if it is missing (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() or python -m baas_amd.build")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    lib = load()
    st = getattr(lib, name)(*args)
    if st != OK:
        raise BaasError(st, name)
    return st
