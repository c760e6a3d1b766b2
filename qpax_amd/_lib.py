"""ctypes binding of libqpx.so (the C ABI declared in include/qpx.h).

The library is built in-tree by ``python -m qpax_amd.build``.  There is no fallback: if the
shared library is missing or fails to load, every solver entry point raises.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqpx.so")

QPX_F32, QPX_F64 = 0, 1
QPX_PLAIN, QPX_FUSED = 0, 1
MODES = {"plain": QPX_PLAIN, "fused": QPX_FUSED}

_vp, _i32, _i64, _f64, _sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); the full list of exported symbols of include/qpx.h
SIGNATURES = {
    "qpx_version": (_i32, []),
    "qpx_last_error": (ctypes.c_char_p, []),
    "qpx_workspace_size": (_sz, [_i64, _i32, _i32, _i32]),
    "qpx_solve_qp": (_i32, [_i32, _i32, _i64, _i32, _i32, _i32] + [_vp] * 6 + [_f64] + [_vp] * 6 + [_vp, _sz, _vp]),
    "qpx_solve_qp_relaxed": (_i32, [_i32, _i32, _i64, _i32, _i32, _i32] + [_vp] * 6 + [_f64, _f64] + [_vp] * 12
                             + [_vp, _sz, _vp]),
    "qpx_diff_qp": (_i32, [_i32, _i32, _i64, _i32, _i32, _i32] + [_vp] * 6 + [_vp] * 5 + [_i32] + [_vp] * 6
                    + [_vp, _sz, _vp]),
    "qpx_solve_qp_elastic": (_i32, [_i32, _i32, _i64, _i32, _i32, _i32] + [_vp] * 7 + [_f64] + [_vp] * 9
                             + [_vp, _sz, _vp]),
    "qpx_solve_qp_elastic_relaxed": (_i32, [_i32, _i32, _i64, _i32, _i32, _i32] + [_vp] * 7 + [_f64, _f64, _i32]
                                     + [_vp] * 18 + [_vp, _sz, _vp]),
    "qpx_diff_qp_elastic": (_i32, [_i32, _i32, _i64, _i32, _i32, _i32] + [_vp] * 7 + [_vp] * 8 + [_i32] + [_vp] * 6
                            + [_vp, _sz, _vp]),
    "qpx_solve_qp_special": (_i32, [_i32, _i64, _i32, _i32] + [_vp] * 4 + [_vp] * 4 + [_vp, _sz, _vp]),
    "qpx_solve_qp_relaxed_noalloc": (_i32, [_i32, _i64, _i32, _i32, _i32] + [_vp] * 6 + [_f64, _f64] + [_vp] * 12
                                     + [_vp]),
    "qpx_noalloc_scratch_bytes": (_sz, []),
    "qpx_initialize": (_i32, [_i32, _i64, _i32, _i32, _i32] + [_vp] * 6 + [_vp] * 4 + [_vp, _sz, _vp]),
    "qpx_initialize_elastic": (_i32, [_i32, _i64, _i32, _i32, _i32] + [_vp] * 7 + [_vp] * 7 + [_vp, _sz, _vp]),
    "qpx_solve_qp_trace": (_i32, [_i32, _i64, _i32, _i32, _i32] + [_vp] * 6 + [_f64, _i32] + [_vp] * 7
                           + [_vp, _sz, _vp]),
    "qpx_special_workspace_size": (_sz, [_i64, _i32, _i32]),
    "qpx_plan_create": (_i32, [_i32, _i64, _vp, _vp, _vp, _vp]),
    "qpx_plan_destroy": (None, [_vp]),
    "qpx_plan_info": (_i32, [_vp, _vp]),
    "qpx_plan_workspace_size": (_sz, [_vp]),
    "qpx_plan_offsets": (_i32, [_vp, _vp]),
    "qpx_plan_bucket": (_i32, [_vp, _i32, _vp]),
    "qpx_plan_solve": (_i32, [_vp] + [_vp] * 6 + [_f64] + [_vp] * 6 + [_vp, _sz, _vp]),
    "qpx_plan_solve_relaxed": (_i32, [_vp] + [_vp] * 6 + [_f64, _f64] + [_vp] * 12 + [_vp, _sz, _vp]),
    "qpx_plan_diff": (_i32, [_vp] + [_vp] * 11 + [_i32] + [_vp] * 6 + [_vp, _sz, _vp]),
    "qpx_plan_tune": (_i32, [_vp] + [_vp] * 6 + [_f64, _f64] + [_vp] * 12 + [_vp, _sz, _vp]),
}
QPX_EXACT_DG, QPX_ELASTIC_CONSISTENT, QPX_TRACE_DEBUG_STEP = 1, 2, 4
TRACE_ITERS, TRACE_COLS = 30, 5

_lib = None


def load():
    """Load libqpx.so once; raise with a build hint if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libqpx.so not found at {LIB_PATH}; build it with `python -m qpax_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = load().qpx_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {status}): {msg}")
