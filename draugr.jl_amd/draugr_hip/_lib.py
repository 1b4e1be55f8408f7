"""Loader of the HIP backend library (``lib/libdraugr_hip.so``) and the ctypes
signatures of its C-ABI (``include/draugr_hip.h``).

There is no fallback: if the library is missing the import fails loudly.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DRAUGR_HIP_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libdraugr_hip.so")

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)

SIGNATURES = {
    "init_julia": (None, [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]),
    "shutdown_julia": (None, [ctypes.c_int]),
    "draugr_amg_last_error": (ctypes.c_char_p, []),
    "draugr_amg_config_from_json": (_i32, [ctypes.c_char_p]),
    "draugr_amg_setup": (_i32, [_i32, _i32, _vp, _vp, _vp, _i32, _i32, _i32]),
    "draugr_amg_resetup": (_i32, [_i32, _i32, _i32, _vp, _vp, _vp, _i32, _i32, _i32, _i32]),
    "draugr_amg_solve": (_i32, [_i32, _i32, _vp, _vp, _i32, _f64, _i32]),
    "draugr_amg_cycle": (_i32, [_i32, _i32, _vp, _vp, _i32]),
    "draugr_amg_free": (_i32, [_i32]),
    "draugr_amg_config_free": (_i32, [_i32]),
    "draugr_hip_device_count": (_i32, []),
    "draugr_hip_setup_host": (_i32, [_i32, _i32, _vp, _vp, _vp, _i32, _i32, _i32]),
    "draugr_hip_upload": (_i32, [_i32]),
    "draugr_hip_hierarchy_info": (_i32, [_i32, _i64p, _i32]),
    "draugr_hip_level_info": (_i32, [_i32, _i32, _i64p, _i32]),
    "draugr_hip_level_export": (_i32, [_i32, _i32, _i32, _vp]),
    "draugr_hip_summary": (_i32, [_i32, ctypes.c_char_p, _i32]),
    "draugr_hip_cycle_device": (_i32, [_i32, _i32, _vp, _vp, _i32, _i32, _vp]),
    "draugr_hip_solve_device": (_i32, [_i32, _i32, _vp, _vp, _i32, _f64, _i32, _vp, _f64p]),
    "draugr_hip_pcg_device": (_i32, [_i32, _i32, _vp, _vp, _i32, _f64, _i32, _vp, _f64p]),
    "draugr_hip_pcg": (_i32, [_i32, _i32, _vp, _vp, _i32, _f64, _i32, _f64p]),
    "draugr_hip_pcg_device_noalloc": (_i32, [_i32, _i32, _vp, _vp, _i32, _f64, _i32, _vp, _f64p]),
    "draugr_hip_resetup_device": (_i32, [_i32, _i64, _vp, _vp]),
    "draugr_hip_level_op": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp]),
    "draugr_hip_level_op_repeat": (_i32, [_i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _i32, _vp]),
    "draugr_hip_precondition_repeat": (_i32, [_i32, _i32, _vp, _i32, _vp]),
    "draugr_hip_export_device_values": (_i32, [_i32, _i32, _i32, _vp]),
    "draugr_hip_cycle_bytes": (_i32, [_i32, _i32, _f64p]),
    "draugr_hip_op_bytes": (_i32, [_i32, _i32, _i32, _i32, _i32, _f64p]),
    "draugr_hip_resident_bytes": (_i32, [_i32, _i64p]),
    "draugr_hip_tune_report": (_i32, [_i32, ctypes.c_char_p, _i32]),
    "draugr_hip_assemble_7pt": (_i32, [_i32, _i32, _vp, _vp, _vp]),
    "draugr_hip_matrix_grad": (_i32, [_i32, _vp, _vp, _vp, _vp]),
    "draugr_hip_grad_logk_7pt": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp]),
    "draugr_hip_batch_setup": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32]),
    "draugr_hip_batch_info": (_i32, [_i32, _i64p, _i32]),
    "draugr_hip_batch_system_info": (_i32, [_i32, _i32, _i64p, _i32]),
    "draugr_hip_batch_cycle_device": (_i32, [_i32, _vp, _vp, _i32, _vp]),
    "draugr_hip_batch_pcg_device": (_i32, [_i32, _vp, _vp, _i32, _f64, _i32, _vp, _vp, _vp]),
    "draugr_hip_batch_free": (_i32, [_i32]),
    "draugr_hip_batch_tune_report": (_i32, [_i32, ctypes.c_char_p, _i32]),
    "draugr_hip_batch_level_info": (_i32, [_i32, _i32, _i64p, _i32]),
    "draugr_hip_batch_jacobi_pass": (_i32, [_i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "draugr_hip_ilu0_factor": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


class DraugrError(RuntimeError):
    """A C-ABI call returned -1; the message is draugr_amg_last_error()."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libdraugr_hip.so not found at %s: build it first (python -c 'import __graft_entry__ as g; "
                "g.build()' or make -C draugr.jl_amd/csrc)" % LIB_PATH)
        # One HIP runtime per process: when torch is present it carries its own
        # libamdhip64.so.7; loading torch first makes the library bind to that
        # same runtime (same soname) instead of a second copy from /opt/rocm.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().draugr_amg_last_error().decode("utf-8", "replace")


def check(ret: int, what: str = "") -> int:
    if ret < 0:
        raise DraugrError("%s failed: %s" % (what, last_error()) if what else last_error())
    return ret
