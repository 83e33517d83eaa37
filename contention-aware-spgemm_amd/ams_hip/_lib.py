"""ctypes binding of libams_hip.so (the C ABI in include/ams_hip.h).

The library must have been built (``make -C contention-aware-spgemm_amd`` or
``__graft_entry__.build()``); there is no fallback path: if it cannot be
loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libams_hip.so")

AMS_OK, AMS_ERR_ARG, AMS_ERR_HIP, AMS_ERR_ALLOC, AMS_ERR_STATE, AMS_ERR_LIMIT = 0, -1, -2, -3, -4, -5
AMS_F64, AMS_F32 = 0, 1
AMS_MODE_HYBRID, AMS_MODE_HASH, AMS_MODE_ESC = 0, 1, 2


class AmsDcsr(C.Structure):
    _fields_ = [("num_row", C.c_int64), ("num_col", C.c_int64), ("num_nnz", C.c_int64),
                ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class AmsParams(C.Structure):
    _fields_ = [("sym_hash_max", C.c_int32), ("num_hash_max", C.c_int32), ("esc_max", C.c_int32),
                ("acc_thres", C.c_double), ("dense_width", C.c_int32), ("bitmap_bits", C.c_int32),
                ("bitmap_ratio", C.c_double), ("dense_density", C.c_double), ("mode", C.c_int32),
                ("wgs_per_cu", C.c_int32), ("esc_sort", C.c_int32),
                ("sym_blocked", C.c_int32), ("ghash_ratio", C.c_double), ("light_max", C.c_int32),
                ("row_bitmaps", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AmsStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("rows_total", "rows_empty", "rows_sym_hash", "rows_sym_bitmap",
                                         "rows_num_hash", "rows_num_esc", "rows_num_dense",
                                         "dense_chunks", "flops", "nnz_c")] + [
        ("flops_route", C.c_int64 * 4), ("nnzc_route", C.c_int64 * 4), ("nnza_route", C.c_int64 * 4),
        ("flops_sym", C.c_int64 * 3), ("nnza_sym", C.c_int64 * 3), ("rows_sym_ghash", C.c_int64),
        ("rows_num_ghash", C.c_int64)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if not isinstance(v, int) else v
        return out


PHASES = ["flops", "classify", "symb_hash", "symb_bitmap", "scan", "bblk", "numc_dense", "numc_hash",
          "numc_esc", "symb_ghash", "numc_ghash", "numc_light"]


# name -> (restype, argtypes); every symbol declared in include/ams_hip.h.
P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
PD = C.POINTER(AmsDcsr)
SIGNATURES = {
    "ams_ctx_create": (C.c_int, [C.POINTER(C.c_void_p), P]),
    "ams_ctx_destroy": (C.c_int, [P]),
    "ams_set_stream": (C.c_int, [P, P]),
    "ams_last_error": (C.c_char_p, [P]),
    "ams_default_params": (None, [C.POINTER(AmsParams)]),
    "ams_set_params": (C.c_int, [P, C.POINTER(AmsParams)]),
    "ams_get_params": (C.c_int, [P, C.POINTER(AmsParams)]),
    "ams_get_stats": (C.c_int, [P, C.POINTER(AmsStats)]),
    "ams_enable_timing": (C.c_int, [P, C.c_int]),
    "ams_get_timings": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(I64)]),
    "ams_row_flops": (C.c_int, [P, PD, PD, P, C.POINTER(I64)]),
    "ams_scan_i64": (C.c_int, [P, P, P, I64]),
    "ams_spgemm_symbolic": (C.c_int, [P, PD, PD, C.c_int, P, C.POINTER(I64)]),
    "ams_spgemm_numeric": (C.c_int, [P, PD, PD, C.c_int, P, P, P]),
    "ams_spgemm": (C.c_int, [P, PD, PD, C.c_int, PD]),
    "ams_batched_plan": (C.c_int, [P, PD, PD, C.c_int, I32, P, P, P, C.POINTER(I64)]),
    "ams_batched_run": (C.c_int, [P, PD, PD, C.c_int, P, P, P, I64, P]),
    "ams_batched_capture": (C.c_int, [P, PD, PD, C.c_int, P, P, P, I64, P, C.POINTER(C.c_void_p)]),
    "ams_graph_launch": (C.c_int, [P, P]),
    "ams_graph_destroy": (C.c_int, [P]),
    "ams_get_error_flags": (C.c_int, [P, C.POINTER(C.c_int32)]),
    "ams_spgemm_fixed": (C.c_int, [P, PD, PD, C.c_int, P, P, P, I64, P]),
    "ams_spgemm_masked_bwd": (C.c_int, [P, PD, PD, PD, C.c_int, P, P]),
    "ams_autotune": (C.c_int, [P, PD, PD, C.c_int, C.c_int, C.POINTER(AmsParams), C.POINTER(D)]),
    "ams_verify_result": (C.c_int, [P, PD, C.c_int, C.POINTER(I64), C.POINTER(I64), C.POINTER(D),
                                    C.POINTER(I64)]),
    "ams_dev_alloc": (C.c_int, [C.POINTER(C.c_void_p), I64]),
    "ams_dev_free": (C.c_int, [P]),
    "ams_memcpy_h2d": (C.c_int, [P, P, P, I64]),
    "ams_memcpy_d2h": (C.c_int, [P, P, P, I64]),
    "ams_sync": (C.c_int, [P]),
    "ams_i64_to_i32": (C.c_int, [P, P, P, I64]),
    "ams_i32_to_i64": (C.c_int, [P, P, P, I64]),
}

_lib = None


class AmsError(RuntimeError):
    pass


def lib():
    """Loads the native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AmsError(f"{LIB_PATH} is missing: build it with `make -C contention-aware-spgemm_amd` "
                           "(no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, ctx=None, what=""):
    if rc != AMS_OK:
        msg = lib().ams_last_error(ctx).decode() if ctx else ""
        raise AmsError(f"{what} failed with code {rc}: {msg}")
