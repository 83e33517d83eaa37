"""Torch-facing front end of the device API.

torch provides device memory and the stream; every computation is a call
into libams_hip.so.  ``DeviceCsr`` holds int64 row pointers, int32 columns
and float64/float32 values on one GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from .csr import Csr


@dataclass
class DeviceCsr:
    num_row: int
    num_col: int
    row_ptr: torch.Tensor  # int64
    col: torch.Tensor      # int32
    val: torch.Tensor      # float64 / float32

    @property
    def num_nnz(self) -> int:
        return int(self.col.numel())

    @property
    def dtype_code(self) -> int:
        return L.AMS_F64 if self.val.dtype == torch.float64 else L.AMS_F32

    def c_struct(self) -> L.AmsDcsr:
        return L.AmsDcsr(self.num_row, self.num_col, self.num_nnz, self.row_ptr.data_ptr(),
                         self.col.data_ptr() if self.num_nnz else None,
                         self.val.data_ptr() if self.num_nnz else None)

    @staticmethod
    def from_host(m: Csr, device="cuda", val_dtype=None) -> "DeviceCsr":
        vd = m.rats.dtype if val_dtype is None else val_dtype
        return DeviceCsr(m.num_row, m.num_col,
                         torch.from_numpy(np.ascontiguousarray(m.row_ptrs, np.int64)).to(device),
                         torch.from_numpy(np.ascontiguousarray(m.cols, np.int32)).to(device),
                         torch.from_numpy(np.ascontiguousarray(m.rats, vd)).to(device))

    def to_host(self) -> Csr:
        return Csr(self.num_row, self.num_col, self.row_ptr.cpu().numpy(), self.col.cpu().numpy(),
                   self.val.cpu().numpy())


class Context:
    """One ams_ctx_t bound to a torch stream (the current one by default)."""

    def __init__(self, stream: torch.cuda.Stream | None = None, params: dict | None = None):
        self._lib = L.lib()
        self._p = C.c_void_p()
        s = stream if stream is not None else torch.cuda.current_stream()
        self.stream = s
        L.check(self._lib.ams_ctx_create(C.byref(self._p), C.c_void_p(s.cuda_stream)), None, "ams_ctx_create")
        if params:
            self.set_params(**params)

    def close(self):
        if self._p:
            self._lib.ams_ctx_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        L.check(rc, self._p, what)

    def params(self) -> dict:
        p = L.AmsParams()
        self._chk(self._lib.ams_get_params(self._p, C.byref(p)), "ams_get_params")
        return p.as_dict()

    def set_params(self, **kw):
        p = L.AmsParams()
        self._chk(self._lib.ams_get_params(self._p, C.byref(p)), "ams_get_params")
        for k, v in kw.items():
            setattr(p, k, v)
        self._chk(self._lib.ams_set_params(self._p, C.byref(p)), "ams_set_params")

    def stats(self) -> dict:
        s = L.AmsStats()
        self._chk(self._lib.ams_get_stats(self._p, C.byref(s)), "ams_get_stats")
        return s.as_dict()

    def enable_timing(self, on: bool = True):
        self._chk(self._lib.ams_enable_timing(self._p, int(on)), "ams_enable_timing")

    def timings(self) -> dict:
        """Per-phase event time (ms) and launch count accumulated since
        enable_timing()."""
        n = len(L.PHASES)
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        self._chk(self._lib.ams_get_timings(self._p, ms, cnt), "ams_get_timings")
        return {name: (ms[i], cnt[i]) for i, name in enumerate(L.PHASES)}

    def row_flops(self, a: DeviceCsr, b: DeviceCsr):
        out = torch.empty(max(a.num_row, 1), dtype=torch.int64, device=a.row_ptr.device)
        tot = C.c_int64()
        sa, sb = a.c_struct(), b.c_struct()
        self._chk(self._lib.ams_row_flops(self._p, C.byref(sa), C.byref(sb), out.data_ptr(), C.byref(tot)),
                  "ams_row_flops")
        return out[:a.num_row], tot.value

    def symbolic(self, a: DeviceCsr, b: DeviceCsr):
        rp = torch.empty(a.num_row + 1, dtype=torch.int64, device=a.row_ptr.device)
        nnz = C.c_int64()
        sa, sb = a.c_struct(), b.c_struct()
        self._chk(self._lib.ams_spgemm_symbolic(self._p, C.byref(sa), C.byref(sb), a.dtype_code, rp.data_ptr(),
                                                C.byref(nnz)), "ams_spgemm_symbolic")
        return rp, nnz.value

    def numeric(self, a: DeviceCsr, b: DeviceCsr, rp: torch.Tensor, nnz: int) -> DeviceCsr:
        dev = a.row_ptr.device
        col = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        val = torch.empty(max(nnz, 1), dtype=a.val.dtype, device=dev)
        sa, sb = a.c_struct(), b.c_struct()
        self._chk(self._lib.ams_spgemm_numeric(self._p, C.byref(sa), C.byref(sb), a.dtype_code, rp.data_ptr(),
                                               col.data_ptr(), val.data_ptr()), "ams_spgemm_numeric")
        return DeviceCsr(a.num_row, b.num_col, rp, col[:nnz], val[:nnz])

    def spgemm(self, a: DeviceCsr, b: DeviceCsr) -> DeviceCsr:
        """C = A * B into freshly allocated torch tensors (symbolic + numeric)."""
        if a.val.dtype != b.val.dtype:
            raise ValueError("A and B must share a value dtype")
        rp, nnz = self.symbolic(a, b)
        return self.numeric(a, b, rp, nnz)

    def spgemm_pooled(self, a: DeviceCsr, b: DeviceCsr) -> L.AmsDcsr:
        """One-shot product into the context pool (valid until the next call)."""
        sa, sb, c = a.c_struct(), b.c_struct(), L.AmsDcsr()
        self._chk(self._lib.ams_spgemm(self._p, C.byref(sa), C.byref(sb), a.dtype_code, C.byref(c)), "ams_spgemm")
        return c

    def masked_bwd(self, a: DeviceCsr, b: DeviceCsr, g: DeviceCsr):
        """dA = (G B^T) on pattern(A), dB = (A^T G) on pattern(B), for G on
        pattern(A B) with ascending columns in every row."""
        da = torch.empty(max(a.num_nnz, 1), dtype=a.val.dtype, device=a.val.device)
        db = torch.empty(max(b.num_nnz, 1), dtype=b.val.dtype, device=b.val.device)
        sa, sb, sg = a.c_struct(), b.c_struct(), g.c_struct()
        self._chk(self._lib.ams_spgemm_masked_bwd(self._p, C.byref(sa), C.byref(sb), C.byref(sg), a.dtype_code,
                                                  da.data_ptr(), db.data_ptr()), "ams_spgemm_masked_bwd")
        return da[:a.num_nnz], db[:b.num_nnz]

    def verify(self, c, dtype_code: int = L.AMS_F64) -> dict:
        """VERIFYRESULT triple (nnz, sum of columns, sum of values) of a
        product, plus the count of rows whose columns are not ascending."""
        sc = c.c_struct() if isinstance(c, DeviceCsr) else c
        if isinstance(c, DeviceCsr):
            dtype_code = c.dtype_code
        nnz, scol, uns, sval = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        self._chk(self._lib.ams_verify_result(self._p, C.byref(sc), dtype_code, C.byref(nnz), C.byref(scol),
                                              C.byref(sval), C.byref(uns)), "ams_verify_result")
        return {"nnz": nnz.value, "sum_col": scol.value, "sum_val": sval.value, "unsorted_rows": uns.value}

    def error_flags(self) -> int:
        """Error flags raised by the kernels of the last run on this
        context (0 = none), read after its stream drains."""
        f = C.c_int32()
        self._chk(self._lib.ams_get_error_flags(self._p, C.byref(f)), "ams_get_error_flags")
        return f.value

    def autotune(self, a: DeviceCsr, b: DeviceCsr, reps: int = 2):
        best, ms = L.AmsParams(), C.c_double()
        sa, sb = a.c_struct(), b.c_struct()
        self._chk(self._lib.ams_autotune(self._p, C.byref(sa), C.byref(sb), a.dtype_code, reps, C.byref(best),
                                         C.byref(ms)), "ams_autotune")
        return best.as_dict(), ms.value

    @property
    def handle(self):
        return self._p
