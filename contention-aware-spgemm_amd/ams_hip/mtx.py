"""MatrixMarket and PaRMAT text files through the library's reader and
writer (mtx2csr / trpl2mtx / txt_transpose of include/ams_dropin.h, the
file semantics of the reference's src/mm/mm_utils.c:451-862).

The reader forces every value to 1.0, drops explicit zeros, mirrors the
off-diagonal entries of symmetric files and sorts by (row, column), as the
reference's does: it reads the structure of a matrix, not its values.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from .csr import Csr


class CsrT(C.Structure):
    """csr_t (src/inc/store.h:49-56)."""
    _fields_ = [("num_row", C.c_int64), ("num_col", C.c_int64), ("num_nnz", C.c_int64),
                ("row_ptrs", C.POINTER(C.c_int64)), ("cols", C.POINTER(C.c_int64)),
                ("rats", C.POINTER(C.c_double))]


TRIPLE = np.dtype([("user_id", np.int64), ("item_id", np.int64), ("rating", np.float64)])

_BOUND = False


def _lib():
    global _BOUND
    lib = L.lib()
    if not _BOUND:
        lib.mtx2csr.argtypes = [C.POINTER(CsrT), C.c_char_p, C.c_void_p]
        lib.mtx2csr.restype = None
        lib.free_csr.argtypes = [C.POINTER(CsrT)]
        lib.free_csr.restype = None
        lib.trpl2mtx.argtypes = [C.c_char_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_bool]
        lib.trpl2mtx.restype = None
        lib.txt_transpose.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
        lib.txt_transpose.restype = None
        _BOUND = True
    return lib


def read_mtx(path: str) -> Csr:
    """A MatrixMarket coordinate file as CSR (int32 columns, values 1.0)."""
    if not os.path.exists(path):
        raise FileNotFoundError(path)  # the C reader would exit the process
    lib = _lib()
    c = CsrT()
    lib.mtx2csr(C.byref(c), os.fsencode(path), None)
    try:
        m, z = c.num_row, c.num_nnz
        rp = np.ctypeslib.as_array(c.row_ptrs, (m + 1,)).copy()
        cols = np.ctypeslib.as_array(c.cols, (z,)).astype(np.int32) if z else np.zeros(0, np.int32)
        vals = np.ctypeslib.as_array(c.rats, (z,)).copy() if z else np.zeros(0)
        return Csr(m, c.num_col, rp, cols, vals)
    finally:
        lib.free_csr(C.byref(c))


def write_mtx(path: str, a: Csr, symmetric: bool = False):
    """Writes the structure of `a` (values written as 1.0; for a symmetric
    file only the entries with row <= column)."""
    rows = np.repeat(np.arange(a.num_row, dtype=np.int64), np.diff(np.asarray(a.row_ptrs, np.int64)))
    t = np.zeros(a.num_nnz, TRIPLE)
    t["user_id"] = rows
    t["item_id"] = np.asarray(a.cols, np.int64)
    t["rating"] = 1.0
    _lib().trpl2mtx(os.fsencode(path), t.ctypes.data_as(C.c_void_p), a.num_row, a.num_col, a.num_nnz, symmetric)


def rmat_ones(src_txt: str, dst_org: str, dst_trs: str):
    """PaRMAT text "<name>_<scale>_<ef>.txt" (0-based edges) -> the symmetric
    all-ones MatrixMarket pair (src/rmat_ones.c:6-37)."""
    for p in (src_txt,):
        if not os.path.exists(p):
            raise FileNotFoundError(p)
    _lib().txt_transpose(os.fsencode(dst_trs), os.fsencode(dst_org), os.fsencode(src_txt), 0)
