"""ctypes front end of the CPU oracle (oracle/ams_oracle.c).

TEST INFRASTRUCTURE ONLY: the correctness oracle and the reported CPU
baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module; the product package never does.

Each wrapper names the reference function it restates (see ams_oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libams_oracle.so")

HYBACC, HASH, ESC = 0, 1, 2
_MODES = {"hybacc": HYBACC, "hash": HASH, "esc": ESC}


class _Csr(C.Structure):
    _fields_ = [("num_row", C.c_int64), ("num_col", C.c_int64), ("num_nnz", C.c_int64),
                ("row_ptrs", C.POINTER(C.c_int64)), ("cols", C.POINTER(C.c_int64)),
                ("rats", C.POINTER(C.c_double))]


class _KeyVal(C.Structure):
    _fields_ = [("key", C.c_int64), ("val", C.c_double)]


class _Runtime(C.Structure):
    _fields_ = [("init_time", C.c_double), ("symb_time", C.c_double), ("mgmt_time", C.c_double),
                ("numc_time", C.c_double), ("total_time", C.c_double)]


class _Pack(C.Structure):
    _fields_ = [("b_roff_maxc", C.POINTER(_KeyVal)), ("b_arr", C.POINTER(_KeyVal)),
                ("total_cnt", C.c_int64), ("high_head_idx", C.c_int64),
                ("high_tail_idx", C.c_int64), ("low_head_idx", C.c_int64),
                ("low_tail_idx", C.c_int64)]


class Params(C.Structure):
    _fields_ = [("thread_num", C.c_int), ("smt_core_num", C.c_int), ("high_thread_num", C.c_int),
                ("ams_size", C.c_int64), ("ams_min_num", C.c_int64), ("ams_prt", C.c_double),
                ("acc_thres", C.c_double), ("smt_symb_thres", C.c_int64),
                ("smt_numc_thres", C.c_int64), ("dn_min", C.c_int64), ("hash_constant", C.c_int64)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P64 = C.POINTER(C.c_int64)
        L.orc_default_params.argtypes = [C.POINTER(Params), C.c_int]
        L.orc_next_pow2.argtypes = [C.c_int64]
        L.orc_next_pow2.restype = C.c_int64
        L.orc_lower_bound.argtypes = [P64, P64, C.c_int64]
        L.orc_lower_bound.restype = P64
        L.orc_scan.argtypes = [P64, P64, C.c_int64]
        L.orc_flp_populate.argtypes = [C.POINTER(_Csr), C.POINTER(_Csr), P64, C.c_int]
        L.orc_get_intprod.argtypes = [C.POINTER(_Csr), C.POINTER(_Csr), C.c_int]
        L.orc_get_intprod.restype = C.c_int64
        L.orc_hash_symb_dnr.argtypes = [C.POINTER(_Csr), C.POINTER(_Csr), P64, P64,
                                        C.POINTER(Params), C.c_int]
        L.orc_pack.argtypes = [P64, P64, P64, C.c_int64, C.POINTER(Params), C.POINTER(_Pack)]
        L.orc_pack.restype = C.c_int64
        L.orc_free_pack.argtypes = [C.POINTER(_Pack)]
        L.orc_pack_mmmv.argtypes = [P64, C.c_int64, C.c_int64, P64, P64, C.c_int64, C.POINTER(Params),
                                    C.POINTER(_Pack)]
        L.orc_pack_mmmv.restype = C.c_int64
        L.orc_radix_sort_keyval.argtypes = [C.POINTER(_KeyVal), C.POINTER(_KeyVal), C.c_int64]
        L.orc_radix_sort_keyval.restype = C.POINTER(_KeyVal)
        L.orc_spgemm.argtypes = [C.POINTER(_Csr), C.POINTER(_Csr), C.POINTER(_Csr),
                                 C.POINTER(Params), C.c_int]
        L.orc_spgemm.restype = _Runtime
        L.orc_free_csr.argtypes = [C.POINTER(_Csr)]
        L.orc_canonicalize.argtypes = [C.POINTER(_Csr), C.c_int]
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class _Held:
    """Keeps the int64/double copies alive while C reads them."""

    def __init__(self, m):
        self.rp = np.ascontiguousarray(m.row_ptrs, np.int64)
        self.cols = np.ascontiguousarray(m.cols, np.int64)
        self.rats = np.ascontiguousarray(m.rats, np.float64)
        self.s = _Csr(m.num_row, m.num_col, int(self.rp[-1]), _p64(self.rp), _p64(self.cols),
                      self.rats.ctypes.data_as(C.POINTER(C.c_double)))


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


def params(threads=None, **over):
    p = Params()
    lib().orc_default_params(C.byref(p), threads or default_threads())
    for k, v in over.items():
        setattr(p, k, v)
    return p


def next_pow2(x):
    """NEXT_POW_2 of src/inc/utils.h:92."""
    return lib().orc_next_pow2(int(x))


def lower_bound(arr, val):
    """lower_bound of src/mm/mm_utils.c:963-974, as an index."""
    a = np.ascontiguousarray(arr, np.int64)
    end = C.cast(C.c_void_p(a.ctypes.data + 8 * len(a)), C.POINTER(C.c_int64))
    r = lib().orc_lower_bound(_p64(a), end, int(val))
    return (C.cast(r, C.c_void_p).value - a.ctypes.data) // 8


def scan(inp, num):
    """scan(out, in, num) of src/mm/mm_utils.c:866-957."""
    i = np.ascontiguousarray(inp, np.int64)
    out = np.zeros(num, np.int64)
    lib().orc_scan(_p64(out), _p64(i), int(num))
    return out


def flp_populate(a, b, threads=None):
    """flp_populate of src/mm/mm_utils.c:1063-1125."""
    ha, hb = _Held(a), _Held(b)
    out = np.zeros(max(a.num_row, 1), np.int64)
    lib().orc_flp_populate(C.byref(ha.s), C.byref(hb.s), _p64(out), threads or default_threads())
    return out[:a.num_row]


def get_intprod(a, b, threads=None):
    """get_intprod of src/mm/mm_utils.c:981-1055."""
    ha, hb = _Held(a), _Held(b)
    return lib().orc_get_intprod(C.byref(ha.s), C.byref(hb.s), threads or default_threads())


def hash_symb(a, b, threads=None, p=None):
    """ams_hash_symb_dnr of src/ams/ams_utils.c:1600-1752: nnz per C row."""
    p = p or params(threads)
    flp = flp_populate(a, b, threads)
    ha, hb = _Held(a), _Held(b)
    f = np.ascontiguousarray(flp, np.int64)
    out = np.zeros(max(a.num_row, 1), np.int64)
    lib().orc_hash_symb_dnr(C.byref(ha.s), C.byref(hb.s), _p64(f), _p64(out), C.byref(p),
                            p.thread_num)
    return out[:a.num_row]


def pack(ps_row_nz, row_flp, p=None):
    """pack_ams_pack of src/ams/ams_utils.c:1410-1596.

    Returns (b_roff_maxc [(first_row, max_flp)], b_arr [(id, cf)], indices dict, max_row_flp).
    """
    p = p or params()
    ps = np.ascontiguousarray(ps_row_nz, np.int64)
    flp = np.ascontiguousarray(row_flp, np.int64)
    ps_flp = scan(flp, len(flp) + 1)
    pk = _Pack()
    mx = lib().orc_pack(_p64(ps), _p64(flp), _p64(ps_flp), len(flp), C.byref(p), C.byref(pk))
    n = pk.total_cnt
    roff = [(pk.b_roff_maxc[i].key, pk.b_roff_maxc[i].val) for i in range(n + 1)]
    arr = [(pk.b_arr[i].key, pk.b_arr[i].val) for i in range(n)]
    idx = dict(total_cnt=n, high_head_idx=pk.high_head_idx, high_tail_idx=pk.high_tail_idx,
               low_head_idx=pk.low_head_idx, low_tail_idx=pk.low_tail_idx)
    lib().orc_free_pack(C.byref(pk))
    return roff, arr, idx, mx


def pack_mmmv(row_ptrs, ps_row_coef, nnz_per_row, coef, p=None):
    """pack_ams_pack_mmmv of src/ams/ams_utils.c:1851-2046 for a CSR with the
    given row pointers.  Returns what pack() returns."""
    p = p or params()
    rp = np.ascontiguousarray(row_ptrs, np.int64)
    ps = np.ascontiguousarray(ps_row_coef, np.int64)
    nz = np.ascontiguousarray(nnz_per_row, np.int64)
    pk = _Pack()
    m = len(rp) - 1
    mx = lib().orc_pack_mmmv(_p64(rp), m, int(rp[-1]), _p64(ps), _p64(nz), int(coef), C.byref(p), C.byref(pk))
    n = pk.total_cnt
    roff = [(pk.b_roff_maxc[i].key, pk.b_roff_maxc[i].val) for i in range(n + 1)]
    arr = [(pk.b_arr[i].key, pk.b_arr[i].val) for i in range(n)]
    idx = dict(total_cnt=n, high_head_idx=pk.high_head_idx, high_tail_idx=pk.high_tail_idx,
               low_head_idx=pk.low_head_idx, low_tail_idx=pk.low_tail_idx)
    lib().orc_free_pack(C.byref(pk))
    return roff, arr, idx, mx


def radix_sort_keyval(keys, vals):
    """radix_sort_keyval_toggling_unrolling of src/mm/mm_utils.c:1273-1375."""
    n = len(keys)
    arr = (_KeyVal * max(n, 1))()
    aux = (_KeyVal * max(n, 1))()
    for i in range(n):
        arr[i].key, arr[i].val = int(keys[i]), float(vals[i])
    res = lib().orc_radix_sort_keyval(arr, aux, n)
    return (np.array([res[i].key for i in range(n)], np.int64),
            np.array([res[i].val for i in range(n)], np.float64))


def spgemm(a, b, mode="hybacc", threads=None, canonical=True, p=None):
    """The whole hybaccspgemm driver (src/ams_hybaccspgemm.c:36-220) and its
    hash-only / ESC-only siblings.  Returns (row_ptrs, cols, rats, runtime dict);
    with canonical=True every row is sorted by column (the raw reference order
    is slot order for hash rows)."""
    p = p or params(threads)
    ha, hb = _Held(a), _Held(b)
    c = _Csr()
    rt = lib().orc_spgemm(C.byref(ha.s), C.byref(hb.s), C.byref(c), C.byref(p), _MODES[mode])
    if canonical:
        lib().orc_canonicalize(C.byref(c), p.thread_num)
    m, nnz = c.num_row, c.num_nnz
    rp = np.ctypeslib.as_array(c.row_ptrs, (m + 1,)).copy()
    cols = np.ctypeslib.as_array(c.cols, (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int64)
    rats = np.ctypeslib.as_array(c.rats, (max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0)
    lib().orc_free_csr(C.byref(c))
    runtime = {k: getattr(rt, k) for k, _ in _Runtime._fields_}
    return rp, cols, rats, runtime
