"""The reference's C entry points (include/ams_dropin.h) called through the C
ABI with the reference's host layout (int64 indices, double values), as a C
caller of libAms would call them, against the CPU oracle."""
import ctypes as C

import numpy as np
import pytest

from ams_hip import gen
from helpers import oracle_product, random_csr

pytestmark = pytest.mark.gpu


class CsrT(C.Structure):
    _fields_ = [("num_row", C.c_int64), ("num_col", C.c_int64), ("num_nnz", C.c_int64),
                ("row_ptrs", C.POINTER(C.c_int64)), ("cols", C.POINTER(C.c_int64)),
                ("rats", C.POINTER(C.c_double))]


class KeyVal(C.Structure):
    _fields_ = [("key", C.c_int64), ("val", C.c_double)]


class Runtime(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("init_time", "symb_time", "mgmt_time", "numc_time", "total_time")]


class Pack(C.Structure):
    _fields_ = [("b_roff_maxc", C.POINTER(KeyVal)), ("b_arr", C.POINTER(KeyVal)), ("total_cnt", C.c_int64),
                ("high_head_idx", C.c_int64), ("high_tail_idx", C.c_int64), ("low_head_idx", C.c_int64),
                ("low_tail_idx", C.c_int64)]


class Bin(C.Structure):
    _fields_ = [("csr", C.POINTER(CsrT)), ("ams_pack", C.POINTER(Pack)), ("row_flp", C.POINTER(C.c_int64)),
                ("ps_row_flp", C.POINTER(C.c_int64)), ("row_nz", C.POINTER(C.c_int64))]


def lib():
    from ams_hip import _lib
    L = C.CDLL(_lib.LIB_PATH)
    P = C.POINTER
    L.hybaccspgemm.restype = Runtime
    L.hybaccspgemm.argtypes = [P(CsrT), P(CsrT), P(CsrT), C.c_void_p]
    L.hashspgemm.restype = Runtime
    L.hashspgemm.argtypes = [P(CsrT), P(CsrT), P(CsrT), C.c_void_p]
    L.hserscspgemm.restype = Runtime
    L.hserscspgemm.argtypes = [P(CsrT), P(CsrT), P(CsrT), C.c_void_p]
    L.free_csr.argtypes = [P(CsrT)]
    L.flp_populate.argtypes = [P(CsrT), P(CsrT), P(C.c_int64)]
    L.get_intprod.argtypes = [P(CsrT), P(CsrT), P(C.c_int64)]
    L.get_intprod.restype = C.c_int64
    L.scan.argtypes = [P(C.c_int64), P(C.c_int64), C.c_int64]
    L.init_ams_bin.argtypes = [P(Bin), P(Pack), P(CsrT), C.c_void_p]
    L.free_ams_bin.argtypes = [P(Bin)]
    L.pack_ams_pack.argtypes = [P(C.c_int64), P(Bin), C.c_void_p]
    L.pack_ams_pack.restype = C.c_int64
    L.free_ams_pack.argtypes = [P(Pack)]
    L.ams_hash_symb_dnr.argtypes = [P(CsrT), P(CsrT), P(C.c_int64), P(Bin)]
    L.ams_hybacc_numc.argtypes = [P(CsrT), P(CsrT), P(CsrT), P(KeyVal), P(KeyVal), C.c_int64, P(C.c_int64),
                                  P(C.c_double), P(Bin)]
    L.ams_dropin_set_pack_params.argtypes = [C.c_int64, C.c_int64, C.c_double]
    L.ams_hash_symb_omp_dnr.argtypes = [P(CsrT), P(CsrT), P(C.c_int64), P(Bin)]
    L.ams_dropin_counters.argtypes = [P(C.c_int64), P(C.c_int64)]
    return L


def counters(L):
    up, sy = C.c_int64(), C.c_int64()
    L.ams_dropin_counters(C.byref(up), C.byref(sy))
    return up.value, sy.value


class Held:
    def __init__(self, m):
        self.rp = np.ascontiguousarray(m.row_ptrs, np.int64)
        self.cols = np.ascontiguousarray(m.cols, np.int64)
        self.rats = np.ascontiguousarray(m.rats, np.float64)
        self.s = CsrT(m.num_row, m.num_col, int(self.rp[-1]), self.rp.ctypes.data_as(C.POINTER(C.c_int64)),
                      self.cols.ctypes.data_as(C.POINTER(C.c_int64)),
                      self.rats.ctypes.data_as(C.POINTER(C.c_double)))


def as_np(ptr, n, dt):
    return np.ctypeslib.as_array(ptr, (max(n, 1),))[:n].astype(dt).copy() if n else np.zeros(0, dt)


@pytest.mark.parametrize("fn", ["hybaccspgemm", "hashspgemm", "hserscspgemm"])
def test_whole_product(gpu_ctx, fn):
    L = lib()
    a = gen.rmat_symmetric(11, 8, seed=3, values="uniform")
    ha = Held(a)
    c = CsrT()
    up0, sy0 = counters(L)
    rt = getattr(L, fn)(C.byref(ha.s), C.byref(ha.s), C.byref(c), None)
    assert counters(L) == (up0 + 1, sy0 + 1), "C = A*A uploads A once and plans once"
    assert rt.total_time > 0
    rp = as_np(c.row_ptrs, c.num_row + 1, np.int64)
    cols = as_np(c.cols, c.num_nnz, np.int64)
    rats = as_np(c.rats, c.num_nnz, np.float64)
    L.free_csr(C.byref(c))
    orp, ocols, orats = oracle_product(a, a)
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(cols, ocols)
    np.testing.assert_allclose(rats, orats, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("symb", ["ams_hash_symb_dnr", "ams_hash_symb_omp_dnr"])
def test_phase_functions(gpu_ctx, symb):
    """The driver sequence of src/ams_hybaccspgemm.c:36-220 written against
    the phase entry points: A and B go to the GPU once each and the symbolic
    phase runs once, its plan reused by the numeric phase."""
    L = lib()
    from oracle import oracle as orc
    a = random_csr(300, 200, 0.05, seed=7)
    b = random_csr(200, 500, 0.03, seed=8, sort=False)
    ha, hb = Held(a), Held(b)
    m = a.num_row
    pack, binr = Pack(), Bin()
    up0, sy0 = counters(L)
    L.init_ams_bin(C.byref(binr), C.byref(pack), C.byref(ha.s), None)
    L.flp_populate(C.byref(ha.s), C.byref(hb.s), binr.row_flp)
    flp = as_np(binr.row_flp, m, np.int64)
    np.testing.assert_array_equal(flp, orc.flp_populate(a, b))
    assert L.get_intprod(C.byref(ha.s), C.byref(hb.s), binr.row_nz) == int(flp.sum())
    L.scan(binr.ps_row_flp, binr.row_flp, m + 1)
    np.testing.assert_array_equal(as_np(binr.ps_row_flp, m + 1, np.int64), np.concatenate([[0], np.cumsum(flp)]))
    getattr(L, symb)(C.byref(ha.s), C.byref(hb.s), None, C.byref(binr))
    row_nz = as_np(binr.row_nz, m, np.int64)
    np.testing.assert_array_equal(row_nz, orc.hash_symb(a, b))
    crp = np.zeros(m + 1, np.int64)
    L.scan(crp.ctypes.data_as(C.POINTER(C.c_int64)), binr.row_nz, m + 1)
    nnz = int(crp[-1])
    L.ams_dropin_set_pack_params(64, 16, 0.85)
    mx = L.pack_ams_pack(crp.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(binr), None)
    p = orc.params(threads=2, ams_min_num=16, ams_size=64, ams_prt=0.85)
    roff, arr, idx, omx = orc.pack(crp, flp, p)
    assert mx == omx and pack.total_cnt == idx["total_cnt"]
    assert [pack.b_roff_maxc[i].key for i in range(pack.total_cnt + 1)] == [r[0] for r in roff]
    assert [pack.b_arr[i].key for i in range(pack.total_cnt)] == [r[0] for r in arr]
    assert (pack.low_head_idx, pack.high_tail_idx) == (idx["low_head_idx"], idx["high_tail_idx"])
    cols = np.zeros(nnz, np.int64)
    rats = np.zeros(nnz)
    c = CsrT(m, b.num_col, nnz, crp.ctypes.data_as(C.POINTER(C.c_int64)), cols.ctypes.data_as(C.POINTER(C.c_int64)),
             rats.ctypes.data_as(C.POINTER(C.c_double)))
    L.ams_hybacc_numc(C.byref(ha.s), C.byref(hb.s), C.byref(c), None, None, mx, None, None, C.byref(binr))
    orp, ocols, orats = oracle_product(a, b)
    np.testing.assert_array_equal(crp, orp)
    np.testing.assert_array_equal(cols, ocols)
    np.testing.assert_allclose(rats, orats, rtol=1e-12, atol=1e-13)
    assert counters(L) == (up0 + 2, sy0 + 1), "one upload of A and of B, one symbolic phase"
    L.free_ams_pack(C.byref(pack))
    L.free_ams_bin(C.byref(binr))
    # after the sequence nothing is kept: the next call uploads again
    L.flp_populate(C.byref(ha.s), C.byref(hb.s), None)
    assert counters(L) == (up0 + 4, sy0 + 1)


def test_numeric_needs_same_operands(gpu_ctx):
    """The numeric phase refuses operands other than those of the symbolic
    call it follows (same shape but other arrays): its routing and chunk
    plan would not describe them."""
    import torch
    from ams_hip import _lib as L
    from ams_hip.device import DeviceCsr
    a = gen.rmat_symmetric(9, 8, seed=1)
    b = gen.rmat_symmetric(9, 8, seed=2)
    da, db = DeviceCsr.from_host(a), DeviceCsr.from_host(b)
    rp, nnz = gpu_ctx.symbolic(da, da)
    with pytest.raises(L.AmsError, match="same A and B"):
        gpu_ctx.numeric(db, db, rp, nnz)
    c = gpu_ctx.numeric(da, da, rp, nnz)
    torch.cuda.synchronize()
    assert gpu_ctx.error_flags() == 0
    assert c.num_nnz == nnz


def test_numeric_reroutes_on_param_change(gpu_ctx):
    """Numeric-routing parameters changed between the phases (here the
    accumulator mode) only re-route the rows; parameters of the symbolic
    geometry (the dense chunk width) require a new symbolic phase."""
    import torch
    from ams_hip import _lib as L
    from ams_hip.device import Context, DeviceCsr
    from helpers import assert_matches_oracle
    a = gen.rmat_symmetric(10, 8, seed=4, values="uniform")
    ctx = Context()
    da = DeviceCsr.from_host(a)
    rp, nnz = ctx.symbolic(da, da)
    assert ctx.stats()["rows_num_esc"] > 0
    ctx.set_params(mode=L.AMS_MODE_HASH)
    c = ctx.numeric(da, da, rp, nnz)
    torch.cuda.synchronize()
    assert ctx.stats()["rows_num_esc"] == 0
    assert_matches_oracle(c.to_host(), a, a, exact=False)
    ctx.set_params(mode=L.AMS_MODE_HYBRID, dense_width=4096)
    with pytest.raises(L.AmsError, match="symbolic parameters changed"):
        ctx.numeric(da, da, rp, nnz)
    ctx.close()
