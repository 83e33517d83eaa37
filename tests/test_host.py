"""Host-side logic without a GPU: generators, CSR helpers, batch stacking."""
import numpy as np
import pytest
import scipy.sparse as sp

from ams_hip import gen
from ams_hip.batched import stack
from ams_hip.csr import Csr
from helpers import random_csr


@pytest.mark.parametrize("abc", [gen.G500, gen.ER])
def test_rmat_structure(abc):
    s, ef = 11, 8
    a = gen.rmat_symmetric(s, ef, abc, seed=3)
    n = 1 << s
    assert a.num_row == a.num_col == n
    assert a.num_nnz == 2 * ef * n                       # exactly ef * 2^s undirected edges, mirrored
    m = a.to_scipy()
    assert (m != m.T).nnz == 0                            # symmetric
    assert m.diagonal().sum() == 0                        # loop-free
    assert a.sorted_rows()                                # duplicate-free, ascending
    b = gen.rmat_symmetric(s, ef, abc, seed=3)
    np.testing.assert_array_equal(a.cols, b.cols)         # seeded
    c = gen.rmat_symmetric(s, ef, abc, seed=4)
    assert not np.array_equal(a.cols, c.cols)


def test_rmat_skew():
    g = gen.rmat_symmetric(12, 16, gen.G500, seed=1)
    e = gen.rmat_symmetric(12, 16, gen.ER, seed=1)
    assert np.diff(g.row_ptrs).max() > 4 * np.diff(e.row_ptrs).max()


def test_uniform_values_symmetric():
    a = gen.rmat_symmetric(9, 4, seed=2, values="uniform")
    m = a.to_scipy()
    assert abs(m - m.T).max() == 0
    assert np.all(np.abs(a.rats) <= 1)


def test_join_incidence_small():
    inc = gen.join_incidence(num_row=3000, num_col=500, cap=200, seed=5)
    deg = np.diff(inc.row_ptrs)
    assert deg.min() >= 1 and deg.max() <= 200
    assert inc.sorted_rows()                              # no repeated column inside a row
    colcnt = np.bincount(inc.cols, minlength=500)
    assert colcnt[0] > colcnt[100] > 0                    # Zipf head


def test_transpose_and_slice():
    a = random_csr(37, 53, 0.1, seed=8)
    t = a.transpose()
    assert (t.to_scipy() != a.to_scipy().T).nnz == 0
    assert t.sorted_rows()
    s = a.row_slice(5, 20)
    assert (s.to_scipy() != a.to_scipy()[5:20]).nnz == 0


def test_stack_is_block_diagonal_product():
    probs = [(random_csr(5, 7, 0.4, seed=i), random_csr(7, 9 + i, 0.3, seed=10 + i)) for i in range(4)]
    A, B, ao, bo, co = stack(probs)
    c = (A.to_scipy() @ B.to_scipy()).tocsr()
    want = sp.block_diag([(a.to_scipy() @ b.to_scipy()) for a, b in probs]).tocsr()
    assert abs(c - want).max() == 0
    assert list(ao) == [0, 5, 10, 15, 20] and list(bo) == [0, 7, 14, 21, 28]
    assert co[-1] == sum(9 + i for i in range(4))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dropin_pack_mmmv_matches_oracle(seed):
    """The drop-in's pack_ams_pack_mmmv (host bookkeeping inside
    libams_hip.so, no GPU involved) builds the oracle's morsel table."""
    import ctypes as C
    from ams_hip import _lib
    from oracle import oracle as orc

    class KeyVal(C.Structure):
        _fields_ = [("key", C.c_int64), ("val", C.c_double)]

    class Pack(C.Structure):
        _fields_ = [("b_roff_maxc", C.POINTER(KeyVal)), ("b_arr", C.POINTER(KeyVal)), ("total_cnt", C.c_int64),
                    ("high_head_idx", C.c_int64), ("high_tail_idx", C.c_int64), ("low_head_idx", C.c_int64),
                    ("low_tail_idx", C.c_int64)]

    class CsrT(C.Structure):
        _fields_ = [("num_row", C.c_int64), ("num_col", C.c_int64), ("num_nnz", C.c_int64),
                    ("row_ptrs", C.POINTER(C.c_int64)), ("cols", C.POINTER(C.c_int64)),
                    ("rats", C.POINTER(C.c_double))]

    L = C.CDLL(_lib.LIB_PATH)
    P64 = C.POINTER(C.c_int64)
    L.pack_ams_pack_mmmv.argtypes = [C.POINTER(CsrT), C.POINTER(Pack), P64, P64, C.c_int64, C.c_void_p]
    L.pack_ams_pack_mmmv.restype = C.c_int64
    L.free_ams_pack.argtypes = [C.POINTER(Pack)]
    L.ams_dropin_set_pack_params.argtypes = [C.c_int64, C.c_int64, C.c_double]
    rng = np.random.default_rng(seed)
    m = 500
    nz = rng.integers(0, 60, m).astype(np.int64)
    nz[rng.integers(0, m, 5)] = 400                       # rows longer than a morsel: empty morsels
    rp = np.concatenate([[0], np.cumsum(nz)]).astype(np.int64)
    coef = int(rng.integers(1, 4))
    ps = coef * rp
    csr = CsrT(m, 10, int(rp[-1]), rp.ctypes.data_as(P64), None, None)
    L.ams_dropin_set_pack_params(256, 64, 0.85)
    pk = Pack()
    mx = L.pack_ams_pack_mmmv(C.byref(csr), C.byref(pk), ps.ctypes.data_as(P64), nz.ctypes.data_as(P64), coef,
                              None)
    p = orc.params(threads=1, ams_min_num=64, ams_size=256, ams_prt=0.85)
    roff, arr, idx, omx = orc.pack_mmmv(rp, ps, nz, coef, p)
    assert mx == omx and pk.total_cnt == idx["total_cnt"]
    assert [(pk.b_roff_maxc[i].key, pk.b_roff_maxc[i].val) for i in range(pk.total_cnt + 1)] == roff
    assert [(pk.b_arr[i].key, pk.b_arr[i].val) for i in range(pk.total_cnt)] == arr
    assert (pk.low_head_idx, pk.high_head_idx, pk.high_tail_idx, pk.low_tail_idx) == (
        idx["low_head_idx"], idx["high_head_idx"], idx["high_tail_idx"], idx["low_tail_idx"])
    assert idx["low_head_idx"] > 0
    L.free_ams_pack(C.byref(pk))
    L.ams_dropin_set_pack_params(524288, 32768, 0.85)
