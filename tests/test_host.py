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
