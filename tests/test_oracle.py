"""Pins the CPU oracle (oracle/ams_oracle.c) before anything is compared with it.

The reference ships no tests, fixtures or golden outputs, so every oracle
function is checked against something independent of its own formula:
closed forms of the all-ones R-MAT products, scipy's SpGEMM, brute force on
tiny inputs, Python's bisect, numpy's stable sort, and hand-computed morsel
tables.  A plausible slip (a dropped product, a swapped operand, an
unstable sort, an off-by-one boundary) fails at least one of these.
"""
import bisect
import itertools
import math

import numpy as np
import pytest
import scipy.sparse as sp

from ams_hip import gen
from ams_hip.csr import Csr
from helpers import random_csr
from oracle import oracle as orc


def sp_product(a, b):
    c = (a.to_scipy() @ b.to_scipy()).tocsr()
    c.sort_indices()
    return c


# ------------------------------------------------------------ utilities

def test_next_pow2_is_smallest_power_not_below():
    xs = list(range(1, 5000)) + [2 ** k + d for k in range(12, 29) for d in (-1, 0, 1)]
    for x in xs:
        want = 1 << (x - 1).bit_length()
        assert orc.next_pow2(x) == want, x


def test_next_pow2_reference_quirk_at_2_29():
    # pow(2, ceil(log(x)/log(2))) overshoots at 2^29: log ratio rounds above 29.
    assert orc.next_pow2(2 ** 29) in (2 ** 29, 2 ** 30)


def test_lower_bound_matches_bisect_left():
    rng = np.random.default_rng(0)
    for _ in range(200):
        arr = np.sort(rng.integers(0, 50, rng.integers(0, 40)))
        for v in range(-2, 53):
            assert orc.lower_bound(arr, v) == bisect.bisect_left(list(arr), v)


def test_scan_matches_cumsum():
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 1000, (1 << 17) + 3):
        x = rng.integers(0, 1000, n)
        out = orc.scan(x, n)
        want = np.concatenate([[0], np.cumsum(x[:-1])])
        np.testing.assert_array_equal(out, want)


# ----------------------------------------------------------- row flops

def test_flops_closed_form_symmetric():
    a = gen.rmat_symmetric(12, 8, seed=5)
    deg = np.diff(a.row_ptrs)
    flp = orc.flp_populate(a, a, threads=4)
    np.testing.assert_array_equal(flp, a.to_scipy() @ deg.astype(np.float64))
    assert orc.get_intprod(a, a, threads=4) == int((deg.astype(np.int64) ** 2).sum())


def test_flops_rectangular():
    a = random_csr(50, 80, 0.1, seed=2)
    b = random_csr(80, 30, 0.2, seed=3)
    flp = orc.flp_populate(a, b)
    degb = np.diff(b.row_ptrs).astype(np.float64)
    np.testing.assert_array_equal(flp, a.to_scipy().astype(bool).astype(float) @ degb)


# ------------------------------------------------------------ symbolic

def test_symbolic_matches_scipy_and_brute_force():
    for seed in range(3):
        a = random_csr(40, 60, 0.08, seed=seed)
        b = random_csr(60, 500, 0.05, seed=seed + 10, sort=False)
        nz = orc.hash_symb(a, b, threads=3)
        brute = []
        for i in range(a.num_row):
            cols = set()
            for p in range(a.row_ptrs[i], a.row_ptrs[i + 1]):
                k = a.cols[p]
                cols.update(b.cols[b.row_ptrs[k]:b.row_ptrs[k + 1]].tolist())
            brute.append(len(cols))
        np.testing.assert_array_equal(nz, brute)
        np.testing.assert_array_equal(nz, np.diff(sp_product(a, b).indptr))


# --------------------------------------------------------------- pack

def test_pack_hand_computed():
    # T = MIN(AMS_MIN_NUM, AMS_SIZE/8) = 4; ps = prefix of nnz per row.
    row_nz = np.array([3, 0, 2, 9, 1, 1], np.int64)
    ps = np.concatenate([[0], np.cumsum(row_nz)])       # [0,3,3,5,14,15,16]
    flp = np.array([6, 0, 2, 20, 1, 3], np.int64)
    p = orc.params(threads=1, ams_min_num=4, ams_size=64, ams_prt=0.5)
    roff, arr, idx, mx = orc.pack(ps, flp, p)
    # total = 16 // 4 = 4; boundaries lower_bound(ps, 4k) = 3, 4, 4; last = 6
    assert idx["total_cnt"] == 4
    assert [r[0] for r in roff] == [0, 3, 4, 4, 6]
    # morsel 2 = rows [4, 4) is empty -> invalid, cf 0, sorted first
    cfs = {0: 8 / 5, 1: 20 / 9, 2: 0.0, 3: 4 / 2}
    assert [a[0] for a in arr] == [2, 0, 3, 1]
    for k, v in arr:
        assert v == pytest.approx(cfs[k])
    assert [r[1] for r in roff[:4]] == [6, 20, 0, 3]
    assert mx == 20
    assert idx["low_head_idx"] == 1
    assert idx["high_head_idx"] == 3
    assert idx["high_tail_idx"] == 2      # int((4 - 1) * 0.5 + 1)
    assert idx["low_tail_idx"] == 1


def test_pack_tiny_input_guard():
    ps = np.array([0, 1, 2], np.int64)
    roff, arr, idx, mx = orc.pack(ps, np.array([1, 1], np.int64), orc.params(threads=1))
    assert idx["total_cnt"] == 1 and [r[0] for r in roff] == [0, 2]


def test_pack_mmmv_hand_computed():
    # rows with 2, 0, 5, 1, 4 nonzeros: row_ptrs [0, 2, 2, 7, 8, 12];
    # coef 2 and ps_row_coef = 2 * row_ptrs; T = MIN(4, 64/8) = 4
    rp = np.array([0, 2, 2, 7, 8, 12], np.int64)
    ps = 2 * rp                                           # [0, 4, 4, 14, 16, 24]
    nz = np.array([2, 0, 5, 1, 4], np.int64)
    p = orc.params(threads=1, ams_min_num=4, ams_size=64, ams_prt=0.5)
    roff, arr, idx, mx = orc.pack_mmmv(rp, ps, nz, 2, p)
    # total = 12 * 2 // 4 = 6; boundaries lower_bound(ps, 4k), k = 1..5:
    # 4 -> 1, 8 -> 3, 12 -> 3, 16 -> 4, 20 -> 5; last = 5
    assert idx["total_cnt"] == 6
    assert [r[0] for r in roff] == [0, 1, 3, 3, 4, 5, 5]
    # values: mean nonzeros per row of each morsel; [3, 3) and [5, 5) empty
    vals = {0: 2 / 1, 1: 5 / 2, 2: 0.0, 3: 1 / 1, 4: 4 / 1, 5: 0.0}
    assert [a[0] for a in arr] == [2, 5, 3, 0, 1, 4]
    for k, v in arr:
        assert v == pytest.approx(vals[k])
    assert [r[1] for r in roff[:6]] == [2, 5, 0, 1, 4, 0]
    assert mx == 5
    assert idx["low_head_idx"] == 2
    assert idx["high_head_idx"] == 5
    assert idx["high_tail_idx"] == 4      # int((6 - 2) * 0.5 + 2)
    assert idx["low_tail_idx"] == 3


# ------------------------------------------------------------ radix sort

def test_radix_sort_is_stable_lsd():
    rng = np.random.default_rng(4)
    for n in (0, 1, 2, 17, 300):
        for hi in (1, 7, 300, 70000, 2 ** 40):
            keys = rng.integers(0, hi, n)
            vals = rng.random(n)
            k2, v2 = orc.radix_sort_keyval(keys, vals)
            order = np.argsort(keys, kind="stable")
            np.testing.assert_array_equal(k2, keys[order])
            np.testing.assert_array_equal(v2, vals[order])


# ------------------------------------------------------------- numeric

@pytest.mark.parametrize("mode", ["hybacc", "hash", "esc"])
def test_all_ones_invariants(mode):
    a = gen.rmat_symmetric(11, 8, seed=6)
    rp, cols, rats, _ = orc.spgemm(a, a, mode, threads=4)
    deg = np.diff(a.row_ptrs)
    assert rats.sum() == float((deg.astype(np.int64) ** 2).sum())   # sum C = P
    c = sp.csr_matrix((rats, cols, rp), shape=(a.num_row, a.num_row))
    np.testing.assert_array_equal(c.diagonal(), deg)                # C_ii = deg(i)
    assert (c != c.T).nnz == 0                                      # symmetric
    assert rats.max() <= deg.max()
    ref = sp_product(a, a)
    np.testing.assert_array_equal(rp, ref.indptr)
    np.testing.assert_array_equal(cols, ref.indices)
    np.testing.assert_array_equal(rats, ref.data)


def brute_sequential(a, b):
    """Dense, loop-order product: C[i,j] = (((p1 + p2) + p3) ...) over A(i,:)
    in storage order, the first product assigned."""
    out = {}
    for i in range(a.num_row):
        for p in range(a.row_ptrs[i], a.row_ptrs[i + 1]):
            k = a.cols[p]
            for q in range(b.row_ptrs[k], b.row_ptrs[k + 1]):
                j = int(b.cols[q])
                prod = float(a.rats[p]) * float(b.rats[q])
                out[(i, j)] = prod if (i, j) not in out else out[(i, j)] + prod
    return out


@pytest.mark.parametrize("mode", ["hybacc", "hash", "esc"])
def test_random_values_bitwise_sequential_order(mode):
    for seed in range(3):
        a = random_csr(25, 40, 0.15, seed=seed)
        b = random_csr(40, 30, 0.2, seed=seed + 5, sort=False)
        rp, cols, rats, _ = orc.spgemm(a, b, mode, threads=2)
        want = brute_sequential(a, b)
        got = {}
        for i in range(a.num_row):
            for q in range(rp[i], rp[i + 1]):
                got[(i, int(cols[q]))] = rats[q]
        assert got.keys() == want.keys()
        for key, v in want.items():
            assert got[key] == v, key   # bit-exact, not approximately


def test_special_cases():
    b = random_csr(30, 45, 0.2, seed=9)
    eye = Csr(30, 30, np.arange(31, dtype=np.int64), np.arange(30, dtype=np.int32), np.ones(30))
    rp, cols, rats, _ = orc.spgemm(eye, b)
    np.testing.assert_array_equal(rp, b.row_ptrs)
    np.testing.assert_array_equal(cols, b.cols)
    np.testing.assert_array_equal(rats, b.rats)
    d = Csr(30, 30, np.arange(31, dtype=np.int64), np.arange(30, dtype=np.int32), np.arange(2.0, 32.0))
    rp, cols, rats, _ = orc.spgemm(d, b)
    rows = np.repeat(np.arange(30), np.diff(b.row_ptrs))
    np.testing.assert_array_equal(rats, b.rats * (rows + 2.0))
    # n = 1: dot products
    x, y = random_csr(12, 20, 0.4, seed=1), random_csr(20, 1, 0.7, seed=2)
    rp, cols, rats, _ = orc.spgemm(x, y)
    dense = x.to_scipy().toarray() @ y.to_scipy().toarray()
    nzr = np.nonzero(np.diff(rp))[0]
    np.testing.assert_allclose(rats, dense[nzr, 0], rtol=1e-14, atol=1e-15)
    # empty A
    z = Csr(4, 20, np.zeros(5, np.int64), np.zeros(0, np.int32), np.zeros(0))
    rp, cols, rats, _ = orc.spgemm(z, y)
    assert len(cols) == 0 and (rp == 0).all()


def test_thread_count_does_not_change_results():
    a = gen.rmat_symmetric(12, 16, seed=8, values="uniform")
    ref = orc.spgemm(a, a, threads=1)
    for t in (2, 3, 8):
        got = orc.spgemm(a, a, threads=t)
        for x, y in zip(ref[:3], got[:3]):
            np.testing.assert_array_equal(x, y)


def test_raw_hash_rows_are_in_slot_order():
    """Without canonicalisation hash rows keep the reference's slot order.
    Rows whose keys all have distinct home slots (key * 2654435761) mod size
    must come out in ascending home slot; every row is a permutation of the
    canonical row."""
    a = random_csr(400, 30, 0.05, seed=3, values="ones")
    b = random_csr(30, 4000, 0.001, seed=4, values="ones")
    rp, cols, rats, _ = orc.spgemm(a, b, "hash", canonical=False, p=orc.params(threads=1))
    rp2, cols2, _, _ = orc.spgemm(a, b, "hash", canonical=True, p=orc.params(threads=1))
    flp = orc.flp_populate(a, b)
    checked = 0
    for i in range(a.num_row):
        row = cols[rp[i]:rp[i + 1]].astype(np.int64)
        assert sorted(row.tolist()) == cols2[rp2[i]:rp2[i + 1]].tolist()
        if len(row) < 2:
            continue
        size = orc.next_pow2(min(flp[i], b.num_col))
        home = (row * 2654435761) & (size - 1)
        if len(set(home.tolist())) == len(home):
            assert (np.diff(home) > 0).all()
            checked += 1
    assert checked > 0
