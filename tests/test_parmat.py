"""PaRMAT-construction generator (ams_hip/parmat.py).  PaRMAT seeds from
std::random_device and time(0), so no output can be matched bit for bit;
these checks pin its construction instead:

* the integer closed form of genEdgeIndex_FP against a literal
  double-precision replay of its halving loop, for every choice sequence;
* the cut lines (row a + c, column a + b) and the independence of the two
  coordinates through bit marginals of a large sample;
* hand-worked splits (inclusive shared boundaries, truncated shares, the b
  part folded into c on the diagonal) and the square plan of the reference's
  dataset flags;
* output invariants (edge count, undirected loop-free rows < columns,
  sorted columns of squares) and the rmat_ones matrix of the text output."""
import os

import numpy as np
import pytest

from ams_hip import mtx, parmat


def literal_halving(lo, hi, bits):
    """PaRMAT/src/utils.cpp:101-121 in doubles, with the draws replaced by
    the given choices (True keeps the upper half)."""
    s, e = float(lo), float(hi)
    it = iter(bits)
    while (e - s) >= 1.0:
        cut_index = (e + s) / 2.0
        if next(it):
            s = cut_index
        else:
            e = cut_index
    assert next(it, None) is None
    return int(np.floor((s + e) / 2.0 + 0.5))


@pytest.mark.parametrize("lo,hi", [(0, 15), (7, 15), (0, 7), (0, 1), (5, 5), (3, 10), (1023, 2047),
                                   (0, (1 << 12) - 1), (100, 1 << 11)])
def test_index_closed_form_matches_literal_loop(lo, hi):
    k = parmat.halvings(lo, hi)
    js = np.arange(1 << k, dtype=np.int64) if k <= 12 else np.random.default_rng(0).integers(0, 1 << k, 4096)
    got = parmat.index_from_halvings(lo, hi, js)
    want = [literal_halving(lo, hi, [(int(j) >> (k - 1 - b)) & 1 for b in range(k)]) for j in js]
    np.testing.assert_array_equal(got, want)
    assert got.min() >= lo and got.max() <= max(hi, lo)


def test_full_range_index_is_the_choice_bits():
    # [0, 2^s - 1]: the rounded midpoint is exactly the integer of the choices
    j = np.arange(1 << 10)
    np.testing.assert_array_equal(parmat.index_from_halvings(0, (1 << 10) - 1, j), j)


@pytest.mark.parametrize("cut", [0.0, 0.25, 0.5, 0.6, 0.76, 1.0, 1 / 3])
def test_below_count(cut):
    r = parmat.RAND_MAX
    k = parmat.below_count(cut)
    assert k == 0 or (k - 1) / r < cut
    assert k == r + 1 or k / r >= cut


def test_cut_lines_and_independence():
    # one square (threads 1), directed, loops and duplicates kept: rows and
    # columns are independent, bit b of a row is 1 with probability 1 - (a + c)
    a, b, c = 0.45, 0.15, 0.30
    n = 60000
    rows, cols = parmat.parmat_edges(256, n, a, b, c, seed=4, threads=1, directed=True, self_loops=True,
                                     duplicates=True)
    tol = 6 * np.sqrt(0.25 / n)
    for bit in range(8):
        rb, cb = (rows >> bit) & 1, (cols >> bit) & 1
        assert abs(rb.mean() - (1 - (a + c))) < tol
        assert abs(cb.mean() - (1 - (a + b))) < tol
        assert abs((rb & cb).mean() - (1 - (a + c)) * (1 - (a + b))) < tol


def sq_tuple(s):
    return (s.x0, s.x1, s.y0, s.y1, s.n, s.h, s.v)


def test_shatter_hand_worked():
    sq = parmat.Square(0, 15, 0, 15, 1000)
    zero = np.zeros(6)
    # 0.57 * 1000 -> 570, 0.19 * 1000 -> 190 twice, d takes the remaining 50;
    # the midpoint 7 belongs to both halves
    assert [sq_tuple(p) for p in parmat.shatter(sq, 0.57, 0.19, 0.19, zero, directed=True)] == [
        (0, 7, 0, 7, 570, 0, 0), (7, 15, 0, 7, 190, 1, 0), (0, 7, 7, 15, 190, 0, 1), (7, 15, 7, 15, 50, 1, 1)]
    # undirected diagonal square: b is dropped, its edges go to c
    assert [sq_tuple(p) for p in parmat.shatter(sq, 0.57, 0.19, 0.19, zero, directed=False)] == [
        (0, 7, 0, 7, 570, 0, 0), (0, 7, 7, 15, 380, 0, 1), (7, 15, 7, 15, 50, 1, 1)]
    # share noise enters the budgets, truncated: 0.57 + 0.001 -> 571
    noisy = np.array([0, 0, 0, 0.001, 0, 0])
    assert [p.n for p in parmat.shatter(sq, 0.57, 0.19, 0.19, noisy, directed=False)] == [571, 380, 49]
    # off the diagonal all four parts stay
    off = parmat.Square(0, 7, 7, 15, 400, 0, 1)
    assert len(parmat.shatter(off, 0.57, 0.19, 0.19, zero, directed=False)) == 4


def test_square_plan():
    zero = np.zeros(6)
    sqs = parmat.plan_squares(1 << 12, 8 << 12, 0.57, 0.19, 0.19, zero, threads=8)
    # three rounds of splits: the 36 squares on or above the diagonal of an 8 x 8 grid
    assert len(sqs) == 36 and sorted({s.h for s in sqs}) == list(range(8))
    assert all(s.h <= s.v for s in sqs) and sum(s.n for s in sqs) == 8 << 12
    assert [(s.h, s.v) for s in sqs] == sorted((s.h, s.v) for s in sqs)
    # fewer than 10000 edges: one worker, no split
    assert len(parmat.plan_squares(1 << 12, 9999, 0.57, 0.19, 0.19, zero, threads=8)) == 1
    # too dense to split (3 * 10000 >= 127 * 127)
    assert len(parmat.plan_squares(128, 10000, 0.57, 0.19, 0.19, zero, threads=8)) == 1
    # a column capacity forces splits before the thread count does
    sqs = parmat.plan_squares(1 << 12, 8 << 12, 0.57, 0.19, 0.19, zero, threads=1, capacity=4000)
    run = {}
    for s in sqs:
        run[s.x1] = run.get(s.x1, 0) + s.n
    assert max(run.values()) <= 4000


def test_output_invariants():
    n = 1 << 12
    rows, cols = parmat.parmat_edges(n, 8 * n, 0.57, 0.19, 0.19, seed=9)
    assert len(rows) == 8 * n
    assert rows.min() >= 0 and cols.max() <= n - 1 and np.all(rows < cols)
    keys = rows * n + cols
    # distinct inside every square; squares share only boundary vertices
    assert len(keys) - len(np.unique(keys)) <= len(keys) // 1000
    # sorted inside each of the 8 columns of squares
    assert np.count_nonzero(np.diff(keys) < 0) <= 7


def test_rmat_ones_of_text_output(tmp_path):
    scale, ef = 10, 8
    n = 1 << scale
    rows, cols = parmat.parmat_edges(n, ef * n, 0.57, 0.19, 0.19, seed=3)
    src = os.path.join(tmp_path, f"g500_{scale}_{ef}.txt")
    parmat.write_txt(src, rows, cols)
    org, trs = os.path.join(tmp_path, "a.org.mtx"), os.path.join(tmp_path, "a.trs.mtx")
    mtx.rmat_ones(src, org, trs)
    got = mtx.read_mtx(org)
    want = parmat.parmat_symmetric(scale, ef, (0.57, 0.19, 0.19), seed=3)
    np.testing.assert_array_equal(got.row_ptrs, want.row_ptrs)
    np.testing.assert_array_equal(got.cols, want.cols)
    np.testing.assert_array_equal(got.rats, want.rats)


@pytest.mark.gpu
def test_gpu_product_of_parmat_matrix(gpu_ctx):
    """The GPU product of a PaRMAT-construction matrix against the oracle
    (all-ones: values bit-exact); squares sharing boundary vertices can
    leave duplicate entries, which both keep."""
    from ams_hip.device import DeviceCsr
    from helpers import assert_matches_oracle
    a = parmat.parmat_symmetric(12, 16, (0.57, 0.19, 0.19), seed=11)
    da = DeviceCsr.from_host(a)
    c = gpu_ctx.spgemm(da, da).to_host()
    assert_matches_oracle(c, a, a, exact=True)
