"""Join-aggregate preparation (ams_hip/joinagg.py): ids by first
appearance, pairs kept once, and the org/trs MatrixMarket pair in the
reference's text format (scripts/join-aggregate_preparation.sh:95-150),
checked on hand-worked pairs, through the library's own reader, and, on the
GPU, the product org x trs of the offline stand-in against the oracle."""
import os

import numpy as np
import pytest

from ams_hip import gen, joinagg, mtx

ROWS = ["u9", "u2", "u9", "u5", "u2"]
COLS = ["x", "y", "x", "z", "x"]
# u9 -> 0, u2 -> 1, u5 -> 2; x -> 0, y -> 1, z -> 2; (u9, x) twice
ORG = ("%%MatrixMarket matrix coordinate real general\n"
       "% binary incidence matrix prepared for contention-aware SpGEMM\n"
       "3 3 4\n1 1 1.0\n2 1 1.0\n2 2 1.0\n3 3 1.0\n")
TRS = ("%%MatrixMarket matrix coordinate real general\n"
       "% binary incidence matrix prepared for contention-aware SpGEMM\n"
       "3 3 4\n1 1 1.0\n1 2 1.0\n2 2 1.0\n3 3 1.0\n")


def test_intern_first_appearance():
    ids, n = joinagg.intern(["b", "a", "b", "c", "a"])
    assert n == 3 and ids.tolist() == [0, 1, 0, 2, 1]
    ids, n = joinagg.intern([])
    assert n == 0 and len(ids) == 0


def test_hand_worked_pair(tmp_path):
    m = joinagg.incidence(ROWS, COLS)
    assert (m.num_row, m.num_col) == (3, 3)
    assert m.row_ptrs.tolist() == [0, 1, 3, 4] and m.cols.tolist() == [0, 0, 1, 2]
    org, trs = joinagg.export_pair(str(tmp_path), "toy", m)
    assert open(org).read() == ORG
    assert open(trs).read() == TRS


def test_reader_round_trip_rectangular(tmp_path):
    rng = np.random.default_rng(5)
    r = rng.integers(0, 50, 400)
    c = rng.zipf(1.5, 400) % 37
    m = joinagg.incidence(r, c)
    org, trs = joinagg.export_pair(str(tmp_path), "rect", m)
    back = mtx.read_mtx(org)
    assert (back.num_row, back.num_col) == (m.num_row, m.num_col)
    np.testing.assert_array_equal(back.row_ptrs, m.row_ptrs)
    np.testing.assert_array_equal(back.cols, m.cols)
    t = mtx.read_mtx(trs)
    mt = m.transpose()
    np.testing.assert_array_equal(t.row_ptrs, mt.row_ptrs)
    np.testing.assert_array_equal(t.cols, mt.cols)


def test_empty_matrix_raises(tmp_path):
    m = joinagg.incidence_ids(np.zeros(0, np.int64), np.zeros(0, np.int64), 2, 2)
    with pytest.raises(RuntimeError):
        joinagg.export_pair(str(tmp_path), "empty", m)


@pytest.mark.gpu
def test_standin_pair_product(tmp_path, gpu_ctx):
    """org x trs of the stand-in pair is the join-aggregate product A^T A."""
    from ams_hip.device import DeviceCsr
    from helpers import assert_matches_oracle
    shape = dict(num_row=3000, num_col=400, cap=200)
    org, trs = joinagg.prepare_standin(str(tmp_path), "ja", seed=7, **shape)
    a = gen.join_incidence(seed=7, **shape)
    ao, at = mtx.read_mtx(org), mtx.read_mtx(trs)
    np.testing.assert_array_equal(ao.cols, a.transpose().cols)
    np.testing.assert_array_equal(at.cols, a.cols)
    c = gpu_ctx.spgemm(DeviceCsr.from_host(ao), DeviceCsr.from_host(at)).to_host()
    assert (c.num_row, c.num_col) == (400, 400)
    assert_matches_oracle(c, ao, at, exact=True)
