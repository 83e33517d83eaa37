"""MatrixMarket / PaRMAT text I/O (include/ams_dropin.h: mtx2csr, trpl2mtx,
txt_transpose) against the oracle restatement of the reference's reader and
writer (oracle/mtx.py), which is itself pinned here by hand-worked files."""
import os

import numpy as np
import pytest

from oracle import mtx as omtx

from ams_hip import gen, mtx
from ams_hip.csr import Csr
from helpers import random_csr

SYM = """%%MatrixMarket matrix coordinate real symmetric
% a comment
3 3 4
1 1 2.5
2 1 1
3 2 0
3 1 -4
"""

PATTERN = """%%MatrixMarket matrix coordinate pattern general
2 4 4
2 4
1 3
2 4
1 1
"""

COMMENT_SYM = """%%MatrixMarket matrix coordinate real general
% this matrix is symmetric
2 2 1
2 1 7.0
"""


def write(tmp_path, name, text):
    p = os.path.join(tmp_path, name)
    with open(p, "w") as f:
        f.write(text)
    return p


def as_tuple(c):
    return (c.num_row, c.num_col, list(c.row_ptrs), list(np.asarray(c.cols, np.int64)), list(c.rats))


@pytest.mark.parametrize("text,want", [
    # diagonal kept once, (2,1) mirrored, the 0 entry dropped, -4 forced to 1
    (SYM, (3, 3, [0, 3, 4, 5], [0, 1, 2, 0, 0], [1.0] * 5)),
    # pattern lines are ones; duplicates kept; sorted by (row, column)
    (PATTERN, (2, 4, [0, 2, 4], [0, 2, 3, 3], [1.0] * 4)),
    # any comment containing "symmetric" makes the file symmetric
    (COMMENT_SYM, (2, 2, [0, 1, 2], [1, 0], [1.0] * 2)),
])
def test_hand_worked_files(tmp_path, text, want):
    m, n, rp, cols, vals, _ = omtx.read_mtx(text)
    assert (m, n, list(rp), list(cols), list(vals)) == want  # pins the oracle
    assert as_tuple(mtx.read_mtx(write(tmp_path, "a.mtx", text))) == want


def test_reader_matches_oracle_on_random_files(tmp_path):
    rng = np.random.default_rng(4)
    for trial in range(6):
        m, n = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        sym = bool(trial % 2) and m == n or trial == 5
        if trial == 5:
            n = m
        z = int(rng.integers(0, 200))
        lines = ["%%MatrixMarket matrix coordinate real " + ("symmetric" if sym else "general"),
                 f"{m} {n} {z}"]
        for _ in range(z):
            i, j = int(rng.integers(1, m + 1)), int(rng.integers(1, n + 1))
            v = float(rng.choice([0.0, 1.0, -2.5, 3e-300]))
            lines.append(f"{i}  {j}\t{v!r}" if rng.random() < 0.7 else f"{i} {j}")
        text = "\n".join(lines) + "\n"
        om, on, orp, ocols, ovals, _ = omtx.read_mtx(text)
        c = mtx.read_mtx(write(tmp_path, f"r{trial}.mtx", text))
        assert (c.num_row, c.num_col) == (om, on)
        np.testing.assert_array_equal(c.row_ptrs, orp)
        np.testing.assert_array_equal(np.asarray(c.cols, np.int64), ocols)
        np.testing.assert_array_equal(c.rats, ovals)


@pytest.mark.parametrize("symmetric", [False, True])
def test_writer_text_matches_oracle_and_round_trips(tmp_path, symmetric):
    a = gen.rmat_symmetric(7, 4, seed=3) if symmetric else random_csr(40, 70, 0.1, seed=5)
    p = os.path.join(tmp_path, "w.mtx")
    mtx.write_mtx(p, a, symmetric=symmetric)
    text = open(p).read()
    assert text == omtx.write_mtx(a.num_row, a.num_col, a.row_ptrs, a.cols, symmetric)
    back = mtx.read_mtx(p)
    np.testing.assert_array_equal(back.row_ptrs, a.row_ptrs)
    np.testing.assert_array_equal(back.cols, a.cols)
    assert np.all(back.rats == 1.0)


def test_rmat_ones_pair(tmp_path):
    """PaRMAT text -> symmetric .org.mtx and .trs.mtx (src/rmat_ones.c)."""
    scale, ef = 6, 4
    lo, hi = gen.rmat_undirected_edges(scale, ef, seed=9)
    rng = np.random.default_rng(1)
    flip = rng.random(len(lo)) < 0.5  # PaRMAT writes either orientation
    src, dst = np.where(flip, hi, lo), np.where(flip, lo, hi)
    txt = write(tmp_path, f"g500_{scale}_{ef}.txt", "".join(f"{s}\t{d}\n" for s, d in zip(src, dst)))
    org, trs = os.path.join(tmp_path, "g.org.mtx"), os.path.join(tmp_path, "g.trs.mtx")
    mtx.rmat_ones(txt, org, trs)
    n, _, rp, cols = omtx.read_parmat(open(txt).read(), scale)
    for path in (org, trs):
        c = mtx.read_mtx(path)
        assert c.num_row == c.num_col == n
        np.testing.assert_array_equal(c.row_ptrs, rp)
        np.testing.assert_array_equal(np.asarray(c.cols, np.int64), cols)
    # the generator's own symmetric matrix is the same structure
    a = gen.rmat_symmetric(scale, ef, seed=9)
    np.testing.assert_array_equal(a.row_ptrs, rp)
    np.testing.assert_array_equal(np.asarray(a.cols, np.int64), cols)


def test_missing_file_raises(tmp_path):
    with pytest.raises(FileNotFoundError):
        mtx.read_mtx(os.path.join(tmp_path, "absent.mtx"))


def test_raw2csr_is_mtx2csr(tmp_path):
    """raw2csr, the name the reference's drivers call (src/mm/mm_utils.c:
    730-732), reads the same CSR as mtx2csr."""
    import ctypes as C
    lib = mtx._lib()
    lib.raw2csr.argtypes = [C.POINTER(mtx.CsrT), C.c_char_p, C.c_void_p]
    lib.raw2csr.restype = None
    p = write(tmp_path, "sym.mtx", SYM)
    c = mtx.CsrT()
    lib.raw2csr(C.byref(c), os.fsencode(p), None)
    try:
        got = (c.num_row, c.num_col, [c.row_ptrs[i] for i in range(c.num_row + 1)],
               [c.cols[i] for i in range(c.num_nnz)], [c.rats[i] for i in range(c.num_nnz)])
    finally:
        lib.free_csr(C.byref(c))
    assert got == as_tuple(mtx.read_mtx(p))
