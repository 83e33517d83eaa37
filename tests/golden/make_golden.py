"""Writes the golden fixtures of this directory (run from the repo root:
python tests/golden/make_golden.py).

The reference ships no golden outputs, so these come from scipy's own
SpGEMM (an independent implementation whose row products accumulate in the
same A-row order) on small seeded inputs:
  g500_s9_ef4_ones   symmetric skewed R-MAT, all-ones, C = A * A
  er_s8_ef4_unif     symmetric uniform R-MAT, U[-1,1] values, C = A * A
  rect_unsorted      40 x 120 times 120 x 900 random, B rows unsorted
Each .npz holds the inputs and C's row pointers, sorted columns and values.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "contention-aware-spgemm_amd"))
sys.path.insert(0, os.path.join(HERE, ".."))

from ams_hip import gen  # noqa: E402
from helpers import random_csr  # noqa: E402


def product(a, b):
    c = (a.to_scipy() @ b.to_scipy()).tocsr()
    c.sort_indices()
    return c


def save(name, a, b):
    c = product(a, b)
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        a_shape=np.array([a.num_row, a.num_col]), a_rp=a.row_ptrs, a_col=a.cols, a_val=a.rats,
                        b_shape=np.array([b.num_row, b.num_col]), b_rp=b.row_ptrs, b_col=b.cols, b_val=b.rats,
                        c_rp=c.indptr.astype(np.int64), c_col=c.indices.astype(np.int32), c_val=c.data)


if __name__ == "__main__":
    a = gen.rmat_symmetric(9, 4, seed=101)
    save("g500_s9_ef4_ones", a, a)
    a = gen.rmat_symmetric(8, 4, gen.ER, seed=102, values="uniform")
    save("er_s8_ef4_unif", a, a)
    save("rect_unsorted", random_csr(40, 120, 0.05, seed=103), random_csr(120, 900, 0.02, seed=104, sort=False))
