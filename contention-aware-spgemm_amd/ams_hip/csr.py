"""Host-side CSR container.

The reference stores a matrix as ``csr_t`` (src/inc/store.h:49-56): int64
``row_ptrs`` (num_row + 1), int64 ``cols`` (num_nnz) and double ``rats``.
On the device this library keeps int64 row pointers (nnz(C) can pass 2^31),
int32 column indices and FP64 or FP32 values; this class is the numpy mirror
of that layout.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Csr:
    num_row: int
    num_col: int
    row_ptrs: np.ndarray  # int64, num_row + 1
    cols: np.ndarray      # int32 (device layout) or int64 (reference layout)
    rats: np.ndarray      # float64 or float32

    @property
    def num_nnz(self) -> int:
        return int(self.row_ptrs[-1]) if len(self.row_ptrs) else 0

    def astype(self, col_dtype=np.int32, val_dtype=None) -> "Csr":
        vd = self.rats.dtype if val_dtype is None else val_dtype
        return Csr(self.num_row, self.num_col,
                   np.ascontiguousarray(self.row_ptrs, dtype=np.int64),
                   np.ascontiguousarray(self.cols, dtype=col_dtype),
                   np.ascontiguousarray(self.rats, dtype=vd))

    def row_slice(self, r0: int, r1: int) -> "Csr":
        """Rows [r0, r1) as a new matrix with rebased row pointers."""
        s, e = int(self.row_ptrs[r0]), int(self.row_ptrs[r1])
        return Csr(r1 - r0, self.num_col, self.row_ptrs[r0:r1 + 1] - s,
                   self.cols[s:e], self.rats[s:e])

    def transpose(self) -> "Csr":
        """Exact transpose; columns of each output row come out ascending."""
        m, n = self.num_row, self.num_col
        rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(self.row_ptrs))
        order = np.lexsort((rows, self.cols.astype(np.int64)))
        cnt = np.bincount(self.cols.astype(np.int64), minlength=n)
        rp = np.zeros(n + 1, np.int64)
        np.cumsum(cnt, out=rp[1:])
        return Csr(n, m, rp, rows[order].astype(self.cols.dtype), self.rats[order])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.rats, self.cols, self.row_ptrs),
                             shape=(self.num_row, self.num_col))

    @staticmethod
    def from_scipy(mat, col_dtype=np.int32, val_dtype=np.float64) -> "Csr":
        mat = mat.tocsr()
        mat.sort_indices()
        return Csr(mat.shape[0], mat.shape[1],
                   mat.indptr.astype(np.int64), mat.indices.astype(col_dtype),
                   mat.data.astype(val_dtype))

    def sorted_rows(self) -> bool:
        """True when every row's columns are strictly ascending."""
        if self.num_nnz < 2:
            return True
        d = np.diff(self.cols.astype(np.int64))
        starts = self.row_ptrs[1:-1]
        starts = starts[(starts > 0) & (starts < self.num_nnz)]
        ok = d > 0
        ok[starts - 1] = True  # row boundaries may step down
        return bool(ok.all())
