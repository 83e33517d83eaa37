"""Join-aggregate inputs as binary incidence matrices, written as the
reference's ``<name>.org.mtx`` / ``<name>.trs.mtx`` pairs.

The reference's preparation (scripts/join-aggregate_preparation.sh:79-285)
turns (row entity, column entity) pairs from downloaded tables into binary
incidence matrices:

* row and column entities get dense ids in order of first appearance
  (``intern``, :101-106);
* a pair is kept once (an ``INSERT OR IGNORE`` into a (r, c) primary key,
  :95-99,119-123);
* ``org`` lists the entries sorted by (r, c) with the dimensions
  rows x columns, ``trs`` the transpose sorted by (c, r) with the dimensions
  swapped, both as ``coordinate real general`` MatrixMarket with a comment
  line and 1-based ``r c 1.0`` entries (:126-150);
* the runner multiplies org x trs (src/ams_hybaccspgemm.c:290-300).

The downloads (ClickBench, Book Genome, IMDb) need the network, which this
build does not have; ``prepare_standin`` writes the synthetic
join-aggregate-shaped incidence instead (850,000 x 20,000, lognormal row
degrees, Zipf columns).  Its product is C = A^T A, so the pair is written
with the column entities as rows (org = A^T, 20,000 x 850,000; trs = A):
the runner's org x trs is then that product.
"""
from __future__ import annotations

import os

import numpy as np

from .csr import Csr

BANNER = "%%MatrixMarket matrix coordinate real general\n"
COMMENT = "% binary incidence matrix prepared for contention-aware SpGEMM\n"


def intern(keys) -> tuple[np.ndarray, int]:
    """Dense ids in order of first appearance: the k-th distinct key seen
    gets id k."""
    keys = np.asarray(keys)
    if len(keys) == 0:
        return np.zeros(0, np.int64), 0
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    rank = np.empty(len(uniq), np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(uniq))
    return rank[inv.reshape(-1)], len(uniq)


def incidence(row_keys, col_keys) -> Csr:
    """Binary incidence matrix of (row entity, column entity) pairs: ids by
    first appearance, each pair once, columns ascending in every row."""
    r, nrows = intern(row_keys)
    c, ncols = intern(col_keys)
    return incidence_ids(r, c, nrows, ncols)


def incidence_ids(r: np.ndarray, c: np.ndarray, nrows: int, ncols: int) -> Csr:
    key = np.unique(np.asarray(r, np.int64) * ncols + np.asarray(c, np.int64))
    rows, cols = key // max(ncols, 1), key % max(ncols, 1)
    rp = np.zeros(nrows + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=nrows), out=rp[1:])
    return Csr(nrows, ncols, rp, cols.astype(np.int32), np.ones(len(key), np.float64))


def _write(path: str, nrows: int, ncols: int, rows: np.ndarray, cols: np.ndarray) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w", encoding="ascii", newline="\n") as f:
        f.write(BANNER)
        f.write(COMMENT)
        f.write(f"{nrows} {ncols} {len(rows)}\n")
        step = 1 << 20
        for s in range(0, len(rows), step):
            f.write("".join(f"{r} {c} 1.0\n" for r, c in zip((rows[s:s + step] + 1).tolist(),
                                                              (cols[s:s + step] + 1).tolist())))
    os.replace(tmp, path)


def export_pair(out_dir: str, name: str, m: Csr) -> tuple[str, str]:
    """Write ``<name>.org.mtx`` (m, entries by (r, c)) and ``<name>.trs.mtx``
    (m^T, entries by (c, r)) as the reference's preparation does."""
    if not m.num_row or not m.num_col or not m.num_nnz:
        raise RuntimeError(f"{name}: empty matrix ({m.num_row} rows, {m.num_col} cols, {m.num_nnz} entries)")
    rows = np.repeat(np.arange(m.num_row, dtype=np.int64), np.diff(np.asarray(m.row_ptrs)))
    cols = np.asarray(m.cols, np.int64)
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    org = os.path.join(out_dir, f"{name}.org.mtx")
    trs = os.path.join(out_dir, f"{name}.trs.mtx")
    _write(org, m.num_row, m.num_col, rows, cols)
    order = np.lexsort((rows, cols))
    _write(trs, m.num_col, m.num_row, cols[order], rows[order])
    return org, trs


def prepare_standin(out_dir: str, name: str = "join_agg", seed: int = 0, **shape) -> tuple[str, str]:
    """The offline join-aggregate pair: org = A^T and trs = A for the
    synthetic incidence A (keyword arguments go to gen.join_incidence)."""
    from .gen import join_incidence
    a = join_incidence(seed=seed, **shape)
    return export_pair(out_dir, name, a.transpose())
