"""Oracle restatement of the reference's MatrixMarket reading and writing
(src/mm/mm_utils.c:451-604 mtx2trpl, :607-642 trpl2mtx, :645-682 trpl2csr,
:757-835 txt2trpl).  Test infrastructure only: it checks the library's
reader and writer and is never on the product path.

Pinned in tests/test_mtx.py by hand-written files whose CSR is worked out
in the test itself (symmetric mirroring without doubling the diagonal,
comment-triggered symmetry, pattern lines, zero skip, forced ones, sorting,
kept duplicates).
"""
import numpy as np


def read_mtx(text: str):
    """mtx2trpl + trpl2csr on the file contents: (num_row, num_col,
    row_ptrs int64, cols int64, vals float64, symmetric)."""
    symmetric = False
    header = None
    rows, cols = [], []
    for line in text.split("\n"):
        if line.startswith("%"):  # mm_utils.c:510-515
            if "symmetric" in line:
                symmetric = True
            continue
        if header is None:  # :518-527
            f = line.split()
            header = (int(f[0]), int(f[1]), int(f[2]))
            continue
        if line == "":
            continue  # the reference's buffer ends at the last newline
        f = line.split()
        v = float(f[2]) if len(f) >= 3 else 1.0  # :545-553
        if v == 0.0:  # :556-558
            continue
        i, j = int(f[0]) - 1, int(f[1]) - 1  # :560-563
        rows.append(i)
        cols.append(j)
        if symmetric and i != j:  # :566-573
            rows.append(j)
            cols.append(i)
    m, n = header[0], header[1]
    r = np.asarray(rows, np.int64)
    c = np.asarray(cols, np.int64)
    order = np.lexsort((c, r))  # compare_triples, :47-55
    r, c = r[order], c[order]
    rp = np.zeros(m + 1, np.int64)
    np.add.at(rp, r + 1, 1)  # trpl2csr, :668-677
    rp = np.cumsum(rp)
    return m, n, rp, c, np.ones(len(c)), symmetric


def write_mtx(m, n, row_ptrs, cols, symmetric):
    """trpl2mtx on a CSR's entries, as text."""
    rows = np.repeat(np.arange(m), np.diff(row_ptrs))
    out = []
    if symmetric:
        keep = rows <= np.asarray(cols)
        out.append("%%MatrixMarket matrix coordinate real symmetric")
        out.append(f"{m} {n} {int(keep.sum())}")
        for i, j in zip(rows[keep], np.asarray(cols)[keep]):
            out.append(f"{i + 1} {j + 1} 1.000000")
    else:
        out.append("%%MatrixMarket matrix coordinate real general")
        out.append(f"{m} {n} {len(cols)}")
        for i, j in zip(rows, cols):
            out.append(f"{i + 1} {j + 1} 1.000000")
    return "\n".join(out) + "\n"


def read_parmat(text: str, scale: int):
    """txt2trpl: 0-based "src dst" edges, each mirrored, ones, sorted."""
    rows, cols = [], []
    for line in text.split("\n"):
        if not line.strip():
            continue
        f = line.split()
        v = float(f[2]) if len(f) >= 3 else 1.0
        if v == 0.0:
            continue
        i, j = int(f[0]), int(f[1])
        rows += [i, j]
        cols += [j, i]
    n = 1 << scale
    r, c = np.asarray(rows, np.int64), np.asarray(cols, np.int64)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return n, n, np.cumsum(rp), c
