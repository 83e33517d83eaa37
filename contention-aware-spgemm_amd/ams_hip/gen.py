"""Seeded synthetic inputs for the SpGEMM workloads.

The reference draws its R-MAT graphs with PaRMAT (PaRMAT/src/utils.cpp:63-172,
GraphGen_sorted.cpp:17-207) and turns them into symmetric all-ones
MatrixMarket pairs with rmat_ones (src/rmat_ones.c:6-37).  PaRMAT seeds from
std::random_device and time(0), so its graphs cannot be reproduced; this
module draws the same distribution from a fixed numpy seed instead:

* ``rmat_symmetric``: R-MAT edges (quadrant probabilities a, b, c, d), self
  loops dropped, undirected duplicates dropped, exactly ``ef * 2**scale``
  distinct undirected edges kept in draw order, then mirrored.  nnz = 2 * ef *
  2**scale, columns ascending within every row.
* ``join_incidence``: the join-aggregate-shaped incidence matrix (lognormal
  row degrees, Zipf column ids drawn without repeats inside a row).
* ``batch_problems``: the uneven batch of small skewed R-MAT problems.
"""
from __future__ import annotations

import numpy as np

from .csr import Csr

G500 = (0.57, 0.19, 0.19)
ER = (0.25, 0.25, 0.25)


def _rmat_batch(rng, scale, count, a, b, c):
    src = np.zeros(count, np.int64)
    dst = np.zeros(count, np.int64)
    ab, abc = a + b, a + b + c
    for _ in range(scale):
        u = rng.random(count, dtype=np.float64)
        row_bit = u >= ab
        col_bit = ((u >= a) & (u < ab)) | (u >= abc)
        src <<= 1
        dst <<= 1
        src |= row_bit
        dst |= col_bit
    return src, dst


def rmat_undirected_edges(scale, ef, abc=G500, seed=0):
    """Exactly ef * 2**scale distinct undirected, loop-free edges (lo < hi)."""
    a, b, c = abc
    n = 1 << scale
    target = ef * n
    if target > n * (n - 1) // 2:
        raise ValueError(f"scale {scale} has only {n * (n - 1) // 2} loop-free edges, {target} asked")
    rng = np.random.default_rng(seed)
    keys = np.empty(0, np.int64)
    batch = target + target // 4 + 1024
    for _ in range(64):
        s, d = _rmat_batch(rng, scale, batch, a, b, c)
        lo, hi = np.minimum(s, d), np.maximum(s, d)
        keep = lo != hi
        keys = np.concatenate([keys, lo[keep] * n + hi[keep]])
        _, first = np.unique(keys, return_index=True)
        if len(first) >= target:
            first.sort()
            keys = keys[first[:target]]
            return keys // n, keys % n
        keys = keys[np.sort(first)]
        deficit = target - len(keys)
        batch = 2 * deficit + 1024
    raise RuntimeError("R-MAT generator did not reach the edge target")


def rmat_symmetric(scale, ef, abc=G500, seed=0, values="ones", val_dtype=np.float64):
    """Symmetric R-MAT adjacency as CSR (int64 row_ptrs, int32 cols).

    values: "ones" (all 1.0), "uniform" (U[-1, 1], equal on both mirrored
    entries) or "normal" (N(0, 1) drawn per stored entry, so A != A^T).
    """
    n = 1 << scale
    lo, hi = rmat_undirected_edges(scale, ef, abc, seed)
    rows = np.concatenate([lo, hi])
    cols = np.concatenate([hi, lo])
    vrng = np.random.default_rng(seed + 7919)
    if values == "ones":
        vals = np.ones(len(rows), val_dtype)
    elif values == "uniform":
        v = vrng.uniform(-1.0, 1.0, len(lo))
        vals = np.concatenate([v, v]).astype(val_dtype)
    elif values == "normal":
        vals = vrng.standard_normal(len(rows)).astype(val_dtype)
    else:
        raise ValueError(values)
    order = np.argsort(rows * n + cols, kind="stable")
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return Csr(n, n, rp, cols.astype(np.int32), vals)


def join_incidence(num_row=850_000, num_col=20_000, mu=2.0, sigma=1.5, cap=5000,
                   zipf_s=1.2, seed=0):
    """Rows: lognormal degrees clipped to [1, cap]; columns: Zipf(zipf_s) ids
    (rank r -> column r - 1) without repeats inside a row; values 1.0.

    Each row keeps its first `degree` distinct draws in draw order; rows
    short of their degree after a round draw again (more each round)."""
    rng = np.random.default_rng(seed)
    deg = np.clip(np.rint(rng.lognormal(mu, sigma, num_row)), 1, cap).astype(np.int64)
    pmf = np.arange(1, num_col + 1, dtype=np.float64) ** (-zipf_s)
    cdf = np.cumsum(pmf)
    cdf /= cdf[-1]
    done = []                                # final keys (row * num_col + col) of finished rows
    todo = np.arange(num_row, dtype=np.int64)
    cur = np.empty(0, np.int64)              # partial keys of the rows in todo, draw order
    factor = 1.25
    for _ in range(64):
        if len(todo) == 0:
            break
        have = np.bincount(cur // num_col, minlength=num_row)[todo] if len(cur) else np.zeros(len(todo), np.int64)
        want = ((deg[todo] - have) * factor).astype(np.int64) + 4
        rows = np.repeat(todo, want)
        cols = np.minimum(np.searchsorted(cdf, rng.random(len(rows)), side="right"), num_col - 1)
        keys = np.concatenate([cur, rows * num_col + cols])
        _, first = np.unique(keys, return_index=True)
        first.sort()
        keys = keys[first]                   # distinct, in draw order
        order = np.argsort(keys // num_col, kind="stable")
        keys = keys[order]
        r = keys // num_col
        start = np.searchsorted(r, r, side="left")
        keep = (np.arange(len(keys)) - start) < deg[r]
        keys, r = keys[keep], r[keep]
        cnt = np.bincount(r, minlength=num_row)
        full = cnt[r] == deg[r]
        done.append(keys[full])
        cur = keys[~full]
        todo = todo[cnt[todo] < deg[todo]]
        factor *= 2.0
    else:
        raise RuntimeError("join incidence generator did not converge")
    keys = np.sort(np.concatenate(done))
    rows, cols = keys // num_col, keys % num_col
    rp = np.zeros(num_row + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=num_row), out=rp[1:])
    return Csr(num_row, num_col, rp, cols.astype(np.int32), np.ones(len(rows), np.float64))


def batch_problems(count=512, seed=0, scales=(8, 14), efs=(2, 4, 8, 16, 32)):
    """The uneven batch: scale ~ U{scales}, ef ~ U{efs}, skewed R-MAT,
    symmetric, values U[-1, 1]."""
    rng = np.random.default_rng(seed)
    sc = rng.integers(scales[0], scales[1] + 1, count)
    ef = rng.choice(np.asarray(efs), count)
    return [rmat_symmetric(int(s), int(e), G500, seed=seed * 100003 + i, values="uniform")
            for i, (s, e) in enumerate(zip(sc, ef))]


def wide_sparse(m=65536, k=65536, n=1 << 24, deg_a=128, deg_b=64, seed=0):
    """A (m x k, about deg_a nonzeros a row) and B (k x n, about deg_b a row)
    with uniform random columns, values 1.0: rows of C = A B have about
    deg_a * deg_b products spread over n columns, so they are long for an
    LDS table yet very sparse next to their column range."""
    rng = np.random.default_rng(seed)

    def rand_csr(rows, cols, deg):
        r = np.repeat(np.arange(rows, dtype=np.int64), deg)
        c = rng.integers(0, cols, rows * deg, dtype=np.int64)
        keys = np.unique(r * cols + c)
        r, c = keys // cols, keys % cols
        rp = np.zeros(rows + 1, np.int64)
        np.cumsum(np.bincount(r, minlength=rows), out=rp[1:])
        return Csr(rows, cols, rp, c.astype(np.int32), np.ones(len(c), np.float64))

    return rand_csr(m, k, deg_a), rand_csr(k, n, deg_b)
