"""Shared checks: compare a GPU product with the CPU oracle."""
from __future__ import annotations

import numpy as np

from ams_hip.csr import Csr


def abs_csr(m: Csr) -> Csr:
    return Csr(m.num_row, m.num_col, m.row_ptrs, m.cols, np.abs(m.rats).astype(np.float64))


def oracle_product(a: Csr, b: Csr, mode="hybacc", threads=None):
    from oracle import oracle as orc
    rp, cols, rats, rt = orc.spgemm(a.astype(np.int64, np.float64), b.astype(np.int64, np.float64),
                                    mode, threads=threads)
    return rp, cols, rats


def assert_matches_oracle(c: Csr, a: Csr, b: Csr, exact: bool, rel=1e-12, mode="hybacc"):
    """row_ptr and (sorted) columns bit-exact; values bit-exact or within
    rel * (|A| |B|)_ij, the summation error bound of the entry."""
    rp, cols, rats = oracle_product(a, b, mode)
    assert c.num_row == a.num_row and c.num_col == b.num_col
    np.testing.assert_array_equal(np.asarray(c.row_ptrs, np.int64), rp)
    np.testing.assert_array_equal(np.asarray(c.cols, np.int64), cols)
    got = np.asarray(c.rats, np.float64)
    if exact:
        np.testing.assert_array_equal(got, rats)
    else:
        _, _, bound = oracle_product(abs_csr(a), abs_csr(b), mode)
        err = np.abs(got - rats)
        lim = rel * bound + 1e-300
        bad = np.nonzero(err > lim)[0]
        assert len(bad) == 0, f"{len(bad)} entries beyond {rel}*|A||B|, worst {err[bad].max()} vs {lim[bad].min()}"
    return rp, cols, rats


def random_csr(m, n, density, seed, values="uniform", sort=True, dtype=np.float64):
    rng = np.random.default_rng(seed)
    mask = rng.random((m, n)) < density
    rows, cols = np.nonzero(mask)
    if values == "ones":
        vals = np.ones(len(rows))
    else:
        vals = rng.uniform(-1, 1, len(rows))
    rp = np.zeros(m + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=m), out=rp[1:])
    cols = cols.astype(np.int32)
    if not sort:
        for i in range(m):  # shuffle inside every row
            s, e = rp[i], rp[i + 1]
            perm = rng.permutation(e - s)
            cols[s:e] = cols[s:e][perm]
            vals[s:e] = vals[s:e][perm]
    return Csr(m, n, rp, cols, vals.astype(dtype))


def oracle_layout(m: Csr) -> Csr:
    """m in the oracle's int64 / float64 layout (convert large operands once)."""
    return m.astype(np.int64, np.float64)


def oracle_rows(a: Csr, b64: Csr, rows, mode="hybacc"):
    """The oracle's product of the listed rows of A with B (b64 already in
    oracle_layout), in one call; returns (cols int64, vals float64) per row."""
    from oracle import oracle as orc
    rows = np.asarray(rows, np.int64)
    rp_a = np.asarray(a.row_ptrs, np.int64)
    cnt = rp_a[rows + 1] - rp_a[rows]
    rp = np.zeros(len(rows) + 1, np.int64)
    np.cumsum(cnt, out=rp[1:])
    idx = np.concatenate([np.arange(rp_a[r], rp_a[r + 1]) for r in rows]) if len(rows) else np.zeros(0, np.int64)
    sub = Csr(len(rows), a.num_col, rp, np.asarray(a.cols)[idx].astype(np.int64),
              np.asarray(a.rats)[idx].astype(np.float64))
    orp, ocols, orats, _ = orc.spgemm(sub, b64, mode)
    return [(ocols[orp[i]:orp[i + 1]], orats[orp[i]:orp[i + 1]]) for i in range(len(rows))]


def device_row(c, rp_host, r):
    """Columns (int64) and values of row r of a device CSR C."""
    s, e = int(rp_host[r]), int(rp_host[r + 1])
    return c.col[s:e].cpu().numpy().astype(np.int64), c.val[s:e].cpu().numpy()


def sample_rows(row_flp, count_heavy, count_random, seed):
    """The count_heavy rows with the most products and count_random others
    drawn at random (seeded), ascending and distinct."""
    flp = np.asarray(row_flp)
    rng = np.random.default_rng(seed)
    heavy = np.argsort(flp, kind="stable")[-count_heavy:]
    rand = rng.choice(len(flp), min(count_random, len(flp)), replace=False)
    return np.unique(np.concatenate([heavy, rand]))
