"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Bar: row pointers and sorted columns bit-exact; values bit-exact for all-ones
inputs (integer counts) and within 1e-12 * (|A| |B|)_ij for random FP64.
Every route (LDS hash, ESC, bitmap symbolic, dense column chunks, multi-pass
bitmap) is forced on small inputs through the routing parameters.
"""
import sys

import numpy as np
import pytest

from ams_hip import gen
from ams_hip.csr import Csr
from helpers import assert_matches_oracle, random_csr

pytestmark = pytest.mark.gpu

# Parameter sets that push every row through one route.
ROUTES = {
    "default": {},
    "hash_only": dict(mode=1),
    "esc_only": dict(mode=2, esc_max=4096),
    "all_dense": dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=128,
                      bitmap_bits=1024, dense_density=0.0),
    "dense_one_pass": dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=4096,
                           bitmap_bits=65536, dense_density=0.0),
    # every dense morsel through the block kernel (no light morsels)
    "all_dense_block": dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=128,
                            bitmap_bits=1024, dense_density=0.0, light_max=0),
    # morsels of up to 8 C entries through the wave kernel, the rest through the block kernel
    "dense_light_mix": dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=1024,
                            bitmap_bits=65536, dense_density=0.0, light_max=8),
    # without the kept row bitmaps: light morsels in hash tables, the block kernel's bitmap per product
    "all_dense_hash_light": dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=128,
                                 bitmap_bits=1024, dense_density=0.0, row_bitmaps=0),
    "dense_mix_no_bitmaps": dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=1024,
                                 bitmap_bits=65536, dense_density=0.0, light_max=8, row_bitmaps=0),
    # multi-pass bitmap symbolic (bitmap_bits below the row width) keeping its row bitmaps
    "multipass_bitmaps_block": dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=256,
                                    bitmap_bits=512, dense_density=0.0, light_max=0),
    "bitmap_then_hash": dict(sym_hash_max=8, num_hash_max=5461, bitmap_ratio=0.0, dense_density=1.0,
                             bitmap_bits=256, dense_width=64),
    "big_tables": dict(sym_hash_max=5461, num_hash_max=5461, acc_thres=0.0),
    # num_hash_max below sym_hash_max: LDS-hash-symbolic rows above it still take LDS
    # tables, bitmap-symbolic rows above it the dense chunks (no ESC)
    "hash_sym_above_num": dict(sym_hash_max=2048, num_hash_max=16, acc_thres=0.0),
    "blocked_sym": dict(sym_blocked=1, sym_hash_max=64, num_hash_max=64, bitmap_ratio=0.0),
    "blocked_sym_dense": dict(sym_blocked=1, sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=128,
                              bitmap_bits=1024, dense_density=0.0),
    "esc_bucket": dict(mode=2, esc_sort=1),
    "esc_radix": dict(mode=2, esc_sort=2),
    "esc_merge": dict(mode=2, esc_sort=3),
    "esc_count": dict(mode=2, esc_sort=4),
    "ghash": dict(sym_hash_max=64, num_hash_max=64, ghash_ratio=1e9, mode=1),
    "ghash_hybrid": dict(sym_hash_max=64, num_hash_max=64, ghash_ratio=1e9),
}


def run(ctx, a, b, params):
    import torch
    from ams_hip.device import DeviceCsr
    from ams_hip._lib import AmsParams, lib
    p = AmsParams()
    lib().ams_default_params(p)
    ctx.set_params(**{**p.as_dict(), **params})
    da = DeviceCsr.from_host(a)
    db = DeviceCsr.from_host(b)
    c = ctx.spgemm(da, db)
    torch.cuda.synchronize()
    return c.to_host()


def small_cases():
    cases = []
    for s, ef in [(6, 2), (8, 4), (10, 8), (11, 16)]:
        cases.append((f"rmat_s{s}_ef{ef}_ones", gen.rmat_symmetric(s, ef, seed=s * 7 + ef), True))
        cases.append((f"rmat_s{s}_ef{ef}_unif", gen.rmat_symmetric(s, ef, seed=s * 11 + ef, values="uniform"), False))
    cases.append(("er_s10_ef8_ones", gen.rmat_symmetric(10, 8, gen.ER, seed=3), True))
    return cases


CASES = small_cases()


@pytest.mark.parametrize("route", list(ROUTES))
@pytest.mark.parametrize("name,a,exact", CASES, ids=[c[0] for c in CASES])
def test_rmat_square(gpu_ctx, route, name, a, exact):
    print(f"[parity] {name} route={route}", flush=True)
    c = run(gpu_ctx, a, a, ROUTES[route])
    assert c.sorted_rows()
    assert_matches_oracle(c, a, a, exact)


@pytest.mark.parametrize("route", ["default", "all_dense", "all_dense_block", "dense_light_mix",
                                   "all_dense_hash_light", "dense_mix_no_bitmaps", "esc_only",
                                   "blocked_sym_dense", "esc_bucket", "esc_merge",
                                   "ghash"])
def test_rectangular_unsorted(gpu_ctx, route):
    a = random_csr(70, 300, 0.05, seed=1)
    b = random_csr(300, 2000, 0.02, seed=2, sort=False)
    print(f"[parity] rect unsorted route={route}", flush=True)
    c = run(gpu_ctx, a, b, ROUTES[route])
    assert c.sorted_rows()
    assert_matches_oracle(c, a, b, exact=False)


@pytest.mark.parametrize("width", [8448, 14848])
def test_wide_chunks(gpu_ctx, width):
    """Chunks wide enough for the 1024-thread dense kernel (14848 columns)
    next to the default 8448-column chunks, on a 16384-column problem."""
    a = gen.rmat_symmetric(14, 2, seed=77)
    print(f"[parity] rmat_s14_ef2 dense_width={width}", flush=True)
    c = run(gpu_ctx, a, a, dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_density=0.0,
                                dense_width=width))
    assert c.sorted_rows()
    assert_matches_oracle(c, a, a, exact=True)


@pytest.mark.parametrize("bitmap_bits", [1 << 20, 128])
def test_heavy_rows_split(gpu_ctx, bitmap_bits):
    """C = A^T A of a join-shaped incidence whose hub columns give rows of
    more than max(2^20, P/2048) products: those rows are counted in pieces
    by several blocks (OR-ed global bitmaps), in one bitmap pass or in
    several (bitmap_bits 128 on 300 columns)."""
    inc = gen.join_incidence(num_row=60000, num_col=300, mu=3.5, sigma=0.5, cap=100, seed=5)
    at = inc.transpose()
    deg = np.diff(np.asarray(inc.row_ptrs)).astype(np.int64)
    rows = np.repeat(np.arange(at.num_row), np.diff(np.asarray(at.row_ptrs)))
    row_p = np.bincount(rows, weights=deg[np.asarray(at.cols)], minlength=at.num_row)
    assert row_p.max() > max(1 << 20, row_p.sum() / 2048), "the case must have a heavy row"
    print(f"[parity] join 60000x300 heavy rows {(row_p > max(1 << 20, row_p.sum() / 2048)).sum()} "
          f"bitmap_bits={bitmap_bits}", flush=True)
    c = run(gpu_ctx, at, inc, dict(bitmap_bits=bitmap_bits, dense_width=128 if bitmap_bits < 1024 else 8448))
    assert c.sorted_rows()
    assert_matches_oracle(c, at, inc, exact=True)


def test_wide_heavy_rows(gpu_ctx):
    """Rows of more than 2^20 products over 150000 columns: bitmap passes of
    more than 2^17 bits (the 1024-thread symbolic blocks) with the heavy
    rows split into pieces, next to light rows of the same problem."""
    rng = np.random.default_rng(31)
    n, kb, nb = 150000, 900, 3000
    cols = np.sort(rng.integers(0, n, size=(nb, kb)), axis=1)
    keep = np.ones_like(cols, dtype=bool)
    keep[:, 1:] = cols[:, 1:] != cols[:, :-1]
    rp = np.concatenate([[0], np.cumsum(keep.sum(axis=1))]).astype(np.int64)
    b = Csr(nb, n, rp, cols[keep].astype(np.int32), np.ones(int(rp[-1])))
    deg = np.diff(rp)
    arows = [np.sort(rng.choice(nb, size=k, replace=False)) for k in (1500, 1400, 40, 3, 700)]
    a_rp = np.concatenate([[0], np.cumsum([len(r) for r in arows])]).astype(np.int64)
    a = Csr(len(arows), nb, a_rp, np.concatenate(arows).astype(np.int32), np.ones(int(a_rp[-1])))
    row_p = np.array([deg[r].sum() for r in arows])
    assert row_p[:2].min() > (1 << 20), "two rows must be heavy"
    print(f"[parity] wide heavy rows: products {row_p.tolist()}", flush=True)
    c = run(gpu_ctx, a, b, dict(sym_hash_max=1, bitmap_ratio=0.0))
    assert c.sorted_rows()
    assert_matches_oracle(c, a, b, exact=True)


def test_wide_bitmap_dynamic_lds(gpu_ctx):
    """Bitmap rows over 700000 columns: a bitmap pass of 83 chunks (21912
    words) is too large for the symbolic kernel's static LDS array, so the
    dynamic-LDS instantiation counts them."""
    rng = np.random.default_rng(71)
    n, kb, nb = 700000, 300, 2000
    cols = np.sort(rng.integers(0, n, size=(nb, kb)), axis=1)
    keep = np.ones_like(cols, dtype=bool)
    keep[:, 1:] = cols[:, 1:] != cols[:, :-1]
    rp = np.concatenate([[0], np.cumsum(keep.sum(axis=1))]).astype(np.int64)
    b = Csr(nb, n, rp, cols[keep].astype(np.int32), np.ones(int(rp[-1])))
    arows = [np.sort(rng.choice(nb, size=k, replace=False)) for k in (200, 150, 5, 90)]
    a_rp = np.concatenate([[0], np.cumsum([len(r) for r in arows])]).astype(np.int64)
    a = Csr(len(arows), nb, a_rp, np.concatenate(arows).astype(np.int32), np.ones(int(a_rp[-1])))
    print("[parity] bitmap rows over 700000 columns", flush=True)
    c = run(gpu_ctx, a, b, dict(sym_hash_max=1, bitmap_ratio=0.0))
    assert c.sorted_rows()
    assert_matches_oracle(c, a, b, exact=True)


@pytest.mark.parametrize("route", ["default", "esc_only", "esc_merge"])
@pytest.mark.parametrize("b_density", [0.0005, 0.002, 0.0055])
def test_esc_long_a_rows(gpu_ctx, route, b_density):
    """ESC rows with 260+ A nonzeros over short B rows: several expansion
    windows per row (64 A nonzeros in the 512-product bin, 128 in the 2048-
    and 4096-product bins, whose blocks have 256 and 512 threads)."""
    a = random_csr(24, 3000, 0.1, seed=41)
    b = random_csr(3000, 2000, b_density, seed=42)
    assert np.diff(a.row_ptrs).min() > 128
    print(f"[parity] esc long A rows route={route} b_density={b_density}", flush=True)
    c = run(gpu_ctx, a, b, ROUTES[route])
    assert c.sorted_rows()
    assert_matches_oracle(c, a, b, exact=False)


@pytest.mark.parametrize("density", [0.05, 0.2, 0.4])
def test_esc_wide_columns(gpu_ctx, density):
    """ESC rows over 2^26 + 7 columns: (column, rank) no longer packs into 32
    bits for any bin, so the 64-bit key sort runs (random FP64 values).  Rows
    of about 300, 1200 and 2400 products: the 1536- and 3072-product bins then
    run the 64-bit kernel with the 2048 and 4096 geometry."""
    rng = np.random.default_rng(17)
    n = (1 << 26) + 7
    a = random_csr(200, 300, density, seed=18)
    rows = np.repeat(np.arange(300), 20)
    cols = np.sort(rng.choice(n, size=(300, 20), replace=False if n > 20 else True), axis=1).reshape(-1)
    rp = np.arange(0, 301 * 20, 20, dtype=np.int64)
    b = Csr(300, n, rp, cols.astype(np.int32), rng.uniform(-1, 1, len(cols)))
    print(f"[parity] esc over 2^26+7 columns, A density {density}", flush=True)
    c = run(gpu_ctx, a, b, dict(mode=2, esc_max=4096))
    assert c.sorted_rows()
    assert_matches_oracle(c, a, b, exact=False)


def test_edge_shapes(gpu_ctx):
    z = Csr(5, 7, np.zeros(6, np.int64), np.zeros(0, np.int32), np.zeros(0))
    b = random_csr(7, 9, 0.5, seed=4)
    c = run(gpu_ctx, z, b, {})
    assert c.num_nnz == 0 and np.all(c.row_ptrs == 0)
    # identity times B is B; diagonal scales rows
    bb = random_csr(40, 50, 0.2, seed=5, values="ones")
    eye = Csr(40, 40, np.arange(41, dtype=np.int64), np.arange(40, dtype=np.int32), np.ones(40))
    c = run(gpu_ctx, eye, bb, {})
    np.testing.assert_array_equal(c.row_ptrs, bb.row_ptrs)
    np.testing.assert_array_equal(c.cols, bb.cols)
    np.testing.assert_array_equal(c.rats, bb.rats)
    d = Csr(40, 40, np.arange(41, dtype=np.int64), np.arange(40, dtype=np.int32), np.arange(1, 41, dtype=np.float64))
    c = run(gpu_ctx, d, bb, {})
    rows = np.repeat(np.arange(40), np.diff(bb.row_ptrs))
    np.testing.assert_array_equal(c.rats, bb.rats * (rows + 1))
    # n = 1: dot products
    x = random_csr(30, 20, 0.3, seed=6)
    y = random_csr(20, 1, 0.6, seed=7)
    c = run(gpu_ctx, x, y, {})
    assert_matches_oracle(c, x, y, exact=False)
    # rows with no products in the middle of heavy ones
    a = random_csr(64, 64, 0.3, seed=8, values="ones")
    r0, r1 = a.row_ptrs[10], a.row_ptrs[20]
    cols = np.concatenate([a.cols[:r0], a.cols[r1:]])
    rats = np.concatenate([a.rats[:r0], a.rats[r1:]])
    rp = a.row_ptrs.copy()
    rp[10:20] = r0
    rp[20:] -= (r1 - r0)
    a = Csr(64, 64, rp, cols, rats)
    for route in ("default", "all_dense"):
        c = run(gpu_ctx, a, a, ROUTES[route])
        assert_matches_oracle(c, a, a, exact=True)


def with_duplicates(m: Csr, seed: int, shuffle: bool) -> Csr:
    """m with about a third of its entries repeated inside their rows (the
    reference keeps such repeats: its reader does not merge them)."""
    rng = np.random.default_rng(seed)
    rp, cols, vals = [0], [], []
    for i in range(m.num_row):
        s, e = int(m.row_ptrs[i]), int(m.row_ptrs[i + 1])
        c, v = list(m.cols[s:e]), list(m.rats[s:e])
        for j in range(e - s):
            if rng.random() < 0.35:
                c.append(m.cols[s + j])
                v.append(rng.uniform(-1, 1) if m.rats.dtype == np.float64 and not np.all(m.rats == 1) else 1.0)
        order = rng.permutation(len(c)) if shuffle else np.argsort(c, kind="stable")
        cols += [c[k] for k in order]
        vals += [v[k] for k in order]
        rp.append(len(cols))
    return Csr(m.num_row, m.num_col, np.array(rp, np.int64), np.array(cols, np.int32), np.array(vals, np.float64))


@pytest.mark.parametrize("route", ["default", "hash_only", "esc_only", "all_dense", "big_tables"])
@pytest.mark.parametrize("shuffle", [False, True])
def test_duplicate_entries(gpu_ctx, route, shuffle):
    """Repeated column entries inside rows of A and B: every copy adds its
    products, and C has each column once."""
    a1 = with_duplicates(gen.rmat_symmetric(10, 4, seed=21), 22, shuffle)
    c = run(gpu_ctx, a1, a1, ROUTES[route])
    assert c.sorted_rows()
    assert_matches_oracle(c, a1, a1, exact=True)
    a2 = with_duplicates(random_csr(150, 150, 0.05, seed=23), 24, shuffle)
    c = run(gpu_ctx, a2, a2, ROUTES[route])
    assert_matches_oracle(c, a2, a2, exact=False)


def test_fp32(gpu_ctx):
    a = gen.rmat_symmetric(10, 8, seed=9, values="normal", val_dtype=np.float32)
    for route in ("default", "all_dense", "esc_only"):
        c = run(gpu_ctx, a, a.transpose(), ROUTES[route])
        assert c.rats.dtype == np.float32
        a64 = a.astype(np.int32, np.float64)
        assert_matches_oracle(Csr(c.num_row, c.num_col, c.row_ptrs, c.cols, c.rats.astype(np.float64)),
                              a64, a64.transpose(), exact=False, rel=1e-5)


def test_config1_full(gpu_ctx):
    """G500 s16 ef16 all-ones, C = A * A^T: the reference's CPU-runnable case."""
    a = gen.rmat_symmetric(16, 16, seed=20261204)
    print("[parity] config1 s16 ef16", flush=True)
    c = run(gpu_ctx, a, a, {})
    assert c.sorted_rows()
    assert_matches_oracle(c, a, a, exact=True)
    st = gpu_ctx.stats()
    print(st, file=sys.stderr)
    assert st["nnz_c"] == c.num_nnz


def test_fp32_chunk_width(gpu_ctx):
    """FP32 dense chunks are 7/4 as wide as FP64 ones (rounded down to 64
    columns): the same wide product runs in fewer FP32 chunks, both exact."""
    a64 = gen.rmat_symmetric(15, 4, seed=61)
    a32 = a64.astype(np.int32, np.float32)
    params = dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_density=0.0, dense_width=8448)
    c64 = run(gpu_ctx, a64, a64, params)
    r64 = gpu_ctx.stats()["dense_chunks"]
    c32 = run(gpu_ctx, a32, a32, params)
    r32 = gpu_ctx.stats()["dense_chunks"]
    n = a64.num_col
    assert r64 == -(-n // 8448) and r32 == -(-n // (8448 * 7 // 4 // 64 * 64)), (r64, r32)
    assert_matches_oracle(c64, a64, a64, exact=True)
    np.testing.assert_array_equal(c32.row_ptrs, c64.row_ptrs)
    np.testing.assert_array_equal(c32.cols, c64.cols)
    np.testing.assert_array_equal(c32.rats, c64.rats.astype(np.float32))


ESC_CASES = [("rmat_s8_ef4", lambda: gen.rmat_symmetric(8, 4, seed=101, values="uniform")),
             ("rmat_s10_ef8", lambda: gen.rmat_symmetric(10, 8, seed=102, values="uniform")),
             ("er_s11_ef8", lambda: gen.rmat_symmetric(11, 8, gen.ER, seed=103, values="uniform")),
             # rows of about 2400 products: the 1536, 2048 and 3072 bins
             ("er_s11_ef24", lambda: gen.rmat_symmetric(11, 24, gen.ER, seed=104, values="uniform")),
             ("rect_unsorted", None)]


@pytest.mark.parametrize("esc_sort", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("name,make", ESC_CASES, ids=[c[0] for c in ESC_CASES])
def test_esc_random_values_bit_exact(gpu_ctx, name, make, esc_sort):
    """Random FP64 values through the ESC route: every row of at most esc_max
    products is summed left to right in A-row order from separately rounded
    products, as the reference's compress does (src/ams/ams_utils.c:1062-1067),
    so its values equal the oracle's bit for bit; a fused multiply-add or any
    other order would change the last bits of most entries.  Longer rows (hash
    or dense routes) are held to the 1e-12 * (|A||B|)_ij bound."""
    import torch
    from ams_hip.device import DeviceCsr
    from helpers import abs_csr, oracle_product
    if make is None:
        a, b = random_csr(90, 400, 0.06, seed=11), random_csr(400, 3000, 0.03, seed=12, sort=False)
    else:
        a = make()
        b = a
    esc_max = 4096
    # ESC for every row of at most esc_max products: no bitmap symbolic below
    # 5461 products and no dense chunks for such rows
    c = run(gpu_ctx, a, b, dict(mode=2, esc_max=esc_max, esc_sort=esc_sort, sym_hash_max=5461,
                                num_hash_max=5461, bitmap_ratio=1e9, dense_density=2.0))
    routed_esc = gpu_ctx.stats()["rows_num_esc"]
    flp, _ = gpu_ctx.row_flops(DeviceCsr.from_host(a), DeviceCsr.from_host(b))
    flp = flp.cpu().numpy()
    torch.cuda.synchronize()
    rp, cols, vals = oracle_product(a, b)
    np.testing.assert_array_equal(np.asarray(c.row_ptrs, np.int64), rp)
    np.testing.assert_array_equal(np.asarray(c.cols, np.int64), cols)
    esc_row = (flp <= esc_max) & (np.diff(rp) > 0)
    assert routed_esc == int(esc_row.sum()), "every row of at most esc_max products took ESC"
    esc_rows = np.repeat(esc_row, np.diff(rp))
    assert esc_rows.mean() > 0.5, "most entries must come from ESC rows"
    np.testing.assert_array_equal(c.rats[esc_rows], vals[esc_rows])
    _, _, bound = oracle_product(abs_csr(a), abs_csr(b))
    assert np.all(np.abs(c.rats - vals) <= 1e-12 * bound + 1e-300)
    print(f"[parity] {name} esc_sort={esc_sort}: {int(esc_rows.sum())} of {len(vals)} entries from ESC rows "
          f"bit-exact", flush=True)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_global_hash_route(gpu_ctx, dtype):
    """Rows too long for an LDS table and sparse next to their column range
    go to global-memory hash tables (symbolic and numeric) when ghash_ratio
    allows; the rows come back sorted and equal to the oracle's."""
    import torch
    from ams_hip.device import DeviceCsr
    from ams_hip._lib import AmsParams, lib
    a = gen.rmat_symmetric(12, 16, seed=61, values="uniform")
    if dtype == np.float32:  # the oracle sees the values the GPU multiplies
        a = Csr(a.num_row, a.num_col, a.row_ptrs, a.cols, a.rats.astype(np.float32).astype(np.float64))
    p = AmsParams()
    lib().ams_default_params(p)
    # rows of more than 64 products and fewer than 40 * words(4096) = 5120
    # take the global-hash symbolic; without ESC (acc_thres 0) and with LDS
    # tables up to 64 entries their numeric phase does too
    gpu_ctx.set_params(**{**p.as_dict(), "ghash_ratio": 40.0, "sym_hash_max": 64, "num_hash_max": 64,
                          "acc_thres": 0.0})
    da = DeviceCsr.from_host(a, val_dtype=dtype)
    c = gpu_ctx.spgemm(da, da)
    torch.cuda.synchronize()
    st = gpu_ctx.stats()
    print(f"[parity] ghash {np.dtype(dtype).name}: rows sym {st['rows_sym_ghash']} num {st['rows_num_ghash']}",
          flush=True)
    assert st["rows_sym_ghash"] > 0 and st["rows_num_ghash"] > 0
    h = c.to_host()
    assert h.sorted_rows()
    if dtype == np.float64:
        assert_matches_oracle(h, a, a, exact=False)
    else:
        # FP32 sums of at most ~16000 terms: |error| <= 16000 * 2^-24 * (|A||B|)_ij
        assert_matches_oracle(h, a, a, exact=False, rel=1e-3)


def test_global_hash_wide_sparse(gpu_ctx):
    """Very wide sparse rows (4,000,000 columns, a few thousand products a
    row): all-ones values, bit-exact against the oracle with the global
    hash on and off."""
    import torch
    from ams_hip.device import DeviceCsr
    from ams_hip._lib import AmsParams, lib
    rng = np.random.default_rng(5)
    m, k, n = 300, 2000, 4_000_000
    a = random_csr(m, k, 0.05, seed=51, values="ones")
    deg = rng.integers(20, 60, k)
    rp = np.concatenate([[0], np.cumsum(deg)]).astype(np.int64)
    cols = np.concatenate([np.sort(rng.choice(n, d, replace=False)) for d in deg]).astype(np.int32)
    b = Csr(k, n, rp, cols, np.ones(len(cols)))
    p = AmsParams()
    lib().ams_default_params(p)
    for ratio in (0.0, 1.0):
        # rows of about 4000 products: LDS tables take them up to 5461, so
        # the global hash needs lower LDS limits
        gpu_ctx.set_params(**{**p.as_dict(), "ghash_ratio": ratio, "sym_hash_max": 2730, "num_hash_max": 2730})
        c = gpu_ctx.spgemm(DeviceCsr.from_host(a), DeviceCsr.from_host(b))
        torch.cuda.synchronize()
        st = gpu_ctx.stats()
        print(f"[parity] wide sparse ghash_ratio={ratio}: ghash rows {st['rows_num_ghash']}", flush=True)
        assert (st["rows_num_ghash"] > 0) == (ratio > 0)
        assert_matches_oracle(c.to_host(), a, b, exact=True)
