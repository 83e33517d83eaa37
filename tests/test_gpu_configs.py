"""Parity at the workloads' own sizes, one test per configuration:

* uniform R-MAT s20 ef16 (a=b=c=d=0.25), C = A A^T, FP64 all-ones;
* join-aggregate C = A^T A of the 850000 x 20000 incidence (lognormal
  degrees, Zipf(1.2) columns), FP64 all-ones;
* the batch of 512 skewed R-MAT problems (scales 8..14, edge factors
  2..32, values U[-1, 1]) replayed from one captured graph;
* the differentiable product of skewed R-MAT s18 ef16 with FP32 N(0, 1)
  values, C = A B with B = A^T, forward and masked backward.

The oracle cannot redo these products whole in seconds, so each test checks
size-independent properties of the whole output (sum(C) = P for all-ones
inputs, ascending columns, nnz(C) = row_ptr[m], no kernel error flags) and
compares sampled rows, the heaviest among them, with the CPU oracle: bit-exact
for all-ones FP64, within 1e-12 * (|A| |B|)_ij for random FP64 and within
gamma(k+1) * (|A| |B|)_ij for FP32 (k = products of the entry,
gamma(n) = n u / (1 - n u), u = 2^-24, the bound of any summation order).
Inputs use the seeds of bench.py's workloads.
"""
import numpy as np
import pytest

from ams_hip import gen
from ams_hip.csr import Csr
from helpers import abs_csr, device_row, oracle_layout, oracle_product, oracle_rows, sample_rows

pytestmark = pytest.mark.gpu

SEED = 20261204  # bench.py's seed
U32 = 2.0 ** -24


def gamma(n):
    n = np.asarray(n, np.float64)
    return n * U32 / (1.0 - n * U32)


def ones_csr(m: Csr) -> Csr:
    return Csr(m.num_row, m.num_col, m.row_ptrs, m.cols, np.ones(m.num_nnz))


def whole_invariants(ctx, c, P, ones):
    ver = ctx.verify(c)
    assert ctx.error_flags() == 0
    assert ver["nnz"] == c.num_nnz == int(c.row_ptr[-1].item())
    assert ver["unsorted_rows"] == 0
    if ones:
        assert ver["sum_val"] == float(P)
    return ver


def sampled_rows_exact(c, a, b, row_flp, heavy, rand, seed):
    """Sampled rows of C = A B bit-exact against the oracle (all-ones)."""
    rows = sample_rows(row_flp, heavy, rand, seed)
    rp = c.row_ptr.cpu().numpy()
    ref = oracle_rows(a, oracle_layout(b), rows)
    for r, (ocols, ovals) in zip(rows.tolist(), ref):
        cols, vals = device_row(c, rp, r)
        assert len(cols) == len(ocols), f"row {r}: {len(cols)} entries, oracle {len(ocols)}"
        np.testing.assert_array_equal(cols, ocols)
        np.testing.assert_array_equal(vals, ovals)
    return rows


def test_er_s20_ef16_full(gpu_ctx):
    import torch
    from ams_hip.device import DeviceCsr
    a = gen.rmat_symmetric(20, 16, gen.ER, seed=SEED + 20)
    print(f"[configs] ER s20 ef16: nnz(A) {a.num_nnz}", flush=True)
    da = DeviceCsr.from_host(a)
    row_flp, P = gpu_ctx.row_flops(da, da)
    c = gpu_ctx.spgemm(da, da)
    torch.cuda.synchronize()
    ver = whole_invariants(gpu_ctx, c, P, ones=True)
    flp = row_flp.cpu().numpy()
    rows = sampled_rows_exact(c, a, a, flp, 8, 32, seed=20)
    print(f"[configs] ER s20 ef16: P {P} nnz(C) {ver['nnz']}, {len(rows)} sampled rows bit-exact "
          f"(max row products {int(flp.max())})", flush=True)
    del c
    torch.cuda.empty_cache()


def test_join_agg_full(gpu_ctx):
    """The Zipf-head columns of A give rows of A^T A with up to about 10^7
    products: those rows are counted in pieces by the heavy-row split."""
    import torch
    from ams_hip.device import DeviceCsr
    inc = gen.join_incidence(seed=SEED)
    at = inc.transpose()
    print(f"[configs] join-aggregate 850000 x 20000: nnz {inc.num_nnz}", flush=True)
    dat, dinc = DeviceCsr.from_host(at), DeviceCsr.from_host(inc)
    row_flp, P = gpu_ctx.row_flops(dat, dinc)
    flp = row_flp.cpu().numpy()
    kpiece = max(1 << 20, P // 2048)
    assert (flp > kpiece).sum() > 0, "the workload has rows the symbolic phase splits"
    c = gpu_ctx.spgemm(dat, dinc)
    torch.cuda.synchronize()
    ver = whole_invariants(gpu_ctx, c, P, ones=True)
    # C = A^T A is symmetric with C_ii = column degree of i
    rp = c.row_ptr.cpu().numpy()
    coldeg = np.diff(np.asarray(at.row_ptrs))
    rows = sampled_rows_exact(c, at, inc, flp, 8, 24, seed=3)
    for r in rows.tolist():
        cols, vals = device_row(c, rp, r)
        if coldeg[r]:
            assert vals[np.searchsorted(cols, r)] == coldeg[r]
    print(f"[configs] join-aggregate: P {P} nnz(C) {ver['nnz']}, {int((flp > kpiece).sum())} split rows, "
          f"{len(rows)} sampled rows bit-exact", flush=True)
    del c
    torch.cuda.empty_cache()


def problem_slice(bp, rp_host, i):
    r0, r1 = int(bp.a_off[i]), int(bp.a_off[i + 1])
    s, e = int(rp_host[r0]), int(rp_host[r1])
    return r0, r1, s, e


def test_batch_512_full(gpu_ctx):
    """Every problem of the batch against its plain one-shot product on the
    GPU (row pointers and columns equal, values within 1e-12 * (|A||B|)_ij
    from a GPU product of |A|), and a seeded subset holding the two largest
    problems against the CPU oracle; on two replays of the graph."""
    import torch
    from ams_hip.batched import BatchedProduct
    from ams_hip.device import Context, DeviceCsr
    probs = gen.batch_problems(512, seed=SEED)
    pairs = [(a, a) for a in probs]
    P = np.array([int((np.diff(a.row_ptrs).astype(np.int64) ** 2).sum()) for a in probs])
    print(f"[configs] batch of 512: P {int(P.sum())}, largest {int(P.max())}", flush=True)
    stream = torch.cuda.Stream()
    ctx = Context(stream=stream)
    bp = BatchedProduct(pairs, ams=ctx)
    plain = Context()
    subset = sample_rows(P, 2, 30, seed=4)
    for rep in range(2):
        with torch.cuda.stream(stream):
            bp.c_val.fill_(float("nan"))
            bp.run()
        bp.synchronize()
        bp.check()
        assert int(bp.nnz.cpu()[0]) == bp.cap
        rp = bp.c_rp.cpu().numpy()
        for i, a in enumerate(probs):
            r0, r1, s, e = problem_slice(bp, rp, i)
            da = DeviceCsr.from_host(a)
            ref = plain.spgemm(da, da)
            bnd = plain.spgemm(*(2 * [DeviceCsr.from_host(abs_csr(a))]))
            torch.cuda.synchronize()
            assert torch.equal(bp.c_rp[r0:r1 + 1] - s, ref.row_ptr), f"problem {i}: row pointers"
            assert torch.equal(bp.c_col[s:e] - int(bp.c_off[i]), ref.col), f"problem {i}: columns"
            err = (bp.c_val[s:e] - ref.val).abs()
            assert bool((err <= 1e-12 * bnd.val + 1e-300).all()), f"problem {i}: values"
        print(f"[configs] replay {rep}: 512 problems equal to the plain products", flush=True)
    rp = bp.c_rp.cpu().numpy()
    for i in subset.tolist():
        a = probs[i]
        r0, r1, s, e = problem_slice(bp, rp, i)
        orp, ocols, ovals = oracle_product(a, a)
        _, _, obnd = oracle_product(abs_csr(a), abs_csr(a))
        np.testing.assert_array_equal(rp[r0:r1 + 1] - s, orp)
        np.testing.assert_array_equal(bp.c_col[s:e].cpu().numpy().astype(np.int64) - int(bp.c_off[i]), ocols)
        vals = bp.c_val[s:e].cpu().numpy()
        assert np.all(np.abs(vals - ovals) <= 1e-12 * obnd + 1e-300), f"problem {i}"
    print(f"[configs] {len(subset)} problems (largest P {int(P[subset].max())}) within 1e-12 of the oracle",
          flush=True)
    bp.close()
    plain.close()
    ctx.close()


def gather_rows(c, vals, rp_host, rows):
    """Rows of a device CSR (pattern of c, values `vals`) as a host Csr of
    len(rows) rows, through one device gather."""
    import torch
    rows = np.asarray(rows, np.int64)
    starts = rp_host[rows]
    lens = rp_host[rows + 1] - starts
    sub_rp = np.zeros(len(rows) + 1, np.int64)
    np.cumsum(lens, out=sub_rp[1:])
    dev = c.col.device
    t_lens = torch.from_numpy(lens).to(dev)
    pos = (torch.arange(int(sub_rp[-1]), device=dev)
           - torch.repeat_interleave(torch.from_numpy(sub_rp[:-1]).to(dev), t_lens)
           + torch.repeat_interleave(torch.from_numpy(starts).to(dev), t_lens))
    return Csr(len(rows), c.num_col, sub_rp, c.col[pos].cpu().numpy().astype(np.int64),
               vals[pos].cpu().numpy().astype(np.float64))


def masked_rows(prod_rp, prod_cols, prod_vals, pattern_cols_per_row):
    """Values of a product's rows at the given column patterns (0 if absent)."""
    out = []
    for i, pc in enumerate(pattern_cols_per_row):
        s, e = prod_rp[i], prod_rp[i + 1]
        pcols, pv = prod_cols[s:e], prod_vals[s:e]
        if len(pcols) == 0:
            out.append(np.zeros(len(pc)))
            continue
        j = np.minimum(np.searchsorted(pcols, pc), len(pcols) - 1)
        out.append(np.where(pcols[j] == pc, pv[j], 0.0))
    return out


def fp32_check(name, got, ref, cnt, bnd):
    """FP32 result against FP64 within gamma(k + 1) * (|A||B|)_ij; returns
    the largest error over the bound's scale and the largest relative error."""
    if len(ref) == 0:
        assert len(got) == 0
        return 0.0, 0.0
    err = np.abs(np.asarray(got, np.float64) - ref)
    lim = gamma(cnt + 1) * bnd
    assert np.all(err <= lim + 1e-30), f"{name}: {int((err > lim).sum())} entries beyond gamma(k+1)*|A||B|"
    big = np.abs(ref) > 1e-6 * max(np.abs(ref).max(), 1e-300)
    rel = float((err[big] / np.abs(ref[big])).max()) if big.any() else 0.0
    return float((err / np.maximum(bnd, 1e-300)).max()), rel


def test_fp32_s18_fwd_bwd(gpu_ctx):
    """Forward C = A B and the masked backward (dA = (G B^T) on pattern(A),
    dB = (A^T G) on pattern(B)) in FP32 against the oracle's FP64 product
    and an oracle-built FP64 gradient on sampled rows."""
    import torch
    from ams_hip.device import DeviceCsr
    a = gen.rmat_symmetric(18, 16, gen.G500, seed=SEED + 5, values="normal", val_dtype=np.float32)
    b = a.transpose()
    print(f"[configs] G500 s18 ef16 FP32: nnz(A) {a.num_nnz}", flush=True)
    A, B = DeviceCsr.from_host(a), DeviceCsr.from_host(b)
    row_flp, P = gpu_ctx.row_flops(A, B)
    c = gpu_ctx.spgemm(A, B)
    gen_t = torch.Generator(device=c.val.device).manual_seed(11)
    g = DeviceCsr(c.num_row, c.num_col, c.row_ptr, c.col,
                  torch.randn(c.num_nnz, dtype=torch.float32, device=c.val.device, generator=gen_t))
    dA, dB = gpu_ctx.masked_bwd(A, B, g)
    torch.cuda.synchronize()
    assert gpu_ctx.error_flags() == 0
    ver = gpu_ctx.verify(c)
    assert ver["unsorted_rows"] == 0 and ver["nnz"] == c.num_nnz
    check_fp32_fwd_bwd(a, b, c, g.val, dA, dB, row_flp.cpu().numpy())
    del c, g, dA, dB
    torch.cuda.empty_cache()


def check_fp32_fwd_bwd(a, b, c, gval, dA, dB, row_flp):
    """Sampled rows of C = A B, dA and dB (FP32, torch tensors; c with
    row_ptr / col / val tensors) against the oracle's FP64 results."""
    rp = c.row_ptr.cpu().numpy()
    a64, b64 = a.astype(np.int64, np.float64), b.astype(np.int64, np.float64)
    # forward, sampled rows of C
    rows = sample_rows(row_flp, 6, 18, seed=8)
    ref = oracle_rows(a64, b64, rows)
    cnt = oracle_rows(ones_csr(a64), oracle_layout(ones_csr(b64)), rows)
    bnd = oracle_rows(abs_csr(a64), oracle_layout(abs_csr(b64)), rows)
    worst = [0.0, 0.0]
    for r, (oc, ov), (_, ok), (_, ob) in zip(rows.tolist(), ref, cnt, bnd):
        cols, vals = device_row(c, rp, r)
        np.testing.assert_array_equal(cols, oc)
        w = fp32_check(f"C row {r}", vals, ov, ok, ob)
        worst = [max(worst[0], w[0]), max(worst[1], w[1])]
    print(f"[configs] forward: {len(rows)} rows, max |err|/(|A||B|) {worst[0]:.3g}, max rel err {worst[1]:.3g}",
          flush=True)
    # dA rows i: (G_i . B^T) = G_i . A on pattern(A_i)
    arp = np.asarray(a.row_ptrs)
    rows_a = sample_rows(np.diff(arp), 4, 12, seed=9)
    gsub = gather_rows(c, gval, rp, rows_a)
    pats = [np.asarray(a.cols[arp[i]:arp[i + 1]], np.int64) for i in rows_a.tolist()]
    prod = oracle_product(gsub, a64)
    pc = oracle_product(ones_csr(gsub), ones_csr(a64))
    pb = oracle_product(abs_csr(gsub), abs_csr(a64))
    da_host = dA.cpu().numpy()
    worst_a = [0.0, 0.0]
    for k, (i, r, n, bb) in enumerate(zip(rows_a.tolist(), masked_rows(*prod, pats), masked_rows(*pc, pats),
                                          masked_rows(*pb, pats))):
        w = fp32_check(f"dA row {i}", da_host[arp[i]:arp[i + 1]], r, n, bb)
        worst_a = [max(worst_a[0], w[0]), max(worst_a[1], w[1])]
    print(f"[configs] dA: {len(rows_a)} rows, max |err|/bound {worst_a[0]:.3g}, max rel err {worst_a[1]:.3g}",
          flush=True)
    # dB rows k: (A^T G)_k = B_k . G on pattern(B_k); only the G rows B_k touches
    brp = np.asarray(b.row_ptrs)
    deg = np.diff(brp)
    gnnz = np.diff(rp)

    # G entries that row k of dB reads: sum of nnz(G_i) over i in B_k
    per_entry = gnnz[np.asarray(b.cols, np.int64)]
    load = np.zeros(b.num_row, np.int64)
    nz = deg > 0
    load[nz] = np.add.reduceat(per_entry, brp[:-1][nz])
    # the highest-degree row whose G rows stay under 3e7 entries, then
    # random rows of degree at most 64
    ok = np.nonzero(load <= 30_000_000)[0]
    hub = int(ok[np.argmax(deg[ok])])
    rows_b = [int(hub)]
    for k in np.random.default_rng(10).permutation(b.num_row).tolist():
        if len(rows_b) >= 16:
            break
        if 0 < deg[k] <= 64 and k != hub:
            rows_b.append(k)
    need = np.unique(np.concatenate([np.asarray(b.cols[brp[k]:brp[k + 1]], np.int64) for k in rows_b]))
    gneed = gather_rows(c, gval, rp, need)
    # G with only the needed rows, as m x n
    full_rp = np.zeros(a.num_row + 1, np.int64)
    lens = np.zeros(a.num_row, np.int64)
    lens[need] = np.diff(gneed.row_ptrs)
    np.cumsum(lens, out=full_rp[1:])
    gfull = Csr(a.num_row, c.num_col, full_rp, gneed.cols, gneed.rats)
    bsel = Csr(len(rows_b), b.num_col, np.concatenate([[0], np.cumsum(deg[rows_b])]).astype(np.int64),
               np.concatenate([b64.cols[brp[k]:brp[k + 1]] for k in rows_b]),
               np.concatenate([b64.rats[brp[k]:brp[k + 1]] for k in rows_b]))
    pats = [np.asarray(b.cols[brp[k]:brp[k + 1]], np.int64) for k in rows_b]
    gf64 = oracle_layout(gfull)
    prod = oracle_product(bsel, gf64)
    pc = oracle_product(ones_csr(bsel), ones_csr(gf64))
    pb = oracle_product(abs_csr(bsel), abs_csr(gf64))
    db_host = dB.cpu().numpy()
    worst_b = [0.0, 0.0]
    for k, r, n, bb in zip(rows_b, masked_rows(*prod, pats), masked_rows(*pc, pats), masked_rows(*pb, pats)):
        w = fp32_check(f"dB row {k}", db_host[brp[k]:brp[k + 1]], r, n, bb)
        worst_b = [max(worst_b[0], w[0]), max(worst_b[1], w[1])]
    print(f"[configs] dB: {len(rows_b)} rows (one of degree {int(deg[hub])} among them), max |err|/bound "
          f"{worst_b[0]:.3g}, max rel err {worst_b[1]:.3g}", flush=True)
    return worst, worst_a, worst_b
