"""Parity at the headline size: G500 R-MAT scale 19, edge factor 32, all-ones
FP64, C = A * A (P = 1.0e11 products, nnz(C) = 2.0e10, C about 245 GB in
HBM).  The oracle cannot redo the whole product, so the checks are:

* size-independent invariants of the whole C: sum(C) = P for all-ones
  inputs, every row's columns strictly ascending, nnz(C) = row_ptr[m];
* 40 sampled rows (the 8 with the most products and 32 at random) against
  the CPU oracle's product of those rows of A with A, bit-exact.
"""
import numpy as np
import pytest

from ams_hip import gen
from helpers import oracle_product

pytestmark = pytest.mark.gpu

SEED = 20261204  # the headline workload's seed (bench.py)


def test_g500_s19_ef32_full(gpu_ctx):
    import torch
    from ams_hip.device import DeviceCsr
    a = gen.rmat_symmetric(19, 32, gen.G500, seed=SEED + 19)
    print(f"[scale] G500 s19 ef32: nnz(A) {a.num_nnz}", flush=True)
    da = DeviceCsr.from_host(a)
    row_flp, P = gpu_ctx.row_flops(da, da)
    c = gpu_ctx.spgemm(da, da)
    torch.cuda.synchronize()
    ver = gpu_ctx.verify(c)
    print(f"[scale] P {P} nnz(C) {ver['nnz']} sum {ver['sum_val']}", flush=True)
    assert ver["nnz"] == c.num_nnz == int(c.row_ptr[-1].item())
    assert ver["unsorted_rows"] == 0
    assert ver["sum_val"] == float(P)
    flp = row_flp.cpu().numpy()
    rng = np.random.default_rng(3)
    rows = np.unique(np.concatenate([np.argsort(flp)[-8:], rng.choice(a.num_row, 32, replace=False)]))
    rp_dev = c.row_ptr.cpu().numpy()
    for r in rows.tolist():
        sub = a.row_slice(r, r + 1)
        orp, ocols, orats = oracle_product(sub, a)
        s, e = int(rp_dev[r]), int(rp_dev[r + 1])
        assert e - s == int(orp[-1]), f"row {r}: {e - s} entries, oracle {orp[-1]}"
        np.testing.assert_array_equal(c.col[s:e].cpu().numpy().astype(np.int64), ocols)
        np.testing.assert_array_equal(c.val[s:e].cpu().numpy(), orats)
    print(f"[scale] {len(rows)} sampled rows bit-exact (max row products {int(flp.max())})", flush=True)
    del c
    torch.cuda.empty_cache()
