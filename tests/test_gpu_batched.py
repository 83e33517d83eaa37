"""Batched form: a stacked batch replayed from one captured hipGraph must
give every problem's product as the oracle does, on every replay."""
import numpy as np
import pytest

from ams_hip import gen
from helpers import assert_matches_oracle, random_csr

pytestmark = pytest.mark.gpu


def small_batch(count, seed):
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        s = int(rng.integers(6, 11))
        ef = int(rng.choice([2, 4, 8, 16]))
        a = gen.rmat_symmetric(s, ef, seed=seed * 1000 + i, values="uniform")
        probs.append((a, a))
    probs.append((random_csr(7, 11, 0.3, seed=1), random_csr(11, 300, 0.05, seed=2)))
    z = random_csr(5, 6, 0.0, seed=3)
    probs.append((z, random_csr(6, 9, 0.5, seed=4)))
    return probs


@pytest.mark.parametrize("params", [{}, dict(sym_hash_max=16, num_hash_max=16, bitmap_ratio=0.0,
                                              dense_width=64, bitmap_bits=256, dense_density=0.0),
                                    dict(sym_blocked=1, sym_hash_max=16, num_hash_max=16, bitmap_ratio=0.0,
                                         dense_width=64, bitmap_bits=256, dense_density=0.0),
                                    dict(sym_hash_max=1, bitmap_ratio=0.0)],
                         ids=["default", "all_dense", "blocked_sym", "bitmap_rows"])
def test_batched_graph_replays(gpu_ctx, params):
    import torch
    from ams_hip.batched import BatchedProduct
    from ams_hip.device import Context
    probs = small_batch(40, seed=5)
    stream = torch.cuda.Stream()
    ctx = Context(stream=stream)
    if params:
        ctx.set_params(**params)
    bp = BatchedProduct(probs, ams=ctx)
    for rep in range(3):
        with torch.cuda.stream(stream):
            bp.c_val.fill_(float("nan"))
            bp.run()
        res = bp.results()
        assert int(bp.nnz.cpu()[0]) == bp.cap
        for (a, b), c in zip(probs, res):
            assert c.sorted_rows()
            assert_matches_oracle(c, a, b, exact=False)
    bp.close()
    ctx.close()


def test_batched_matches_plain(gpu_ctx):
    import torch
    from ams_hip.batched import BatchedProduct
    from ams_hip.device import DeviceCsr
    probs = small_batch(20, seed=9)
    bp = BatchedProduct(probs)
    bp.run()
    res = bp.results()
    for (a, b), c in zip(probs, res):
        p = gpu_ctx.spgemm(DeviceCsr.from_host(a), DeviceCsr.from_host(b)).to_host()
        np.testing.assert_array_equal(c.row_ptrs, p.row_ptrs)
        np.testing.assert_array_equal(c.cols, p.cols)
        np.testing.assert_allclose(c.rats, p.rats, rtol=1e-12, atol=1e-13)
    torch.cuda.synchronize()
    bp.close()


def test_batched_autotune_keeps_results(gpu_ctx):
    """The batched autotuner times planned runs over its search space, leaves
    the best parameters on the context, and the batch planned with them
    still matches the oracle."""
    import torch
    from ams_hip.batched import BatchedProduct, Stacked, autotune_batched
    from ams_hip.device import Context
    probs = small_batch(12, seed=21)
    stream = torch.cuda.Stream()
    ctx = Context(stream=stream)
    stacked = Stacked(probs)
    best, ms = autotune_batched(probs, ctx, reps=1, stacked=stacked)
    assert ms > 0
    assert ctx.params() == best
    bp = BatchedProduct(probs, ams=ctx, stacked=stacked)
    bp.run()
    for (a, b), c in zip(probs, bp.results()):
        assert c.sorted_rows()
        assert_matches_oracle(c, a, b, exact=False)
    bp.close()
    ctx.close()


def test_batched_heavy_rows(gpu_ctx):
    """A batch holding a product with rows of more than max(2^20, P/2048)
    products (a join-shaped C = A^T A over 300 columns) between small
    ones: the heavy rows are counted in pieces by the batch's narrow bitmap
    blocks, each problem at its own column offset, on every replay."""
    import torch
    from ams_hip.batched import BatchedProduct
    from ams_hip.device import Context
    inc = gen.join_incidence(num_row=60000, num_col=300, mu=3.5, sigma=0.5, cap=100, seed=5)
    probs = small_batch(6, seed=31)
    probs.insert(3, (inc.transpose(), inc))
    stream = torch.cuda.Stream()
    ctx = Context(stream=stream)
    bp = BatchedProduct(probs, ams=ctx)
    for rep in range(2):
        with torch.cuda.stream(stream):
            bp.c_val.fill_(float("nan"))
            bp.run()
        res = bp.results()
        for (a, b), c in zip(probs, res):
            assert c.sorted_rows()
            assert_matches_oracle(c, a, b, exact=False)
    bp.close()
    ctx.close()


def test_batched_mixed_widths(gpu_ctx):
    """A batch that stacks one problem wider than 16384 columns (R-MAT s15,
    32768 columns) with narrow s8-s10 problems: the narrow problems' light
    bitmap rows are counted by the wave kernel, which must run although the
    widest problem is too wide for it; every product against the oracle."""
    import torch
    from ams_hip.batched import BatchedProduct
    from ams_hip.device import Context
    probs = []
    for i, (s, ef) in enumerate([(8, 16), (9, 32), (15, 4), (10, 16), (8, 32)]):
        a = gen.rmat_symmetric(s, ef, seed=700 + i, values="uniform")
        probs.append((a, a))
    stream = torch.cuda.Stream()
    ctx = Context(stream=stream)
    ctx.set_params(bitmap_ratio=1.0)
    bp = BatchedProduct(probs, ams=ctx)
    print(f"[batched] mixed widths s8..s15 cap {bp.cap}", flush=True)
    for rep in range(2):
        with torch.cuda.stream(stream):
            bp.c_val.fill_(float("nan"))
            bp.run()
        res = bp.results()
        assert int(bp.nnz.cpu()[0]) == bp.cap
        for (a, b), c in zip(probs, res):
            assert c.sorted_rows()
            assert_matches_oracle(c, a, b, exact=False)
    bp.close()
    ctx.close()


def test_stale_graph_rejected(gpu_ctx):
    """A captured batch replays only while the context's scratch stays put:
    a larger product on the same context reallocates buffers and a new plan
    rewrites the problem tables, after which the old graph is refused."""
    import torch
    from ams_hip import _lib as L
    from ams_hip.batched import BatchedProduct
    from ams_hip.device import Context, DeviceCsr
    probs = small_batch(6, seed=41)
    stream = torch.cuda.Stream()
    ctx = Context(stream=stream)
    bp = BatchedProduct(probs, ams=ctx)
    bp.run()
    bp.results()
    big = gen.rmat_symmetric(14, 16, seed=5)
    with torch.cuda.stream(stream):
        d = DeviceCsr.from_host(big)
    stream.synchronize()
    ctx.spgemm_pooled(d, d)
    with pytest.raises(L.AmsError, match="stale"):
        bp.run()
    bp2 = BatchedProduct(probs, ams=ctx)  # re-plans and captures afresh
    bp2.run()
    for (a, b), c in zip(probs, bp2.results()):
        assert_matches_oracle(c, a, b, exact=False)
    with pytest.raises(L.AmsError, match="stale"):
        bp.run()
    bp.close()
    bp2.close()
    ctx.close()
