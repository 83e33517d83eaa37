"""Differentiable SpGEMM: the masked-SpGEMM backward against (1) a reference
gradient built from the CPU oracle's products, dA = (G B^T) on pattern(A)
and dB = (A^T G) on pattern(B), and (2) FP64 central finite differences of
L = sum(W * C) on a scale-12 instance over sampled nonzeros of A and B."""
import numpy as np
import pytest

from ams_hip import gen
from ams_hip.csr import Csr
from helpers import oracle_product, random_csr

pytestmark = pytest.mark.gpu


def masked(prod_rp, prod_cols, prod_vals, pattern: Csr):
    """Values of a product at the entries of `pattern` (0 where absent)."""
    out = np.zeros(pattern.num_nnz)
    for i in range(pattern.num_row):
        s, e = prod_rp[i], prod_rp[i + 1]
        lut = dict(zip(prod_cols[s:e].tolist(), prod_vals[s:e].tolist()))
        for q in range(pattern.row_ptrs[i], pattern.row_ptrs[i + 1]):
            out[q] = lut.get(int(pattern.cols[q]), 0.0)
    return out


def reference_grads(a: Csr, b: Csr, g: Csr):
    bt = b.transpose()
    at = a.transpose()
    rp, cols, vals = oracle_product(g, bt)
    da = masked(rp, cols, vals, a)
    rp, cols, vals = oracle_product(at, g)
    db = masked(rp, cols, vals, b)
    return da, db


def run_bwd(ctx, a, b, params, dtype=np.float64, seed=0):
    import torch
    from ams_hip.device import DeviceCsr
    from ams_hip._lib import AmsParams, lib
    p = AmsParams()
    lib().ams_default_params(p)
    ctx.set_params(**{**p.as_dict(), **params})
    da = DeviceCsr.from_host(a, val_dtype=dtype)
    db = DeviceCsr.from_host(b, val_dtype=dtype)
    c = ctx.spgemm(da, db)
    rng = np.random.default_rng(seed)
    gval = torch.from_numpy(rng.standard_normal(c.num_nnz).astype(dtype)).to(c.val.device)
    g = DeviceCsr(c.num_row, c.num_col, c.row_ptr, c.col, gval)
    dA, dB = ctx.masked_bwd(da, db, g)
    torch.cuda.synchronize()
    gh = Csr(c.num_row, c.num_col, c.row_ptr.cpu().numpy(), c.col.cpu().numpy(), gval.cpu().numpy().astype(np.float64))
    return dA.cpu().numpy(), dB.cpu().numpy(), gh


ROUTES = {
    "default": {},
    "all_dense": dict(sym_hash_max=1, num_hash_max=1, bitmap_ratio=0.0, dense_width=128, bitmap_bits=1024,
                      dense_density=0.0),
    "small_tables": dict(sym_hash_max=64, num_hash_max=64),
}


@pytest.mark.parametrize("route", list(ROUTES))
def test_backward_matches_oracle_gradient(gpu_ctx, route):
    a = gen.rmat_symmetric(10, 8, seed=21, values="normal")
    b = a.transpose()
    dA, dB, g = run_bwd(gpu_ctx, a, b, ROUTES[route])
    ra, rb = reference_grads(a, b, g)
    # bound: the same sums with absolute values
    aa = Csr(a.num_row, a.num_col, a.row_ptrs, a.cols, np.abs(a.rats))
    ab = Csr(b.num_row, b.num_col, b.row_ptrs, b.cols, np.abs(b.rats))
    ag = Csr(g.num_row, g.num_col, g.row_ptrs, g.cols, np.abs(g.rats))
    ba, bb = reference_grads(aa, ab, ag)
    assert np.all(np.abs(dA - ra) <= 1e-12 * ba + 1e-300)
    assert np.all(np.abs(dB - rb) <= 1e-12 * bb + 1e-300)


def test_backward_rectangular(gpu_ctx):
    a = random_csr(60, 90, 0.08, seed=3)
    b = random_csr(90, 700, 0.03, seed=4)
    for route in ("default", "all_dense"):
        dA, dB, g = run_bwd(gpu_ctx, a, b, ROUTES[route])
        ra, rb = reference_grads(a, b, g)
        np.testing.assert_allclose(dA, ra, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(dB, rb, rtol=1e-11, atol=1e-12)


def test_backward_fp32_error_vs_fp64(gpu_ctx):
    a = gen.rmat_symmetric(10, 16, seed=5, values="normal")
    b = a.transpose()
    d32a, d32b, g = run_bwd(gpu_ctx, a, b, {}, dtype=np.float32, seed=1)
    ra, rb = reference_grads(a, b, Csr(g.num_row, g.num_col, g.row_ptrs, g.cols, g.rats))
    scale_a = np.abs(ra).max()
    scale_b = np.abs(rb).max()
    assert np.abs(d32a - ra).max() <= 1e-4 * scale_a
    assert np.abs(d32b - rb).max() <= 1e-4 * scale_b


def test_autograd_finite_differences_s12(gpu_ctx):
    """FP64 central differences of L = sum(W * C) over 1000 sampled nonzeros
    (500 of A, 500 of B) of a scale-12 skewed R-MAT, B = A^T as its own input.
    L is affine in any single entry of A or B, so the central difference has
    no truncation error for any step; the step only divides the rounding of
    L (about log2(nnz) * 1e-16 * sum |W C|, ~1e-8 here), hence a large step."""
    import torch
    from ams_hip import autograd
    from ams_hip.device import DeviceCsr
    a = gen.rmat_symmetric(12, 16, seed=31, values="normal")
    b = a.transpose()
    A = DeviceCsr.from_host(a)
    B = DeviceCsr.from_host(b)
    A.val.requires_grad_(True)
    B.val.requires_grad_(True)
    C = autograd.spgemm(A, B, gpu_ctx)
    w = torch.randn(C.num_nnz, dtype=torch.float64, device=C.val.device,
                    generator=torch.Generator(device=C.val.device).manual_seed(7))
    loss = (C.val * w).sum()
    loss.backward()
    ga, gb = A.val.grad.detach().clone(), B.val.grad.detach().clone()

    def loss_at(av, bv):
        with torch.no_grad():
            Cp = gpu_ctx.spgemm(DeviceCsr(A.num_row, A.num_col, A.row_ptr, A.col, av),
                                DeviceCsr(B.num_row, B.num_col, B.row_ptr, B.col, bv))
            assert torch.equal(Cp.col, C.col)
            return float((Cp.val * w).sum())

    rng = np.random.default_rng(0)
    eps = 0.5
    worst = 0.0
    for which, n in (("a", A.num_nnz), ("b", B.num_nnz)):
        for q in rng.choice(n, 500, replace=False):
            av = A.val.detach().clone()
            bv = B.val.detach().clone()
            tgt = av if which == "a" else bv
            x0 = float(tgt[q])
            tgt[q] = x0 + eps
            lp = loss_at(av, bv)
            tgt[q] = x0 - eps
            lm = loss_at(av, bv)
            fd = (lp - lm) / (2 * eps)
            an = float((ga if which == "a" else gb)[q])
            err = abs(fd - an) / max(1.0, abs(an))
            worst = max(worst, err)
    assert worst < 1e-7, worst
