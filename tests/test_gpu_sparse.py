"""torch sparse CSR front end and the ams:: torch.library operators
(ams_hip/sparse.py): products against the CPU oracle to the usual bar
(row_ptr and columns bit-exact, values bit-exact on all-ones inputs and
within 1e-12 * (|A||B|)_ij otherwise), zero-copy hand-over of int32-indexed
tensors, gradients that agree with the autograd.Function form to the same
bound, torch.library's opcheck, and forward + backward under torch.compile."""
import numpy as np
import pytest

from ams_hip import gen
from ams_hip.csr import Csr
from helpers import assert_matches_oracle, random_csr

pytestmark = pytest.mark.gpu


def torch_csr(m: Csr, index_dtype, requires_grad=False):
    import torch
    t = torch.sparse_csr_tensor(torch.from_numpy(np.asarray(m.row_ptrs)).to(index_dtype),
                                torch.from_numpy(np.asarray(m.cols)).to(index_dtype),
                                torch.from_numpy(np.asarray(m.rats)), size=(m.num_row, m.num_col))
    t = t.to("cuda")
    return t.requires_grad_() if requires_grad else t


def host(t) -> Csr:
    return Csr(t.shape[0], t.shape[1], t.crow_indices().cpu().numpy().astype(np.int64),
               t.col_indices().cpu().numpy().astype(np.int32), t.values().detach().cpu().numpy())


@pytest.mark.parametrize("index_dtype", ["int32", "int64"])
@pytest.mark.parametrize("case", ["rmat_ones", "random", "rect_unsorted"])
def test_mm_matches_oracle(gpu_ctx, case, index_dtype):
    import torch
    from ams_hip import sparse
    idt = getattr(torch, index_dtype)
    if case == "rmat_ones":
        a = b = gen.rmat_symmetric(11, 8, gen.G500, seed=3)
        exact = True
    elif case == "random":
        a = random_csr(300, 300, 0.03, seed=5)
        b = random_csr(300, 300, 0.03, seed=6)
        exact = False
    else:
        a = random_csr(257, 130, 0.05, seed=7, sort=False)
        b = random_csr(130, 411, 0.04, seed=8, sort=False)
        exact = False
    print(f"[sparse] {case} {index_dtype}", flush=True)
    c = sparse.mm(torch_csr(a, idt), torch_csr(b, idt))
    assert c.layout == torch.sparse_csr and c.crow_indices().dtype == torch.int32
    assert tuple(c.shape) == (a.num_row, b.num_col)
    assert_matches_oracle(host(c), a, b, exact=exact)


def test_zero_copy_and_index_dtypes(gpu_ctx):
    import torch
    from ams_hip import sparse
    a = random_csr(64, 64, 0.1, seed=1)
    t = torch_csr(a, torch.int32)
    d = sparse.from_torch(t)
    assert d.col.data_ptr() == t.col_indices().data_ptr()
    assert d.val.data_ptr() == t.values().data_ptr()
    assert d.row_ptr.dtype == torch.int64
    c = gpu_ctx.spgemm(d, d)
    out = sparse.to_torch(c)
    assert out.col_indices().data_ptr() == c.col.data_ptr()
    assert out.values().data_ptr() == c.val.data_ptr()
    out64 = sparse.to_torch(c, torch.int64)
    assert out64.crow_indices().dtype == torch.int64 and out64.col_indices().dtype == torch.int64
    assert torch.equal(out64.col_indices().to(torch.int32), c.col)


def test_empty_and_mismatch(gpu_ctx):
    import torch
    from ams_hip import sparse
    z = Csr(5, 7, np.zeros(6, np.int64), np.zeros(0, np.int32), np.zeros(0))
    b = random_csr(7, 3, 0.5, seed=2)
    c = sparse.mm(torch_csr(z, torch.int32), torch_csr(b, torch.int32))
    assert tuple(c.shape) == (5, 3) and c.values().numel() == 0
    assert torch.equal(c.crow_indices().cpu(), torch.zeros(6, dtype=torch.int32))
    with pytest.raises(ValueError):
        sparse.mm(torch_csr(b, torch.int32), torch_csr(b, torch.int32))
    with pytest.raises(TypeError):
        sparse.from_torch(torch.zeros(3, 3, device="cuda"))


def grads_function(a: Csr, b: Csr, w_seed=0):
    """Gradients of L = sum(w * C.val) through the autograd.Function form."""
    import torch
    from ams_hip import autograd
    from ams_hip.device import DeviceCsr
    da, db = DeviceCsr.from_host(a), DeviceCsr.from_host(b)
    da.val.requires_grad_()
    db.val.requires_grad_()
    c = autograd.spgemm(da, db, gpu_ctx_default())
    w = torch.from_numpy(np.random.default_rng(w_seed).uniform(-1, 1, c.num_nnz)).cuda()
    (c.val * w).sum().backward()
    return c, w, da.val.grad.cpu().numpy(), db.val.grad.cpu().numpy()


def gpu_ctx_default():
    from ams_hip import sparse
    return sparse.context()


def dense(m: Csr, vals=None):
    out = np.zeros((m.num_row, m.num_col))
    rows = np.repeat(np.arange(m.num_row), np.diff(np.asarray(m.row_ptrs)))
    out[rows, np.asarray(m.cols)] += m.rats if vals is None else vals
    return out, rows


def assert_within(got, want, bound, rel=1e-12):
    """Hash and dense rows sum with LDS atomics, so two runs may differ in
    the last bits: |got - want| <= rel * (the entry's |.||.| bound)."""
    err = np.abs(np.asarray(got) - np.asarray(want))
    assert np.all(err <= rel * bound + 1e-300), f"worst {err.max()}"


def test_sparse_autograd_equals_function(gpu_ctx):
    import torch
    from ams_hip import sparse
    a = random_csr(200, 150, 0.04, seed=11)
    b = random_csr(150, 180, 0.04, seed=12)
    c_ref, w, ga, gb = grads_function(a, b)
    ta, tb = torch_csr(a, torch.int32, True), torch_csr(b, torch.int32, True)
    c = sparse.mm(ta, tb)
    assert torch.equal(c.crow_indices().to(torch.int64), c_ref.row_ptr)
    assert torch.equal(c.col_indices(), c_ref.col)
    ad, arows = dense(a)
    bd, brows = dense(b)
    hc = host(c)
    gd, crows = dense(hc, np.abs(w.cpu().numpy()))
    cbound = (np.abs(ad) @ np.abs(bd))[crows, hc.cols]
    assert_within(c.values().detach().cpu().numpy(), c_ref.val.detach().cpu().numpy(), cbound)
    (c.values() * w).sum().backward()
    # dA = (G B^T) on pattern(A), dB = (A^T G) on pattern(B)
    assert_within(ta.grad.values().cpu().numpy(), ga, (gd @ np.abs(bd).T)[arows, a.cols])
    assert_within(tb.grad.values().cpu().numpy(), gb, (np.abs(ad).T @ gd)[brows, b.cols])
    # the gradient keeps the operand's pattern
    assert torch.equal(ta.grad.col_indices(), ta.col_indices())


def op_args(a: Csr, b: Csr, requires_grad=False):
    import torch
    from ams_hip.device import DeviceCsr
    da, db = DeviceCsr.from_host(a), DeviceCsr.from_host(b)
    return (da.row_ptr, da.col, da.val.requires_grad_(requires_grad), db.row_ptr, db.col,
            db.val.requires_grad_(requires_grad), b.num_col)


def test_opcheck(gpu_ctx):
    import torch
    from ams_hip import sparse
    a = random_csr(40, 30, 0.1, seed=21)
    b = random_csr(30, 50, 0.1, seed=22)
    for rg in (False, True):
        torch.library.opcheck(sparse.spgemm_op, op_args(a, b, rg),
                              test_utils=("test_schema", "test_autograd_registration", "test_faketensor"))


def test_compile_forward_backward(gpu_ctx):
    import torch
    from ams_hip import sparse
    a = random_csr(120, 90, 0.06, seed=31)
    b = random_csr(90, 110, 0.06, seed=32)

    def loss(a_rp, a_col, a_val, b_rp, b_col, b_val, n):
        _, _, v = sparse.spgemm_op(a_rp, a_col, a_val, b_rp, b_col, b_val, n)
        return (v * v).sum()

    args = op_args(a, b, True)
    eager = loss(*args)
    eager.backward()
    ga, gb = args[2].grad.clone(), args[5].grad.clone()
    args[2].grad = None
    args[5].grad = None
    compiled = torch.compile(loss, backend="aot_eager", fullgraph=True)
    out = compiled(*args)
    out.backward()
    # atomically accumulated rows may differ in the last bits between runs
    torch.testing.assert_close(out.detach(), eager.detach(), rtol=1e-12, atol=0)
    torch.testing.assert_close(args[2].grad, ga, rtol=1e-12, atol=1e-13)
    torch.testing.assert_close(args[5].grad, gb, rtol=1e-12, atol=1e-13)
