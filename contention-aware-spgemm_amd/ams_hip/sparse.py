"""torch.sparse_csr_tensor front end and the ``ams::`` torch.library operators.

    c = sparse.mm(a, b)                 # a, b: torch sparse CSR on one GPU
    c.values()                          # differentiable in a's and b's values

Layout at the boundary.  The library takes int64 row pointers, int32 columns
and float64/float32 values in HBM.  A sparse CSR tensor with int32 indices
hands over its columns and values without a copy; only its row pointers are
widened (m + 1 entries).  int64 column indices are narrowed once (a copy of
nnz entries), since every kernel reads 32-bit columns.  The product comes
back with int32 indices when nnz(C) < 2^31 (row pointers narrowed, columns
without a copy) and with int64 indices otherwise, because torch needs one
index dtype for both arrays.

Operators.  ``ams::spgemm`` (values of C = A * B plus its row pointers and
columns) and ``ams::spgemm_masked_bwd`` (dA = (G B^T) on pattern(A),
dB = (A^T G) on pattern(B)) are torch.library custom ops with fake
(meta) implementations, so they trace under torch.compile; the forward's
output length is data-dependent and is declared as an unbacked size.
``ams::spgemm`` carries the masked backward as its autograd formula.

The reference has no torch interface; its C entry point is hybaccspgemm
(src/ams_hybaccspgemm.c:36-220), which the same library calls serve.
Every computation is a call into libams_hip.so; CPU tensors are rejected.
"""
from __future__ import annotations

import torch

from .device import Context, DeviceCsr

_contexts: dict[tuple[int, int], Context] = {}


def context() -> Context:
    """The library context bound to the current stream of the current GPU
    (one per stream, created on first use and kept for the process)."""
    s = torch.cuda.current_stream()
    key = (s.device.index, s.cuda_stream)
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = _contexts[key] = Context(s)
    return ctx


def _on_gpu(*ts: torch.Tensor):
    for t in ts:
        if not t.is_cuda:
            raise ValueError("ams::spgemm runs on the GPU only; got a tensor on " + str(t.device))


def _check_operands(a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val):
    _on_gpu(a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val)
    for rp in (a_row_ptr, b_row_ptr):
        if rp.dtype != torch.int64 or rp.dim() != 1 or rp.numel() < 1:
            raise ValueError("row pointers must be a non-empty 1-D int64 tensor")
    for c in (a_col, b_col):
        if c.dtype != torch.int32 or c.dim() != 1:
            raise ValueError("columns must be a 1-D int32 tensor")
    if a_val.dtype != b_val.dtype or a_val.dtype not in (torch.float64, torch.float32):
        raise ValueError("values must be float64 or float32, the same for A and B")
    if a_val.numel() != a_col.numel() or b_val.numel() != b_col.numel():
        raise ValueError("values and columns differ in length")


def _csr(rp: torch.Tensor, col: torch.Tensor, val: torch.Tensor, num_col: int) -> DeviceCsr:
    return DeviceCsr(rp.numel() - 1, num_col, rp.contiguous(), col.contiguous(), val.contiguous())


@torch.library.custom_op("ams::spgemm", mutates_args=())
def spgemm_op(a_row_ptr: torch.Tensor, a_col: torch.Tensor, a_val: torch.Tensor, b_row_ptr: torch.Tensor,
              b_col: torch.Tensor, b_val: torch.Tensor, num_col: int) -> tuple[torch.Tensor, torch.Tensor,
                                                                              torch.Tensor]:
    """C = A * B for A (m x k) and B (k x num_col) in the library's CSR
    layout; returns C's (row_ptr int64, col int32, val), columns ascending
    in every row."""
    _check_operands(a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val)
    k = b_row_ptr.numel() - 1
    a = _csr(a_row_ptr, a_col, a_val, k)
    b = _csr(b_row_ptr, b_col, b_val, num_col)
    with torch.cuda.device(a_row_ptr.device):
        c = context().spgemm(a, b)
    return c.row_ptr, c.col, c.val


@spgemm_op.register_fake
def _spgemm_fake(a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val, num_col):
    nnz = torch.library.get_ctx().new_dynamic_size()
    return (a_row_ptr.new_empty(a_row_ptr.shape[0]), a_col.new_empty(nnz), a_val.new_empty(nnz))


@torch.library.custom_op("ams::spgemm_masked_bwd", mutates_args=())
def masked_bwd_op(a_row_ptr: torch.Tensor, a_col: torch.Tensor, a_val: torch.Tensor, b_row_ptr: torch.Tensor,
                  b_col: torch.Tensor, b_val: torch.Tensor, c_row_ptr: torch.Tensor, c_col: torch.Tensor,
                  grad: torch.Tensor, num_col: int) -> tuple[torch.Tensor, torch.Tensor]:
    """(dA, dB) on the patterns of A and B for the gradient ``grad`` on
    the pattern of C = A * B (columns ascending in every row)."""
    _check_operands(a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val)
    _on_gpu(c_row_ptr, c_col, grad)
    if grad.dtype != a_val.dtype or grad.numel() != c_col.numel():
        raise ValueError("grad must have C's length and the operands' dtype")
    k = b_row_ptr.numel() - 1
    a = _csr(a_row_ptr, a_col, a_val, k)
    b = _csr(b_row_ptr, b_col, b_val, num_col)
    g = _csr(c_row_ptr, c_col, grad, num_col)
    with torch.cuda.device(a_row_ptr.device):
        return context().masked_bwd(a, b, g)


@masked_bwd_op.register_fake
def _masked_bwd_fake(a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val, c_row_ptr, c_col, grad, num_col):
    return torch.empty_like(a_val), torch.empty_like(b_val)


def _setup_context(ctx, inputs, output):
    a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val, num_col = inputs
    c_row_ptr, c_col, _ = output
    ctx.save_for_backward(a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val, c_row_ptr, c_col)
    ctx.num_col = num_col


def _backward(ctx, _g_row_ptr, _g_col, g_val):
    a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val, c_row_ptr, c_col = ctx.saved_tensors
    if g_val is None:
        g_val = torch.zeros(c_col.shape[0], dtype=a_val.dtype, device=a_val.device)
    da, db = masked_bwd_op(a_row_ptr, a_col, a_val, b_row_ptr, b_col, b_val, c_row_ptr, c_col,
                           g_val.contiguous(), ctx.num_col)
    return None, None, da, None, None, db, None


spgemm_op.register_autograd(_backward, setup_context=_setup_context)


def from_torch(t: torch.Tensor) -> DeviceCsr:
    """Library CSR view of a 2-D torch sparse CSR tensor on a GPU: columns
    and values without a copy when the indices are int32, row pointers
    widened to int64.  The values keep their autograd history."""
    if t.layout != torch.sparse_csr:
        raise TypeError(f"expected a sparse CSR tensor, got layout {t.layout}")
    if t.dim() != 2:
        raise ValueError("only 2-D sparse CSR tensors (no batch or dense dimensions)")
    _on_gpu(t)
    m, n = t.shape
    if n >= 2**31:
        raise ValueError("column count must fit int32")
    crow, col, val = t.crow_indices(), t.col_indices(), t.values()
    if val.dtype not in (torch.float64, torch.float32):
        raise TypeError("values must be float64 or float32")
    rp = crow if crow.dtype == torch.int64 else crow.to(torch.int64)
    ci = col if col.dtype == torch.int32 else col.to(torch.int32)
    return DeviceCsr(m, n, rp.contiguous(), ci.contiguous(), val.contiguous())


def to_torch(c: DeviceCsr, index_dtype: torch.dtype | None = None) -> torch.Tensor:
    """Sparse CSR tensor of a library CSR.  int32 indices (columns without
    a copy) unless nnz(C) needs int64 or ``index_dtype`` asks for it."""
    want64 = index_dtype == torch.int64 or c.num_nnz >= 2**31
    if want64:
        crow, col = c.row_ptr, c.col.to(torch.int64)
    else:
        crow, col = c.row_ptr.to(torch.int32), c.col
    return torch.sparse_csr_tensor(crow, col, c.val, size=(c.num_row, c.num_col), check_invariants=False)


def mm(a: torch.Tensor, b: torch.Tensor, index_dtype: torch.dtype | None = None) -> torch.Tensor:
    """C = A * B for torch sparse CSR tensors A (m x k) and B (k x n) on one
    GPU, through ``ams::spgemm``.  Gradients reach the values of A and B."""
    if a.shape[1] != b.shape[0]:
        raise ValueError(f"inner dimensions differ: {tuple(a.shape)} x {tuple(b.shape)}")
    da, db = from_torch(a), from_torch(b)
    rp, col, val = spgemm_op(da.row_ptr, da.col, da.val, db.row_ptr, db.col, db.val, db.num_col)
    return to_torch(DeviceCsr(da.num_row, db.num_col, rp, col, val), index_dtype)
