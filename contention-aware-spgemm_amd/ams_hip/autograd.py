"""Differentiable SpGEMM: values of C = A * B as a function of the values of
A and B, with the masked-SpGEMM backward of the C ABI
(ams_spgemm_masked_bwd).  The patterns are fixed; gradients flow to
``a.val`` and ``b.val``.

    c = spgemm(a, b)            # a.val / b.val may require grad
    loss = (c.val * w).sum()
    loss.backward()             # a.val.grad = (G B^T) on pattern(A), ...

The reference has no differentiable form; this wraps the same library calls
as the plain product.
"""
from __future__ import annotations

import torch

from .device import Context, DeviceCsr


def default_context() -> Context:
    """The library context of the current stream (one per stream), so the
    product runs in order with the caller's work on that stream."""
    from .sparse import context
    return context()


class _SpGEMM(torch.autograd.Function):
    @staticmethod
    def forward(fctx, a_val, b_val, a, b, ams):
        A = DeviceCsr(a.num_row, a.num_col, a.row_ptr, a.col, a_val.detach())
        B = DeviceCsr(b.num_row, b.num_col, b.row_ptr, b.col, b_val.detach())
        C = ams.spgemm(A, B)
        fctx.ams = ams
        fctx.A, fctx.B = A, B
        fctx.c_row_ptr, fctx.c_col = C.row_ptr, C.col
        fctx.mark_non_differentiable(C.row_ptr, C.col)
        return C.val, C.row_ptr, C.col

    @staticmethod
    def backward(fctx, grad_val, _grad_rp, _grad_col):
        G = DeviceCsr(fctx.A.num_row, fctx.B.num_col, fctx.c_row_ptr, fctx.c_col, grad_val.contiguous())
        da, db = fctx.ams.masked_bwd(fctx.A, fctx.B, G)
        return da, db, None, None, None


def spgemm(a: DeviceCsr, b: DeviceCsr, ams: Context | None = None) -> DeviceCsr:
    """C = A * B; C.val carries the autograd graph back to a.val and b.val."""
    ams = ams or default_context()
    val, rp, col = _SpGEMM.apply(a.val, b.val, a, b, ams)
    return DeviceCsr(a.num_row, b.num_col, rp, col, val)
