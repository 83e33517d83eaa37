"""Host-side checks of the ams:: torch.library operators (no GPU needed):
their fake implementations give the output shapes and dtypes torch.compile
traces with, and tensors off the GPU are rejected before any library call."""
import pytest


def fake_operands(mode):
    import torch
    with mode:
        rp = torch.empty(5, dtype=torch.int64, device="cuda")
        col = torch.empty(7, dtype=torch.int32, device="cuda")
        val = torch.empty(7, dtype=torch.float64, device="cuda")
    return rp, col, val


def test_fake_shapes():
    import torch
    from torch._subclasses.fake_tensor import FakeTensorMode
    from torch.fx.experimental.symbolic_shapes import ShapeEnv

    from ams_hip import sparse  # noqa: F401  (registers the operators)
    mode = FakeTensorMode(shape_env=ShapeEnv())
    rp, col, val = fake_operands(mode)
    with mode:
        c_rp, c_col, c_val = torch.ops.ams.spgemm(rp, col, val, rp, col, val, 4)
        assert c_rp.shape == (5,) and c_rp.dtype == torch.int64
        assert c_col.dtype == torch.int32 and c_val.dtype == torch.float64
        assert c_col.shape == c_val.shape and not isinstance(c_col.shape[0], int)  # data-dependent nnz
        da, db = torch.ops.ams.spgemm_masked_bwd(rp, col, val, rp, col, val, c_rp, c_col, c_val, 4)
        assert da.shape == val.shape and db.shape == val.shape and da.dtype == torch.float64


def test_cpu_tensors_rejected():
    import torch

    from ams_hip import sparse
    rp = torch.zeros(2, dtype=torch.int64)
    col = torch.zeros(0, dtype=torch.int32)
    val = torch.zeros(0, dtype=torch.float64)
    with pytest.raises(ValueError, match="GPU only"):
        sparse.spgemm_op(rp, col, val, rp, col, val, 1)
    with pytest.raises(ValueError, match="GPU only"):
        sparse.masked_bwd_op(rp, col, val, rp, col, val, rp, col, val, 1)
    t = torch.sparse_csr_tensor(rp.to(torch.int32), col, val, size=(1, 1))
    with pytest.raises(ValueError, match="GPU only"):
        sparse.from_torch(t)
