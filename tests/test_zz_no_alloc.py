"""ams_spgemm_fixed, the product that allocates nothing (all scratch in the
library's fixed per-GPU array), called through the C ABI and checked against
the one-shot ams_spgemm on the same inputs: row pointers and columns
bit-exact, values bit-exact for all-ones inputs and within
1e-12 * (|A| |B|)_ij for random FP64.  Two calls in a row on one context,
the second smaller, so the second runs over the first one's leftovers in
the shared array."""
import ctypes as C

import numpy as np
import pytest

from ams_hip import gen
from helpers import abs_csr, oracle_product

pytestmark = pytest.mark.gpu


def fixed_product(ctx, da, db, cap=None):
    """Calls ams_spgemm_fixed; returns (rc, row_ptr, col, val, nnz) with the
    device outputs copied to the host (col/val cut to nnz when they fit)."""
    import torch
    from ams_hip import _lib as L
    dev = da.row_ptr.device
    m = da.num_row
    rp = torch.full((m + 1,), -7, dtype=torch.int64, device=dev)
    nnz = torch.full((1,), -7, dtype=torch.int64, device=dev)
    cap = 0 if cap is None else cap
    col = torch.full((max(cap, 1),), -7, dtype=torch.int32, device=dev)
    val = torch.full((max(cap, 1),), float("nan"), dtype=da.val.dtype, device=dev)
    sa, sb = da.c_struct(), db.c_struct()
    rc = L.lib().ams_spgemm_fixed(ctx.handle, C.byref(sa), C.byref(sb), da.dtype_code, rp.data_ptr(),
                                  col.data_ptr() if cap else None, val.data_ptr() if cap else None, cap,
                                  nnz.data_ptr())
    torch.cuda.synchronize()
    n = int(nnz.cpu()[0]) if rc == 0 else -1
    return rc, rp.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy(), n


def check_against_one_shot(ctx, a, exact):
    import torch
    from ams_hip.device import DeviceCsr
    da = DeviceCsr.from_host(a)
    torch.cuda.synchronize()
    # capacity query: no entries written, nnz reported
    rc, rp0, col0, _, nnz = fixed_product(ctx, da, da, cap=0)
    assert rc == 0, ctx._lib.ams_last_error(ctx.handle)
    assert ctx.error_flags() == 0
    assert (col0 == -7).all()
    rc, rp, col, val, nnz2 = fixed_product(ctx, da, da, cap=nnz)
    assert rc == 0, ctx._lib.ams_last_error(ctx.handle)
    assert nnz2 == nnz and ctx.error_flags() == 0
    np.testing.assert_array_equal(rp, rp0)
    ref = ctx.spgemm(da, da).to_host()
    assert nnz == ref.num_nnz
    np.testing.assert_array_equal(rp, np.asarray(ref.row_ptrs, np.int64))
    np.testing.assert_array_equal(col[:nnz], np.asarray(ref.cols))
    if exact:
        np.testing.assert_array_equal(val[:nnz], np.asarray(ref.rats))
    else:
        _, _, bound = oracle_product(abs_csr(a), abs_csr(a))
        err = np.abs(val[:nnz] - np.asarray(ref.rats))
        assert (err <= 1e-12 * bound + 1e-300).all(), float(err.max())


def test_fixed_matches_one_shot_twice(gpu_ctx):
    from ams_hip.device import Context
    ctx = Context()
    a1 = gen.rmat_symmetric(14, 16, seed=91, values="uniform")
    a2 = gen.rmat_symmetric(10, 8, seed=92)
    print(f"[no_alloc] call 1: G500 s14 ef16 uniform, nnz(A) {a1.num_nnz}", flush=True)
    check_against_one_shot(ctx, a1, exact=False)
    print(f"[no_alloc] call 2: G500 s10 ef8 ones, nnz(A) {a2.num_nnz}", flush=True)
    check_against_one_shot(ctx, a2, exact=True)
    ctx.close()


def test_fixed_rejects_oversized(gpu_ctx):
    """A product whose scratch exceeds the fixed array is refused before
    any work is queued (here: 2^20 empty rows over 2^20 columns)."""
    import torch
    from ams_hip import _lib as L
    from ams_hip.device import Context, DeviceCsr
    ctx = Context()
    n = 1 << 20
    z = DeviceCsr(n, n, torch.zeros(n + 1, dtype=torch.int64, device="cuda"),
                  torch.zeros(0, dtype=torch.int32, device="cuda"),
                  torch.zeros(0, dtype=torch.float64, device="cuda"))
    rc, *_ = fixed_product(ctx, z, z, cap=0)
    assert rc == L.AMS_ERR_LIMIT
    assert b"fixed scratch" in ctx._lib.ams_last_error(ctx.handle)
    ctx.close()
