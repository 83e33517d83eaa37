"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from
scipy's SpGEMM) against the oracle on the CPU and the HIP path on the GPU:
row pointers, sorted columns and values must match bit for bit (the
fixtures' values accumulate in the same A-row order as the reference's)."""
import glob
import os

import numpy as np
import pytest

from ams_hip.csr import Csr

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def load(path):
    z = np.load(path)
    a = Csr(int(z["a_shape"][0]), int(z["a_shape"][1]), z["a_rp"], z["a_col"], z["a_val"])
    b = Csr(int(z["b_shape"][0]), int(z["b_shape"][1]), z["b_rp"], z["b_col"], z["b_val"])
    return a, b, z["c_rp"], z["c_col"], z["c_val"]


def test_fixtures_present():
    assert len(FIXTURES) >= 3


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
@pytest.mark.parametrize("mode", ["hybacc", "hash", "esc"])
def test_oracle_matches_golden(path, mode):
    from oracle import oracle as orc
    a, b, crp, ccol, cval = load(path)
    rp, cols, rats, _ = orc.spgemm(a, b, mode, threads=2)
    np.testing.assert_array_equal(rp, crp)
    np.testing.assert_array_equal(cols, ccol)
    np.testing.assert_array_equal(rats, cval)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_matches_golden(gpu_ctx, path):
    import torch
    from ams_hip.device import DeviceCsr
    a, b, crp, ccol, cval = load(path)
    c = gpu_ctx.spgemm(DeviceCsr.from_host(a), DeviceCsr.from_host(b)).to_host()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(c.row_ptrs, crp)
    np.testing.assert_array_equal(c.cols, ccol)
    if np.all(a.rats == 1.0) and np.all(b.rats == 1.0):
        np.testing.assert_array_equal(c.rats, cval)
    else:
        np.testing.assert_allclose(c.rats, cval, rtol=1e-12, atol=1e-14)
