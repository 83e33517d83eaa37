"""Autotuner: it times candidate routing parameters on the given product,
installs the fastest, and the product under the tuned parameters still
matches the oracle."""
import numpy as np
import pytest

from ams_hip import gen
from helpers import assert_matches_oracle

pytestmark = pytest.mark.gpu


def test_autotune_installs_valid_params():
    import torch
    from ams_hip.device import Context, DeviceCsr
    a = gen.rmat_symmetric(11, 16, seed=41)
    ctx = Context()
    A = DeviceCsr.from_host(a)
    before = ctx.params()
    best, ms = ctx.autotune(A, A, reps=1)
    assert ms > 0
    after = ctx.params()
    assert after == best
    assert 1 <= best["sym_hash_max"] <= 5461 and 1 <= best["num_hash_max"] <= 5461
    assert best["dense_width"] % 64 == 0 and best["esc_sort"] in (0, 1, 2, 3, 4)
    assert set(before) == set(after)
    c = ctx.spgemm(A, A).to_host()
    torch.cuda.synchronize()
    assert_matches_oracle(c, a, a, exact=True)
    ctx.close()
