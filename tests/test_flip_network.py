"""The flip (mirror) bitonic network of the ESC kernels, replayed on the CPU.

`reg_flip_sort` (contention-aware-spgemm_amd/csrc/ams_device.hpp) sorts T*IPT
32-bit keys held IPT per thread (element e = tid*IPT + r) with the network of
the next power of two: merge size KS opens with a mirror stage (partner
e ^ (KS - 1)) followed by half-cleaners (partner e ^ J), and the lower
element of every pair takes the minimum.  Two skip rules make it cheap on
padded rows: a wave whose 64*IPT slots are all padding (0xffffffff, at and
past nact = f rounded up to whole waves) runs no stage, and a stage that
crosses waves leaves a wave's keys as they are when its partner wave lies in
the padding suffix or past T*IPT (the virtual last quarter of the 1536- and
3072-key bins).  This replays that schedule, with both rules, index by index
and checks the result against numpy's sort for every block shape the ESC
bins use and for fills from one key to full.  The GPU kernels themselves are
checked against the oracle by the -m gpu parity tests.
"""
import numpy as np
import pytest

WAVE = 64


def next_pow2(x):
    p = 1
    while p < x:
        p *= 2
    return p


def flip_sort(keys, T, IPT, f):
    N = T * IPT
    NV = next_pow2(N)
    WK = WAVE * IPT
    v = keys.copy()
    nact = (f + WK - 1) // WK * WK

    def stage(KS, J, mir):
        in_thread = (KS <= IPT) if mir else (J < IPT)
        in_wave = (KS <= WK) if mir else (J < WK)
        nv = v.copy()
        for e in range(N):
            w = e // WK
            if w * WK >= nact:  # a wave of padding only: skips the stage
                continue
            p = e ^ (KS - 1) if mir else e ^ J
            if in_thread or in_wave:
                assert p // WK == w  # the partner is a register or a lane of this wave
            else:
                pw = (w ^ (KS // WK - 1)) if mir else (w ^ (J // WK))
                assert p // WK == pw  # one partner wave for the whole wave
                if pw * WK >= nact:  # partner wave in the padding suffix (or virtual)
                    assert e < p
                    continue
            nv[e] = min(v[e], v[p]) if e < p else max(v[e], v[p])
        v[:] = nv

    KS = 2
    while KS <= NV:
        stage(KS, KS // 2, True)
        J = KS // 4
        while J >= 1:
            stage(KS, J, False)
            J //= 2
        KS *= 2
    return v


# (T, IPT): the ESC bins 64 (64x1), 128 (64x2), 256 (64x4), 512 (64x8),
# 1024 (128x8), 1536 (192x8), 2048 (256x8), 3072 (384x8)
SHAPES = [(64, 1), (64, 2), (64, 4), (64, 8), (128, 8), (192, 8), (256, 8), (384, 8)]


@pytest.mark.parametrize("T,IPT", SHAPES)
def test_flip_network_sorts_padded_rows(T, IPT):
    rng = np.random.default_rng(T * 31 + IPT)
    N = T * IPT
    fills = sorted({1, 2, N // 3, N // 2 + 1, N - 1, N, *[int(x) for x in rng.integers(1, N + 1, 3)]})
    for f in fills:
        # products sit in the kernel's strided slots: product t -> wave w, register s, lane l
        k = np.full(N, 0xFFFFFFFF, dtype=np.uint64)
        t = np.arange(f)
        w, rem = t // (IPT * WAVE), t % (IPT * WAVE)
        s, lane = rem // WAVE, rem % WAVE
        k[(w * WAVE + lane) * IPT + s] = rng.permutation(1 << 22)[:f]
        assert np.array_equal(flip_sort(k, T, IPT, f), np.sort(k)), (T, IPT, f)


def test_mirror_partners_are_single_dpp_patterns():
    """The mirror stages inside 16 lanes pair lane i with i ^ 1, 3, 7, 15:
    quad_perm [1,0,3,2], quad_perm [3,2,1,0], row_half_mirror and row_mirror."""
    lanes = np.arange(16)
    assert np.array_equal(lanes ^ 1, [(i // 4) * 4 + [1, 0, 3, 2][i % 4] for i in lanes])
    assert np.array_equal(lanes ^ 3, [(i // 4) * 4 + [3, 2, 1, 0][i % 4] for i in lanes])
    assert np.array_equal(lanes ^ 7, [(i // 8) * 8 + 7 - i % 8 for i in lanes])
    assert np.array_equal(lanes ^ 15, 15 - lanes)
