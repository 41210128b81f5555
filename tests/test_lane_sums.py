"""CPU check of the 16-lane block sum of the level-0 SAP (lanes.h seg16_sum): partners
from DPP row rotations by 8, 4, 2, 1 (lane i receives lane (i - K) mod 16) give, in
every lane, exactly the bits of the butterfly with partners lane ^ 8, ^ 4, ^ 2, ^ 1
(the order of wave_sum(v, 16)), because after the step with rotation K the row is
K-periodic.  Both are emulated here with IEEE float32 and float64 adds (round to
nearest even, as on the GPU), each lane adding (own, partner)."""
import numpy as np
import pytest


def xor_butterfly(v):
    lanes = np.arange(16)
    for k in (8, 4, 2, 1):
        v = v + v[..., lanes ^ k]
    return v


def rotation_butterfly(v):
    lanes = np.arange(16)
    for k in (8, 4, 2, 1):
        v = v + v[..., (lanes - k) % 16]
    return v


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_rotation_sum_equals_xor_butterfly(dtype):
    rng = np.random.default_rng(16)
    # magnitudes over many decades, so that the rounding of every add matters
    v = (rng.standard_normal((20000, 16)) * 10.0 ** rng.uniform(-6, 6, (20000, 16))).astype(dtype)
    a, b = xor_butterfly(v), rotation_butterfly(v)
    assert a.dtype == dtype and b.dtype == dtype
    assert np.array_equal(a.view(np.uint32 if dtype == np.float32 else np.uint64),
                          b.view(np.uint32 if dtype == np.float32 else np.uint64))
    # every lane of a row holds the same bits
    assert np.array_equal(b, np.repeat(b[:, :1], 16, axis=1))


def test_sum_order_matters():
    """The check above is not vacuous: a sum in another order (here left to right) differs
    from the butterfly's in the last bits for some rows."""
    rng = np.random.default_rng(17)
    v = (rng.standard_normal((20000, 16)) * 10.0 ** rng.uniform(-6, 6, (20000, 16))).astype(np.float32)
    seq = np.zeros(v.shape[0], np.float32)
    for i in range(16):
        seq = seq + v[:, i]
    assert np.any(seq != xor_butterfly(v)[:, 0])
