"""The FP32 flux denominators are divided as a double product with the
double reciprocal, rounded once (csrc/common.h div_const).  That equals
IEEE float division x / 60 and x / 12 for every float input (an
exhaustive sweep of all 2^32 bit patterns found no difference); here a
strided sweep of the bit patterns, seeded random inputs and the special
values, checked bit for bit (the device step's own parity tests then pin
the kernels that use it)."""
import numpy as np
import pytest


def _check(x, c):
    with np.errstate(all="ignore"):
        ref = x / np.float32(c)
        alt = (x.astype(np.float64) * (1.0 / c)).astype(np.float32)
    same = (ref.view(np.uint32) == alt.view(np.uint32)) | (np.isnan(ref)
                                                            & np.isnan(alt))
    return int((~same).sum())


@pytest.mark.parametrize("c", [60, 12])
def test_double_product_division_matches_float_division(c):
    bits = (np.arange(1 << 24, dtype=np.uint64) * 257).astype(np.uint32)
    rng = np.random.default_rng(c)
    rnd = rng.integers(0, 1 << 32, 1 << 22, dtype=np.uint64).astype(np.uint32)
    edge = np.array([0, 0x80000000, 1, 0x80000001, 0x007fffff, 0x00800000,
                     0x7f7fffff, 0x7f800000, 0xff800000, 0x7fc00000],
                    dtype=np.uint32)
    for b in (bits, rnd, edge):
        assert _check(b.view(np.float32), c) == 0
