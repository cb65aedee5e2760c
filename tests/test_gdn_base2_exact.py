"""The premise of the fused gdn kernel's base-2 path (csrc/gdn.hip): with
b = 2 every term of gdn_integer_to_float's sum (gdn_integer_to_float.c:34-38:
recip_0 = RN(1/b), recip_{t+1} = RN(recip_t / b), v += recip_t * y_t in order)
is 0 or 2^-(t+1), every partial sum is a multiple of 2^-tnew below 1, and so
for tnew <= 53 the sum equals Y 2^-tnew, Y the digits as the bit word
sum_t y_t 2^(tnew-1-t).  Checked in float64 on random and edge digit words."""
import numpy as np
import pytest


def reference_sum(y):
    v = 0.0
    recip = 1.0 / 2.0
    for yt in y:
        v = v + recip * float(yt)
        recip = recip / 2.0
    return v


def word(y):
    Y = 0
    for yt in y:
        Y = (Y << 1) | int(yt)
    return Y


@pytest.mark.parametrize("tnew", [1, 2, 20, 52, 53])
def test_base2_sum_is_the_bit_word(tnew):
    rng = np.random.default_rng(tnew)
    samples = [np.ones(tnew, dtype=np.uint64), np.zeros(tnew, dtype=np.uint64)]
    samples += [rng.integers(0, 2, size=tnew, dtype=np.uint64) for _ in range(2000)]
    for y in samples:
        want = reference_sum(y)
        got = float(np.ldexp(np.float64(word(y)), -tnew))
        assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (tnew, y)


def test_a_reversed_word_is_caught():
    """A plausible mistake (digit t at bit t instead of bit tnew - 1 - t) fails."""
    rng = np.random.default_rng(7)
    bad = 0
    for _ in range(200):
        y = rng.integers(0, 2, size=52, dtype=np.uint64)
        wrong = word(y[::-1])
        bad += float(np.ldexp(np.float64(wrong), -52)) != reference_sum(y)
    assert bad > 0
