"""The MFMA FFT's twiddle index arithmetic (csrc/mfma_fft.hip), on the CPU.

A radix-16 level on tile bits [ta, ta+4) of a 2^T tile multiplies output k of
the column whose tile bits above the level are c (b = T - ta - 4 of them) by
w_{2^(b+4)}^x with x = brev_b(c) k mod 2^(b+4) (level_twiddle).  A lane holds
the outputs k = 4 g + i (i < 4) of one column and computes their indices from
one bit reversal per column (out_twiddles): forward x = rb 4g + rb i; inverse
(the conjugate twiddles of the next level, on tile bits [ta-4, ta)) x =
brev_b'(c') k' with the outputs' bits the low bits of c', so x = R0 k' +
brev2(g) k' 2^(b'-4) + brev2(i) k' 2^(b'-2).  Both reductions are checked here
against the direct formula for every tile size, level, column and output.
"""
import pytest


def brev(v, nbits):
    r = 0
    for _ in range(nbits):
        r = (r << 1) | (v & 1)
        v >>= 1
    return r


def direct(T, ta, t):
    """level_twiddle's index for tile position t of the level on bits [ta, ta+4)."""
    k = (t >> ta) & 15
    b = T - ta - 4
    rb = brev(t >> (ta + 4), b) if b > 0 else 0
    return (rb * k) & ((16 << b) - 1)


def reduced_forward(T, ta, base, g, i):
    b = T - ta - 4
    rb = brev(base >> (ta + 4), b) if b > 0 else 0
    return (rb * 4 * g + i * rb) & ((16 << b) - 1)


def reduced_inverse(T, ta, base, g, i):
    b = T - ta  # bits above the next level (tile bits [ta - 4, ta))
    kp = (base >> (ta - 4)) & 15
    x0 = brev(base >> ta, b) * kp + ((brev(g, 2) * kp) << (b - 4))
    return (x0 + brev(i, 2) * (kp << (b - 2))) & ((16 << b) - 1)


@pytest.mark.parametrize("T", range(8, 14))
def test_strength_reduced_indices_equal_the_direct_ones(T):
    checked = 0
    for ta in range(0, T - 3, 4):
        for col in range(1 << (T - 4)):
            lo = col & ((1 << ta) - 1)
            base = ((col >> ta) << (ta + 4)) | lo
            for g in range(4):
                for i in range(4):
                    t = base + ((4 * g + i) << ta)
                    assert reduced_forward(T, ta, base, g, i) == direct(T, ta, t), (T, ta, col, g, i)
                    if ta >= 4:
                        assert reduced_inverse(T, ta, base, g, i) == direct(T, ta - 4, t), (T, ta, col, g, i)
                    checked += 1
    assert checked == (T // 4) * 16 * (1 << (T - 4))


def test_a_wrong_reversal_is_caught():
    """The check fails for a plausible mistake: the output digit's two bits
    reversed the other way round in the inverse reduction."""
    T, ta = 12, 8
    bad = 0
    for col in range(1 << (T - 4)):
        base = ((col >> ta) << (ta + 4)) | (col & ((1 << ta) - 1))
        for g in range(4):
            for i in range(4):
                b = T - ta
                kp = (base >> (ta - 4)) & 15
                x0 = brev(base >> ta, b) * kp + ((g * kp) << (b - 4))  # g instead of brev2(g)
                x = (x0 + brev(i, 2) * (kp << (b - 2))) & ((16 << b) - 1)
                bad += x != direct(T, ta - 4, base + ((4 * g + i) << ta))
    assert bad > 0
