"""Host-side checks of the batch's gdn inputs and of the early exit of its
sums (no GPU): the exported LMS matrices and permutations have the reference
draws' shapes and domains (rand_funcs.py:80-153), the permutations are
uniform, and stopping the reference's sequential sum
(gdn_integer_to_float.c:34-38) at the thresholds leaves every result
unchanged."""
import math

import numpy as np
import pytest

from qmctoolscl_amd import batch as B


def test_first_primes():
    p = B.first_primes(256)
    assert p[:8] == [2, 3, 5, 7, 11, 13, 17, 19] and p[-1] == 1619 and len(set(p)) == 256
    assert all(all(x % q for q in range(2, int(math.isqrt(x)) + 1)) for x in p)


@pytest.mark.parametrize("m,r,j0,dj", [(5, 3, 0, 6), (12, 2, 30, 4), (22, 1, 250, 6)])
def test_gdn_inputs_shapes_and_domains(m, r, j0, dj):
    bases, S, perms = B.gdn_inputs(77, m, r, j0=j0, dj=dj)
    assert bases.tolist() == [B.first_primes(j0 + dj)[j0:]]
    assert S.shape == (r, dj, 52, m) and perms.shape == (r, dj, 52, int(bases.max()))
    for jj in range(dj):
        b = int(bases[0, jj])
        for l in range(r):
            Slj = S[l, jj].astype(np.int64)
            assert np.all(np.triu(Slj, 1) == 0)                      # zero above the diagonal
            diag = np.diagonal(Slj)                                   # eye(52, m)
            assert np.all((diag >= 1) & (diag < b))
            assert np.all(np.tril(Slj, -1) < b)
            for t in range(52):
                row = perms[l, jj, t]
                assert sorted(row[:b].tolist()) == list(range(b))    # a permutation of the digits
                assert np.all(row[b:] == 0)


def test_gdn_inputs_are_column_independent():
    """A column's inputs do not depend on which column range is exported."""
    _, S1, P1 = B.gdn_inputs(5, 9, 2, j0=0, dj=5)
    _, S2, P2 = B.gdn_inputs(5, 9, 2, j0=3, dj=1)
    assert np.array_equal(S1[:, 3], S2[:, 0])
    assert np.array_equal(P1[:, 3, :, :7], P2[:, 0, :, :7])


def test_gdn_permutations_are_uniform():
    """Base 3: the six permutations are equally likely over independent (l, t)
    levels (chi-square, 5 degrees of freedom); base 7: every value is equally
    likely at every position."""
    counts = {}
    for seed in (1, 2, 3):
        _, _, P = B.gdn_inputs(seed, 4, 40, j0=1, dj=1)
        for row in P[:, 0, :, :3].reshape(-1, 3):
            counts[tuple(row.tolist())] = counts.get(tuple(row.tolist()), 0) + 1
    assert len(counts) == 6
    tot = sum(counts.values())
    chi2 = sum((c - tot / 6) ** 2 / (tot / 6) for c in counts.values())
    assert chi2 < 25.7, (chi2, counts)  # p > 1e-4
    _, _, P = B.gdn_inputs(9, 4, 60, j0=3, dj=1)
    rows = P[:, 0, :, :7].reshape(-1, 7)
    tot = rows.shape[0]
    for pos in range(7):
        c = np.bincount(rows[:, pos].astype(np.int64), minlength=7)
        chi2 = float(np.sum((c - tot / 7) ** 2 / (tot / 7)))
        assert chi2 < 27.9, (pos, chi2)  # 6 degrees of freedom, p > 1e-4


def _ref_sum(rec, y):
    v = 0.0
    for t in range(52):
        v += rec[t] * float(y[t])
    return v


def _early_sum(rec, thr, tmin, y):
    v = 0.0
    for t in range(52):
        if t >= tmin and v >= thr[t]:
            break
        v += rec[t] * float(y[t])
    return v


def test_gdn_base_constants():
    for j, b in enumerate(B.first_primes(256)):
        rec, thr, levels, tmin = B.gdn_base(j)
        want = 1.0 / b
        for t in range(52):
            assert rec[t] == want  # gdn_integer_to_float.c:34-38
            want /= b
        assert levels in (8, 10, 12, 16, 24, 48, 52)
        if b > 2:
            assert float(b) ** levels >= 2.0 ** 74
        assert all(thr[t] >= 1.0 for t in range(tmin)) and (tmin == 52 or thr[tmin] < 1.0)


@pytest.mark.parametrize("j", list(range(0, 256, 5)) + [1, 2, 3, 255])
def test_gdn_threshold_makes_later_terms_vanish(j):
    """The exit lemma at its boundary: a sum at thr_t, or one ulp above it, is
    unchanged by the largest later term RN(rec_s (b - 1)), s >= t; one binade
    lower it is not always unchanged (the threshold is not loose by a binade)."""
    b = B.first_primes(256)[j]
    rec, thr, _, tmin = B.gdn_base(j)
    tight = 0
    for t in range(tmin, 52):
        for v in (thr[t], np.nextafter(thr[t], 2.0)):
            for s in range(t, 52):
                assert v + rec[s] * float(b - 1) == v, (t, s)
        lower = thr[t] / 2
        if lower + rec[t] * float(b - 1) != lower:
            tight += 1
    assert tight > 0 or tmin == 52


@pytest.mark.parametrize("j", [1, 2, 3, 10, 31, 100, 255])
def test_gdn_early_exit_equals_full_sum(j):
    """Random digits, digits with long runs of leading zeros (small sums that
    need deep levels), all b - 1 (the largest terms), and sums landing just
    above a threshold: the early exit returns the reference loop's double."""
    b = B.first_primes(256)[j]
    rec, thr, _, tmin = B.gdn_base(j)
    rng = np.random.default_rng(j)
    cases = [rng.integers(0, b, 52) for _ in range(3000)]
    for z in range(1, 30):
        y = rng.integers(0, b, 52)
        y[:z] = 0
        cases.append(y)
    cases.append(np.full(52, b - 1))
    cases.append(np.zeros(52, dtype=np.int64))
    for t in range(1, 12):
        y = np.zeros(52, dtype=np.int64)
        y[t] = 1
        y[t + 1:] = b - 1
        cases.append(y)
    for y in cases:
        assert _ref_sum(rec, y) == _early_sum(rec, thr, tmin, y), y.tolist()
