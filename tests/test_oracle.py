"""CPU tests that pin the oracle before it is trusted as the checker:
golden vectors of the reference doctests, the reference C backend built from
its own sources (when present), committed digests of that backend's outputs,
and closed forms / invariants chosen so that a dropped term, a wrong index or a
transposed operand fails at least one of them."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import _clib
import cases
import replay
from oracle_api import OracleAPI
from qmctoolscl_amd._api import KERNEL_NAMES

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def api():
    return OracleAPI()


def test_doctest_replay_oracle(api):
    chk = replay.replay(api)
    assert len(chk.results) >= 36
    assert chk.failures() == []


@pytest.mark.parametrize("name", KERNEL_NAMES)
@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_matches_reference_build(name, seed):
    lib = _clib.reference()
    if lib is None:
        pytest.skip("reference C backend not built here (make -C oracle ref)")
    a = cases.make_case(name, seed=seed, size="small")
    b = cases.copy_case(a)
    _clib.call_oracle(name, a)
    _clib.call_ref_abi(lib, name, b)
    for o in cases.outputs_of(name):
        assert np.array_equal(a[o].view(np.uint64), b[o].view(np.uint64)), (name, o)


def test_oracle_matches_reference_digests():
    with open(os.path.join(HERE, "golden", "ref_digests.json")) as f:
        table = json.load(f)
    for key, outs in table["digests"].items():
        name, seed = key.split("/")
        c = cases.make_case(name, seed=int(seed), size=table["size"])
        _clib.call_oracle(name, c)
        for o, dg in outs.items():
            assert _clib.digest(c[o]) == dg, key + ":" + o


# ----------------------------------------------------------------- closed forms

def _phi2(k):
    """radical inverse of k in base 2 as an exact fraction"""
    v, den = 0, 1
    while k:
        den *= 2
        v = 2 * v + (k & 1)
        k >>= 1
    return Fraction(v, den)


def _frac_exact(q):
    return q - (q.numerator // q.denominator)


def test_lattice_closed_form():
    # x = frac(g * phi2(Gray(k))) exactly, since g*phi2 < 2^53 for these sizes
    r, n, d, n_start = 2, 64, 3, 0
    g = np.array([[1, 5, 77], [3, 11, 1023]], dtype=np.uint64)
    nat = np.empty((r, n, d))
    gray = np.empty((r, n, d))
    lin = np.empty((r, n, d))
    _clib.call_oracle("lat_gen_natural", dict(r=r, n=n, d=d, n_start=n_start, g=g, x=nat))
    _clib.call_oracle("lat_gen_gray", dict(r=r, n=n, d=d, n_start=n_start, g=g, x=gray))
    _clib.call_oracle("lat_gen_linear", dict(r=r, n=n, d=d, g=g, x=lin))
    for l in range(r):
        for i in range(n):
            gk = i ^ (i >> 1)
            for j in range(d):
                gj = int(g[l, j])
                assert nat[l, i, j] == float(_frac_exact(gj * _phi2(i)))
                assert gray[l, i, j] == float(_frac_exact(gj * _phi2(gk)))
                assert lin[l, i, j] == float(_frac_exact(Fraction(gj * i, n)))
    # natural[Gray(i)] == gray[i]
    idx = np.arange(n) ^ (np.arange(n) >> 1)
    assert np.array_equal(nat[:, idx, :], gray)


def test_lat_shift_is_frac_of_sum():
    r_x, r, n, d = 1, 3, 16, 4
    rng = np.random.default_rng(3)
    x = rng.random((r_x, n, d))
    s = rng.random((r, d))
    xr = np.empty((r, n, d))
    _clib.call_oracle("lat_shift_mod_1", dict(r=r, n=n, d=d, r_x=r_x, x=x, shifts=s, xr=xr))
    y = x[0][None] + s[:, None, :]
    assert np.array_equal(xr, y - np.trunc(y))


def _xor_cols(C, k):
    v, b = 0, 0
    while k:
        if k & 1:
            v ^= int(C[b])
        k >>= 1
        b += 1
    return v


def test_dnb2_brute_force_and_gray_relation():
    rng = np.random.default_rng(5)
    r, d, mmax = 2, 3, 7
    C = rng.integers(0, 2**64, size=(r, d, mmax), dtype=np.uint64)
    for n_start, n in [(0, 128), (32, 96), (1, 3), (16, 16)]:
        nat = np.zeros((r, n, d), dtype=np.uint64)
        gray = np.zeros((r, n, d), dtype=np.uint64)
        _clib.call_oracle("dnb2_gen_natural", dict(r=r, n=n, d=d, n_start=n_start, mmax=mmax, C=C, xb=nat))
        _clib.call_oracle("dnb2_gen_gray", dict(r=r, n=n, d=d, n_start=n_start, mmax=mmax, C=C, xb=gray))
        for l in range(r):
            for i in range(n):
                k = i + n_start
                for j in range(d):
                    assert int(nat[l, i, j]) == _xor_cols(C[l, j], k)
                    assert int(gray[l, i, j]) == _xor_cols(C[l, j], k ^ (k >> 1))
    # XOR linearity in C
    C2 = rng.integers(0, 2**64, size=(r, d, mmax), dtype=np.uint64)
    outs = []
    for CC in (C, C2, C ^ C2):
        o = np.zeros((r, 64, d), dtype=np.uint64)
        _clib.call_oracle("dnb2_gen_natural", dict(r=r, n=64, d=d, n_start=0, mmax=mmax, C=CC, xb=o))
        outs.append(o)
    assert np.array_equal(outs[0] ^ outs[1], outs[2])


def test_dnb2_lms_identity_and_gf2_product():
    rng = np.random.default_rng(7)
    d, mmax, t = 3, 9, 20
    C = rng.integers(0, 2**t, size=(1, d, mmax), dtype=np.uint64)
    eye = np.array([[1 << (t - 1 - i) for i in range(t)]] * d, dtype=np.uint64)[None]
    out = np.zeros((1, d, mmax), dtype=np.uint64)
    _clib.call_oracle("dnb2_linear_matrix_scramble", dict(r=1, d=d, mmax=mmax, r_C=1, tmax_new=t, S=eye, C=C, C_lms=out))
    assert np.array_equal(out, C)
    # random S: compare with a dense GF(2) matrix product
    S = rng.integers(0, 2**t, size=(1, d, t), dtype=np.uint64)
    _clib.call_oracle("dnb2_linear_matrix_scramble", dict(r=1, d=d, mmax=mmax, r_C=1, tmax_new=t, S=S, C=C, C_lms=out))
    bits = lambda v: np.array([(int(v) >> (t - 1 - i)) & 1 for i in range(t)])
    for j in range(d):
        Sm = np.array([bits(S[0, j, i]) for i in range(t)])
        for k in range(mmax):
            col = (Sm @ bits(C[0, j, k])) % 2
            assert int(out[0, j, k]) == int("".join(map(str, col)), 2)


def test_dnb2_interlace_round_trip_and_bit_placement():
    rng = np.random.default_rng(9)
    r, d_alpha, alpha, mmax, tmax = 2, 3, 2, 5, 30
    d = d_alpha * alpha
    C = rng.integers(0, 2**tmax, size=(r, d, mmax), dtype=np.uint64)
    Ca = np.zeros((r, d_alpha, mmax), dtype=np.uint64)
    _clib.call_oracle("dnb2_interlace", dict(r=r, d_alpha=d_alpha, mmax=mmax, d=d, tmax=tmax,
                                             tmax_alpha=alpha * tmax, alpha=alpha, C=C, C_alpha=Ca))
    back = np.zeros_like(C)
    _clib.call_oracle("dnb2_undo_interlace", dict(r=r, d=d, mmax=mmax, d_alpha=d_alpha, tmax=tmax,
                                                  tmax_alpha=alpha * tmax, alpha=alpha, C_alpha=Ca, C=back))
    assert np.array_equal(back, C)
    # top bit of C_alpha[l,0,k] is the top bit of C[l,0,k]; the next is the top bit of C[l,1,k]
    for k in range(mmax):
        assert (int(Ca[0, 0, k]) >> 59) & 1 == (int(C[0, 0, k]) >> 29) & 1
        assert (int(Ca[0, 0, k]) >> 58) & 1 == (int(C[0, 1, k]) >> 29) & 1


def test_integer_to_float_is_scaled_conversion():
    xb = np.array([[[0, 1, 2**63, 2**64 - 1, 2**64 - 1024, 2**64 - 1025]]], dtype=np.uint64)
    x = np.empty(xb.shape)
    _clib.call_oracle("dnb2_integer_to_float", dict(r=1, n=1, d=6, tmaxes=np.array([64], dtype=np.uint64), xb=xb, x=x))
    assert list(x.ravel()) == [0.0, 2.0**-64, 0.5, 1.0, 1.0, float(np.float64(2**64 - 2048)) / 2**64]


def _radical_inverse(k, b, t):
    digs = []
    for _ in range(t):
        digs.append(k % b)
        k //= b
    return digs


def test_gdn_halton_identity_gives_radical_inverse_digits():
    bases = np.array([[2, 3, 5, 7]], dtype=np.uint64)
    mmax = tmax = 8
    C = np.tile(np.eye(mmax, dtype=np.uint64)[None, None], (1, 4, 1, 1))
    n, n_start = 50, 3
    xdig = np.zeros((1, n, 4, tmax), dtype=np.uint64)
    _clib.call_oracle("gdn_gen_natural", dict(r=1, n=n, d=4, r_b=1, mmax=mmax, tmax=tmax, n_start=n_start,
                                              bases=bases, C=C, xdig=xdig))
    x = np.empty((1, n, 4))
    _clib.call_oracle("gdn_integer_to_float", dict(r=1, n=n, d=4, r_b=1, tmax=tmax, bases=bases, xdig=xdig, x=x))
    for i in range(n):
        for j, b in enumerate([2, 3, 5, 7]):
            assert list(xdig[0, i, j]) == _radical_inverse(i + n_start, b, tmax)
            v = sum(Fraction(dg, b ** (t + 1)) for t, dg in enumerate(_radical_inverse(i + n_start, b, tmax)))
            assert abs(x[0, i, j] - float(v)) < 1e-15
    # base 2 with identity C reproduces the base-2 net with identity columns
    n = 64  # natural order needs an aligned range
    C2 = np.array([[[1 << (7 - k) for k in range(8)]]], dtype=np.uint64)
    xb = np.zeros((1, n, 1), dtype=np.uint64)
    _clib.call_oracle("dnb2_gen_natural", dict(r=1, n=n, d=1, n_start=0, mmax=8, C=C2, xb=xb))
    xd = np.zeros((1, n, 1, 8), dtype=np.uint64)
    _clib.call_oracle("gdn_gen_natural", dict(r=1, n=n, d=1, r_b=1, mmax=8, tmax=8, n_start=0,
                                              bases=np.array([[2]], dtype=np.uint64),
                                              C=np.eye(8, dtype=np.uint64)[None, None], xdig=xd))
    as_int = (xd[0, :, 0, :] << np.arange(7, -1, -1, dtype=np.uint64)).sum(axis=1)
    assert np.array_equal(as_int, xb[0, :, 0])


def test_gdn_same_base_equals_general():
    c = cases.make_case("gdn_gen_natural_same_base", seed=4, size="small")
    a = cases.copy_case(c)
    _clib.call_oracle("gdn_gen_natural_same_base", a)
    d = c["d"]
    g = dict(r=c["r"], n=c["n"], d=d, r_b=1, mmax=c["mmax"], tmax=c["tmax"], n_start=c["n_start"],
             bases=np.full((1, d), c["b"], dtype=np.uint64), C=c["C"], xdig=c["xdig"].copy())
    _clib.call_oracle("gdn_gen_natural", g)
    assert np.array_equal(a["xdig"], g["xdig"])


def test_gdn_lms_brute_force():
    c = cases.make_case("gdn_linear_matrix_scramble", seed=3, size="small")
    _clib.call_oracle("gdn_linear_matrix_scramble", c)
    S, C, bases = c["S"].astype(object), c["C"].astype(object), c["bases"]
    for l in range(c["r"]):
        for j in range(c["d"]):
            b = int(bases[l % c["r_b"], j])
            for k in range(c["mmax"]):
                for t in range(c["tmax_new"]):
                    want = sum((S[l, j, t, cc] * C[l % c["r_C"], j, k, cc]) % b for cc in range(c["tmax"]))
                    assert int(c["C_lms"][l, j, k, t]) == want


def test_transforms_against_numpy():
    rng = np.random.default_rng(2)
    from scipy.linalg import hadamard
    for m in (1, 2, 5, 10):
        nn = 2**m
        x = rng.standard_normal((2, 3, nn))
        y = x.copy()
        _clib.call_oracle("fwht_1d_radix2", dict(d1=2, d2=3, n_half=nn // 2, x=y))
        assert np.allclose(y, x @ hadamard(nn).T / np.sqrt(nn), atol=1e-12)
        ir = np.array([int(format(i, "0%db" % m)[::-1], 2) for i in range(nn)])
        xr, xi = rng.standard_normal((2, 3, nn)), rng.standard_normal((2, 3, nn))
        a, b = xr.copy(), xi.copy()
        tw1, tw2 = np.empty(nn), np.empty(nn)
        _clib.call_oracle("fft_bro_1d_radix2", dict(d1=2, d2=3, n_half=nn // 2, twiddler=tw1, twiddlei=tw2, xr=a, xi=b))
        assert np.allclose(a + 1j * b, np.fft.fft((xr + 1j * xi)[..., ir], norm="ortho"), atol=1e-12)
        assert np.allclose(tw1 + 1j * tw2, np.exp(-2j * np.pi * np.arange(nn) / nn), atol=1e-15)
        _clib.call_oracle("ifft_bro_1d_radix2", dict(d1=2, d2=3, n_half=nn // 2, twiddler=tw1, twiddlei=tw2, xr=a, xi=b))
        assert np.allclose(a + 1j * b, xr + 1j * xi, atol=1e-12)
        # ifft alone: natural in, bit-reversed out
        a, b = xr.copy(), xi.copy()
        _clib.call_oracle("ifft_bro_1d_radix2", dict(d1=2, d2=3, n_half=nn // 2, twiddler=tw1, twiddlei=tw2, xr=a, xi=b))
        assert np.allclose((a + 1j * b)[..., ir], np.fft.ifft(xr + 1j * xi, norm="ortho"), atol=1e-12)


def test_vectorized_draws_match_sequential_draws():
    """The product's NUS draws all of a tree's random values with one
    vectorized numpy call per (l, j); this pins that it consumes the stream
    exactly like the reference's one-scalar-call-per-leaf loop."""
    widths = [8, 7, 3, 0, 7, 1, 33, 0, 63, 12, 31, 32, 2, 0, 5]
    a = np.random.Generator(np.random.SFC64(np.random.SeedSequence(7).spawn(3)[2]))
    b = np.random.Generator(np.random.SFC64(np.random.SeedSequence(7).spawn(3)[2]))
    seq = [int(a.integers(0, 2 ** w)) for w in widths]
    vec = b.integers(0, np.uint64(1) << np.array(widths, dtype=np.uint64), dtype=np.uint64)
    assert seq == [int(v) for v in vec]
    assert a.integers(0, 2**40) == b.integers(0, 2**40)  # streams stay in step
    # permutation rows: a loop of rng.permutation(b) == rng.permuted over a tiled arange
    a = np.random.Generator(np.random.SFC64(11))
    b = np.random.Generator(np.random.SFC64(11))
    rows = [a.permutation(np.uint64(13)) for _ in range(9)]
    tiled = b.permuted(np.tile(np.arange(13, dtype=np.uint64), (9, 1)), axis=1)
    assert np.array_equal(np.array(rows), tiled)
    # ... and the product's 16-bit rows take the same draws as 64-bit ones
    for base, k in ((2, 40), (13, 9), (257, 5), (65536, 2)):
        a = np.random.Generator(np.random.SFC64(base))
        b = np.random.Generator(np.random.SFC64(base))
        rows = [a.permutation(np.uint64(base)) for _ in range(k)]
        tiled = b.permuted(np.tile(np.arange(base, dtype=np.uint16), (k, 1)), axis=1)
        assert np.array_equal(np.array(rows, dtype=np.uint64), tiled.astype(np.uint64)), base
        assert a.integers(0, 2**40) == b.integers(0, 2**40)
