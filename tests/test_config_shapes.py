"""The benchmark configurations' own shapes on the GPU, against the CPU oracle.

Config 0 (dnb2, n = 2^16, d = 8, r = 4, LMS + 64-bit shift + float) runs at
full size through both the drop-in chain and the fused kernel.  Config 2
(Halton-type net, first 32 primes, n = 3^12, d = 32, r = 8, LMS to 52 digits,
digit permutations, float) runs at full size on the GPU -- 57 GB of digits per
copy for the drop-in chain -- and sampled row blocks are checked against the
oracle, which restarts at any n_start.  The float64 transforms run at 2^24,
the longest config-4 length, and the gdn nested uniform scramble on the 32
prime bases."""
import os
import sys

import numpy as np
import pytest

import _clib
import qmctoolscl_amd as q
from qmctoolscl_amd import rand_funcs as rf

sys.path.insert(0, os.path.join(_clib.REPO, "oracle"))
import nus_oracle  # noqa: E402

pytestmark = pytest.mark.gpu

PRIMES32 = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101,
            103, 107, 109, 113, 127, 131]
U = np.uint64


@pytest.fixture(scope="module")
def torch():
    t = pytest.importorskip("torch")
    if not t.cuda.is_available():
        pytest.skip("no GPU visible")
    return t


def to_t(torch, a):
    return torch.from_numpy(a.view(np.int64) if a.dtype == np.uint64 else a).cuda()


def test_config0_full_size(torch):
    """dnb2_linear_matrix_scramble -> dnb2_gen_natural -> dnb2_digital_shift
    -> dnb2_integer_to_float at n = 2^16, d = 8, r = 4, tmax = 32 (upper
    triangular unit-diagonal generating matrices): the drop-ins, the fused
    generation kernel and the one-launch chain, bit for bit against the
    oracle."""
    from qmctoolscl_amd import fused
    rng = np.random.Generator(np.random.SFC64(0))
    n, d, r, t = 1 << 16, 8, 4, 32
    C = np.empty((1, d, t), dtype=np.uint64)
    for j in range(d):
        for k in range(t):
            above = rng.integers(0, 2, size=k, dtype=np.uint64)
            C[0, j, k] = (1 << (t - 1 - k)) + sum(int(b) << (t - 1 - c) for c, b in enumerate(above))
    S = rf.dnb2_get_linear_scramble_matrix(rng, r, d, t, t, False)
    sh = rf.dnb2_get_digital_shifts(rng, r, d, 64)
    lsh, tm = np.full(r, 32, np.uint64), np.full(r, 64, np.uint64)
    Cl = np.empty((r, d, t), dtype=np.uint64)
    _clib.call_oracle("dnb2_linear_matrix_scramble", dict(r=r, d=d, mmax=t, r_C=1, tmax_new=t, S=S, C=C, C_lms=Cl))
    xb = np.empty((r, n, d), dtype=np.uint64)
    _clib.call_oracle("dnb2_gen_natural", dict(r=r, n=n, d=d, n_start=0, mmax=t, C=Cl, xb=xb))
    xrb = np.empty_like(xb)
    _clib.call_oracle("dnb2_digital_shift", dict(r=r, n=n, d=d, r_x=r, lshifts=lsh, xb=xb, shiftsb=sh, xrb=xrb))
    x = np.empty((r, n, d))
    _clib.call_oracle("dnb2_integer_to_float", dict(r=r, n=n, d=d, tmaxes=tm, xb=xrb, x=x))
    # drop-ins on device tensors
    tC, tS, tsh, tl, ttm = (to_t(torch, a) for a in (C, S, sh, lsh, tm))
    tCl = torch.empty((r, d, t), dtype=torch.int64, device="cuda")
    txb = torch.empty((r, n, d), dtype=torch.int64, device="cuda")
    txrb = torch.empty_like(txb)
    tx = torch.empty((r, n, d), dtype=torch.float64, device="cuda")
    q.dnb2_linear_matrix_scramble(U(r), U(d), U(t), U(1), U(t), tS, tC, tCl)
    q.dnb2_gen_natural(U(r), U(n), U(d), U(0), U(t), tCl, txb)
    q.dnb2_digital_shift(U(r), U(n), U(d), U(r), tl, txb, tsh, txrb)
    q.dnb2_integer_to_float(U(r), U(n), U(d), ttm, txrb, tx)
    assert np.array_equal(txrb.cpu().numpy().view(np.uint64), xrb)
    assert np.array_equal(tx.cpu().numpy().view(np.uint64), x.view(np.uint64))
    # fused generation + shift + float
    fx = torch.empty_like(tx)
    fxrb = torch.empty_like(txrb)
    fused.dnb2_gen_shift_float(1, r, n, d, r, 0, t, tCl, tl, tsh, ttm, fxrb, fx)
    torch.cuda.synchronize()
    assert np.array_equal(fxrb.cpu().numpy().view(np.uint64), xrb)
    assert np.array_equal(fx.cpu().numpy().view(np.uint64), x.view(np.uint64))
    # the whole chain in one launch
    fx.zero_()
    fxrb.zero_()
    fused.dnb2_lms_gen_shift_float(1, r, n, d, 1, 0, t, t, tS, tC, tl, tsh, ttm, fxrb, fx)
    torch.cuda.synchronize()
    assert np.array_equal(fxrb.cpu().numpy().view(np.uint64), xrb)
    assert np.array_equal(fx.cpu().numpy().view(np.uint64), x.view(np.uint64))


def _config2_inputs():
    rng = np.random.Generator(np.random.SFC64(2))
    r, d, mmax, tnew = 8, 32, 20, 52
    bases = np.array([PRIMES32], dtype=np.uint64)
    C = rf.gdn_get_halton_generating_matrix(1, d, mmax)
    S = rf.gdn_get_linear_scramble_matrix(rng, r, d, mmax, tnew, 1, bases)
    perms = rf.gdn_get_digital_permutations(rng, r, d, tnew, 1, bases)
    C_lms = np.empty((r, d, mmax, tnew), dtype=np.uint64)
    _clib.call_oracle("gdn_linear_matrix_scramble", dict(r=r, d=d, mmax=mmax, r_C=1, r_b=1, tmax=mmax,
                                                         tmax_new=tnew, bases=bases, S=S, C=C, C_lms=C_lms))
    return r, d, mmax, tnew, bases, C, S, perms, C_lms


def _oracle_rows(start, rows, r, d, mmax, tnew, bases, C_lms, perms):
    xd = np.empty((r, rows, d, tnew), dtype=np.uint64)
    _clib.call_oracle("gdn_gen_natural", dict(r=r, n=rows, d=d, r_b=1, mmax=mmax, tmax=tnew, n_start=start,
                                              bases=bases, C=C_lms, xdig=xd))
    xp = np.empty_like(xd)
    _clib.call_oracle("gdn_digital_permutation", dict(r=r, n=rows, d=d, r_x=r, r_b=1, tmax=tnew, tmax_new=tnew,
                                                      bmax=131, perms=perms, xdig=xd, xdig_new=xp))
    x = np.empty((r, rows, d))
    _clib.call_oracle("gdn_integer_to_float", dict(r=r, n=rows, d=d, r_b=1, tmax=tnew, bases=bases, xdig=xp, x=x))
    return xp, x


STARTS = (0, 3 ** 7, 300001, 3 ** 12 - 256)


def test_config2_full_size_dropin_chain(torch):
    """gdn_linear_matrix_scramble -> gdn_gen_natural -> gdn_digital_permutation
    -> gdn_integer_to_float at full size on the GPU (two 56.6 GB digit
    arrays); four row blocks of 256 points (every replication and column, all
    52 digits) against the oracle."""
    r, d, mmax, tnew, bases, C, S, perms, C_lms = _config2_inputs()
    n = 3 ** 12
    tb, tS, tC, tp = (to_t(torch, a) for a in (bases, S, C, perms))
    tCl = torch.empty((r, d, mmax, tnew), dtype=torch.int64, device="cuda")
    q.gdn_linear_matrix_scramble(U(r), U(d), U(mmax), U(1), U(1), U(mmax), U(tnew), tb, tS, tC, tCl)
    assert np.array_equal(tCl.cpu().numpy().view(np.uint64), C_lms)
    xd = torch.empty((r, n, d, tnew), dtype=torch.int64, device="cuda")
    q.gdn_gen_natural(U(r), U(n), U(d), U(1), U(mmax), U(tnew), U(0), tb, tCl, xd)
    xp = torch.empty_like(xd)
    q.gdn_digital_permutation(U(r), U(n), U(d), U(r), U(1), U(tnew), U(tnew), U(131), tp, xd, xp)
    del xd
    x = torch.empty((r, n, d), dtype=torch.float64, device="cuda")
    q.gdn_integer_to_float(U(r), U(n), U(d), U(1), U(tnew), tb, xp, x)
    torch.cuda.synchronize()
    rows = 256
    for start in STARTS:
        want_p, want_x = _oracle_rows(start, rows, r, d, mmax, tnew, bases, C_lms, perms)
        assert np.array_equal(xp[:, start:start + rows].cpu().numpy().view(np.uint64), want_p), start
        assert np.array_equal(x[:, start:start + rows].cpu().numpy().view(np.uint64), want_x.view(np.uint64)), start
    del xp
    torch.cuda.empty_cache()


def test_config2_full_size_fused(torch):
    """The same chain through the fused kernel (digits never stored), full
    size, the same row blocks against the oracle."""
    from qmctoolscl_amd import fused
    r, d, mmax, tnew, bases, C, S, perms, C_lms = _config2_inputs()
    n = 3 ** 12
    x = torch.empty((r, n, d), dtype=torch.float64, device="cuda")
    fused.gdn_gen_scramble_float(1, r, n, d, 1, mmax, tnew, tnew, 0, 131, to_t(torch, bases), to_t(torch, C_lms),
                                 to_t(torch, perms), x)
    torch.cuda.synchronize()
    rows = 256
    for start in STARTS:
        _, want_x = _oracle_rows(start, rows, r, d, mmax, tnew, bases, C_lms, perms)
        assert np.array_equal(x[:, start:start + rows].cpu().numpy().view(np.uint64), want_x.view(np.uint64)), start
    assert float(x.min()) >= 0.0 and float(x.max()) < 1.0


def test_fp64_transforms_at_2e24(torch):
    """fwht_1d_radix2 bit-exact and fft/ifft_bro_1d_radix2 within 1e-12 of the
    oracle at n = 2^24 (the passes a 2^24 transform takes past the LDS tile)."""
    rng = np.random.default_rng(24)
    n = 1 << 24
    x = rng.standard_normal((1, 1, n))
    want = x.copy()
    _clib.call_oracle("fwht_1d_radix2", dict(d1=1, d2=1, n_half=n // 2, x=want))
    tx = to_t(torch, x.copy())
    q.fwht_1d_radix2(U(1), U(1), U(n // 2), tx)
    assert np.array_equal(tx.cpu().numpy().view(np.uint64), want.view(np.uint64))
    del tx, want
    for name in ("fft_bro_1d_radix2", "ifft_bro_1d_radix2"):
        xr, xi = rng.standard_normal((1, 1, n)), rng.standard_normal((1, 1, n))
        wr, wi = xr.copy(), xi.copy()
        _clib.call_oracle(name, dict(d1=1, d2=1, n_half=n // 2, twiddler=np.empty(n), twiddlei=np.empty(n), xr=wr,
                                     xi=wi))
        txr, txi = to_t(torch, xr), to_t(torch, xi)
        twr = torch.empty(n, dtype=torch.float64, device="cuda")
        twi = torch.empty_like(twr)
        getattr(q, name)(U(1), U(1), U(n // 2), twr, twi, txr, txi)
        scale = max(1.0, float(np.max(np.abs(wr))), float(np.max(np.abs(wi))))
        err = max(float(np.max(np.abs(txr.cpu().numpy() - wr))), float(np.max(np.abs(txi.cpu().numpy() - wi))))
        assert err / scale <= 1e-12, (name, err / scale)


def test_gdn_nus_on_the_32_prime_bases():
    """Nested uniform scrambling of Halton digits in the first 32 prime bases
    (config 2's bases), 300 points in two calls that grow the same trees,
    against the oracle's tree algorithm."""
    r, d, tmax, tnew = 2, 32, 6, 12
    bases = np.array([PRIMES32], dtype=np.uint64)
    C = rf.gdn_get_halton_generating_matrix(1, d, tmax)
    xdig = np.empty((1, 300, d, tmax), dtype=np.uint64)
    _clib.call_oracle("gdn_gen_natural", dict(r=1, n=300, d=d, r_b=1, mmax=tmax, tmax=tmax, n_start=0, bases=bases,
                                              C=C, xdig=xdig))

    def rngs():
        seeds = np.random.SeedSequence(31).spawn(r * d)
        return np.array([np.random.Generator(np.random.SFC64(s)) for s in seeds]).reshape(r, d)
    ro, rp = rngs(), rngs()
    roots_o = np.array([nus_oracle.NodeB() for _ in range(r * d)]).reshape(r, d)
    roots_p = np.array([q.NUSNode_gdn() for _ in range(r * d)]).reshape(r, d)
    for part in (xdig[:, :120], xdig[:, 120:]):
        part = np.ascontiguousarray(part)
        want = np.zeros((r, part.shape[1], d, tnew), dtype=np.uint64)
        got = np.zeros_like(want)
        nus_oracle.gdn_nus(r, part.shape[1], d, 1, 1, tmax, tnew, ro, roots_o, bases, part, want)
        q.gdn_nested_uniform_scramble(r, part.shape[1], d, 1, 1, tmax, tnew, rp, roots_p, bases, part, got)
        assert np.array_equal(got, want)
