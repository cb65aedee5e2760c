"""Generalized Halton points (halton_qrng; qmctoolscl/c_funcs/halton_qrng.c).

CPU: the oracle's restatement against the reference C compiled from its own
sources (oracle/_ref, where present), the committed digests of that library's
outputs, and what the mathematics fixes: unscrambled base-2 points are exact
radical inverses, unscrambled base-b points are the van der Corput values,
and the scrambled points of the first b^K indices fall one per interval
[m/b^K, (m+1)/b^K).  GPU: the HIP kernel (device entry point and host drop-in)
against the oracle, bit for bit."""
import ctypes
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import _clib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "halton_ref.json")

PRIMES = []
_v = 2
while len(PRIMES) < 360:
    if all(_v % p for p in PRIMES if p * p <= _v):
        PRIMES.append(_v)
    _v += 1

# (n, d, n0, generalized, seed): n0 = 0 takes the reference's separate point-0 path
CASES = [(1, 1, 0, 0, 1), (17, 3, 0, 1, 2), (1000, 7, 0, 1, 3), (1000, 7, 0, 0, 4), (513, 5, 12345, 1, 5),
         (300, 12, 7, 0, 6), (64, 360, 0, 1, 7), (257, 2, 1 << 20, 1, 8), (100, 4, 2147483547, 1, 9)]


def inputs(n, d, n0, generalized, seed):
    rng = np.random.Generator(np.random.SFC64(seed))
    randu = rng.random((d, 32))
    dvec = (np.arange(d, dtype=np.int32) if d == 360 else rng.integers(0, 360, size=d).astype(np.int32))
    return randu, dvec


def call(lib, sym, n, d, n0, generalized, randu, dvec):
    res = np.full((d, n), np.nan)
    fn = getattr(lib, sym)
    fn.restype = None
    fn(ctypes.c_int(n), ctypes.c_int(d), ctypes.c_int(n0), ctypes.c_int(generalized),
       res.ctypes.data_as(ctypes.c_void_p), randu.ctypes.data_as(ctypes.c_void_p),
       dvec.ctypes.data_as(ctypes.c_void_p))
    return res


def oracle(*case):
    return call(_clib.oracle(), "oracle_halton_qrng", *case)


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_reference_build(case):
    ref = _clib.reference()
    if ref is None:
        pytest.skip("reference C backend not built here (oracle/_ref)")
    randu, dvec = inputs(*case)
    a = oracle(*case[:4], randu, dvec)
    b = call(ref, "halton_qrng", *case[:4], randu, dvec)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_oracle_matches_committed_reference_digests():
    table = json.load(open(GOLDEN))
    for key, want in table["digests"].items():
        case = tuple(int(v) for v in key.split("/"))
        randu, dvec = inputs(*case)
        assert _clib.digest(oracle(*case[:4], randu, dvec)) == want, key


def test_unscrambled_base2_is_the_radical_inverse():
    n = 1 << 12
    res = oracle(n, 1, 0, 0, np.zeros((1, 32)), np.zeros(1, dtype=np.int32))
    want = np.array([int(format(i, "032b")[::-1], 2) / 2.0 ** 32 for i in range(n)])
    assert np.array_equal(res[0], want)


@pytest.mark.parametrize("j", [1, 2, 9, 359])
def test_unscrambled_base_b_is_van_der_corput(j):
    b = PRIMES[j]
    n0, n = 5, 200
    res = oracle(n, 1, n0, 0, np.zeros((1, 32)), np.array([j], dtype=np.int32))[0]
    for i in range(n):
        k, v, scale = n0 + i, Fraction(0), Fraction(1, b)
        while k:
            v += (k % b) * scale
            k //= b
            scale /= b
        assert abs(Fraction(res[i]) - v) <= Fraction(32, 2 ** 52) * max(v, Fraction(1, b ** 6)), (b, i)


@pytest.mark.parametrize("j,K", [(0, 10), (1, 6), (3, 3), (30, 2)])
def test_scrambled_points_stratify(j, K):
    b = PRIMES[j]
    n = b ** K
    rng = np.random.default_rng(j)
    res = oracle(n, 1, 0, 1, rng.random((1, 32)), np.array([j], dtype=np.int32))[0]
    cells = np.floor(res * n).astype(np.int64)
    assert sorted(cells.tolist()) == list(range(n))


# ----------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_device_and_host_match_oracle(case):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import qmctoolscl_amd as q
    n, d, n0, gen, _ = case
    randu, dvec = inputs(*case)
    want = oracle(n, d, n0, gen, randu, dvec)
    res = torch.empty((d, n), dtype=torch.float64, device="cuda")
    q.halton_qrng(n, d, n0, gen, res, torch.from_numpy(randu).cuda(), torch.from_numpy(dvec).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(res.cpu().numpy().view(np.uint64), want.view(np.uint64)), case
    host = np.full((d, n), np.nan)
    q.halton_qrng_c(n, d, n0, gen, host, randu, dvec)
    assert np.array_equal(host.view(np.uint64), want.view(np.uint64)), case


@pytest.mark.gpu
def test_gpu_large_run_matches_oracle_samples():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import qmctoolscl_amd as q
    n, d = 1 << 20, 32
    randu, dvec = inputs(n, d, 0, 1, 11)
    res = torch.empty((d, n), dtype=torch.float64, device="cuda")
    q.halton_qrng(n, d, 0, 1, res, torch.from_numpy(randu).cuda(), torch.from_numpy(dvec).cuda())
    torch.cuda.synchronize()
    got = res.cpu().numpy()
    for n0 in (0, 4096, 777777, n - 1000):
        want = oracle(1000, d, n0, 1, randu, dvec)
        assert np.array_equal(got[:, n0:n0 + 1000].view(np.uint64), want.view(np.uint64)), n0


@pytest.mark.gpu
def test_gpu_errors():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import qmctoolscl_amd as q
    res = np.empty((1, 4))
    with pytest.raises(q.QMCError, match="QMC_ERR_RANGE"):   # past the 360 tabulated primes
        q.halton_qrng_c(4, 1, 0, 1, res, np.zeros((1, 32)), np.array([360], dtype=np.int32))
    with pytest.raises(q.QMCError, match="QMC_ERR_RANGE"):   # beyond the reference's int indices
        q.halton_qrng_c(4, 1, 2147483645, 1, res, np.zeros((1, 32)), np.array([0], dtype=np.int32))
    q.halton_qrng_c(4, 1, 0, 1, res, np.zeros((1, 32)), np.array([0], dtype=np.int32))  # next call unaffected
