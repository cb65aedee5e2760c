"""qmc_gdn_gen_scramble_float_noalloc_dev (the fused gdn entry point that
allocates no device memory) against qmc_gdn_gen_scramble_float_dev, through
the C-ABI: bit-identical outputs on consecutive calls that share its fixed
workspace, the second smaller than the first, and a clear error for a call
whose workspace would exceed the fixed size."""
import ctypes

import numpy as np
import pytest

import qmctoolscl_amd as q
from qmctoolscl_amd import _lib
from qmctoolscl_amd import rand_funcs as rf

pytestmark = pytest.mark.gpu

PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103,
          107, 109, 113, 127, 131]
SCRATCH_WORDS = 1 << 20


@pytest.fixture(scope="module")
def torch():
    t = pytest.importorskip("torch")
    if not t.cuda.is_available():
        pytest.skip("no GPU visible")
    return t


def _fn(name):
    fn = getattr(_lib.lib(), name)
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int] + [ctypes.c_uint64] * 9 + [ctypes.c_void_p] * 5
    return fn


def _case(torch, seed, mode, r, n, d, mmax, tmax, tnew, n_start):
    rng = np.random.Generator(np.random.SFC64(seed))
    bases = np.array([PRIMES[:d]], dtype=np.uint64)
    bmax = int(bases.max())
    C = rf.gdn_get_halton_generating_matrix(1, d, mmax)
    S = rf.gdn_get_linear_scramble_matrix(rng, r, d, mmax, tmax, 1, bases)
    dev = lambda a: torch.from_numpy(a.view(np.int64).copy()).cuda()
    tb, tS, tC = dev(bases), dev(S), dev(C)
    Cl = torch.empty((r, d, mmax, tmax), dtype=torch.int64, device="cuda")
    q.gdn_linear_matrix_scramble(np.uint64(r), np.uint64(d), np.uint64(mmax), np.uint64(1), np.uint64(1),
                                 np.uint64(mmax), np.uint64(tmax), tb, tS, tC, Cl)
    if mode == 1:
        tab = dev(rf.gdn_get_digital_permutations(rng, r, d, tnew, 1, bases))
    else:
        tab = dev(rf.gdn_get_digital_shifts(rng, r, d, tnew, 1, bases))
    return dict(mode=mode, r=r, n=n, d=d, mmax=mmax, tmax=tmax, tnew=tnew, n_start=n_start, bmax=bmax, bases=tb,
                C=Cl, tab=tab)


def _run(torch, fn, c, x):
    s = torch.cuda.current_stream()
    return fn(c["mode"], c["r"], c["n"], c["d"], 1, c["mmax"], c["tmax"], c["tnew"], c["n_start"], c["bmax"],
              c["bases"].data_ptr(), c["C"].data_ptr(), c["tab"].data_ptr(), x.data_ptr(), s.cuda_stream)


@pytest.mark.parametrize("mode", [1, 2])
def test_noalloc_matches_the_allocating_entry_point(torch, mode):
    ref, noalloc = _fn("qmc_gdn_gen_scramble_float_dev"), _fn("qmc_gdn_gen_scramble_float_noalloc_dev")
    # the first call fills most of the workspace (r*d*mmax*tmax + r*d = 266496
    # words, config 2's shape with fewer points); the second, smaller call
    # finds the first one's table there
    cases = [_case(torch, 7 + mode, mode, 8, 3 ** 9, 32, 20, 52, 52, 0),
             _case(torch, 9 + mode, mode, 3, 5000, 5, 12, 40, 52, 1234)]
    for c in cases:
        print("noalloc case: mode %d r %d n %d d %d" % (mode, c["r"], c["n"], c["d"]), flush=True)
        want = torch.empty((c["r"], c["n"], c["d"]), dtype=torch.float64, device="cuda")
        got = torch.full_like(want, -1.0)
        assert _run(torch, ref, c, want) == 0, _lib.lib().qmc_last_error()
        assert _run(torch, noalloc, c, got) == 0, _lib.lib().qmc_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.int64), want.view(torch.int64))
        assert float(got.min()) >= 0.0 and float(got.max()) < 1.0


def test_noalloc_rejects_a_workspace_above_its_size(torch):
    """r*d*mmax*tmax + r*d = 16*64*20*52 + 1024 words > 2^20: refused before
    any launch, and the allocating entry point still runs the same call."""
    noalloc, ref = _fn("qmc_gdn_gen_scramble_float_noalloc_dev"), _fn("qmc_gdn_gen_scramble_float_dev")
    r, d, mmax, tmax = 16, 64, 20, 52
    assert r * d * mmax * tmax + r * d > SCRATCH_WORDS
    bases = torch.full((1, d), 3, dtype=torch.int64, device="cuda")
    C = torch.zeros((r, d, mmax, tmax), dtype=torch.int64, device="cuda")
    shifts = torch.zeros((r, d, tmax), dtype=torch.int64, device="cuda")
    x = torch.empty((r, 16, d), dtype=torch.float64, device="cuda")
    c = dict(mode=2, r=r, n=16, d=d, mmax=mmax, tmax=tmax, tnew=tmax, n_start=0, bmax=3, bases=bases, C=C,
             tab=shifts)
    rc = _run(torch, noalloc, c, x)
    assert rc == 1 and b"scratch" in _lib.lib().qmc_last_error()
    assert _run(torch, ref, c, x) == 0
    torch.cuda.synchronize()
    assert float(x.abs().max()) == 0.0  # zero matrices and shifts give the origin
