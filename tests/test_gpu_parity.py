"""GPU parity of the HIP kernels against the CPU oracle, the reference's
doctest golden vectors and digests of the reference C backend's outputs.

Bar: bit-exact for every integer output and for the double outputs of the
lattice, integer_to_float and FWHT kernels; FFT/IFFT within 1e-12 relative
(max-abs error over max-abs value) of the oracle, whose only difference is
the twiddle table's cos/sin."""
import json
import os

import numpy as np
import pytest

import _clib
import cases
import replay
import qmctoolscl_amd as q
from qmctoolscl_amd import rand_funcs as rf
from qmctoolscl_amd._api import KERNEL_NAMES, SIGNATURES

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FFT_RTOL = 1e-12
FLOAT_EXACT = {"lat_gen_linear", "lat_gen_natural", "lat_gen_gray", "lat_shift_mod_1",
               "dnb2_integer_to_float", "gdn_integer_to_float", "fwht_1d_radix2"}


@pytest.fixture(scope="module", autouse=True)
def gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


def to_t(a):
    import torch
    if a.dtype == np.uint64:
        return torch.from_numpy(a.view(np.int64).copy()).cuda()
    return torch.from_numpy(a.copy()).cuda()


def from_t(t, like_u64):
    a = t.cpu().numpy()
    return a.view(np.uint64) if like_u64 else a


def run_product(name, case, path):
    names = [s[0] for s in SIGNATURES[name]]
    if path == "host":
        getattr(q, name)(*[case[k] for k in names])
        return case
    args = [to_t(case[k]) if isinstance(case[k], np.ndarray) else np.uint64(case[k]) for k in names]
    getattr(q, name)(*args)
    out = dict(case)
    for k, a in zip(names, args):
        if isinstance(case[k], np.ndarray):
            out[k] = from_t(a, case[k].dtype == np.uint64)
    return out


def compare(name, got, want):
    for o in cases.outputs_of(name):
        g, w = got[o], want[o]
        if name in ("fft_bro_1d_radix2", "ifft_bro_1d_radix2"):
            scale = max(1.0, float(np.max(np.abs(w))))
            err = float(np.max(np.abs(g - w))) / scale
            assert err <= FFT_RTOL, (name, o, err)
        else:
            assert np.array_equal(g.view(np.uint64), w.view(np.uint64)), (name, o)


@pytest.mark.parametrize("path", ["host", "device"])
def test_doctest_replay_gpu(path):
    if path == "host":
        chk = replay.replay(q, nus=hasattr(q, "dnb2_nested_uniform_scramble"))
    else:
        def fd(t):
            return from_t(t, str(t.dtype) in ("torch.int64", "torch.uint64"))
        chk = replay.replay(q, to_device=to_t, from_device=fd, nus=hasattr(q, "dnb2_nested_uniform_scramble"))
    assert len(chk.results) >= 34
    assert chk.failures() == []


@pytest.mark.parametrize("name", KERNEL_NAMES)
@pytest.mark.parametrize("size", ["tiny", "small"])
@pytest.mark.parametrize("path", ["host", "device"])
def test_kernel_matches_oracle(name, size, path):
    for seed in (3, 4):
        want = cases.make_case(name, seed=seed, size=size)
        got = cases.copy_case(want)
        _clib.call_oracle(name, want)
        got = run_product(name, got, path)
        compare(name, got, want)


def test_kernels_match_reference_digests():
    with open(os.path.join(HERE, "golden", "ref_digests.json")) as f:
        table = json.load(f)
    for key, outs in table["digests"].items():
        name, seed = key.split("/")
        if name in ("fft_bro_1d_radix2", "ifft_bro_1d_radix2"):
            continue  # twiddles come from a different cos/sin; covered with a tolerance above
        got = run_product(name, cases.make_case(name, seed=int(seed), size=table["size"]), "device")
        for o, dg in outs.items():
            assert _clib.digest(got[o]) == dg, key + ":" + o


# (tmax, tmax_new, wide): the LDS permutation kernel at even and odd digit
# counts, with tail rows past tmax; wide tables (entries of 2^32 or more)
# read the global table
@pytest.mark.parametrize("tmax,tnew,wide", [(52, 52, False), (6, 8, False), (6, 8, True), (5, 8, False),
                                            (6, 7, False)])
def test_gdn_permutation_lds_forms(tmax, tnew, wide):
    rng = np.random.default_rng(tmax * 100 + tnew)
    r, n, d, r_x = 2, 3000, 5, 1
    bases = np.array([[3, 5, 7, 11, 13]], dtype=np.uint64)
    bmax = 13
    perms = np.zeros((r, d, tnew, bmax), dtype=np.uint64)
    for l in range(r):
        for j in range(d):
            for t in range(tnew):
                perms[l, j, t] = rng.permutation(bmax)
    if wide:
        perms[:, 2, ::2, :] += np.uint64(2**40)
    xdig = (rng.random((r_x, n, d, tmax)) * bases[0][None, None, :, None]).astype(np.uint64)
    case = dict(r=r, n=n, d=d, r_x=r_x, r_b=1, tmax=tmax, tmax_new=tnew, bmax=bmax, perms=perms, xdig=xdig,
                xdig_new=np.zeros((r, n, d, tnew), dtype=np.uint64))
    want = dict(case, xdig_new=case["xdig_new"].copy())
    _clib.call_oracle("gdn_digital_permutation", want)
    got = run_product("gdn_digital_permutation", dict(case, xdig_new=case["xdig_new"].copy()), "device")
    assert np.array_equal(got["xdig_new"], want["xdig_new"])


# (n, d, r_g, r): the generators write replications in groups of 8 (grid y,
# looped past 65535 groups); the cases cover one group, partial groups, r_g
# not dividing the group start, shifts writing more than 8 replications from
# one x slab, and 75000 groups
@pytest.mark.parametrize("n,d,r_g,r", [(1 << 12, 7, 2, 6), (1 << 10, 13, 3, 19), (1 << 9, 6, 11, 20),
                                       (1 << 9, 5, 1, 20), (1 << 8, 3, 2, 37), (2, 1, 600000, 600000)])
def test_fused_lattice_equals_chain(n, d, r_g, r):
    import torch
    from qmctoolscl_amd import fused
    from oracle_api import OracleAPI
    orc = OracleAPI()
    rng = np.random.default_rng(1)
    g = (2 * rng.integers(0, 2**19, size=(r_g, d), dtype=np.uint64) + 1)
    shifts = rng.random((r, d))
    for order, fn in ((0, "lat_gen_linear"), (1, "lat_gen_natural"), (2, "lat_gen_gray")):
        args = (np.uint64(r_g), np.uint64(n), np.uint64(d)) + (() if order == 0 else (np.uint64(0),))
        chain = []
        for api in (q, orc):
            x = np.empty((r_g, n, d))
            getattr(api, fn)(*args, g, x)
            xr = np.empty((r, n, d))
            api.lat_shift_mod_1(np.uint64(r), np.uint64(n), np.uint64(d), np.uint64(r_g), x, shifts, xr)
            chain.append(xr.view(np.uint64))
        assert np.array_equal(chain[0], chain[1]), order
        out = torch.empty((r, n, d), dtype=torch.float64, device="cuda")
        fused.lat_gen_shift(order, r, n, d, r_g, 0, to_t(g), to_t(shifts), out)
        assert np.array_equal(out.cpu().numpy().view(np.uint64), chain[1]), order


def test_fused_dnb2_equals_chain():
    import torch
    from qmctoolscl_amd import fused
    rng = np.random.default_rng(2)
    r_x, r, n, d, mmax = 2, 4, 1 << 10, 5, 32
    C = rng.integers(0, 2**32, size=(r_x, d, mmax), dtype=np.uint64)
    lsh = np.array([32, 31], dtype=np.uint64)
    sh = rng.integers(0, 2**64, size=(r, d), dtype=np.uint64)
    tm = np.full(r, 64, dtype=np.uint64)
    for natural in (1, 0):
        xb = np.empty((r_x, n, d), dtype=np.uint64)
        fn = q.dnb2_gen_natural if natural else q.dnb2_gen_gray
        fn(np.uint64(r_x), np.uint64(n), np.uint64(d), np.uint64(0), np.uint64(mmax), C, xb)
        xrb = np.empty((r, n, d), dtype=np.uint64)
        q.dnb2_digital_shift(np.uint64(r), np.uint64(n), np.uint64(d), np.uint64(r_x), lsh, xb, sh, xrb)
        x = np.empty((r, n, d))
        q.dnb2_integer_to_float(np.uint64(r), np.uint64(n), np.uint64(d), tm, xrb, x)
        o_xrb = torch.empty((r, n, d), dtype=torch.int64, device="cuda")
        o_x = torch.empty((r, n, d), dtype=torch.float64, device="cuda")
        fused.dnb2_gen_shift_float(natural, r, n, d, r_x, 0, mmax, to_t(C), to_t(lsh), to_t(sh), to_t(tm),
                                   o_xrb, o_x)
        assert np.array_equal(from_t(o_xrb, True), xrb)
        assert np.array_equal(o_x.cpu().numpy().view(np.uint64), x.view(np.uint64))


@pytest.mark.parametrize("natural,n,n_start,r_C,tmax_new,lsh,d,mm", [
    (1, 1 << 10, 0, 3, 32, 32, 5, None),     # config 0's chain, one matrix set per replication
    (0, 1 << 10, 0, 1, 32, 32, 5, None),     # Gray order, the matrices shared by every replication
    (1, 1 << 9, 1 << 9, 3, 64, 0, 3, None),  # a later closed block, 64-bit LMS, no left shift
    (0, 1000, 37, 2, 48, 16, 9, None),       # ragged Gray range, r_C between 1 and r
    (1, 1 << 10, 1 << 10, 3, 32, 32, 2, None),
    (1, 1 << 8, 0, 1, 64, 0, 32, None),      # d * (2 mmax + tmax_new) near the LDS limit (60-bit columns)
    # prefix XORs as a lane scan (mmax dividing 64, d mmax <= 256): 16 and 64 columns
    (1, 1 << 12, 0, 3, 16, 48, 16, 16),
    (0, 1 << 11, 5, 2, 64, 0, 4, 64),
    (1, 1 << 10, 0, 1, 64, 0, 4, 64),
    # more dimensions x index bits than threads: the columns in strided loops
    (1, 1 << 10, 0, 2, 32, 32, 32, None),
    (0, 3000, 100, 3, 64, 0, 24, None),
])
def test_fused_dnb2_lms_equals_oracle_chain(natural, n, n_start, r_C, tmax_new, lsh, d, mm):
    """qmc_dnb2_lms_gen_shift_float_dev against the oracle's four-call chain
    (LMS, generation, digital shift with r_x = r, conversion), bit for bit.
    (Natural order over a range that Gray codes do not map onto itself makes
    the reference write outside its output, so the oracle is not run there.)"""
    import torch
    from qmctoolscl_amd import fused
    rng = np.random.Generator(np.random.SFC64(n + d))
    r, mmax = 3, mm if mm else 32 if tmax_new < 64 or lsh else 60
    C = rng.integers(0, 2**mmax if mmax < 64 else 2**63, size=(r_C, d, mmax), dtype=np.uint64)
    S = rf.dnb2_get_linear_scramble_matrix(rng, r, d, mmax, tmax_new, False)
    sh = rf.dnb2_get_digital_shifts(rng, r, d, 64)
    lshifts = np.full(r, lsh, dtype=np.uint64)
    tm = np.full(r, 64, dtype=np.uint64)
    Cl = np.empty((r, d, mmax), dtype=np.uint64)
    _clib.call_oracle("dnb2_linear_matrix_scramble", dict(r=r, d=d, mmax=mmax, r_C=r_C, tmax_new=tmax_new, S=S,
                                                          C=C, C_lms=Cl))
    xb = np.zeros((r, n, d), dtype=np.uint64)
    _clib.call_oracle("dnb2_gen_natural" if natural else "dnb2_gen_gray",
                      dict(r=r, n=n, d=d, n_start=n_start, mmax=mmax, C=Cl, xb=xb))
    xrb = np.empty_like(xb)
    _clib.call_oracle("dnb2_digital_shift", dict(r=r, n=n, d=d, r_x=r, lshifts=lshifts, xb=xb, shiftsb=sh, xrb=xrb))
    x = np.empty((r, n, d))
    _clib.call_oracle("dnb2_integer_to_float", dict(r=r, n=n, d=d, tmaxes=tm, xb=xrb, x=x))
    o_xrb = torch.zeros((r, n, d), dtype=torch.int64, device="cuda")
    o_x = torch.zeros((r, n, d), dtype=torch.float64, device="cuda")
    fused.dnb2_lms_gen_shift_float(natural, r, n, d, r_C, n_start, mmax, tmax_new, to_t(S), to_t(C),
                                   to_t(lshifts), to_t(sh), to_t(tm), o_xrb, o_x)
    torch.cuda.synchronize()
    assert np.array_equal(from_t(o_xrb, True), xrb)
    assert np.array_equal(o_x.cpu().numpy().view(np.uint64), x.view(np.uint64))


def test_full_size_lattice_rows_match_oracle():
    """lattice config: n = 2^20, d = 64, r = 16 shifts.  The oracle restarts
    at any n_start, so sampled blocks of rows of the full GPU output are
    checked exactly without running the oracle at full size."""
    import torch
    from qmctoolscl_amd import fused
    rng = np.random.Generator(np.random.SFC64(0))
    n, d, r = 1 << 20, 64, 16
    g = 2 * rng.integers(0, 2**19, size=(1, d), dtype=np.uint64) + 1
    shifts = rng.random((r, d))
    out = torch.empty((r, n, d), dtype=torch.float64, device="cuda")
    fused.lat_gen_shift(2, r, n, d, 1, 0, to_t(g), to_t(shifts), out)
    torch.cuda.synchronize()
    for start in (0, 4096, 777777, n - 1024):
        rows = 1024
        x = np.empty((1, rows, d))
        _clib.call_oracle("lat_gen_gray", dict(r=1, n=rows, d=d, n_start=start, g=g, x=x))
        xr = np.empty((r, rows, d))
        _clib.call_oracle("lat_shift_mod_1", dict(r=r, n=rows, d=d, r_x=1, x=x, shifts=shifts, xr=xr))
        got = out[:, start:start + rows, :].cpu().numpy()
        assert np.array_equal(got.view(np.uint64), xr.view(np.uint64)), start
    # every value lies in [0, 1)
    assert float(out.min()) >= 0.0 and float(out.max()) < 1.0


def test_errors_are_raised():
    g = np.ones((1, 2), dtype=np.uint64)
    x = np.empty((1, 4, 2))
    # the reference's int shift is undefined at k >= 2^31: rejected
    with pytest.raises(q.QMCError, match="QMC_ERR_RANGE"):
        q.lat_gen_gray(np.uint64(1), np.uint64(4), np.uint64(2), np.uint64(2**31), g, x)
    # natural order needing more than mmax columns
    C = np.ones((1, 2, 3), dtype=np.uint64)
    xb = np.empty((1, 16, 2), dtype=np.uint64)
    with pytest.raises(q.QMCError, match="QMC_ERR_RANGE"):
        q.dnb2_gen_natural(np.uint64(1), np.uint64(16), np.uint64(2), np.uint64(0), np.uint64(3), C, xb)
    # a base of 1 (the reference loops forever)
    bases = np.array([[1, 3]], dtype=np.uint64)
    Cg = np.ones((1, 2, 2, 2), dtype=np.uint64)
    xd = np.empty((1, 3, 2, 2), dtype=np.uint64)
    with pytest.raises(q.QMCError, match="QMC_ERR_RANGE"):
        q.gdn_gen_natural(np.uint64(1), np.uint64(3), np.uint64(2), np.uint64(1), np.uint64(2), np.uint64(2),
                          np.uint64(0), bases, Cg, xd)
    # a failed call leaves the next one unaffected
    q.lat_gen_gray(np.uint64(1), np.uint64(4), np.uint64(2), np.uint64(0), g, x)


def test_tensor_path_reports_status_bits_to_its_own_call(gpu):
    """A tensor call with a base of 1 raises (the bits it sets are read after
    its stream synchronizes), and the next, valid numpy call does not see
    them; with wait=False the bits stay for qmc_device_status."""
    torch = gpu
    U = np.uint64
    bases = to_t(np.array([[1, 3]], dtype=np.uint64))
    Cg = to_t(np.ones((1, 2, 2, 2), dtype=np.uint64))
    xd = torch.empty((1, 3, 2, 2), dtype=torch.int64, device="cuda")
    with pytest.raises(q.QMCError, match="QMC_ERR_RANGE"):
        q.gdn_gen_natural(U(1), U(3), U(2), U(1), U(2), U(2), U(0), bases, Cg, xd)
    g = np.ones((1, 2), dtype=np.uint64)
    x = np.empty((1, 4, 2))
    q.lat_gen_gray(U(1), U(4), U(2), U(0), g, x)
    q.gdn_gen_natural(U(1), U(3), U(2), U(1), U(2), U(2), U(0), bases, Cg, xd, wait=False)
    torch.cuda.synchronize()
    import ctypes
    bits = ctypes.c_uint(0)
    q._lib.check(q._lib.lib().qmc_device_status(ctypes.byref(bits), 1))
    assert bits.value != 0
    # a valid tensor call after the cleared word
    ok = to_t(np.array([[2, 3]], dtype=np.uint64))
    q.gdn_gen_natural(U(1), U(3), U(2), U(1), U(2), U(2), U(0), ok, Cg, xd)


def test_fused_entry_points_reject_undersized_tensors(gpu):
    import ctypes
    from qmctoolscl_amd import fused
    torch = gpu
    g = torch.ones((1, 8), dtype=torch.int64, device="cuda")
    sh = torch.zeros((4, 8), dtype=torch.float64, device="cuda")
    xr = torch.empty((4, 16, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(ctypes.ArgumentError, match="xr holds"):
        fused.lat_gen_shift(1, 4, 32, 8, 1, 0, g, sh, xr)        # needs 4*32*8
    with pytest.raises(ctypes.ArgumentError, match="shifts"):
        fused.lat_gen_shift(1, 4, 16, 8, 1, 0, g, sh[:3], xr)
    with pytest.raises(ctypes.ArgumentError, match="float64"):
        fused.lat_gen_shift(1, 4, 16, 8, 1, 0, g, sh.float(), xr)
    fused.lat_gen_shift(1, 4, 16, 8, 1, 0, g, sh, xr)
    torch.cuda.synchronize()


def test_empty_inputs_are_no_ops():
    g = np.ones((1, 2), dtype=np.uint64)
    x = np.empty((1, 0, 2))
    q.lat_gen_linear(np.uint64(1), np.uint64(0), np.uint64(2), g, x)
    xb = np.empty((0, 4, 2), dtype=np.uint64)
    q.dnb2_gen_gray(np.uint64(0), np.uint64(4), np.uint64(2), np.uint64(0), np.uint64(8),
                    np.ones((0, 2, 8), dtype=np.uint64), xb)


@pytest.mark.parametrize("m,V", [(12, 3), (13, 2), (14, 2), (16, 2), (18, 1), (20, 1), (25, 1)])
def test_large_transforms_match_oracle(m, V):
    """Lengths past one LDS tile exercise the register passes (and the
    stage-major twiddle table; 2^25: register groups of 4, 4 and 4 stages and
    one of a single stage above the 12-bit tile): FWHT still bit-exact,
    FFT/IFFT within 1e-12."""
    rng = np.random.default_rng(m)
    n = 1 << m
    x = rng.standard_normal((1, V, n))
    want = x.copy()
    _clib.call_oracle("fwht_1d_radix2", dict(d1=1, d2=V, n_half=n // 2, x=want))
    got = run_product("fwht_1d_radix2", dict(d1=1, d2=V, n_half=n // 2, x=x.copy()), "device")["x"]
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), m
    for name in ("fft_bro_1d_radix2", "ifft_bro_1d_radix2"):
        case = dict(d1=1, d2=V, n_half=n // 2, twiddler=np.empty(n), twiddlei=np.empty(n),
                    xr=rng.standard_normal((1, V, n)), xi=rng.standard_normal((1, V, n)))
        want = cases.copy_case(case)
        _clib.call_oracle(name, want)
        got = run_product(name, cases.copy_case(case), "device")
        compare(name, got, want)


@pytest.mark.parametrize("r,n,d,r_b,tmax,wide", [
    (2, 1000, 7, 1, 52, False),   # reciprocal table in LDS, a partial last chunk
    (3, 333, 5, 2, 33, True),     # odd tmax (odd word count per chunk), digits >= 2^32
    (1, 257, 200, 1, 40, False),  # table too large for LDS: global reciprocals
    (2, 130, 3, 1, 64, True),     # the largest staged row
    (2, 100, 3, 1, 65, False),    # longer rows: one lane per row
    (1, 64, 1, 1, 1, False),
])
def test_gdn_integer_to_float_shapes(r, n, d, r_b, tmax, wide):
    """The staged conversion (rows through LDS) at the shapes that select its
    variants, bit for bit against the oracle."""
    rng = np.random.default_rng(tmax)
    primes = np.array([2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53], dtype=np.uint64)
    bases = rng.choice(primes, size=(r_b, d)).astype(np.uint64)
    xdig = rng.integers(0, 131, size=(r, n, d, tmax), dtype=np.uint64)
    if wide:
        xdig[:, ::7, :, ::3] = rng.integers(2**32, 2**63, size=xdig[:, ::7, :, ::3].shape, dtype=np.uint64)
    want = np.empty((r, n, d))
    _clib.call_oracle("gdn_integer_to_float", dict(r=r, n=n, d=d, r_b=r_b, tmax=tmax, bases=bases, xdig=xdig,
                                                   x=want))
    for path in ("host", "device"):
        case = dict(r=np.uint64(r), n=np.uint64(n), d=np.uint64(d), r_b=np.uint64(r_b), tmax=np.uint64(tmax),
                    bases=bases.copy(), xdig=xdig.copy(), x=np.empty((r, n, d)))
        got = run_product("gdn_integer_to_float", case, path)["x"]
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), path


def _gen_case(r, n, d, r_b, mmax, tmax, n_start, bases, seed, big_col=None):
    rng = np.random.default_rng(seed)
    C = np.empty((r, d, mmax, tmax), dtype=np.uint64)
    for j in range(d):
        for rb in range(r):
            b = int(bases[rb % r_b, j])
            C[rb, j] = rng.integers(0, 3 * b, size=(mmax, tmax), dtype=np.uint64)  # entries past b-1 too
    if big_col is not None:
        C[:, big_col, :, ::5] = rng.integers(2**32, 2**62, size=C[:, big_col, :, ::5].shape, dtype=np.uint64)
    return dict(r=r, n=n, d=d, r_b=r_b, mmax=mmax, tmax=tmax, n_start=n_start, bases=bases, C=C,
                xdig=np.empty((r, n, d, tmax), dtype=np.uint64))


@pytest.mark.parametrize("shape", [
    # r, n, d, r_b, mmax, tmax, n_start, big_col: blocks of 256 points, carries across them, two block columns
    (2, 700, 50, 2, 12, 52, 1000, None),
    # a column with C >= 2^32 and base-2 columns whose indices outgrow mmax (direct digits)
    (2, 300, 6, 1, 5, 7, 3, 4),
    # indices past 2^32 (64-bit digit and carry paths)
    (1, 300, 3, 1, 40, 5, 2**32 - 100, None),
    (3, 1, 4, 3, 3, 2, 0, None),
])
def test_gdn_gen_point_walks(shape):
    """gdn_gen_natural's point walks (prefix table of C) and its per-digit
    fallback, bit for bit against the oracle at shapes that reach each path."""
    r, n, d, r_b, mmax, tmax, n_start, big_col = shape
    rng = np.random.default_rng(n)
    primes = np.array([2, 3, 5, 7, 11, 13, 17, 131, 257, 65537, 2**31 - 1], dtype=np.uint64)
    bases = rng.choice(primes, size=(r_b, d)).astype(np.uint64)
    if big_col is not None:
        bases[:, 0] = 2  # j = 0 reads past its mmax rows into column 1 (inside C)
        bases[:, -1] = 131  # the last column's reads stay inside C
    want = _gen_case(r, n, d, r_b, mmax, tmax, n_start, bases, 7, big_col)
    _clib.call_oracle("gdn_gen_natural", want)
    for path in ("host", "device"):
        got = run_product("gdn_gen_natural", _gen_case(r, n, d, r_b, mmax, tmax, n_start, bases, 7, big_col), path)
        assert np.array_equal(got["xdig"], want["xdig"]), path
    # one common base (where every index has at most mmax digits: past that the
    # reference reads beyond C)
    for b in (2, 3, 131):
        if b ** mmax <= n_start + n - 1:
            continue
        case = _gen_case(r, n, d, 1, mmax, tmax, n_start, np.full((1, d), b, dtype=np.uint64), 8)
        args = {k: case[k] for k in ("r", "n", "d", "mmax", "tmax", "n_start", "C", "xdig")}
        args["b"] = b
        want = cases.copy_case(args)
        _clib.call_oracle("gdn_gen_natural_same_base", want)
        got = run_product("gdn_gen_natural_same_base", cases.copy_case(args), "device")
        assert np.array_equal(got["xdig"], want["xdig"]), b


def _oracle_gdn_chain(mode, r, n, d, r_b, mmax, tmax, tnew, n_start, bmax, bases, C, tab):
    xdig = np.empty((r, n, d, tmax), dtype=np.uint64)
    _clib.call_oracle("gdn_gen_natural", dict(r=r, n=n, d=d, r_b=r_b, mmax=mmax, tmax=tmax, n_start=n_start,
                                              bases=bases, C=C, xdig=xdig))
    if mode == 0:
        y = xdig
    else:
        y = np.empty((r, n, d, tnew), dtype=np.uint64)
        if mode == 1:
            _clib.call_oracle("gdn_digital_permutation", dict(r=r, n=n, d=d, r_x=r, r_b=r_b, tmax=tmax, tmax_new=tnew,
                                                              bmax=bmax, perms=tab, xdig=xdig, xdig_new=y))
        else:
            _clib.call_oracle("gdn_digital_shift", dict(r=r, n=n, d=d, r_x=r, r_b=r_b, tmax=tmax, tmax_new=tnew,
                                                        bases=bases, shifts=tab, xdig=xdig, xdig_new=y))
    x = np.empty((r, n, d))
    _clib.call_oracle("gdn_integer_to_float", dict(r=r, n=n, d=d, r_b=r_b, tmax=tnew, bases=bases, xdig=y, x=x))
    return x


@pytest.mark.parametrize("mode,r,n,d,r_b,mmax,tmax,tnew,n_start,kind", [
    (1, 2, 20000, 7, 1, 10, 20, 20, 0, "halton"),     # two point ranges, LMS'd Halton matrices, permutations
    (1, 3, 3000, 5, 3, 12, 9, 14, 123, "random"),     # tail rows (tmax < tmax_new), r_b = r, n_start
    (2, 2, 2500, 6, 1, 12, 20, 26, 7, "random"),      # digital shifts (u64 sums that wrap)
    (0, 1, 5000, 4, 1, 14, 33, 33, 0, "random"),      # no scramble, 64-register rows
    (1, 2, 700, 4, 1, 6, 5, 8, 0, "wide"),            # C >= 2^32 column, perms >= 2^32, bases past b^mmax
    (1, 1, 600, 3, 1, 3, 4, 6, 0, "bigb"),            # permutation rows too long for LDS (b = 1999)
    (1, 2, 3000, 5, 1, 8, 52, 52, 0, "halton"),       # 52-register rows (the 52-digit layout)
    (1, 1, 1000, 3, 1, 8, 55, 56, 0, "halton"),       # 56-register rows
    (2, 1, 700, 2, 1, 10, 54, 54, 3, "random"),       # 56-register rows, shifts (2^10 > every index)
    (2, 2, 3000, 4, 2, 8, 60, 62, 5, "smallshift"),   # shifts < b (branch-free path), 64-register rows, tail
    (1, 1, 1, 1, 1, 3, 1, 1, 0, "random"),            # one point, one digit
    (0, 1, 700, 1, 1, 12, 12, 12, 0, "base2"),        # base 2 only
    (1, 1, 500, 3, 1, 40, 6, 8, 2**32 - 250, "random"),  # indices across 2^32 (64-bit digit and carry paths)
    (1, 2, 1500, 4, 1, 8, 10, 30, 0, "random"),       # value table of 16 rows, tail rows 16..29 past it
    (1, 2, 1200, 2, 1, 6, 6, 8, 0, "b593"),           # value table too large (b = 593): u32 rows in LDS
    (2, 1, 900, 3, 1, 8, 12, 40, 0, "random"),        # shift table of 16 rows, raw shifts 16..39 past it
    (2, 1, 800, 2, 1, 6, 60, 60, 0, "b131"),          # b = 131 above the 126 rows a 64-digit shift table holds
    (1, 2, 2000, 6, 1, 12, 52, 52, 0, "zerolead"),    # sums still 0 after 20 levels: every point past the
    (1, 1, 3000, 3, 1, 12, 40, 52, 0, "zerolead"),    # walked levels (digits from the index), tail past tmax
    (1, 2, 3000, 12, 1, 8, 52, 52, 0, "halton"),      # bases 29..37 in the 16-register instance beside the 52
    (1, 1, 2000, 12, 1, 10, 30, 52, 0, "halton"),     # the same with digits 16..29 past its registers, tail rows
    # base 2 with maps of {0, 1} onto itself: the digits as one bit word, the exact sum in one step
    (1, 2, 3000, 5, 1, 8, 52, 52, 0, "haltonperm"),
    (1, 1, 2000, 4, 1, 12, 30, 52, 3, "haltonperm"),  # tail rows (the image of digit 0)
    (1, 2, 1500, 3, 1, 8, 53, 53, 0, "haltonperm"),   # 53 digits, the longest exact sum
    (1, 1, 1500, 3, 1, 8, 54, 54, 0, "haltonperm"),   # 54 digits: the level walk
    (2, 2, 2500, 3, 1, 12, 40, 48, 7, "b2shift"),     # shifts, raw tail shifts of 0 or 1
])
def test_fused_gdn_equals_chain(mode, r, n, d, r_b, mmax, tmax, tnew, n_start, kind):
    """qmc_gdn_gen_scramble_float_dev against the oracle's three-call chain,
    bit for bit."""
    import torch
    from qmctoolscl_amd import fused
    rng = np.random.default_rng(mode * 100 + n)
    primes = np.array([2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37], dtype=np.uint64)
    if kind == "bigb":
        bases = np.full((r_b, d), 1999, dtype=np.uint64)
    elif kind == "base2":
        bases = np.full((r_b, d), 2, dtype=np.uint64)
    elif kind in ("b593", "b131"):
        bases = np.full((r_b, d), int(kind[1:]), dtype=np.uint64)
    else:
        halton = kind in ("halton", "haltonperm", "b2shift")
        bases = np.stack([primes[:d]] * r_b) if halton else rng.choice(primes, size=(r_b, d))
        bases = bases.astype(np.uint64)
    C = np.empty((r, d, mmax, tmax), dtype=np.uint64)
    for l in range(r):
        for j in range(d):
            b = int(bases[l % r_b, j])
            if kind in ("halton", "haltonperm", "b2shift"):  # identity scrambled by a random lower-triangular S
                C[l, j] = 0
                for m in range(min(mmax, tmax)):
                    C[l, j, m, m] = rng.integers(1, b)
                    C[l, j, m, m + 1:] = rng.integers(0, b, size=tmax - m - 1)
            else:
                C[l, j] = rng.integers(0, 3 * b, size=(mmax, tmax), dtype=np.uint64)
    if kind == "wide":
        C[:, 1, :, ::3] = rng.integers(2**32, 2**62, size=C[:, 1, :, ::3].shape, dtype=np.uint64)
        bases[:, 0] = 2  # 2^6 < 700: column 0 reads C rows past mmax (inside C); last column keeps a big base
        bases[:, -1] = 37
    bmax = int(bases.max())
    if mode == 1:
        tab = np.empty((r, d, tnew, bmax), dtype=np.uint64)
        for l in range(r):
            for j in range(d):
                for t in range(tnew):
                    tab[l, j, t] = rng.permutation(bmax).astype(np.uint64)
        if kind == "wide":
            tab[:, :, ::2, 1] += np.uint64(2**40)
        if kind == "haltonperm":  # each level's map is a permutation of the column's digits 0..b-1
            for l in range(r):
                for j in range(d):
                    b = int(bases[l % r_b, j])
                    for t in range(tnew):
                        tab[l, j, t, :b] = rng.permutation(b).astype(np.uint64)
                        tab[l, j, t, b:] = np.arange(b, bmax, dtype=np.uint64)
        if kind == "zerolead":
            # every digit of the first 20 levels maps to 0 (a table the reference
            # accepts): the sums stay 0 there, so no point can stop before level 20
            tab[:, :, :20, :] = 0
    elif mode == 2 and kind == "smallshift":
        tab = rng.integers(0, 2**64 - 1, size=(r, d, tnew), dtype=np.uint64)  # raw past tmax
        for l in range(r):
            for j in range(d):
                tab[l, j, :tmax] = rng.integers(0, int(bases[l % r_b, j]), size=tmax, dtype=np.uint64)
    elif mode == 2 and kind == "b2shift":
        tab = rng.integers(0, 2, size=(r, d, tnew), dtype=np.uint64)  # raw tail shifts of 0 or 1
        for l in range(r):
            for j in range(d):
                tab[l, j, :tmax] = rng.integers(0, int(bases[l % r_b, j]), size=tmax, dtype=np.uint64)
    elif mode == 2:
        tab = rng.integers(0, 2**64 - 1, size=(r, d, tnew), dtype=np.uint64)
    else:
        tab = None
    want = _oracle_gdn_chain(mode, r, n, d, r_b, mmax, tmax, tnew, n_start, bmax, bases, C, tab)
    x = torch.empty((r, n, d), dtype=torch.float64, device="cuda")
    fused.gdn_gen_scramble_float(mode, r, n, d, r_b, mmax, tmax, tnew, n_start, bmax, to_t(bases), to_t(C),
                                 to_t(tab) if tab is not None else None, x)
    got = x.cpu().numpy()
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
