"""Replays the reference's doctest session (docs/index.md) against any object
with the qmctoolscl call surface (the CPU oracle adapter or qmctoolscl_amd)
and compares every displayed array with the transcribed golden vectors in
tests/golden/doctest_vectors.json.

The inputs are the doctests' own: literal arrays (restated here) and draws
from numpy.random.Generator(SFC64(7)) consumed in document order, so the
random helpers are exercised too.  Integer outputs must match exactly; float
outputs to the doctests' print precision (8 decimals)."""
import contextlib
import io
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "doctest_vectors.json")


def golden():
    with open(GOLDEN) as f:
        data = json.load(f)
    arrays = {}
    for a in data["arrays"]:
        dt = np.uint64 if a["dtype"] == "uint64" else np.float64
        arrays[a["line"]] = (a["name"], np.array(a["values"], dtype=dt).reshape(a["shape"]))
    return arrays, data["dnb2_S_printed"]


class Checker:
    def __init__(self):
        self.arrays, self.s_printed = golden()
        self.results = []

    def __call__(self, line, got):
        name, want = self.arrays[line]
        got = np.asarray(got)
        if want.dtype == np.uint64:
            ok = got.shape == want.shape and np.array_equal(got.astype(np.uint64), want)
            err = 0 if ok else int(np.count_nonzero(got.astype(np.uint64) != want)) if got.shape == want.shape else -1
        else:
            ok = got.shape == want.shape and np.allclose(got, want, rtol=0, atol=1e-8)
            err = float(np.max(np.abs(got - want))) if got.shape == want.shape else -1
        self.results.append((line, name, bool(ok), err))

    def failures(self):
        return [r for r in self.results if not r[2]]


def _u(v):
    return np.uint64(v)


def replay(api, to_device=None, from_device=None, nus=True):
    """Run the doctest session on `api`.  `to_device`/`from_device` move arrays
    to and from the buffers `api` works on (identity for numpy).  nus=False
    skips the two nested-uniform-scrambling blocks (they draw from their own
    generators, so the shared stream is unaffected)."""
    td = to_device or (lambda a: a)
    fd = from_device or (lambda a: a)
    chk = Checker()
    rng = np.random.Generator(np.random.SFC64(7))

    # ---- lattices (docs/index.md:64-208)
    g = np.array([[1, 433461, 315689, 441789, 501101],
                  [1, 182667, 469891, 498753, 110745]], dtype=np.uint64)
    r, n, d = _u(2), _u(8), _u(5)
    x = td(np.empty((2, 8, 5)))
    api.lat_gen_linear(r, n, d, td(g), x)
    chk(73, fd(x))
    n_start, n = _u(2), _u(6)
    x = td(np.empty((2, 6, 5)))
    api.lat_gen_natural(r, n, d, n_start, td(g), x)
    chk(110, fd(x))
    api.lat_gen_gray(r, n, d, n_start, td(g), x)
    chk(140, fd(x))
    r_x = r
    r = _u(2 * r_x)
    shifts = api.lat_get_shifts(rng, r, d)
    xr = td(np.empty((4, 6, 5)))
    api.lat_shift_mod_1(r, n, d, r_x, x, td(shifts), xr)
    chk(181, fd(xr))

    # ---- base-2 digital nets (docs/index.md:228-677)
    C_lsb = np.array([[1, 2, 4, 8], [1, 3, 5, 15], [1, 3, 6, 9], [1, 3, 4, 10]], dtype=np.uint64)
    r, d, mmax, tmax = _u(1), _u(4), _u(4), 4
    tmaxes = np.tile(np.uint64(tmax), int(r))
    C = td(np.empty((4, 4), dtype=np.uint64))
    api.dnb2_gmat_lsb_to_msb(r, d, mmax, td(tmaxes), td(C_lsb), C)
    chk(241, fd(C))
    r_C = r
    r = _u(2 * r_C)
    tmax_new = _u(6)
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        S = api.dnb2_get_linear_scramble_matrix(rng, r, d, tmax, tmax_new, np.uint8(True))
    printed = buf.getvalue().splitlines()
    rows = [s.strip() for s in printed if s.strip() and set(s.strip()) <= {"0", "1"}]
    chk.results.append((chk.s_printed["line"], "S(printed)",
                        printed[0] == chk.s_printed["header"] and rows == chk.s_printed["rows"], 0))
    C_lms = td(np.empty((2, 4, 4), dtype=np.uint64))
    api.dnb2_linear_matrix_scramble(r, d, mmax, r_C, tmax_new, td(S), C, C_lms)
    chk(339, fd(C_lms))
    alpha = _u(2)
    d_alpha = _u(d // alpha)
    tmax = tmax_new
    tmax_alpha = min(alpha * tmax_new, 64)
    C_alpha = td(np.empty((2, 2, 4), dtype=np.uint64))
    api.dnb2_interlace(r, d_alpha, mmax, d, tmax, tmax_alpha, alpha, C_lms, C_alpha)
    chk(375, fd(C_alpha))
    C_lms_cp = td(np.empty((2, 4, 4), dtype=np.uint64))
    api.dnb2_undo_interlace(r, d, mmax, d_alpha, tmax, tmax_alpha, alpha, C_alpha, C_lms_cp)
    chk.results.append((401, "undo_interlace==C_lms", bool(np.array_equal(fd(C_lms_cp), fd(C_lms))), 0))
    C, d = C_alpha, d_alpha
    n_start, n = _u(2), _u(14)
    xb = td(np.empty((2, 14, 2), dtype=np.uint64))
    api.dnb2_gen_natural(r, n, d, n_start, mmax, C, xb)
    chk(425, fd(xb))
    api.dnb2_gen_gray(r, n, d, n_start, mmax, C, xb)
    chk(472, fd(xb))
    tmax = tmax_alpha
    tmax_new = _u(64)
    lshifts = np.tile(tmax_new - tmax, int(r))
    r_x = r
    r = _u(2 * r_x)
    shiftsb = api.dnb2_get_digital_shifts(rng, r, d, tmax_new)
    chk(532, shiftsb)
    xrb = td(np.empty((4, 14, 2), dtype=np.uint64))
    api.dnb2_digital_shift(r, n, d, r_x, td(lshifts), xb, td(shiftsb), xrb)
    chk(539, fd(xrb))
    x = td(np.empty((4, 14, 2)))
    tmaxes_new = np.tile(tmax_new, int(r))
    api.dnb2_integer_to_float(r, n, d, td(tmaxes_new), xrb, x)
    chk(617, fd(x))

    # ---- base-2 nested uniform scrambling (docs/index.md:682-736)
    if nus:
        _replay_nus2(api, chk, td, fd)
    _replay_gdn(api, chk, td, fd, rng, nus)
    return chk


def _replay_nus2(api, chk, td, fd):
    r, n_start, n, d = _u(1), _u(0), _u(8), _u(4)
    C = np.array([[8, 4, 2, 1], [8, 12, 10, 15], [8, 12, 6, 9], [8, 12, 2, 5]], dtype=np.uint64)
    mmax = _u(C.shape[1])
    tmax = _u(np.ceil(np.log2(np.max(C))))
    xb = td(np.empty((1, 8, 4), dtype=np.uint64))
    api.dnb2_gen_natural(r, n, d, n_start, mmax, td(C), xb)
    tmax_new = _u(2 * tmax)
    r_x = r
    r = _u(2 * r_x)
    seeds = np.random.SeedSequence(7).spawn(int(r * d))
    rngs = np.array([np.random.Generator(np.random.SFC64(s)) for s in seeds]).reshape(int(r), int(d))
    roots = np.array([api.NUSNode_dnb2() for _ in range(int(r * d))]).reshape(int(r), int(d))
    xrb = td(np.zeros((2, 8, 4), dtype=np.uint64))
    api.dnb2_nested_uniform_scramble(r, n, d, r_x, tmax, tmax_new, rngs, roots, xb, xrb)
    chk(719, fd(xrb))


def _replay_gdn(api, chk, td, fd, rng, nus):
    # ---- generalized digital nets (docs/index.md:761-1416)
    bases = np.array([[2, 3], [5, 7]], dtype=np.uint64)
    r_C = r_b = _u(bases.shape[0])
    d = _u(bases.shape[1])
    mmax = _u(5)
    tmax = mmax
    C = api.gdn_get_halton_generating_matrix(r_C, d, mmax)
    tmax_new = _u(2 * tmax)
    r = _u(2 * r_C)
    S = api.gdn_get_linear_scramble_matrix(rng, r, d, tmax, tmax_new, r_b, bases)
    chk(788, S)
    C_lms = td(np.empty((4, 2, 5, 10), dtype=np.uint64))
    api.gdn_linear_matrix_scramble(r, d, mmax, r_C, r_b, tmax, tmax_new, td(bases), td(S), td(C), C_lms)
    chk(881, fd(C_lms))
    tmax = tmax_new
    C = C_lms
    n_start, n = _u(2), _u(6)
    xdig = td(np.empty((4, 6, 2, 10), dtype=np.uint64))
    api.gdn_gen_natural(r, n, d, r_b, mmax, tmax, n_start, td(bases), C, xdig)
    chk(957, fd(xdig))
    r_x = r
    tmax_new = _u(12)
    shifts = api.gdn_get_digital_shifts(rng, r, d, tmax_new, r_b, bases)
    chk(1065, shifts)
    xdig_new = td(np.empty((4, 6, 2, 12), dtype=np.uint64))
    api.gdn_digital_shift(r, n, d, r_x, r_b, tmax, tmax_new, td(bases), td(shifts), xdig, xdig_new)
    chk(1079, fd(xdig_new))
    perms = api.gdn_get_digital_permutations(rng, r, d, tmax_new, r_b, bases)
    chk(1185, perms)
    bmax = bases.max().astype(np.uint64)
    api.gdn_digital_permutation(r, n, d, r_x, r_b, tmax, tmax_new, bmax, td(perms), xdig, xdig_new)
    chk(1294, fd(xdig_new))
    x = td(np.empty((4, 6, 2)))
    api.gdn_integer_to_float(r, n, d, r_b, tmax_new, td(bases), xdig_new, x)
    chk(1388, fd(x))

    # ---- base-b nested uniform scrambling (docs/index.md:1421-1552)
    if nus:
        _replay_nusb(api, chk, td, fd)
    _replay_tail(api, chk, td, fd, rng)


def _replay_nusb(api, chk, td, fd):
    r, n_start, n = _u(1), _u(0), _u(9)
    bases = np.array([[2, 3, 5]], dtype=np.uint64)
    r_b = _u(1)
    d = _u(3)
    mmax = _u(3)
    tmax = mmax
    C = np.tile(np.eye(int(mmax), dtype=np.uint64)[None, None, :, :], (int(r), int(d), 1, 1))
    xdig = td(np.empty((1, 9, 3, 3), dtype=np.uint64))
    api.gdn_gen_natural(r, n, d, r_b, mmax, tmax, n_start, td(bases), td(C), xdig)
    r_x = r
    r = _u(2)
    tmax_new = _u(8)
    seeds = np.random.SeedSequence(7).spawn(int(r * d))
    rngs = np.array([np.random.Generator(np.random.SFC64(s)) for s in seeds]).reshape(int(r), int(d))
    roots = np.array([api.NUSNode_gdn() for _ in range(int(r * d))]).reshape(int(r), int(d))
    xrdig = td(np.zeros((2, 9, 3, 8), dtype=np.uint64))
    api.gdn_nested_uniform_scramble(r, n, d, r_x, r_b, tmax, tmax_new, rngs, roots, bases, xdig, xrdig)
    chk(1457, fd(xrdig))
    xr = td(np.empty((2, 9, 3)))
    api.gdn_integer_to_float(r, n, d, r_b, tmax_new, td(bases), xrdig, xr)
    chk(1532, fd(xr))


def _replay_tail(api, chk, td, fd, rng):
    # ---- base-b interlacing (docs/index.md:1570-1667)
    r, d, alpha = _u(2), _u(6), _u(3)
    d_alpha = _u(d // alpha)
    mmax, tmax = _u(3), _u(4)
    tmax_alpha = _u(alpha * tmax)
    C = rng.integers(0, 5, (r, d, mmax, tmax), dtype=np.uint64)
    chk(1578, C)
    C_alpha = td(np.empty((2, 2, 3, 12), dtype=np.uint64))
    api.gdn_interlace(r, d_alpha, mmax, d, tmax, tmax_alpha, alpha, td(C), C_alpha)
    chk(1629, fd(C_alpha))
    C_cp = td(np.empty((2, 6, 3, 4), dtype=np.uint64))
    api.gdn_undo_interlace(r, d, mmax, d_alpha, tmax, tmax_alpha, alpha, C_alpha, C_cp)
    chk.results.append((1666, "gdn undo_interlace==C", bool(np.array_equal(fd(C_cp), C)), 0))

    # ---- same base (docs/index.md:1686-1732)
    n_start, n, b = _u(0), _u(9), _u(3)
    C = np.array([[[[1, 1, 0, 0], [1, 1, 1, 1]], [[2, 1, 0, 0], [2, 1, 2, 1]]]], dtype=np.uint64)
    r, d, mmax, tmax = (_u(s) for s in C.shape)
    xdig = td(np.empty((1, 9, 2, 4), dtype=np.uint64))
    api.gdn_gen_natural_same_base(r, n, d, mmax, tmax, n_start, b, td(C), xdig)
    chk(1706, fd(xdig))

    # ---- transforms (docs/index.md:1777-1927)
    d1 = d2 = _u(1)
    xr0 = np.array([1, 0, 1, 0, 0, 1, 1, 0], dtype=np.double)
    xi0 = np.array([0, 1, 0, 1, 1, 0, 0, 1], dtype=np.double)
    xr, xi = td(xr0.copy()), td(xi0.copy())
    twr, twi = td(np.empty(8)), td(np.empty(8))
    n_half = _u(4)
    api.fft_bro_1d_radix2(d1, d2, n_half, twr, twi, xr, xi)
    chk(1787, fd(xr))
    chk(1790, fd(xi))
    api.ifft_bro_1d_radix2(d1, d2, n_half, twr, twi, xr, xi)
    chk.results.append((1794, "ifft(fft(x))", bool(np.allclose(fd(xr), xr0, atol=1e-8)
                                                  and np.allclose(fd(xi), xi0, atol=1e-8)), 0))
    d1, d2, m = 5, 7, 8
    nn = 2**m
    ir = np.array([int(format(i, "0%db" % m)[::-1], 2) for i in range(nn)])
    xr0 = rng.uniform(0, 1, (d1, d2, nn)).astype(np.double)
    xi0 = rng.uniform(0, 1, (d1, d2, nn)).astype(np.double)
    want = np.fft.fft((xr0 + 1j * xi0)[:, :, ir], norm="ortho")
    xr, xi = td(xr0.copy()), td(xi0.copy())
    twr, twi = td(np.empty(nn)), td(np.empty(nn))
    api.fft_bro_1d_radix2(_u(d1), _u(d2), _u(nn // 2), twr, twi, xr, xi)
    chk.results.append((1818, "fft_bro vs numpy", bool(np.allclose(fd(xr) + 1j * fd(xi), want, atol=1e-8)), 0))
    api.ifft_bro_1d_radix2(_u(d1), _u(d2), _u(nn // 2), twr, twi, xr, xi)
    chk.results.append((1821, "ifft_bro round trip", bool(np.allclose(fd(xr) + 1j * fd(xi), xr0 + 1j * xi0, atol=1e-8)), 0))
    x = td(np.array([1, 0, 1, 0, 0, 1, 1, 0], dtype=np.double))
    api.fwht_1d_radix2(_u(1), _u(1), _u(4), x)
    chk(1904, fd(x))
    x0 = rng.uniform(0, 1, (d1, d2, nn)).astype(np.double)
    from scipy.linalg import hadamard
    want = x0 @ hadamard(nn).T / np.sqrt(nn)
    x = td(x0.copy())
    api.fwht_1d_radix2(_u(d1), _u(d2), _u(nn // 2), x)
    chk.results.append((1922, "fwht vs Hadamard/sqrt(n)", bool(np.allclose(fd(x), want, atol=1e-8)), 0))
    api.fwht_1d_radix2(_u(d1), _u(d2), _u(nn // 2), x)
    chk.results.append((1925, "fwht(fwht(x))", bool(np.allclose(fd(x), x0, atol=1e-8)), 0))
