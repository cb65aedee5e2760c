"""Seeded, valid inputs for every kernel, shared by the parity tests and by the
fixture writer (tests/golden/make_golden.py).

`make_case(name, seed, size)` returns an ordered dict argname -> value in the
reference's positional order (qmctoolscl_amd._api.SIGNATURES): Python ints for
scalars, fresh numpy arrays for pointers.  Output arrays are pre-filled with a
sentinel so that rows a kernel leaves untouched are compared too.

Sizes: "tiny" (brute-force checks), "small" (seconds on the CPU oracle),
"medium" (digests of the reference C backend).
"""
from collections import OrderedDict

import numpy as np

from qmctoolscl_amd._api import SIGNATURES

PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71,
          73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131]
SENTINEL_U64 = np.uint64(0xDEADBEEFCAFEF00D)
SENTINEL_F64 = -7.25

_SIZES = {
    "tiny": dict(r=2, n_log=3, d=3, mmax=5, nh_log=3, batch=2),
    "small": dict(r=3, n_log=9, d=5, mmax=12, nh_log=8, batch=6),
    "medium": dict(r=4, n_log=12, d=9, mmax=16, nh_log=11, batch=10),
}


def _u64(rng, shape, bits=64):
    if bits >= 64:
        return rng.integers(0, 2**64, size=shape, dtype=np.uint64)
    return rng.integers(0, 2**bits, size=shape, dtype=np.uint64)


def _out(shape, kind):
    if kind == "pu64":
        return np.full(shape, SENTINEL_U64, dtype=np.uint64)
    return np.full(shape, SENTINEL_F64, dtype=np.float64)


def _aligned_range(rng, n_log):
    """(n_start, n) with n_start and n_start+n each 0 or a power of two."""
    if rng.random() < 0.5:
        return 0, 2**n_log
    a = int(rng.integers(0, n_log))
    return 2**a, 2**n_log - 2**a


def make_case(name, seed=0, size="small"):
    rng = np.random.default_rng([seed, sum(map(ord, name))])
    z = _SIZES[size]
    r, d = z["r"], z["d"]
    n = 2 ** z["n_log"]
    c = OrderedDict()

    if name == "lat_gen_linear":
        c.update(r=r, n=n + 3, d=d)  # n need not be a power of two here
        g = 2 * rng.integers(0, 2**19, size=(r, d), dtype=np.uint64) + 1
        g[0, 0] = 1
        if r > 1:
            g[1, :] = _u64(rng, d)  # huge generators exercise the u64 -> double rounding
        c["g"] = g
        c["x"] = _out((r, n + 3, d), "pf64")
    elif name in ("lat_gen_natural", "lat_gen_gray"):
        if name == "lat_gen_natural":
            n_start, nn = _aligned_range(rng, z["n_log"])
        else:
            n_start, nn = int(rng.integers(0, 2**z["n_log"])), n - 5
        c.update(r=r, n=nn, d=d, n_start=n_start)
        c["g"] = 2 * rng.integers(0, 2**19, size=(r, d), dtype=np.uint64) + 1
        c["x"] = _out((r, nn, d), "pf64")
    elif name == "lat_shift_mod_1":
        r_x = max(1, r // 2)
        rr = 2 * r_x + 1
        c.update(r=rr, n=n, d=d, r_x=r_x)
        x = rng.random((r_x, n, d))
        x.ravel()[:7] = [0.0, 0.5, 1 - 2**-53, 2**-60, 0.75, 0.25, 0.999]
        c["x"] = x
        sh = rng.random((rr, d))
        sh.ravel()[:3] = [0.5, 0.25, 1 - 2**-53]
        c["shifts"] = sh
        c["xr"] = _out((rr, n, d), "pf64")
    elif name == "dnb2_gmat_lsb_to_msb":
        mmax = z["mmax"]
        tmaxes = rng.integers(mmax, 65, size=r, dtype=np.uint64)
        tmaxes[0] = 64
        c.update(r=r, d=d, mmax=mmax)
        c["tmaxes"] = tmaxes
        C = np.empty((r, d, mmax), dtype=np.uint64)
        for l in range(r):
            C[l] = _u64(rng, (d, mmax), int(tmaxes[l]))
        c["C_lsb"] = C
        c["C_msb"] = _out((r, d, mmax), "pu64")
    elif name == "dnb2_linear_matrix_scramble":
        mmax = z["mmax"]
        r_C = max(1, r // 2)
        rr = 2 * r_C
        tmax_new = int(rng.integers(mmax, 65))
        c.update(r=rr, d=d, mmax=mmax, r_C=r_C, tmax_new=tmax_new)
        c["S"] = _u64(rng, (rr, d, tmax_new))
        c["C"] = _u64(rng, (r_C, d, mmax))
        c["C_lms"] = _out((rr, d, mmax), "pu64")
    elif name in ("dnb2_gen_natural", "dnb2_gen_gray"):
        mmax = z["mmax"]
        if name == "dnb2_gen_natural":
            n_start, nn = _aligned_range(rng, z["n_log"])
        else:
            n_start, nn = int(rng.integers(0, 2**z["n_log"])), n - 3
        c.update(r=r, n=nn, d=d, n_start=n_start, mmax=mmax)
        c["C"] = _u64(rng, (r, d, mmax))
        c["xb"] = _out((r, nn, d), "pu64")
    elif name == "dnb2_digital_shift":
        r_x = max(1, r // 2)
        rr = 3 * r_x
        c.update(r=rr, n=n, d=d, r_x=r_x)
        c["lshifts"] = rng.integers(0, 64, size=r_x, dtype=np.uint64)
        c["xb"] = _u64(rng, (r_x, n, d))
        c["shiftsb"] = _u64(rng, (rr, d))
        c["xrb"] = _out((rr, n, d), "pu64")
    elif name == "dnb2_integer_to_float":
        tmaxes = rng.integers(1, 65, size=r, dtype=np.uint64)
        tmaxes[0] = 64
        c.update(r=r, n=n, d=d)
        c["tmaxes"] = tmaxes
        xb = _u64(rng, (r, n, d))
        # values next to the u64 -> double rounding boundary (the 1.0 quirk)
        xb.ravel()[:6] = [2**64 - 1, 2**64 - 1024, 2**64 - 1025, 2**63 + 1, 2**53 + 1, 0]
        c["xb"] = xb
        c["x"] = _out((r, n, d), "pf64")
    elif name in ("dnb2_interlace", "dnb2_undo_interlace"):
        mmax = z["mmax"]
        alpha = int(rng.integers(2, 4))
        d_alpha = max(1, d // alpha)
        dd = d_alpha * alpha
        tmax = int(rng.integers(8, 64 // alpha + 1))
        tmax_alpha = min(alpha * tmax, 64)
        if name == "dnb2_interlace":
            c.update(r=r, d_alpha=d_alpha, mmax=mmax, d=dd, tmax=tmax,
                     tmax_alpha=tmax_alpha, alpha=alpha)
            c["C"] = _u64(rng, (r, dd, mmax), tmax)
            c["C_alpha"] = _out((r, d_alpha, mmax), "pu64")
        else:
            c.update(r=r, d=dd, mmax=mmax, d_alpha=d_alpha, tmax=tmax,
                     tmax_alpha=tmax_alpha, alpha=alpha)
            c["C_alpha"] = _u64(rng, (r, d_alpha, mmax), tmax_alpha)
            c["C"] = _out((r, dd, mmax), "pu64")
    elif name in ("gdn_gen_natural", "gdn_gen_natural_same_base"):
        mmax = 6
        tmax = int(rng.integers(4, 12))
        if name == "gdn_gen_natural":
            r_b = max(1, r // 2)
            bases = np.array(rng.choice(PRIMES[:12], size=(r_b, d)), dtype=np.uint64)
            bmin = int(bases.min())
            nn = min(n, bmin**mmax - 7)
            c.update(r=r, n=nn, d=d, r_b=r_b, mmax=mmax, tmax=tmax, n_start=7)
            c["bases"] = bases
            bb = np.array([[bases[l % r_b, j] for j in range(d)] for l in range(r)], dtype=np.uint64)
            C = (rng.random((r, d, mmax, tmax)) * (bb[:, :, None, None] * tmax)).astype(np.uint64)
        else:
            b = int(rng.choice([3, 5, 7]))
            nn = min(n, b**mmax - 5)
            c.update(r=r, n=nn, d=d, mmax=mmax, tmax=tmax, n_start=5, b=b)
            C = rng.integers(0, b * tmax, size=(r, d, mmax, tmax), dtype=np.uint64)
        c["C"] = C
        c["xdig"] = _out((r, nn, d, tmax), "pu64")
    elif name == "gdn_linear_matrix_scramble":
        mmax, tmax = 5, 6
        tmax_new = int(rng.integers(tmax, 14))
        r_C = max(1, r // 2)
        r_b = r_C
        rr = 2 * r_C
        bases = np.array(rng.choice(PRIMES, size=(r_b, d)), dtype=np.uint64)
        c.update(r=rr, d=d, mmax=mmax, r_C=r_C, r_b=r_b, tmax=tmax, tmax_new=tmax_new)
        c["bases"] = bases
        c["S"] = rng.integers(0, 1000, size=(rr, d, tmax_new, tmax), dtype=np.uint64)
        c["C"] = rng.integers(0, 1000, size=(r_C, d, mmax, tmax), dtype=np.uint64)
        c["C_lms"] = _out((rr, d, mmax, tmax_new), "pu64")
    elif name in ("gdn_digital_permutation", "gdn_digital_shift", "gdn_integer_to_float"):
        tmax = int(rng.integers(3, 10))
        tmax_new = tmax + int(rng.integers(0, 5))
        r_x = max(1, r // 2)
        r_b = r_x
        rr = 2 * r_x if name != "gdn_integer_to_float" else r_x
        bases = np.array(rng.choice(PRIMES, size=(r_b, d)), dtype=np.uint64)
        bb = np.array([[bases[l % r_b, j] for j in range(d)] for l in range(rr)], dtype=np.uint64)
        if name == "gdn_integer_to_float":
            c.update(r=rr, n=n, d=d, r_b=r_b, tmax=tmax_new)
            c["bases"] = bases
            c["xdig"] = (rng.random((rr, n, d, tmax_new)) * bb[:, None, :, None]).astype(np.uint64)
            c["x"] = _out((rr, n, d), "pf64")
            return c
        bbx = bb[:r_x]
        xdig = (rng.random((r_x, n, d, tmax)) * bbx[:, None, :, None]).astype(np.uint64)
        if name == "gdn_digital_permutation":
            bmax = int(bases.max())
            perms = np.zeros((rr, d, tmax_new, bmax), dtype=np.uint64)
            for l in range(rr):
                for j in range(d):
                    b = int(bb[l, j])
                    for t in range(tmax_new):
                        perms[l, j, t, :b] = rng.permutation(b)
            c.update(r=rr, n=n, d=d, r_x=r_x, r_b=r_b, tmax=tmax, tmax_new=tmax_new, bmax=bmax)
            c["perms"] = perms
        else:
            c.update(r=rr, n=n, d=d, r_x=r_x, r_b=r_b, tmax=tmax, tmax_new=tmax_new)
            c["bases"] = bases
            c["shifts"] = (rng.random((rr, d, tmax_new)) * bb[:, :, None]).astype(np.uint64)
        c["xdig"] = xdig
        c["xdig_new"] = _out((rr, n, d, tmax_new), "pu64")
    elif name in ("gdn_interlace", "gdn_undo_interlace"):
        mmax = 4
        alpha = int(rng.integers(2, 4))
        d_alpha = max(1, d // alpha)
        dd = d_alpha * alpha
        tmax = int(rng.integers(3, 8))
        tmax_alpha = alpha * tmax
        if name == "gdn_interlace":
            c.update(r=r, d_alpha=d_alpha, mmax=mmax, d=dd, tmax=tmax,
                     tmax_alpha=tmax_alpha, alpha=alpha)
            c["C"] = rng.integers(0, 7, size=(r, dd, mmax, tmax), dtype=np.uint64)
            c["C_alpha"] = _out((r, d_alpha, mmax, tmax_alpha), "pu64")
        else:
            c.update(r=r, d=dd, mmax=mmax, d_alpha=d_alpha, tmax=tmax,
                     tmax_alpha=tmax_alpha, alpha=alpha)
            c["C_alpha"] = rng.integers(0, 7, size=(r, d_alpha, mmax, tmax_alpha), dtype=np.uint64)
            c["C"] = _out((r, dd, mmax, tmax), "pu64")
    elif name == "fwht_1d_radix2":
        nh = 2 ** z["nh_log"]
        d1, d2 = 2, max(1, z["batch"] // 2)
        c.update(d1=d1, d2=d2, n_half=nh)
        c["x"] = rng.standard_normal((d1, d2, 2 * nh))
    elif name in ("fft_bro_1d_radix2", "ifft_bro_1d_radix2"):
        nh = 2 ** z["nh_log"]
        d1, d2 = 2, max(1, z["batch"] // 2)
        c.update(d1=d1, d2=d2, n_half=nh)
        c["twiddler"] = _out(2 * nh, "pf64")
        c["twiddlei"] = _out(2 * nh, "pf64")
        c["xr"] = rng.standard_normal((d1, d2, 2 * nh))
        c["xi"] = rng.standard_normal((d1, d2, 2 * nh))
    else:
        raise KeyError(name)
    assert list(c) == [s[0] for s in SIGNATURES[name]], (name, list(c))
    return c


def copy_case(case):
    return OrderedDict((k, v.copy() if isinstance(v, np.ndarray) else v) for k, v in case.items())


def outputs_of(name):
    return [s[0] for s in SIGNATURES[name] if len(s) > 2 and s[2] != "in"]
