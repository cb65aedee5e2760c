"""Host-side draws of the random scrambling inputs (shifts, linear scrambling
matrices, digit permutations).

These are the harness side of the path: they consume a caller-supplied
numpy.random.Generator in exactly the order of the reference's helpers
(qmctoolscl/rand_funcs.py:4-153), so the GPU kernels receive the same
matrices, shifts and permutations as the reference C backend would.  Same
names, arguments and return values as the reference.
"""
import numpy as np

_TOP = np.uint64(1 << 63)


def random_tbit_uint64s(rng, t, shape):
    """Array of the given shape of uniform t-bit unsigned integers, 0 <= t <= 64
    (rand_funcs.py:4-20).  For t == 64 the reference draws signed 64-bit values
    over the full range and moves them up by 2^63; flipping the top bit of the
    two's-complement pattern is the same map."""
    assert 0 <= t <= 64, "t must be between 0 and 64"
    if t < 64:
        return rng.integers(0, 1 << int(t), shape, dtype=np.uint64)
    s = rng.integers(-(1 << 63), 1 << 63, shape, dtype=np.int64)
    return np.asarray(s).view(np.uint64) ^ _TOP


def random_uint64_permutations(rng, n, b):
    """n permutations of 0..b-1 as rows of an (n, b) uint64 array, one
    rng.permutation(b) call per row in row order (rand_funcs.py:22-32)."""
    out = np.empty((n, b), dtype=np.uint64)
    for row in range(int(n)):
        out[row] = rng.permutation(b)
    return out


def lat_get_shifts(rng, r, d):
    """Uniform [0,1) shifts of shape (r, d) (rand_funcs.py:34-41)."""
    return rng.random((r, d))


def dnb2_get_digital_shifts(rng, r, d, tmax_new):
    """tmax_new-bit digital shifts of shape (r, d) (rand_funcs.py:43-51)."""
    return random_tbit_uint64s(rng, tmax_new, (r, d))


def dnb2_get_linear_scramble_matrix(rng, r, d, tmax, tmax_new, print_mats):
    """Scrambling matrices for base-2 LMS, shape (r, d, tmax_new), each row an
    MSB-first integer (rand_funcs.py:53-78).  Row t < tmax keeps t random bits
    above a unit diagonal bit (lower triangular); later rows are fully random.
    As in the reference, tmax_new < tmax is rejected by numpy broadcasting."""
    tmin = int(min(tmax_new, tmax))
    S = random_tbit_uint64s(rng, tmin, (r, d, tmax_new))
    drop = np.arange(tmin, 0, -1, dtype=np.uint64)  # low bits cleared in row t: tmin - t
    head = S[:, :, :tmin]
    head >>= drop
    head <<= drop
    head += np.uint64(1) << np.arange(int(tmax) - 1, -1, -1, dtype=np.uint64)
    if print_mats:
        print("S with shape (r=%d, d=%d, tmax_new=%d)" % (r, d, tmax_new))
        for l in range(r):
            print("l = %d" % l)
            for j in range(d):
                print("    j = %d" % j)
                for t in range(tmax_new):
                    bits = bin(S[l, j, t])[2:]
                    print("        " + "0" * int(tmax - len(bits)) + bits)
    return S


def gdn_get_linear_scramble_matrix(rng, r, d, tmax, tmax_new, r_b, bases):
    """Scrambling matrices for base-b LMS, shape (r, d, tmax_new, tmax):
    strictly-lower entries uniform in [0,b), diagonal uniform in [1,b), drawn
    per (l, j) in that order (rand_funcs.py:80-104)."""
    S = np.empty((r, d, tmax_new, tmax), dtype=np.uint64)
    b2 = np.atleast_2d(bases)
    below = np.tri(int(tmax_new), int(tmax), k=-1, dtype=bool)
    n_below = below.sum()
    diag = np.eye(tmax_new, tmax, dtype=bool)
    for l in range(r):
        row_b = b2[int(l % r_b)]
        for j in range(d):
            b = row_b[j]
            blk = np.zeros((tmax_new, tmax), dtype=np.uint64)
            blk[below] = rng.integers(0, b, n_below)
            blk[diag] = rng.integers(1, b, tmax)
            S[l, j] = blk
    return S


def gdn_get_halton_generating_matrix(r, d, mmax):
    """Identity generating matrices (Halton), shape (r, d, mmax, mmax)
    (rand_funcs.py:106-113)."""
    return np.tile(np.eye(mmax, dtype=np.uint64)[None, None, :, :], (int(r), int(d), 1, 1))


def gdn_get_digital_shifts(rng, r, d, tmax_new, r_b, bases):
    """Digit shifts uniform in [0, b), shape (r, d, tmax_new), drawn per (l, j)
    (rand_funcs.py:115-132)."""
    shifts = np.empty((r, d, tmax_new), dtype=np.uint64)
    b2 = np.atleast_2d(bases)
    for l in range(r):
        row_b = b2[int(l % r_b)]
        for j in range(d):
            shifts[l, j] = rng.integers(0, row_b[j], tmax_new, dtype=np.uint64)
    return shifts


def gdn_get_digital_permutations(rng, r, d, tmax_new, r_b, bases):
    """Digit permutations, shape (r, d, tmax_new, bmax), rows padded with zeros
    past b, one rng.permutation(b) per (l, j, t) (rand_funcs.py:134-153)."""
    b2 = np.atleast_2d(bases)
    bmax = b2.max()
    perms = np.zeros((r, d, tmax_new, bmax), dtype=np.uint64)
    for l in range(r):
        row_b = b2[int(l % r_b)]
        for j in range(d):
            b = row_b[j]
            for t in range(tmax_new):
                perms[l, j, t, :b] = rng.permutation(b)
    return perms
