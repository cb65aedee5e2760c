"""Fast QMC kernel methods (csrc/kernel_methods.hip): kernel columns from g or
C, their FFT / FWHT, and K y, K^-1 y through the transforms.

CPU: the chain the product implements -- column of the Gram matrix, its
transform, pointwise product, inverse transform (docs/index.md:1825-1883,
1929-1977) -- is checked against dense Gram matrices built from the oracle's
points: K @ y and solve(K, y).  GPU: the columns bit for bit against the same
formulas on the oracle's points, kt against the oracle's transforms of that
column, and K y / K^-1 y against the dense products."""
import math

import numpy as np
import pytest

import _clib

U = np.uint64


def bitrev(m):
    i = np.arange(1 << m, dtype=np.int64)
    r = np.zeros_like(i)
    for b in range(m):
        r |= ((i >> b) & 1) << (m - 1 - b)
    return r


def lattice_points(g, n):
    r, d = g.shape
    x = np.empty((r, n, d))
    _clib.call_oracle("lat_gen_natural", dict(r=r, n=n, d=d, n_start=0, g=g, x=x))
    return x


def net_points(C, n):
    r, d, mmax = C.shape
    xb = np.empty((r, n, d), dtype=np.uint64)
    _clib.call_oracle("dnb2_gen_natural", dict(r=r, n=n, d=d, n_start=0, mmax=mmax, C=C, xb=xb))
    return xb


def bern(x, alpha):
    b = (x - 1.0) * x
    return b + (1.0 / 6.0) if alpha == 1 else b * b - (1.0 / 30.0)


def bitlen(X):
    X = X.astype(np.uint64)
    bl = np.zeros(X.shape, dtype=np.int64)
    for s in (32, 16, 8, 4, 2, 1):
        hi = (X >> U(s)) != 0
        bl += np.where(hi, s, 0)
        X = np.where(hi, X >> U(s), X)
    return bl + (X != 0)


def omega(X, tmax):
    X = np.asarray(X, dtype=np.uint64)
    return np.where(X != 0, 1.0 - np.ldexp(3.0, bitlen(X) - 1 - tmax), 1.0)


def lat_column(x, coef, alpha):
    """k[l, i] = prod_j (1 + coef[l, j] B(x[l, i, j])), the kernel's steps in order"""
    p = np.ones(x.shape[:2])
    for j in range(x.shape[2]):
        p = p * (1.0 + coef[:, j:j + 1] * bern(x[:, :, j], alpha))
    return p


def net_column(xb, coef, tmax):
    p = np.ones(xb.shape[:2])
    for j in range(xb.shape[2]):
        p = p * (1.0 + coef[:, j:j + 1] * omega(xb[:, :, j], tmax))
    return p


def lat_gram(x, coef, alpha):
    n, d = x.shape
    K = np.ones((n, n))
    for j in range(d):
        K *= 1.0 + coef[j] * bern(np.mod(x[:, None, j] - x[None, :, j], 1.0), alpha)
    return K


def net_gram(xb, coef, tmax):
    n, d = xb.shape
    K = np.ones((n, n))
    for j in range(d):
        K *= 1.0 + coef[j] * omega(xb[:, None, j] ^ xb[None, :, j], tmax)
    return K


def fft_bro(re, im):
    """oracle fft_bro_1d_radix2 over the rows of re + i im"""
    re, im = re.copy(), im.copy()
    V, n = re.shape
    _clib.call_oracle("fft_bro_1d_radix2", dict(d1=1, d2=V, n_half=n // 2, twiddler=np.empty(n),
                                                twiddlei=np.empty(n), xr=re, xi=im))
    return re, im


def ifft_bro(re, im):
    re, im = re.copy(), im.copy()
    V, n = re.shape
    _clib.call_oracle("ifft_bro_1d_radix2", dict(d1=1, d2=V, n_half=n // 2, twiddler=np.empty(n),
                                                 twiddlei=np.empty(n), xr=re, xi=im))
    return re, im


def fwht(x):
    x = x.copy()
    V, n = x.shape
    _clib.call_oracle("fwht_1d_radix2", dict(d1=1, d2=V, n_half=n // 2, x=x))
    return x


def lat_inputs(r, d, alpha, seed):
    rng = np.random.Generator(np.random.SFC64(seed))
    g = 2 * rng.integers(0, 2 ** 19, size=(r, d), dtype=np.uint64) + U(1)
    eta = (2 * math.pi) ** (2 * alpha) * (-1) ** (alpha + 1) / math.factorial(2 * alpha)
    gamma = 1.0 / (1.0 + np.arange(d)) ** 2
    coef = np.tile(gamma * eta, (r, 1)) * (1.0 + 0.1 * np.arange(r))[:, None]
    return g, coef


def net_inputs(r, d, seed, t=32):
    rng = np.random.Generator(np.random.SFC64(seed))
    C = np.empty((r, d, t), dtype=np.uint64)
    for l in range(r):
        for j in range(d):
            for k in range(t):   # column k: bit t-1-k set, random bits above it (MSB first)
                above = int(rng.integers(0, 2 ** k)) if k else 0
                C[l, j, k] = (1 << (t - 1 - k)) | (above << (t - k))
    gamma = 0.9 / (1.0 + np.arange(d))
    coef = np.tile(gamma, (r, 1)) * (1.0 - 0.1 * np.arange(r))[:, None]
    return C, coef


# ------------------------------------------------------------- CPU: the chain

@pytest.mark.parametrize("alpha", [1, 2])
def test_lattice_chain_equals_dense_gram(alpha):
    m, d = 8, 5
    n = 1 << m
    g, coef = lat_inputs(1, d, alpha, 3)
    x = lattice_points(g, n)
    k = lat_column(x, coef, alpha)
    K = lat_gram(x[0], coef[0], alpha)
    assert np.array_equal(K[:, 0], k[0])          # x_0 = 0: the column is K's first column
    kr, ki = fft_bro(k, np.zeros_like(k))
    rng = np.random.default_rng(alpha)
    y = rng.standard_normal((1, n))
    fr, fi = fft_bro(y, np.zeros_like(y))
    pr = (fr * kr - fi * ki) * np.sqrt(n)
    pi = (fr * ki + fi * kr) * np.sqrt(n)
    ur, ui = ifft_bro(pr, pi)
    want = K @ y[0]
    assert np.max(np.abs(ur[0] - want)) <= 1e-11 * np.max(np.abs(want)) and np.max(np.abs(ui)) <= 1e-11 * np.max(np.abs(want))
    den = (kr * kr + ki * ki) * np.sqrt(n)
    vr, vi = ifft_bro((fr * kr + fi * ki) / den, (fi * kr - fr * ki) / den)
    sol = np.linalg.solve(K, y[0])
    assert np.max(np.abs(vr[0] - sol)) <= 1e-9 * np.max(np.abs(sol))


def test_net_chain_equals_dense_gram():
    m, d, t = 8, 4, 32
    n = 1 << m
    C, coef = net_inputs(1, d, 5, t)
    xb = net_points(C, n)
    k = net_column(xb, coef, t)
    K = net_gram(xb[0], coef[0], t)
    assert np.array_equal(K[:, 0], k[0])
    kt = fwht(k)
    y = np.random.default_rng(0).standard_normal((1, n))
    u = fwht(fwht(y) * kt * np.sqrt(n))
    want = K @ y[0]
    assert np.max(np.abs(u[0] - want)) <= 1e-11 * np.max(np.abs(want))
    v = fwht(fwht(y) / (kt * np.sqrt(n)))
    sol = np.linalg.solve(K, y[0])
    assert np.max(np.abs(v[0] - sol)) <= 1e-9 * np.max(np.abs(sol))
    # the order-1 digital kernel is positive definite here: every eigenvalue sqrt(n) kt > 0
    assert kt.min() > 0


def test_omega_closed_form():
    t = 8
    X = np.arange(1 << t, dtype=np.uint64)
    x = X / 2.0 ** t
    want = np.where(X > 0, 6.0 * (1.0 / 6.0 - 2.0 ** (np.floor(np.log2(np.maximum(x, 1e-300))) - 1)), 1.0)
    assert np.allclose(omega(X, t), want, rtol=0, atol=1e-15)


# -------------------------------------------------------------------- GPU

def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


@pytest.mark.gpu
@pytest.mark.parametrize("alpha,m,d,r", [(1, 16, 64, 2), (2, 12, 7, 3), (1, 1, 3, 1), (2, 20, 8, 1)])
def test_gpu_lattice_column_and_fft(alpha, m, d, r):
    torch = _gpu()
    from qmctoolscl_amd import kernel_methods as km
    n = 1 << m
    g, coef = lat_inputs(r, d, alpha, m)
    want = lat_column(lattice_points(g, n), coef, alpha)
    gt = torch.from_numpy(g.view(np.int64)).cuda()
    ct = torch.from_numpy(coef).cuda()
    k = km.lat_kernel_column(alpha, gt, ct, n)
    torch.cuda.synchronize()
    assert np.array_equal(k.cpu().numpy().view(np.uint64), want.view(np.uint64))
    kr, ki = km.lat_kernel_fft(alpha, gt, ct, n)
    wr, wi = fft_bro(want, np.zeros_like(want))
    scale = max(1.0, float(np.max(np.abs(wr))))
    err = max(float(np.max(np.abs(kr.cpu().numpy() - wr))), float(np.max(np.abs(ki.cpu().numpy() - wi)))) / scale
    assert err <= 1e-12, err


@pytest.mark.gpu
@pytest.mark.parametrize("m,d,r", [(16, 64, 2), (10, 5, 3), (4, 2, 1), (20, 8, 1)])
def test_gpu_net_column_and_fwht(m, d, r):
    torch = _gpu()
    from qmctoolscl_amd import kernel_methods as km
    n, t = 1 << m, 32
    C, coef = net_inputs(r, d, m, t)
    want = net_column(net_points(C, n), coef, t)
    Ct = torch.from_numpy(C.view(np.int64)).cuda()
    ct = torch.from_numpy(coef).cuda()
    k = km.dnb2_kernel_column(Ct, ct, n, t)
    torch.cuda.synchronize()
    assert np.array_equal(k.cpu().numpy().view(np.uint64), want.view(np.uint64))
    kt = km.dnb2_kernel_fwht(Ct, ct, n, t)
    assert np.array_equal(kt.cpu().numpy().view(np.uint64), fwht(want).view(np.uint64))   # FWHT is bit-exact


@pytest.mark.gpu
@pytest.mark.parametrize("alpha", [1, 2])
def test_gpu_lattice_apply_matches_dense(alpha):
    torch = _gpu()
    from qmctoolscl_amd import kernel_methods as km
    m, d, r = 9, 6, 2
    n = 1 << m
    g, coef = lat_inputs(r, d, alpha, 40 + alpha)
    x = lattice_points(g, n)
    gt, ct = torch.from_numpy(g.view(np.int64)).cuda(), torch.from_numpy(coef).cuda()
    kr, ki = km.lat_kernel_fft(alpha, gt, ct, n)
    y = np.random.default_rng(alpha).standard_normal((r, n))
    for mode in (0, 1):
        yr, yi = torch.from_numpy(y.copy()).cuda(), torch.zeros((r, n), dtype=torch.float64, device="cuda")
        km.lat_kernel_apply(mode, kr, ki, yr, yi)
        for l in range(r):
            K = lat_gram(x[l], coef[l], alpha)
            want = K @ y[l] if mode == 0 else np.linalg.solve(K, y[l])
            tol = (1e-11 if mode == 0 else 1e-8) * np.max(np.abs(want))
            assert np.max(np.abs(yr[l].cpu().numpy() - want)) <= tol, (mode, l)
            assert np.max(np.abs(yi[l].cpu().numpy())) <= tol, (mode, l)


@pytest.mark.gpu
def test_gpu_net_apply_matches_dense():
    torch = _gpu()
    from qmctoolscl_amd import kernel_methods as km
    m, d, r, t = 9, 5, 2, 32
    n = 1 << m
    C, coef = net_inputs(r, d, 77, t)
    xb = net_points(C, n)
    Ct, ct = torch.from_numpy(C.view(np.int64)).cuda(), torch.from_numpy(coef).cuda()
    kt = km.dnb2_kernel_fwht(Ct, ct, n, t)
    y = np.random.default_rng(9).standard_normal((r, n))
    for mode in (0, 1):
        yt = torch.from_numpy(y.copy()).cuda()
        km.dnb2_kernel_apply(mode, kt, yt)
        for l in range(r):
            K = net_gram(xb[l], coef[l], t)
            want = K @ y[l] if mode == 0 else np.linalg.solve(K, y[l])
            tol = (1e-11 if mode == 0 else 1e-8) * np.max(np.abs(want))
            assert np.max(np.abs(yt[l].cpu().numpy() - want)) <= tol, (mode, l)
