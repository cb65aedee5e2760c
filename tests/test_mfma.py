"""Matrix-core FWHT: operand layouts of the bf16 MFMA builtin, and accuracy
against the float64 transform.  Tolerance: relative L2 error <= 2e-6 and max
error <= 2e-6 of the largest output, for N(0,1) inputs and for a smooth
periodic integrand evaluated at lattice points."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L2_TOL = 2e-6
MAX_TOL = 2e-6


@pytest.fixture(scope="module")
def torch():
    t = pytest.importorskip("torch")
    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


def test_mfma_operand_layout(torch):
    from qmctoolscl_amd.mfma import mfma_probe
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, size=(16, 32)).astype(np.float32)   # exact in bf16
    B = rng.integers(-4, 5, size=(32, 16)).astype(np.float32)
    a = np.empty((64, 8), np.float32)
    b = np.empty((64, 8), np.float32)
    for lane in range(64):
        m, g = lane % 16, lane // 16
        a[lane] = A[m, 8 * g:8 * g + 8]
        b[lane] = B[8 * g:8 * g + 8, m]
    d = mfma_probe(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()).cpu().numpy()
    D = A @ B
    want = np.empty((64, 4), np.float32)
    for lane in range(64):
        n, g = lane % 16, lane // 16
        want[lane] = D[4 * g:4 * g + 4, n]
    assert np.array_equal(d, want)


def _fwht64(x):
    y = x.copy()
    n = y.shape[-1]
    h = 1
    while h < n:
        y = y.reshape(-1, n // (2 * h), 2, h)
        a, b = y[:, :, 0, :].copy(), y[:, :, 1, :].copy()
        y[:, :, 0, :], y[:, :, 1, :] = a + b, a - b
        y = y.reshape(-1, n)
        h *= 2
    return y / np.sqrt(n)


@pytest.mark.parametrize("m,batch", [(8, 3), (9, 3), (10, 4), (12, 5), (14, 2), (15, 2), (17, 2), (20, 1), (22, 1),
                                     (24, 1)])
def test_fwht_mfma_accuracy(torch, m, batch):
    from qmctoolscl_amd.mfma import fwht_mfma, relative_errors
    rng = np.random.default_rng(m)
    n = 1 << m
    x = rng.standard_normal((batch, n))
    want = _fwht64(x)
    t = torch.from_numpy(x.copy()).cuda()
    fwht_mfma(1, batch, n // 2, t)
    l2, mx = relative_errors(t.cpu().numpy(), want)
    assert l2 <= L2_TOL and mx <= MAX_TOL, (m, l2, mx)


def test_fwht_mfma_smooth_integrand(torch):
    """f(x) = prod_j (1 + 0.5 cos(2 pi x_j)) at lattice points: large DC term,
    rapidly decaying coefficients."""
    from qmctoolscl_amd.mfma import fwht_mfma, relative_errors
    m, d = 16, 4
    n = 1 << m
    g = np.array([1, 182667, 469891, 498753])
    i = np.arange(n)
    pts = np.mod(np.outer(i, g) / n, 1.0)
    f = np.prod(1 + 0.5 * np.cos(2 * np.pi * pts), axis=1)[None]
    want = _fwht64(f)
    t = torch.from_numpy(f.copy()).cuda()
    fwht_mfma(1, 1, n // 2, t)
    l2, mx = relative_errors(t.cpu().numpy(), want)
    assert l2 <= L2_TOL and mx <= MAX_TOL, (l2, mx)


def _bitrev(m):
    i = np.arange(1 << m, dtype=np.int64)
    r = np.zeros_like(i)
    for b in range(m):
        r |= ((i >> b) & 1) << (m - 1 - b)
    return r


@pytest.mark.parametrize("m,batch", [(8, 3), (9, 3), (10, 3), (11, 2), (12, 3), (13, 2), (14, 2), (16, 2), (19, 1),
                                     (20, 1), (21, 1), (22, 1), (24, 1)])
def test_fft_bro_mfma_accuracy(torch, m, batch):
    """Every pass shape: one tile (2^8 .. 2^13), two passes, three passes
    with a radix-2/4 or a radix-16 top level."""
    from qmctoolscl_amd.mfma import fft_bro_mfma, relative_errors
    rng = np.random.default_rng(100 + m)
    n = 1 << m
    xr = rng.standard_normal((batch, n))
    xi = rng.standard_normal((batch, n))
    ir = _bitrev(m)
    want = np.fft.fft((xr + 1j * xi)[:, ir], norm="ortho")
    tr, ti = torch.from_numpy(xr.copy()).cuda(), torch.from_numpy(xi.copy()).cuda()
    fft_bro_mfma(1, batch, n // 2, tr, ti)
    got = tr.cpu().numpy() + 1j * ti.cpu().numpy()
    l2, mx = relative_errors(np.stack([got.real, got.imag]), np.stack([want.real, want.imag]))
    assert l2 <= L2_TOL and mx <= MAX_TOL, (m, l2, mx)


@pytest.mark.parametrize("m,batch", [(8, 3), (9, 3), (11, 2), (12, 3), (13, 2), (14, 2), (16, 2), (20, 1), (21, 1),
                                     (22, 1), (24, 1)])
def test_ifft_bro_mfma_accuracy(torch, m, batch):
    """ifft_bro_1d_radix2's transform (natural in, bit-reversed out) on the
    matrix cores, every pass shape, and the round trip with fft_bro_mfma."""
    from qmctoolscl_amd.mfma import fft_bro_mfma, ifft_bro_mfma, relative_errors
    rng = np.random.default_rng(200 + m)
    n = 1 << m
    xr = rng.standard_normal((batch, n))
    xi = rng.standard_normal((batch, n))
    want = np.fft.ifft(xr + 1j * xi, norm="ortho")[:, _bitrev(m)]
    tr, ti = torch.from_numpy(xr.copy()).cuda(), torch.from_numpy(xi.copy()).cuda()
    ifft_bro_mfma(1, batch, n // 2, tr, ti)
    got = tr.cpu().numpy() + 1j * ti.cpu().numpy()
    l2, mx = relative_errors(np.stack([got.real, got.imag]), np.stack([want.real, want.imag]))
    assert l2 <= L2_TOL and mx <= MAX_TOL, (m, l2, mx)
    fft_bro_mfma(1, batch, n // 2, tr, ti)
    back = tr.cpu().numpy() + 1j * ti.cpu().numpy()
    l2, mx = relative_errors(np.stack([back.real, back.imag]), np.stack([xr, xi]))
    assert l2 <= 2 * L2_TOL and mx <= 2 * MAX_TOL, (m, l2, mx)
