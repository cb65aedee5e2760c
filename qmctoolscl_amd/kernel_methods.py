"""Fast QMC kernel methods on GPU tensors (csrc/kernel_methods.hip).

The Gram matrix K of a product kernel on natural-order points is diagonalized
by a fast transform of its first column k: the bit-reversed FFT for rank-1
lattices (shift-invariant kernels), the Walsh-Hadamard transform for base-2
digital nets (digitally shift-invariant kernels), as the reference's
documentation uses its fft_bro_1d_radix2 / ifft_bro_1d_radix2 /
fwht_1d_radix2 (docs/index.md:1825-1883, 1929-1977).  The column is evaluated
from the generating vector g or matrices C without writing the points.

lattice, natural order (lat_gen_natural with n_start = 0), r generating vectors:
    k_i = prod_j (1 + coef_j B(x_ij)),  B = B_2 (alpha = 1) or B_4 (alpha = 2)
    lat_kernel_coef(gamma, alpha) gives coef_j for Fourier weights gamma_j |h|^-2alpha
base-2 net, natural order (dnb2_gen_natural, n_start = 0), r matrix sets:
    k_i = prod_j (1 + coef_j w(X_ij)),  w(0) = 1, w(X) = 1 - 3 * 2^(bitlen(X) - 1 - tmax)

lat_kernel_column / dnb2_kernel_column -> k (r x n float64)
lat_kernel_fft  -> (kr, ki) = fft_bro(k) (orthonormal)
dnb2_kernel_fwht -> kt = fwht(k) (orthonormal)
lat_kernel_apply / dnb2_kernel_apply(mode, ...) overwrite y with K y (mode 0)
    or K^-1 y (mode 1): ifft(fft(y) kt) sqrt(n), ifft(fft(y) / kt) / sqrt(n).
"""
import ctypes
import math

from . import _lib


def _declare(L):
    if getattr(L, "_km_declared", False):
        return
    P, U, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    for name, args in (("qmc_lat_kernel_column_dev", [I, U, U, U, P, P, P, P]),
                       ("qmc_dnb2_kernel_column_dev", [U, U, U, U, I, P, P, P, P]),
                       ("qmc_lat_kernel_fft_dev", [I, U, U, U, P, P, P, P, P, P, P]),
                       ("qmc_dnb2_kernel_fwht_dev", [U, U, U, U, I, P, P, P, P]),
                       ("qmc_lat_kernel_apply_dev", [I, U, U, P, P, P, P, P, P, P]),
                       ("qmc_dnb2_kernel_apply_dev", [I, U, U, P, P, P])):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    L._km_declared = True


def _L():
    L = _lib.lib()
    _declare(L)
    return L


def _stream(t):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _need(t, dtype, name, numel=None):
    import torch
    want = {"f64": torch.float64, "i64": torch.int64}[dtype]
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous() and t.dtype == want):
        raise ctypes.ArgumentError("%s must be a contiguous %s GPU tensor" % (name, want))
    if numel is not None and t.numel() < numel:
        raise ctypes.ArgumentError("%s needs %d entries" % (name, numel))


def lat_kernel_coef(gamma, alpha):
    """coef_j = gamma_j (2 pi)^(2 alpha) (-1)^(alpha+1) / (2 alpha)!"""
    eta = (2 * math.pi) ** (2 * alpha) * (-1) ** (alpha + 1) / math.factorial(2 * alpha)
    return [float(g) * eta for g in gamma]


def lat_kernel_column(alpha, g, coef, n, out=None):
    import torch
    r, d = g.shape
    _need(g, "i64", "g"), _need(coef, "f64", "coef", r * d)
    k = out if out is not None else torch.empty((r, n), dtype=torch.float64, device=g.device)
    _need(k, "f64", "k", r * n)
    _lib.check(_L().qmc_lat_kernel_column_dev(int(alpha), r, int(n), d, g.data_ptr(), coef.data_ptr(),
                                              k.data_ptr(), _stream(g)))
    return k


def dnb2_kernel_column(C, coef, n, tmax, out=None):
    import torch
    r, d, mmax = C.shape
    _need(C, "i64", "C"), _need(coef, "f64", "coef", r * d)
    k = out if out is not None else torch.empty((r, n), dtype=torch.float64, device=C.device)
    _need(k, "f64", "k", r * n)
    _lib.check(_L().qmc_dnb2_kernel_column_dev(r, int(n), d, mmax, int(tmax), C.data_ptr(), coef.data_ptr(),
                                               k.data_ptr(), _stream(C)))
    return k


def lat_kernel_fft(alpha, g, coef, n):
    import torch
    r, d = g.shape
    _need(g, "i64", "g"), _need(coef, "f64", "coef", r * d)
    kr = torch.empty((r, n), dtype=torch.float64, device=g.device)
    ki = torch.empty_like(kr)
    twr = torch.empty(n, dtype=torch.float64, device=g.device)
    twi = torch.empty_like(twr)
    _lib.check(_L().qmc_lat_kernel_fft_dev(int(alpha), r, int(n), d, g.data_ptr(), coef.data_ptr(), twr.data_ptr(),
                                           twi.data_ptr(), kr.data_ptr(), ki.data_ptr(), _stream(g)))
    return kr, ki


def dnb2_kernel_fwht(C, coef, n, tmax):
    import torch
    r, d, mmax = C.shape
    _need(C, "i64", "C"), _need(coef, "f64", "coef", r * d)
    kt = torch.empty((r, n), dtype=torch.float64, device=C.device)
    _lib.check(_L().qmc_dnb2_kernel_fwht_dev(r, int(n), d, mmax, int(tmax), C.data_ptr(), coef.data_ptr(),
                                             kt.data_ptr(), _stream(C)))
    return kt


def lat_kernel_apply(mode, kr, ki, yr, yi):
    import torch
    r, n = kr.shape
    for name, t in (("kr", kr), ("ki", ki), ("yr", yr), ("yi", yi)):
        _need(t, "f64", name, r * n)
    twr = torch.empty(n, dtype=torch.float64, device=kr.device)
    twi = torch.empty_like(twr)
    _lib.check(_L().qmc_lat_kernel_apply_dev(int(mode), r, n, kr.data_ptr(), ki.data_ptr(), twr.data_ptr(),
                                             twi.data_ptr(), yr.data_ptr(), yi.data_ptr(), _stream(kr)))
    return yr, yi


def dnb2_kernel_apply(mode, kt, y):
    r, n = kt.shape
    _need(kt, "f64", "kt", r * n), _need(y, "f64", "y", r * n)
    _lib.check(_L().qmc_dnb2_kernel_apply_dev(int(mode), r, n, kt.data_ptr(), y.data_ptr(), _stream(kt)))
    return y
