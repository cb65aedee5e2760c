"""Reduced-precision fast transforms on the matrix cores (GPU tensors only).

fwht_mfma(d1, d2, n_half, x): the orthonormal Walsh-Hadamard transform of
fwht_1d_radix2 (same arguments, in place on a float64 tensor of d1*d2*2*n_half
values), computed with 16-point Hadamard tiles on fp32 MFMA
(v_mfma_f32_16x16x4f32) with fp32 accumulation.  Not bit-exact: relative L2
error about 2e-7 against the float64 transform (see tests/test_mfma.py); use
fwht_1d_radix2 for the bit-exact result.  Lengths 2^8 .. 2^24.

fft_bro_mfma(d1, d2, n_half, xr, xi): fft_bro_1d_radix2's transform (input
bit-reversed, output natural, orthonormal) with radix-16 DFT tiles on MFMA.

ifft_bro_mfma(d1, d2, n_half, xr, xi): ifft_bro_1d_radix2's transform
(natural input, bit-reversed output), the same network transposed.
"""
import ctypes

import numpy as np

from . import _lib


def _declare(L):
    if getattr(L, "_mfma_declared", False):
        return
    L.qmc_fwht_mfma_dev.restype = ctypes.c_int
    L.qmc_fwht_mfma_dev.argtypes = [ctypes.c_uint64] * 3 + [ctypes.c_void_p] * 3
    L.qmc_ifft_bro_mfma_dev.restype = ctypes.c_int
    L.qmc_ifft_bro_mfma_dev.argtypes = [ctypes.c_uint64] * 3 + [ctypes.c_void_p] * 4
    L.qmc_fft_bro_mfma_dev.restype = ctypes.c_int
    L.qmc_fft_bro_mfma_dev.argtypes = [ctypes.c_uint64] * 3 + [ctypes.c_void_p] * 4
    L.qmc_mfma_probe_dev.restype = ctypes.c_int
    L.qmc_mfma_probe_dev.argtypes = [ctypes.c_void_p] * 4
    L._mfma_declared = True


def _L():
    L = _lib.lib()
    _declare(L)
    return L


def _stream(t, stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream(t.device)
    return ctypes.c_void_p(s.cuda_stream)


def fwht_mfma(d1, d2, n_half, x, work=None, stream=None):
    if not (x.is_cuda and x.is_contiguous()) or str(x.dtype) != "torch.float64":
        raise ctypes.ArgumentError("x must be a contiguous float64 GPU tensor")
    if work is not None and not (work.is_cuda and work.is_contiguous() and str(work.dtype) == "torch.float32"):
        raise ctypes.ArgumentError("work must be a contiguous float32 GPU tensor")
    rc = _L().qmc_fwht_mfma_dev(int(d1), int(d2), int(n_half), x.data_ptr(),
                                work.data_ptr() if work is not None else None, _stream(x, stream))
    _lib.check(rc)


def fft_bro_mfma(d1, d2, n_half, xr, xi, work=None, stream=None):
    """In-place orthonormal FFT of bit-reversed input (fft_bro_1d_radix2's
    transform) on the matrix cores: radix-16 DFT tiles (FP64-generated, 3 x
    bf16 split) on MFMA with fp32 accumulation and fp32 inter-level twiddles.
    xr, xi: contiguous float64 GPU tensors of d1*d2*2*n_half values; lengths
    2^8 .. 2^28; relative L2 error ~1e-7 vs float64."""
    for t in (xr, xi):
        if not (t.is_cuda and t.is_contiguous()) or str(t.dtype) != "torch.float64":
            raise ctypes.ArgumentError("xr, xi must be contiguous float64 GPU tensors")
    if work is not None and not (work.is_cuda and work.is_contiguous() and str(work.dtype) == "torch.float32"):
        raise ctypes.ArgumentError("work must be a contiguous float32 GPU tensor")
    rc = _L().qmc_fft_bro_mfma_dev(int(d1), int(d2), int(n_half), xr.data_ptr(), xi.data_ptr(),
                                   work.data_ptr() if work is not None else None, _stream(xr, stream))
    _lib.check(rc)


def ifft_bro_mfma(d1, d2, n_half, xr, xi, work=None, stream=None):
    """In-place orthonormal inverse FFT with natural-order input and
    bit-reversed output (ifft_bro_1d_radix2's transform) on the matrix cores:
    the conjugate transpose of fft_bro_mfma's network, same accuracy."""
    for t in (xr, xi):
        if not (t.is_cuda and t.is_contiguous()) or str(t.dtype) != "torch.float64":
            raise ctypes.ArgumentError("xr, xi must be contiguous float64 GPU tensors")
    if work is not None and not (work.is_cuda and work.is_contiguous() and str(work.dtype) == "torch.float32"):
        raise ctypes.ArgumentError("work must be a contiguous float32 GPU tensor")
    rc = _L().qmc_ifft_bro_mfma_dev(int(d1), int(d2), int(n_half), xr.data_ptr(), xi.data_ptr(),
                                    work.data_ptr() if work is not None else None, _stream(xr, stream))
    _lib.check(rc)


def mfma_probe(a, b):
    """D = A(16x32) @ B(32x16) through one bf16 MFMA, operands given lane-major
    as the kernels lay them out; returns the lane-major (64, 4) result."""
    import torch
    d = torch.empty((64, 4), dtype=torch.float32, device=a.device)
    rc = _L().qmc_mfma_probe_dev(a.data_ptr(), b.data_ptr(), d.data_ptr(), _stream(a, None))
    _lib.check(rc)
    torch.cuda.synchronize()
    return d


def relative_errors(got, want):
    """(relative L2 error, max abs error / max abs value) of got vs want."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    l2 = float(np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300))
    mx = float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300))
    return l2, mx
