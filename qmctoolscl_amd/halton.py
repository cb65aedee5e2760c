"""Generalized Halton points: the reference's halton_qrng
(qmctoolscl/c_funcs/halton_qrng.c), bound as in its pdm_build.py:85-94.

halton_qrng_c(n, d, n0, generalized, res, randu_d_32, dvec)
    The C symbol with the reference's ctypes prototype (c_int x 4, then
    C-contiguous float64 res and randu_d_32 and int32 dvec numpy arrays).
    res (d x n) receives res[j, i] = point n0 + i in dimension j; randu_d_32
    (d x 32) holds the uniforms of the random digit shift; dvec selects the
    primes (dvec[j] < 360, base = the dvec[j]-th prime, 0 -> 2).
    generalized = 1 applies the generalized Halton digit factors.  Unlike the
    reference, a failed call raises QMCError, and no state is kept between
    calls.

halton_qrng(n, d, n0, generalized, res, randu_d_32, dvec)
    The same on numpy arrays or on GPU tensors (float64 res and randu_d_32,
    int32 dvec, all on the device: no host copies).
"""
import ctypes

import numpy as np

from . import _lib


def _errcheck(result, func, args):
    _lib.check()
    return result


def _declare(L):
    if getattr(L, "_halton_declared", False):
        return
    fn = L.halton_qrng
    fn.restype = None
    fn.argtypes = [ctypes.c_int] * 4 + [np.ctypeslib.ndpointer(ctypes.c_double, flags="C_CONTIGUOUS"),
                                        np.ctypeslib.ndpointer(ctypes.c_double, flags="C_CONTIGUOUS"),
                                        np.ctypeslib.ndpointer(ctypes.c_int, flags="C_CONTIGUOUS")]
    fn.errcheck = _errcheck
    L.qmc_halton_qrng_dev.restype = ctypes.c_int
    L.qmc_halton_qrng_dev.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 4
    L._halton_declared = True


def _L():
    L = _lib.lib()
    _declare(L)
    return L


class _Lazy:
    """halton_qrng_c resolves the library on first call, so that importing the
    package works without a GPU or a built library."""

    def __call__(self, *args):
        return _L().halton_qrng(*args)

    @property
    def argtypes(self):
        return _L().halton_qrng.argtypes


halton_qrng_c = _Lazy()


def halton_qrng(n, d, n0, generalized, res, randu_d_32, dvec, stream=None):
    if hasattr(res, "is_cuda") and res.is_cuda:
        import torch
        for name, t, dt in (("res", res, torch.float64), ("randu_d_32", randu_d_32, torch.float64),
                            ("dvec", dvec, torch.int32)):
            if not (hasattr(t, "is_cuda") and t.is_cuda and t.is_contiguous() and t.dtype == dt):
                raise ctypes.ArgumentError("%s must be a contiguous %s GPU tensor" % (name, dt))
        if res.numel() < int(n) * int(d) or randu_d_32.numel() < 32 * int(d) or dvec.numel() < int(d):
            raise ctypes.ArgumentError("halton_qrng: res needs d*n, randu_d_32 d*32 and dvec d entries")
        s = stream if stream is not None else torch.cuda.current_stream(res.device)
        rc = _L().qmc_halton_qrng_dev(int(n), int(d), int(n0), int(generalized), res.data_ptr(),
                                      randu_d_32.data_ptr(), dvec.data_ptr(), ctypes.c_void_p(s.cuda_stream))
        _lib.check(rc)
        return res
    for name, a, dt, size in (("res", res, np.float64, int(n) * int(d)),
                              ("randu_d_32", randu_d_32, np.float64, 32 * int(d)),
                              ("dvec", dvec, np.int32, int(d))):
        if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous and a.size >= size):
            raise ctypes.ArgumentError("%s must be a C-contiguous %s array of >= %d entries" % (name, dt, size))
    halton_qrng_c(int(n), int(d), int(n0), int(generalized), res, randu_d_32, dvec)
    return res
