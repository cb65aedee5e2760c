"""Loads libqmctoolscl_amd.so (built in place by qmctoolscl_amd.build) and
declares the ctypes prototypes of its C ABI (include/qmctoolscl_amd.h).

There is no fallback: if the library is missing or no GPU is visible, calls
raise.  The library is loaded on first use so that importing the package (for
its signature table or random helpers) works on machines without a GPU."""
import ctypes
import os

from ._api import SIGNATURES

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqmctoolscl_amd.so")

QMC_ERRORS = {1: "QMC_ERR_ARG", 2: "QMC_ERR_RANGE", 3: "QMC_ERR_HIP", 4: "QMC_ERR_NOMEM"}

_lib = None

_C = {"u64": ctypes.c_uint64, "pu64": ctypes.c_void_p, "pf64": ctypes.c_void_p}


class QMCError(RuntimeError):
    """A status code from the HIP library."""

    def __init__(self, code, msg):
        super().__init__("%s (%s): %s" % (QMC_ERRORS.get(code, "error %d" % code), code, msg))
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("qmctoolscl_amd: the HIP library %s is not built; run "
                          "`python -m qmctoolscl_amd.build` (hipcc, gfx950)" % LIB_PATH)
    # PyTorch-ROCm wheels bundle their own libamdhip64 under the same soname.
    # Loaded after torch, the library binds to that runtime, so the streams and
    # device pointers torch hands over are valid in it; loaded first, the
    # process would hold two HIP runtimes and torch's streams would be foreign.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    L.qmc_last_error.restype = ctypes.c_char_p
    L.qmc_last_error_code.restype = ctypes.c_int
    L.qmc_version.restype = ctypes.c_char_p
    L.qmc_init.restype = ctypes.c_int
    L.qmc_device_status.restype = ctypes.c_int
    L.qmc_device_status.argtypes = [ctypes.POINTER(ctypes.c_uint), ctypes.c_int]
    L.qmc_status_reset_dev.restype = ctypes.c_int
    L.qmc_status_reset_dev.argtypes = [ctypes.c_void_p]
    for name, sig in SIGNATURES.items():
        kinds = [_C[s[1]] for s in sig]
        host = getattr(L, name)
        host.restype = None
        host.argtypes = kinds[:3] + [ctypes.c_uint64] * 3 + kinds[3:]
        dev = getattr(L, name + "_dev")
        dev.restype = ctypes.c_int
        dev.argtypes = kinds + [ctypes.c_void_p]
    L.qmc_lat_gen_shift_dev.restype = ctypes.c_int
    L.qmc_lat_gen_shift_dev.argtypes = ([ctypes.c_int] + [ctypes.c_uint64] * 5 + [ctypes.c_void_p] * 4)
    L.qmc_dnb2_gen_shift_float_dev.restype = ctypes.c_int
    L.qmc_dnb2_gen_shift_float_dev.argtypes = ([ctypes.c_int] + [ctypes.c_uint64] * 6 + [ctypes.c_void_p] * 7)
    _lib = L
    return L


def check_status(what):
    """Read and clear the current device's status bits (after the caller's
    stream has synchronized); raise QMCError(QMC_ERR_RANGE) if any is set."""
    L = lib()
    bits = ctypes.c_uint(0)
    check(L.qmc_device_status(ctypes.byref(bits), 1))
    if bits.value:
        raise QMCError(2, "%s: input outside the reference's defined domain (status 0x%x): a base below 2 or "
                          "an index that reads past an input array" % (what, bits.value))


def check(rc=None):
    """Raise QMCError for a nonzero status (or, for the void drop-ins, the
    thread's last error code)."""
    L = lib()
    code = L.qmc_last_error_code() if rc is None else rc
    if code:
        raise QMCError(code, L.qmc_last_error().decode(errors="replace"))
