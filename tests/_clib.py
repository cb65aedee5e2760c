"""ctypes plumbing shared by the tests: call a function of the CPU oracle
(oracle/liboracle.so), of the reference C backend built from its own sources
(oracle/_ref/libqmcref.so), or of this project's HIP library through its
host-pointer drop-in symbols (same prototypes as the reference C backend)."""
import ctypes
import hashlib
import os

import numpy as np

from qmctoolscl_amd._api import SIGNATURES

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libqmcref.so")

_cache = {}


def load(path):
    if path not in _cache:
        _cache[path] = ctypes.CDLL(path)
    return _cache[path]


def oracle():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
    return load(ORACLE_SO)


def reference():
    """The reference C backend, or None where it was not built (GPU box)."""
    return load(REF_SO) if os.path.exists(REF_SO) else None


def _argv(name, case):
    argv = []
    for spec in SIGNATURES[name]:
        v = case[spec[0]]
        if spec[1] == "u64":
            argv.append(ctypes.c_uint64(int(v)))
        else:
            want = np.uint64 if spec[1] == "pu64" else np.float64
            assert isinstance(v, np.ndarray) and v.dtype == want and v.flags.c_contiguous, spec
            argv.append(ctypes.c_void_p(v.ctypes.data))
    return argv


def call_oracle(name, case, lib=None):
    lib = lib or oracle()
    fn = getattr(lib, "oracle_" + name)
    fn.restype = None
    fn(*_argv(name, case))


def call_ref_abi(lib, name, case):
    """Call a symbol with the reference's C prototype: the three sizes, the
    three batch sizes (set equal to the sizes, as qmctoolscl/util.py:107
    does), then the remaining arguments."""
    fn = getattr(lib, name)
    fn.restype = None
    argv = _argv(name, case)
    fn(*(argv[:3] + [ctypes.c_uint64(int(case[s[0]])) for s in SIGNATURES[name][:3]] + argv[3:]))


def digest(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()
