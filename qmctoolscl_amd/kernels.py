"""The qmctoolscl call surface for the 23 kernels, running on the GPU.

Every function takes the reference's positional arguments (its C prototype
without the three batch sizes, qmctoolscl/util.py:99-114), writes its
outputs in place and returns (t_perf, t_process) like the reference wrapper.

Array arguments may be
  * numpy arrays (C-contiguous uint64 / float64, as the reference's ndpointer
    argtypes require): the call goes through the library's host drop-in
    symbols, which stage the arrays to the GPU and back; or
  * torch tensors on a ROCm device (uint64 data as torch.uint64 or
    torch.int64, doubles as torch.float64), zero-copy through the *_dev entry
    points on torch's current stream.

Keyword arguments of the reference (backend, global_size, local_size, wait,
context, queue, program, platform_id, device_id, profile) are accepted;
backend must be one of "c", "cl", "hip", "gpu" (any case), others raise
AssertionError as in qmctoolscl/util.py:52-54.  wait=False skips the final
stream synchronization for tensor arguments, and with it the check of the
device status bits that kernels set on inputs the reference leaves undefined
(read them with qmc_device_status); with wait=True such a call raises
QMCError (QMC_ERR_RANGE), as the numpy path does.  An array smaller than
the elements the call touches raises ctypes.ArgumentError before any launch.
There is no CPU path.
"""
import ctypes
import time

import numpy as np

from . import _lib
from ._api import SIGNATURES, required_elements

_BACKENDS = ("c", "cl", "hip", "gpu")

_NP_DTYPE = {"pu64": np.uint64, "pf64": np.float64}


def _parse_backend(kwargs):
    if "backend" in kwargs:
        b = str(kwargs["backend"]).lower()
        assert b in _BACKENDS, "backend must be one of %s" % (_BACKENDS,)
    return kwargs.get("wait", True)


def _is_torch(v):
    return type(v).__module__.startswith("torch") and hasattr(v, "data_ptr")


def _torch_ok(v, kind):
    import torch
    if not v.is_cuda:
        raise ctypes.ArgumentError("tensor arguments must live on a GPU device")
    if not v.is_contiguous():
        raise ctypes.ArgumentError("tensor arguments must be contiguous")
    ok = {"pu64": (torch.int64, getattr(torch, "uint64", torch.int64)), "pf64": (torch.float64,)}[kind]
    if v.dtype not in ok:
        raise ctypes.ArgumentError("expected %s data, got %s" % ("uint64" if kind == "pu64" else "float64", v.dtype))


def _np_ok(v, kind):
    if not isinstance(v, np.ndarray):
        raise ctypes.ArgumentError("expected a numpy array or a GPU tensor, got %s" % type(v).__name__)
    if v.dtype != _NP_DTYPE[kind] or not v.flags.c_contiguous:
        raise ctypes.ArgumentError("array must be C-contiguous %s, got %s" % (_NP_DTYPE[kind].__name__, v.dtype))


def _preprocess_lat_gen_natural(r, n, d, n_start, *rest):
    # qmctoolscl/util.py:84-86
    n_start, n = int(n_start), int(n)
    if not ((n_start == 0 or np.log2(n_start) % 1 == 0) and ((n + n_start) == 0 or np.log2(n + n_start) % 1 == 0)):
        raise Exception("lat_gen_natural requires n_start and n+n_start are either 0 or powers of 2")


_PREPROCESS = {"lat_gen_natural": _preprocess_lat_gen_natural}


def _make(name):
    sig = SIGNATURES[name]
    nargs = len(sig)

    def fn(*args, **kwargs):
        wait = _parse_backend(kwargs)
        if len(args) != nargs:
            raise TypeError("%s() takes %d positional arguments (%s), got %d"
                            % (name, nargs, ", ".join(s[0] for s in sig), len(args)))
        t0_perf, t0_proc = time.perf_counter(), time.process_time()
        if name in _PREPROCESS:
            _PREPROCESS[name](*args)
        arrays = [(s, a) for s, a in zip(sig, args) if s[1] != "u64"]
        on_gpu = [_is_torch(a) for _, a in arrays]
        need = required_elements(name, {s[0]: int(a) for s, a in zip(sig, args) if s[1] == "u64"})
        for s, a in arrays:
            have = a.numel() if _is_torch(a) else getattr(a, "size", None)
            if have is not None and have < need[s[0]]:
                raise ctypes.ArgumentError("%s: %s holds %d elements, the call touches %d"
                                           % (name, s[0], have, need[s[0]]))
        L = _lib.lib()
        if all(on_gpu):
            import torch
            for s, a in arrays:
                _torch_ok(a, s[1])
            dev = arrays[0][1].device
            with torch.cuda.device(dev):
                stream = torch.cuda.current_stream(dev)
                sp = ctypes.c_void_p(stream.cuda_stream)
                cargs = [int(a) if s[1] == "u64" else a.data_ptr() for s, a in zip(sig, args)]
                if wait:  # the status bits this call sets are read back below
                    _lib.check(L.qmc_status_reset_dev(sp))
                rc = getattr(L, name + "_dev")(*cargs, sp)
                _lib.check(rc)
                if wait:
                    stream.synchronize()
                    _lib.check_status(name)
        elif not any(on_gpu):
            for s, a in arrays:
                _np_ok(a, s[1])
            scal = [int(a) for s, a in zip(sig[:3], args[:3])]
            cargs = [int(a) if s[1] == "u64" else a.ctypes.data for s, a in zip(sig, args)]
            getattr(L, name)(*(cargs[:3] + scal + cargs[3:]))
            _lib.check()
        else:
            raise ctypes.ArgumentError("%s: pass either all numpy arrays or all GPU tensors" % name)
        return time.perf_counter() - t0_perf, time.process_time() - t0_proc

    fn.__name__ = name
    fn.__qualname__ = name
    fn.__doc__ = _doc(name)
    return fn


def _doc(name):
    lines = ["%s on the GPU (same arguments and in-place outputs as qmctoolscl.%s)." % (name, name), "", "Args:"]
    for s in SIGNATURES[name]:
        kind = {"u64": "np.uint64", "pu64": "array of np.uint64", "pf64": "array of np.double"}[s[1]]
        lines.append("    %s (%s)%s" % (s[0], kind, "" if len(s) < 3 or s[2] == "in" else " [%s]" % s[2]))
    return "\n".join(lines)


for _name in SIGNATURES:
    globals()[_name] = _make(_name)

__all__ = list(SIGNATURES)
