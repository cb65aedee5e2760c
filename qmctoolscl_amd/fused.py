"""Fused pipelines on GPU tensors (no single reference symbol; each equals a
chain of reference calls, as the reference doctests chain them, in one pass
over HBM).  Tensors must be contiguous and on a ROCm device; uint64 data may
be held in torch.int64 tensors.  Calls are enqueued on `stream` (default:
torch's current stream) and do not synchronize."""
import ctypes

from . import _lib


def _ptr(t):
    if t is None:
        return None
    if not (t.is_cuda and t.is_contiguous()):
        raise ctypes.ArgumentError("expected a contiguous GPU tensor")
    return t.data_ptr()


def _stream(stream, like):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream(like.device)
    return ctypes.c_void_p(s.cuda_stream)


def lat_gen_shift(order, r, n, d, r_g, n_start, g, shifts, xr, stream=None):
    """xr = lat_shift_mod_1(r, n, d, r_g, lat_gen_<order>(r_g, n, d, [n_start,] g), shifts).
    order: 0 linear, 1 natural, 2 gray.  g: (r_g, d) uint64, shifts: (r, d)
    float64, xr: (r, n, d) float64."""
    rc = _lib.lib().qmc_lat_gen_shift_dev(int(order), int(r), int(n), int(d), int(r_g), int(n_start),
                                          _ptr(g), _ptr(shifts), _ptr(xr), _stream(stream, xr))
    _lib.check(rc)


def dnb2_gen_shift_float(natural, r, n, d, r_x, n_start, mmax, C, lshifts, shiftsb, tmaxes, xrb=None, x=None,
                         stream=None):
    """Base-2 net generation (natural or Gray order, r_x generating matrices
    C of shape (r_x, d, mmax)), digital shift to r replications
    (lshifts: (r_x,), shiftsb: (r, d)) and conversion to double (tmaxes:
    (r,)) in one pass; writes xrb (r, n, d) uint64 and/or x (r, n, d) float64."""
    like = xrb if xrb is not None else x
    rc = _lib.lib().qmc_dnb2_gen_shift_float_dev(int(bool(natural)), int(r), int(n), int(d), int(r_x),
                                                 int(n_start), int(mmax), _ptr(C), _ptr(lshifts), _ptr(shiftsb),
                                                 _ptr(tmaxes), _ptr(xrb), _ptr(x), _stream(stream, like))
    _lib.check(rc)


def gdn_gen_scramble_float(mode, r, n, d, r_b, mmax, tmax, tmax_new, n_start, bmax, bases, C, tab, x, stream=None):
    """x = gdn_integer_to_float(r, n, d, r_b, tmax_new, bases, S(gdn_gen_natural(r, n, d, r_b, mmax, tmax,
    n_start, bases, C))) with S = the identity (mode 0, tmax_new == tmax),
    gdn_digital_permutation(r, n, d, r, r_b, tmax, tmax_new, bmax, tab, .) (mode 1)
    or gdn_digital_shift(r, n, d, r, r_b, tmax, tmax_new, bases, tab, .) (mode 2),
    in one pass that materializes no digits.  bases: (r_b, d), C: (r, d, mmax,
    tmax), tab: perms (r, d, tmax_new, bmax) or shifts (r, d, tmax_new),
    x: (r, n, d) float64."""
    L = _lib.lib()
    fn = L.qmc_gdn_gen_scramble_float_dev
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int] + [ctypes.c_uint64] * 9 + [ctypes.c_void_p] * 5
    rc = fn(int(mode), int(r), int(n), int(d), int(r_b), int(mmax), int(tmax), int(tmax_new), int(n_start),
            int(bmax), _ptr(bases), _ptr(C), _ptr(tab) if tab is not None else None, _ptr(x), _stream(stream, x))
    _lib.check(rc)
