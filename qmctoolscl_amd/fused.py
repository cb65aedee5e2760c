"""Fused pipelines on GPU tensors (no single reference symbol; each equals a
chain of reference calls, as the reference doctests chain them, in one pass
over HBM).  Tensors must be contiguous and on a ROCm device; uint64 data may
be held in torch.int64 tensors.  Calls are enqueued on `stream` (default:
torch's current stream) and do not synchronize."""
import ctypes

from . import _lib


def _ptr(t, need=0, f64=False, what="tensor"):
    """Device pointer of a contiguous GPU tensor holding at least `need`
    elements of float64 (f64) or 64-bit integers."""
    if t is None:
        return None
    import torch
    if not (getattr(t, "is_cuda", False) and t.is_contiguous()):
        raise ctypes.ArgumentError("%s: expected a contiguous GPU tensor" % what)
    ok = (torch.float64,) if f64 else (torch.int64, getattr(torch, "uint64", torch.int64))
    if t.dtype not in ok:
        raise ctypes.ArgumentError("%s: expected %s data, got %s" % (what, "float64" if f64 else "64-bit integer",
                                                                     t.dtype))
    if t.numel() < need:
        raise ctypes.ArgumentError("%s holds %d elements, the call touches %d" % (what, t.numel(), need))
    return t.data_ptr()


def _stream(stream, like):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream(like.device)
    return ctypes.c_void_p(s.cuda_stream)


def lat_gen_shift(order, r, n, d, r_g, n_start, g, shifts, xr, stream=None):
    """xr = lat_shift_mod_1(r, n, d, r_g, lat_gen_<order>(r_g, n, d, [n_start,] g), shifts).
    order: 0 linear, 1 natural, 2 gray.  g: (r_g, d) uint64, shifts: (r, d)
    float64, xr: (r, n, d) float64."""
    r, n, d, r_g = int(r), int(n), int(d), int(r_g)
    rc = _lib.lib().qmc_lat_gen_shift_dev(int(order), r, n, d, r_g, int(n_start),
                                          _ptr(g, r_g * d, what="g"), _ptr(shifts, r * d, True, "shifts"),
                                          _ptr(xr, r * n * d, True, "xr"), _stream(stream, xr))
    _lib.check(rc)


def dnb2_gen_shift_float(natural, r, n, d, r_x, n_start, mmax, C, lshifts, shiftsb, tmaxes, xrb=None, x=None,
                         stream=None):
    """Base-2 net generation (natural or Gray order, r_x generating matrices
    C of shape (r_x, d, mmax)), digital shift to r replications
    (lshifts: (r_x,), shiftsb: (r, d)) and conversion to double (tmaxes:
    (r,)) in one pass; writes xrb (r, n, d) uint64 and/or x (r, n, d) float64."""
    like = xrb if xrb is not None else x
    r, n, d, r_x, mmax = int(r), int(n), int(d), int(r_x), int(mmax)
    rx = min(r_x, r)  # nets the call reads
    rc = _lib.lib().qmc_dnb2_gen_shift_float_dev(int(bool(natural)), r, n, d, r_x, int(n_start), mmax,
                                                 _ptr(C, rx * d * mmax, what="C"), _ptr(lshifts, rx, what="lshifts"),
                                                 _ptr(shiftsb, r * d, what="shiftsb"), _ptr(tmaxes, r, what="tmaxes"),
                                                 _ptr(xrb, r * n * d, what="xrb"), _ptr(x, r * n * d, True, "x"),
                                                 _stream(stream, like))
    _lib.check(rc)


def dnb2_lms_gen_shift_float(natural, r, n, d, r_C, n_start, mmax, tmax_new, S, C, lshifts, shiftsb, tmaxes,
                             xrb=None, x=None, stream=None):
    """dnb2_linear_matrix_scramble(r, d, mmax, r_C, tmax_new, S, C) -> dnb2_gen_<order>(r) ->
    dnb2_digital_shift(r, r_x = r) -> dnb2_integer_to_float(r) in one launch.  S: (r, d, tmax_new),
    C: (r_C, d, mmax), lshifts: (r,), shiftsb: (r, d), tmaxes: (r,); writes xrb (r, n, d) uint64
    and/or x (r, n, d) float64."""
    L = _lib.lib()
    fn = L.qmc_dnb2_lms_gen_shift_float_dev
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int] + [ctypes.c_uint64] * 7 + [ctypes.c_void_p] * 8
    like = xrb if xrb is not None else x
    r, n, d, r_C, mmax, tmax_new = (int(v) for v in (r, n, d, r_C, mmax, tmax_new))
    rc = fn(int(bool(natural)), r, n, d, r_C, int(n_start), mmax, tmax_new, _ptr(S, r * d * tmax_new, what="S"),
            _ptr(C, r_C * d * mmax, what="C"), _ptr(lshifts, r, what="lshifts"), _ptr(shiftsb, r * d, what="shiftsb"),
            _ptr(tmaxes, r, what="tmaxes"), _ptr(xrb, r * n * d, what="xrb"), _ptr(x, r * n * d, True, "x"),
            _stream(stream, like))
    _lib.check(rc)


def gdn_gen_scramble_float(mode, r, n, d, r_b, mmax, tmax, tmax_new, n_start, bmax, bases, C, tab, x, stream=None,
                           no_alloc=False):
    """x = gdn_integer_to_float(r, n, d, r_b, tmax_new, bases, S(gdn_gen_natural(r, n, d, r_b, mmax, tmax,
    n_start, bases, C))) with S = the identity (mode 0, tmax_new == tmax),
    gdn_digital_permutation(r, n, d, r, r_b, tmax, tmax_new, bmax, tab, .) (mode 1)
    or gdn_digital_shift(r, n, d, r, r_b, tmax, tmax_new, bases, tab, .) (mode 2),
    in one pass that materializes no digits.  bases: (r_b, d), C: (r, d, mmax,
    tmax), tab: perms (r, d, tmax_new, bmax) or shifts (r, d, tmax_new),
    x: (r, n, d) float64.  no_alloc=True takes the entry point that allocates
    no device memory (its fixed workspace holds r*d*mmax*tmax + r*d <= 2^20
    words; calls overlapping on one device race on it)."""
    L = _lib.lib()
    fn = L.qmc_gdn_gen_scramble_float_noalloc_dev if no_alloc else L.qmc_gdn_gen_scramble_float_dev
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int] + [ctypes.c_uint64] * 9 + [ctypes.c_void_p] * 5
    r, n, d, r_b, mmax, tmax, tmax_new, bmax = (int(v) for v in (r, n, d, r_b, mmax, tmax, tmax_new, bmax))
    need_tab = {0: 0, 1: r * d * tmax_new * bmax, 2: r * d * tmax_new}.get(int(mode), 0)
    rc = fn(int(mode), r, n, d, r_b, mmax, tmax, tmax_new, int(n_start), bmax, _ptr(bases, r_b * d, what="bases"),
            _ptr(C, r * d * mmax * tmax, what="C"), _ptr(tab, need_tab, what="tab") if tab is not None else None,
            _ptr(x, r * n * d, True, "x"), _stream(stream, x))
    _lib.check(rc)
