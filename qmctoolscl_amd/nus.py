"""Nested uniform scrambling with persistent trees, on the GPU.

Same call surface as the reference (qmctoolscl/rand_funcs.py:155-356):

    dnb2_nested_uniform_scramble(r, n, d, r_x, tmax, tmax_new, rngs, root_nodes, xb, xrb)
    gdn_nested_uniform_scramble(r, n, d, r_x, r_b, tmax, tmax_new, rngs, root_nodes, bases, xdig, xrdig)

with `rngs` an (r, d) array of numpy Generators and `root_nodes` an (r, d)
array of NUSNode_dnb2() / NUSNode_gdn() that keep each tree between calls.

Each tree is grown on the host by the library's C++ builder in point order,
which records how many random bits (base 2) or permutation rows (base b)
every new leaf needs.  Base 2: the draws are made with the tree's Generator in
that order -- one vectorized numpy call per tree, which consumes the stream
exactly as the reference's one call per leaf does.  Base b with numpy's SFC64
(the generators the reference's documentation uses): the builder replays the
generator itself on host threads, one tree per thread, and keeps one
generator state per leaf instead of its rows (a reference leaf holds
tmax_new - depth rows of b entries); the Generator objects are left in the
state the reference's draws would leave them.  Other bit generators: the rows
are drawn with numpy and all of them kept.  The GPU then walks the finished
trees to scramble every point in parallel, regenerating leaf rows from their
states.  Outputs equal the reference's for the same Generators, bit for bit.
"""
import os
import ctypes
import time

import numpy as np

from . import _lib
from .rand_funcs import random_tbit_uint64s, random_uint64_permutations

_u64p = ctypes.c_void_p


def _declare(L):
    if getattr(L, "_nus_declared", False):
        return
    L.qmc_nus2_tree_new.restype = ctypes.c_void_p
    L.qmc_nus2_tree_new.argtypes = [ctypes.c_int]
    L.qmc_nus2_tree_free.argtypes = [ctypes.c_void_p]
    L.qmc_nus2_insert.restype = ctypes.c_int64
    L.qmc_nus2_insert.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _u64p]
    L.qmc_nus2_add_draws.restype = ctypes.c_int
    L.qmc_nus2_add_draws.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64]
    L.qmc_nus2_apply_dev.restype = ctypes.c_int
    L.qmc_nus2_apply_dev.argtypes = ([ctypes.c_void_p] + [ctypes.c_uint64] * 4 + [ctypes.c_int] * 2 +
                                     [_u64p, _u64p, ctypes.c_void_p])
    L.qmc_nusb_tree_new.restype = ctypes.c_void_p
    L.qmc_nusb_tree_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    L.qmc_nusb_tree_free.argtypes = [ctypes.c_void_p]
    L.qmc_nusb_insert.restype = ctypes.c_int64
    L.qmc_nusb_insert.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _u64p]
    L.qmc_nusb_add_rows.restype = ctypes.c_int
    L.qmc_nusb_add_rows.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    L.qmc_nusb_tree_rows_needed.restype = ctypes.c_uint64
    L.qmc_nusb_tree_rows_needed.argtypes = [ctypes.c_void_p]
    for f in ("qmc_nusb_tree_rows", "qmc_nusb_tree_states", "qmc_nusb_tree_nodes", "qmc_nusb_tree_bytes"):
        getattr(L, f).restype = ctypes.c_uint64
        getattr(L, f).argtypes = [ctypes.c_void_p]
    L.qmc_nusb_root_rows.restype = ctypes.c_int
    L.qmc_nusb_root_rows.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.qmc_nusb_grow.restype = ctypes.c_int
    L.qmc_nusb_grow.argtypes = ([ctypes.c_void_p, ctypes.c_uint64, _u64p] + [ctypes.c_uint64] * 4 +
                                [_u64p, ctypes.c_void_p, ctypes.c_int])
    L.qmc_nus_sfc64_permutations.restype = None
    L.qmc_nus_sfc64_permutations.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
    L.qmc_nusb_apply_dev.restype = ctypes.c_int
    L.qmc_nusb_apply_dev.argtypes = ([ctypes.c_void_p] + [ctypes.c_uint64] * 5 + [ctypes.c_int] * 2 +
                                     [_u64p, _u64p, _u64p, ctypes.c_void_p])
    L._nus_declared = True


def _lib_nus():
    L = _lib.lib()
    _declare(L)
    return L


class _NativeTree:
    """Owns one native tree handle."""

    def __init__(self, handle, free):
        self.handle = handle
        self._free = free

    def __del__(self):
        if self.handle:
            self._free(self.handle)
            self.handle = None


class NUSNode_dnb2(object):
    """Root of one base-2 NUS tree (rand_funcs.py:155-160).  Create empty and
    pass in `root_nodes`; the tree lives in native memory and grows across
    calls.  `shift_bits` and `xb` report the root's leaf state while the root
    is still a leaf (None before first use), as the reference's attributes do."""

    def __init__(self, shift_bits=None, xb=None, left_b2=None, right_b2=None):
        if shift_bits is not None or xb is not None or left_b2 is not None or right_b2 is not None:
            raise NotImplementedError("NUSNode_dnb2: only empty root nodes are supported; trees are grown natively")
        self._tree = None
        self._tmax_new = None
        self.shift_bits = None
        self.xb = None
        self.left_b2 = None
        self.right_b2 = None


class NUSNode_gdn(object):
    """Root of one base-b NUS tree (rand_funcs.py:268-272); see NUSNode_dnb2."""

    def __init__(self, perm=None, xdig=None, children=None):
        if perm is not None or xdig is not None or children is not None:
            raise NotImplementedError("NUSNode_gdn: only empty root nodes are supported; trees are grown natively")
        self._tree = None
        self._key = None
        self.perm = None
        self.xdig = None
        self.children = None


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def _host_u64(a):
    if _is_torch(a):
        return a.detach().cpu().numpy().view(np.uint64)
    if a.dtype != np.uint64 or not a.flags.c_contiguous:
        raise ctypes.ArgumentError("expected a C-contiguous uint64 array")
    return a


class _Device:
    """Device views of the input/output: zero-copy for GPU tensors, staged for numpy."""

    def __init__(self, src, dst):
        import torch
        self.torch = torch
        self.dst = dst
        if _is_torch(dst):
            self.dev = dst.device
        else:
            self.dev = torch.device("cuda", torch.cuda.current_device())
        self.src_t = src if _is_torch(src) else torch.from_numpy(np.ascontiguousarray(src).view(np.int64)).to(self.dev)
        self.dst_t = dst if _is_torch(dst) else torch.empty(dst.shape, dtype=torch.int64, device=self.dev)

    def finish(self):
        if not _is_torch(self.dst):
            self.dst[...] = self.dst_t.cpu().numpy().view(np.uint64)


def dnb2_nested_uniform_scramble(r, n, d, r_x, tmax, tmax_new, rngs, root_nodes, xb, xrb):
    """Nested uniform scramble of a base-2 digital net (rand_funcs.py:162-266).
    xb: (r_x, n, d) uint64, xrb: (r, n, d) uint64 out.  Returns (t_perf, t_process)."""
    t0_perf, t0_proc = time.perf_counter(), time.process_time()
    r, n, d, r_x, tmax, tmax_new = (int(v) for v in (r, n, d, r_x, tmax, tmax_new))
    if not (1 <= tmax_new <= 64 and 0 <= tmax <= tmax_new):
        raise ValueError("dnb2_nested_uniform_scramble: need 0 <= tmax <= tmax_new <= 64")
    L = _lib_nus()
    host = _host_u64(xb).reshape(-1, n, d)
    tdelta = tmax_new - tmax
    handles = (ctypes.c_void_p * (r * d))()
    widths = np.empty(max(n, 1), dtype=np.uint32)
    for l in range(r):
        col_base = host[l % r_x]
        for j in range(d):
            node = root_nodes[l, j]
            if not isinstance(node, NUSNode_dnb2):
                raise TypeError("root_nodes must hold NUSNode_dnb2 objects")
            rng = rngs[l, j]
            if node._tree is None:
                h = L.qmc_nus2_tree_new(tmax_new)
                if not h:
                    _lib.check()
                node._tree = _NativeTree(h, L.qmc_nus2_tree_free)
                node._tmax_new = tmax_new
                root = random_tbit_uint64s(rng, tmax_new, 1)  # rand_funcs.py:199
                L.qmc_nus2_add_draws(h, root.ctypes.data, 1)
                node.shift_bits, node.xb = np.uint64(root[0]), np.uint64(0)
            elif node._tmax_new != tmax_new:
                raise ValueError("root node was grown with tmax_new=%d" % node._tmax_new)
            h = node._tree.handle
            col = np.ascontiguousarray(col_base[:, j])
            k = L.qmc_nus2_insert(h, col.ctypes.data, n, 1, tdelta, widths.ctypes.data)
            if k:
                vals = rng.integers(0, np.uint64(1) << widths[:k].astype(np.uint64), dtype=np.uint64)
                vals = np.ascontiguousarray(vals, dtype=np.uint64)
                _lib.check(L.qmc_nus2_add_draws(h, vals.ctypes.data, k))
            handles[l * d + j] = h
    io = _Device(xb, xrb)
    stream = io.torch.cuda.current_stream(io.dev)
    with io.torch.cuda.device(io.dev):
        rc = L.qmc_nus2_apply_dev(handles, r, n, d, r_x, tmax, tmax_new, io.src_t.data_ptr(),
                                  io.dst_t.data_ptr(), ctypes.c_void_p(stream.cuda_stream))
    _lib.check(rc)
    io.finish()
    return time.perf_counter() - t0_perf, time.process_time() - t0_proc


def _sfc64_state(rng):
    """numpy's SFC64 state as the library's 5 words, or None for other bit generators."""
    bg = rng.bit_generator
    if not isinstance(bg, np.random.SFC64):
        return None
    st = bg.state
    w = np.empty(5, dtype=np.uint64)
    w[:4] = np.asarray(st["state"]["state"], dtype=np.uint64)
    w[4] = (int(st["has_uint32"]) << 32) | (int(st["uinteger"]) & 0xFFFFFFFF)
    return w


def _set_sfc64_state(rng, w):
    rng.bit_generator.state = {"bit_generator": "SFC64", "state": {"state": np.array(w[:4], dtype=np.uint64)},
                               "has_uint32": int(w[4] >> np.uint64(32)), "uinteger": int(w[4] & np.uint64(0xFFFFFFFF))}


def sfc64_permutations(rng, count, b):
    """`count` rows of Generator.permutation(b) drawn by the library's SFC64
    replay from rng's state (rng advanced as numpy's draws would advance it);
    the routine the base-b builder uses, exposed for checks."""
    w = _sfc64_state(rng)
    if w is None:
        raise TypeError("sfc64_permutations needs a Generator over numpy's SFC64")
    out = np.empty((int(count), int(b)), dtype=np.uint16)
    _lib_nus().qmc_nus_sfc64_permutations(w.ctypes.data, int(b), int(count), out.ctypes.data)
    _set_sfc64_state(rng, w)
    return out


def _host_threads():
    return max(1, min(16, os.cpu_count() or 1))


def gdn_nested_uniform_scramble(r, n, d, r_x, r_b, tmax, tmax_new, rngs, root_nodes, bases, xdig, xrdig):
    """Nested uniform scramble of a base-b digital net (rand_funcs.py:274-356).
    xdig: (r_x, n, d, tmax) uint64, xrdig: (r, n, d, tmax_new) uint64 out,
    bases: (r_b, d).  Returns (t_perf, t_process)."""
    t0_perf, t0_proc = time.perf_counter(), time.process_time()
    r, n, d, r_x, r_b, tmax, tmax_new = (int(v) for v in (r, n, d, r_x, r_b, tmax, tmax_new))
    if not (0 <= tmax <= tmax_new):
        raise ValueError("gdn_nested_uniform_scramble: need tmax <= tmax_new")
    L = _lib_nus()
    host = np.ascontiguousarray(_host_u64(xdig).reshape(-1, n, d, tmax))
    if host.shape[0] < r_x:
        raise ctypes.ArgumentError("xdig holds %d replications, r_x = %d" % (host.shape[0], r_x))
    b2 = np.atleast_2d(_host_u64(bases) if not isinstance(bases, np.ndarray) else bases)
    nt = r * d
    handles = (ctypes.c_void_p * max(nt, 1))()
    states = np.zeros((max(nt, 1), 5), dtype=np.uint64)
    replay = np.zeros(max(nt, 1), dtype=np.uint8)
    drawn = []  # trees whose rows numpy draws (bit generators other than SFC64)
    fresh = []
    for l in range(r):
        for j in range(d):
            k = l * d + j
            node = root_nodes[l, j]
            if not isinstance(node, NUSNode_gdn):
                raise TypeError("root_nodes must hold NUSNode_gdn objects")
            rng = rngs[l, j]
            b = int(b2[l % r_b, j])
            w = _sfc64_state(rng)
            if node._tree is None:
                h = L.qmc_nusb_tree_new(tmax, tmax_new, b)
                if not h:
                    _lib.check()
                node._tree = _NativeTree(h, L.qmc_nusb_tree_free)
                node._key = (tmax, tmax_new, b)
                node.xdig = np.zeros(tmax_new, dtype=np.uint64)
                fresh.append(k)
                if w is None:  # rand_funcs.py:316, drawn now as the reference does
                    rows = random_uint64_permutations(rng, tmax_new, b)
                    rows16 = np.ascontiguousarray(rows, dtype=np.uint16)  # kept alive through the call
                    _lib.check(L.qmc_nusb_add_rows(h, rows16.ctypes.data, tmax_new))
                    node.perm = rows
            elif node._key != (tmax, tmax_new, b):
                raise ValueError("root node was grown with (tmax, tmax_new, b) = %s" % (node._key,))
            if w is None:
                drawn.append(k)
            else:
                states[k] = w
                replay[k] = 1
            handles[k] = node._tree.handle
    if nt:
        _lib.check(L.qmc_nusb_grow(handles, nt, host.ctypes.data, r, n, d, r_x, states.ctypes.data,
                                   replay.ctypes.data, _host_threads()))
    for k in range(nt):
        l, j = divmod(k, d)
        node = root_nodes[l, j]
        if replay[k]:
            _set_sfc64_state(rngs[l, j], states[k])
            if k in fresh:
                b = node._key[2]
                rows = np.empty((tmax_new, b), dtype=np.uint16)
                _lib.check(L.qmc_nusb_root_rows(node._tree.handle, rows.ctypes.data))
                node.perm = rows.astype(np.uint64)
    for k in drawn:
        l, j = divmod(k, d)
        h = root_nodes[l, j]._tree.handle
        total = int(L.qmc_nusb_tree_rows_needed(h))
        if total:
            bi = root_nodes[l, j]._key[2]
            # Generator.permuted shuffles each row with the same draws whatever
            # the dtype, so 16-bit rows equal the reference's 64-bit ones.
            rows = rngs[l, j].permuted(np.tile(np.arange(bi, dtype=np.uint16), (total, 1)), axis=1)
            rows = np.ascontiguousarray(rows)
            _lib.check(L.qmc_nusb_add_rows(h, rows.ctypes.data, total))
    io = _Device(host if not _is_torch(xdig) else xdig, xrdig)
    import torch
    bases_t = torch.from_numpy(np.ascontiguousarray(b2, dtype=np.uint64).view(np.int64)).to(io.dev)
    stream = torch.cuda.current_stream(io.dev)
    with torch.cuda.device(io.dev):
        rc = L.qmc_nusb_apply_dev(handles, r, n, d, r_x, r_b, tmax, tmax_new, bases_t.data_ptr(),
                                  io.src_t.data_ptr(), io.dst_t.data_ptr(), ctypes.c_void_p(stream.cuda_stream))
    _lib.check(rc)
    io.finish()
    return time.perf_counter() - t0_perf, time.process_time() - t0_proc
