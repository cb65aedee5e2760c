"""Ragged batches of generation requests in one persistent, graph-replayed
kernel, with an autotuner for its tile size and grid.

A request is (method, m, d, r, seed): r x 2^m x d float64 points in natural
order, methods
    0 lattice + r shifts mod 1          (lat_gen_natural -> lat_shift_mod_1)
    1 dnb2, LMS + 64-bit digital shift  (dnb2_linear_matrix_scramble ->
                                         dnb2_gen_natural -> dnb2_digital_shift
                                         -> dnb2_integer_to_float)
    2 dnb2 + hash-based nested uniform scramble
    3 Halton-type base-b net (identity generating matrices), LMS to 52 digits,
      a uniform random permutation of the digit values per level, float
                                        (gdn_linear_matrix_scramble ->
                                         gdn_gen_natural ->
                                         gdn_digital_permutation ->
                                         gdn_integer_to_float)
with all random inputs drawn counter-based from the seed on the device (see
csrc/batch.hip).  Output goes to a caller-provided float64 ring buffer; every
replication slab lands contiguously at offset `slab_offset(q, l)`.

    plan = Batch(methods, m, d, r, seeds, points_per_tile=1 << 16)
    plan.run(ring)                # one persistent launch, replayed from a hipGraph
    plan.run_per_call(ring)       # the same points, one launch per request
"""
import ctypes
import json
import os
import time

import numpy as np

from . import _lib

_vp = ctypes.c_void_p


def _declare(L):
    if getattr(L, "_batch_declared", False):
        return
    L.qmc_batch_plan.restype = ctypes.c_void_p
    L.qmc_batch_plan.argtypes = [ctypes.c_int] + [_vp] * 5 + [ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                                               ctypes.c_uint64, ctypes.c_int]
    L.qmc_batch_plan_tiles.restype = ctypes.c_void_p
    L.qmc_batch_plan_tiles.argtypes = [ctypes.c_int] + [_vp] * 6 + [ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
    L.qmc_batch_free.argtypes = [_vp]
    L.qmc_batch_total_tiles.restype = ctypes.c_uint64
    L.qmc_batch_total_tiles.argtypes = [_vp]
    L.qmc_batch_points.restype = ctypes.c_uint64
    L.qmc_batch_points.argtypes = [_vp]
    L.qmc_batch_slabs.restype = ctypes.c_uint64
    L.qmc_batch_slabs.argtypes = [_vp]
    L.qmc_batch_set_digest.argtypes = [_vp, _vp]
    L.qmc_batch_set_gdn_cus.restype = ctypes.c_int
    L.qmc_batch_set_gdn_cus.argtypes = [_vp, ctypes.c_int]
    L.qmc_batch_slab_offset.restype = ctypes.c_uint64
    L.qmc_batch_slab_offset.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    for f in ("qmc_batch_run_dev", "qmc_batch_graph_run_dev", "qmc_batch_run_per_call_dev"):
        getattr(L, f).restype = ctypes.c_int
        getattr(L, f).argtypes = [_vp, _vp, _vp]
    L.qmc_batch_lattice_inputs.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp]
    L.qmc_batch_dnb2_inputs.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    L.qmc_batch_gdn_inputs.restype = ctypes.c_int
    L.qmc_batch_gdn_inputs.argtypes = [ctypes.c_uint64] + [ctypes.c_int] * 4 + [_vp] * 3
    L.qmc_batch_gdn_base.restype = ctypes.c_int
    L.qmc_batch_gdn_base.argtypes = [ctypes.c_int] + [_vp] * 4
    L.qmc_owen_hash64.restype = ctypes.c_uint64
    L.qmc_owen_hash64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L._batch_declared = True


def _L():
    L = _lib.lib()
    _declare(L)
    return L


def sample_requests(count, seed, m_range=(8, 22), d_range=(1, 256), r_range=(1, 64), methods=(0, 1, 2, 3)):
    """Requests with method uniform over `methods`, m, d, r uniform over the
    inclusive ranges, and per-request seeds from SeedSequence(seed).spawn."""
    rng = np.random.Generator(np.random.SFC64(seed))
    meth = rng.choice(np.array(methods, dtype=np.int32), size=count).astype(np.int32)
    m = rng.integers(m_range[0], m_range[1] + 1, size=count).astype(np.int32)
    d = rng.integers(d_range[0], d_range[1] + 1, size=count).astype(np.int32)
    r = rng.integers(r_range[0], r_range[1] + 1, size=count).astype(np.int32)
    seeds = np.array([s.generate_state(1, np.uint64)[0] for s in np.random.SeedSequence(seed).spawn(count)],
                     dtype=np.uint64)
    return meth, m, d, r, seeds


class Batch:
    """A plan.  points_per_tile: one tile size (points) for the lattice and
    dnb2 methods, or a sequence of three (lattice, dnb2 LMS, dnb2 hash NUS);
    gdn_points_per_tile (0 = the lattice's) sizes the gdn tiles."""

    def __init__(self, methods, m, d, r, seeds, points_per_tile=1 << 16, grid=0, ring_elems=None,
                 gdn_points_per_tile=0, gdn_cus=0, gdn_level_cap=0):
        self.methods = np.ascontiguousarray(methods, dtype=np.int32)
        self.m = np.ascontiguousarray(m, dtype=np.int32)
        self.d = np.ascontiguousarray(d, dtype=np.int32)
        self.r = np.ascontiguousarray(r, dtype=np.int32)
        self.seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        slabs = (np.int64(1) << self.m.astype(np.int64)) * self.d
        self.points = int(np.sum(slabs * self.r))
        self.ring_elems = int(ring_elems) if ring_elems is not None else self.points
        tp = [int(v) for v in points_per_tile] if np.ndim(points_per_tile) else [int(points_per_tile)] * 3
        if len(tp) != 3:
            raise ValueError("points_per_tile: one size or three (lattice, dnb2 LMS, dnb2 hash NUS)")
        self.points_per_tile = tp[0] if len(set(tp)) == 1 else tp
        self.gdn_points_per_tile = int(gdn_points_per_tile) or tp[0]
        self.tile_points = np.array(tp + [self.gdn_points_per_tile], dtype=np.uint64)
        self.grid = int(grid)
        L = _L()
        h = L.qmc_batch_plan_tiles(len(self.m), self.methods.ctypes.data, self.m.ctypes.data, self.d.ctypes.data,
                                   self.r.ctypes.data, self.seeds.ctypes.data, self.tile_points.ctypes.data,
                                   self.grid, self.ring_elems, int(gdn_level_cap))
        if not h:
            _lib.check()
        self._h = h
        self._free = L.qmc_batch_free
        self.gdn_cus = 0
        if gdn_cus:
            self.set_gdn_cus(gdn_cus)

    def __del__(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None

    @property
    def total_tiles(self):
        return int(_L().qmc_batch_total_tiles(self._h))

    @property
    def slabs(self):
        return int(_L().qmc_batch_slabs(self._h))

    def digests(self, ring, per_call=False, graph=True, stream=None):
        """Run once and return the per-slab digests (wrap-around sums of the
        stored float64 bit patterns), as a uint64 numpy array of length slabs."""
        import torch
        dg = torch.zeros(max(self.slabs, 1), dtype=torch.int64, device=ring.device)
        _lib.check(_L().qmc_batch_set_digest(self._h, dg.data_ptr()))
        try:
            if per_call:
                self.run_per_call(ring, stream)
            else:
                self.run(ring, graph=graph, stream=stream)
            torch.cuda.synchronize(ring.device)
        finally:
            _lib.check(_L().qmc_batch_set_digest(self._h, None))
        return dg[:self.slabs].cpu().numpy().view(np.uint64)

    def set_gdn_cus(self, n):
        """CUs (persistent workgroups) of the gdn class; the lattice and dnb2
        class runs on the others beside it."""
        _lib.check(_L().qmc_batch_set_gdn_cus(self._h, int(n)))
        self.gdn_cus = int(n)

    def slab_offset(self, q, l):
        return int(_L().qmc_batch_slab_offset(self._h, int(q), int(l)))

    def _check_ring(self, ring):
        if not (ring.is_cuda and ring.is_contiguous() and str(ring.dtype) == "torch.float64"):
            raise ctypes.ArgumentError("ring must be a contiguous float64 GPU tensor")
        if ring.numel() < self.ring_elems:
            raise ValueError("ring holds %d doubles, the plan needs %d" % (ring.numel(), self.ring_elems))

    def _stream(self, ring, stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream(ring.device)
        return _vp(s.cuda_stream)

    def run(self, ring, graph=True, stream=None):
        self._check_ring(ring)
        f = _L().qmc_batch_graph_run_dev if graph else _L().qmc_batch_run_dev
        _lib.check(f(self._h, ring.data_ptr(), self._stream(ring, stream)))

    def run_per_call(self, ring, stream=None):
        self._check_ring(ring)
        _lib.check(_L().qmc_batch_run_per_call_dev(self._h, ring.data_ptr(), self._stream(ring, stream)))

    def view(self, ring, q):
        """(r, n, d) view of request q's points (when its slabs are consecutive)."""
        n, d, r = 1 << int(self.m[q]), int(self.d[q]), int(self.r[q])
        o = self.slab_offset(q, 0)
        return ring[o:o + r * n * d].view(r, n, d)


def lattice_inputs(seed, m, d, r):
    g = np.empty((1, d), dtype=np.uint64)
    sh = np.empty((r, d), dtype=np.float64)
    _L().qmc_batch_lattice_inputs(int(seed), int(m), int(d), int(r), g.ctypes.data, sh.ctypes.data)
    return g, sh


def dnb2_inputs(seed, method, d, r):
    C = np.empty((1, d, 32), dtype=np.uint64)
    S = np.empty((r, d, 32), dtype=np.uint64)
    sh = np.empty((r, d), dtype=np.uint64)
    _L().qmc_batch_dnb2_inputs(int(seed), int(method), int(d), int(r), C.ctypes.data, S.ctypes.data, sh.ctypes.data)
    return C, S, sh


def first_primes(k):
    """The first k primes (the bases of a gdn request's columns)."""
    out, c = [], 2
    while len(out) < k:
        if all(c % p for p in out if p * p <= c):
            out.append(c)
        c += 1
    return out


def gdn_inputs(seed, m, r, j0=0, dj=1):
    """Inputs of a gdn request for columns [j0, j0 + dj) in the reference's
    layouts: bases (1, dj), S (r, dj, 52, m) -- the LMS matrices, tmax = m --
    and perms (r, dj, 52, bmax); the generating matrices are m x m identities."""
    bmax = first_primes(j0 + dj)[-1]
    bases = np.empty((1, dj), dtype=np.uint64)
    S = np.empty((r, dj, 52, m), dtype=np.uint64)
    perms = np.empty((r, dj, 52, bmax), dtype=np.uint64)
    _lib.check(_L().qmc_batch_gdn_inputs(int(seed), int(m), int(j0), int(dj), int(r), bases.ctypes.data,
                                          S.ctypes.data, perms.ctypes.data))
    return bases, S, perms


def gdn_base(j):
    """(rec, thr, levels, tmin) of column j's base: the reciprocals of the
    reference's sum, the early-exit thresholds, the levels kept in LDS and the
    first level that can end a sum."""
    rec = np.empty(52)
    thr = np.empty(52)
    lv = ctypes.c_int()
    tm = ctypes.c_int()
    _lib.check(_L().qmc_batch_gdn_base(int(j), rec.ctypes.data, thr.ctypes.data, ctypes.byref(lv), ctypes.byref(tm)))
    return rec, thr, lv.value, tm.value


def owen_hash64(x, seed):
    return int(_L().qmc_owen_hash64(int(x), int(seed)))


# ------------------------------------------------------------------ autotuner

DEFAULT_TILES = (1 << 13, 1 << 14, 1 << 15, 1 << 16, 1 << 17)
DEFAULT_GRIDS_PER_CU = (4, 8)
DEFAULT_GDN_TILES = (1 << 15, 1 << 16, 1 << 17)
DEFAULT_GDN_SHARES = (0.5, 0.625, 0.75, 0.875, 1.0)


def _cache_path():
    base = os.environ.get("QMC_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "qmctoolscl_amd"))
    return os.path.join(base, "batch_autotune.json")


def _time_plan(plan, ring, reps):
    import torch
    plan.run(ring)  # warm: captures the graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.run(ring)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def autotune(methods, m, d, r, seeds, ring, tiles=DEFAULT_TILES, grids_per_cu=DEFAULT_GRIDS_PER_CU,
             gdn_tiles=DEFAULT_GDN_TILES, gdn_shares=DEFAULT_GDN_SHARES, reps=2, cache=True):
    """Pick the batch's tunables by timing it, one dimension at a time:
    1. each store-bound method's tile size (rows of all d columns per tile: the
       split of its n x d slabs; lattice, dnb2 LMS, dnb2 hash NUS) on that
       method's requests alone;
    2. the persistent grid of the lattice/dnb2 class on those requests with
       their chosen tiles;
    3. the gdn tile size (rows of a column group: the n split of gdn slabs,
       whose d split follows the 128-byte lines) on the gdn requests alone;
    4. the split of the CUs between the two classes on the whole batch (the
       gdn kernel's share; the store-bound kernel runs on the rest beside it).
    Returns {"points_per_tile": [lattice, dnb2, nus], "grid",
    "gdn_points_per_tile", "gdn_cus", "ms", "table"}, cached per (device,
    batch signature).  The searched shapes replace the reference's global-size
    sweeps (gs_run.py:42-88) for this path."""
    import torch
    methods = np.asarray(methods, np.int32)
    m, d, r, seeds = (np.asarray(a) for a in (m, d, r, seeds))
    dev = ring.device
    props = torch.cuda.get_device_properties(dev)
    cus = props.multi_processor_count
    sig = "%s|%d|%d|%s|v2" % (props.name, len(m), int(np.sum((np.int64(1) << m.astype(np.int64)) * d * r)),
                              hash(seeds.tobytes() + methods.tobytes()))
    path = _cache_path()
    if cache and os.path.exists(path):
        try:
            with open(path) as f:
                hit = json.load(f).get(sig)
            if hit:
                return hit
        except (OSError, ValueError):
            pass
    table = []
    n_ring = ring.numel()
    names = ("lattice", "dnb2_lms", "dnb2_nus")
    c0 = methods != 3
    tp = [tiles[0]] * 3
    grid0 = grids_per_cu[0] * cus
    for k in range(3):
        sel = methods == k
        if not sel.any():
            continue
        best = None
        for p in tiles:
            plan = Batch(methods[sel], m[sel], d[sel], r[sel], seeds[sel], points_per_tile=p, grid=grid0,
                         ring_elems=n_ring)
            ms = _time_plan(plan, ring, reps)
            table.append({"class": names[k], "points_per_tile": p, "grid": grid0, "ms": round(ms, 3)})
            if best is None or ms < best:
                best, tp[k] = ms, p
            del plan
    best0 = grid0
    if c0.any():
        ms0 = None
        for g in grids_per_cu:
            plan = Batch(methods[c0], m[c0], d[c0], r[c0], seeds[c0], points_per_tile=tp, grid=g * cus,
                         ring_elems=n_ring)
            ms = _time_plan(plan, ring, reps)
            table.append({"class": "lattice+dnb2", "points_per_tile": list(tp), "grid": g * cus, "ms": round(ms, 3)})
            if ms0 is None or ms < ms0:
                ms0, best0 = ms, g * cus
            del plan
    bestg = gdn_tiles[0]
    if (~c0).any():
        msg = None
        for p in gdn_tiles:
            plan = Batch(methods[~c0], m[~c0], d[~c0], r[~c0], seeds[~c0], points_per_tile=p, ring_elems=n_ring)
            ms = _time_plan(plan, ring, reps)
            table.append({"class": "gdn", "gdn_points_per_tile": p, "ms": round(ms, 3)})
            if msg is None or ms < msg:
                msg, bestg = ms, p
            del plan
    plan = Batch(methods, m, d, r, seeds, points_per_tile=tp, grid=best0, ring_elems=n_ring,
                 gdn_points_per_tile=bestg)
    best = None
    shares = gdn_shares if (c0.any() and (~c0).any()) else (1.0,)
    for sh in shares:
        n_cus = max(1, min(cus, int(round(sh * cus))))
        plan.set_gdn_cus(n_cus)
        ms = _time_plan(plan, ring, reps)
        table.append({"class": "both", "gdn_cus": n_cus, "ms": round(ms, 3)})
        if best is None or ms < best["ms"]:
            best = {"points_per_tile": list(tp), "grid": best0, "gdn_points_per_tile": bestg, "gdn_cus": n_cus,
                    "ms": round(ms, 3)}
    del plan
    best["table"] = table
    if cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            data = {}
            if os.path.exists(path):
                with open(path) as f:
                    data = json.load(f)
            data[sig] = best
            with open(path, "w") as f:
                json.dump(data, f)
        except (OSError, ValueError):
            pass
    return best
