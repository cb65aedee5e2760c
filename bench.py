"""Benchmark of the QMC point-generation and transform hot path on one MI355X.

Headline workload (one step): a ragged batch of 512 generation requests --
method uniform over {lattice + shifts, dnb2 LMS + digital shift, dnb2 with a
hash-based nested uniform scramble, Halton-type gdn LMS + uniform digit
permutations (52 digits, the reference's sequential float)}, n = 2^m with m
uniform in [8, 22], d in [1, 256], r in [1, 64] -- 1.36e12 float64 points
(10.9 TB) streamed through a 32 GiB ring by the persistent batch kernels,
replayed from a hipGraph; random inputs counter-based, built on the device
before timing.  value = points/s (Gpts/s); tile sizes and the CU split
between the two kernel classes come from the autotuner.  Beside it: the same
requests with one launch per request, and, on a bounded ragged sub-batch, the
reference's own call sequence per request through the drop-in kernels.

Further sections (same JSON line, under "sections"):
  lattice     rank-1 lattice, n=2^20, d=64, r=16 shifts mod 1, fused kernel vs
              the two drop-ins (the reference's chain), with its own roofline
  dnb2        base-2 net, n=2^16, d=8, r=4: LMS + natural order + 64-bit shift
              + float, fused vs the four drop-ins; nested uniform scrambling
  gdn         Halton-type net, first 32 primes, n=3^12, d=32, r=8, LMS to 52
              digits, permutations, float: the drop-in chain and the fused
              kernel; NUS on the first 3^7 points; dnb2 interlace
  transforms  FWHT / FFT / IFFT, float64 drop-ins (bit-exact / 1e-12) and the
              matrix-core forms, at 4096 x 2^12, 64 x 2^20, 4 x 2^24
  kernels     kernel-matrix columns + transforms (lattice and net kernels)
  halton      generalized Halton, n=2^20, d=32

Roofline: the headline's algorithmic bytes (8 B written per point) over the
duration of one graph replay (the two persistent batch kernels, running
side by side) from HIP events on the launching stream, against 8 TB/s.
CPU baseline: the reference C backend built from its own sources
(oracle/_ref; this project's C restatement of it where absent), one core as
the reference runs it (qmctoolscl/util.py:107), on bounded samples of each
configuration, with the scale stated.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--sections ...] [--quick-batch]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_MFMA_PEAK_TFLOPS = 157.3
BATCH_TRAFFIC = os.path.join(HERE, "profiles", "batch_512_ragged_hbm_traffic.json")
LATTICE_TRAFFIC = os.path.join(HERE, "profiles", "lat_gen_shift_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    def __init__(self, torch, stream):
        self.torch = torch
        self.stream = stream

    def __call__(self, fn, steps, warmup):
        """Average ms per call from HIP events recorded on the launching stream."""
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        start.record(self.stream)
        for _ in range(steps):
            fn()
        end.record(self.stream)
        torch.cuda.synchronize()
        return start.elapsed_time(end) / steps, time.perf_counter() - t0


def t64(torch, a, dev):
    return torch.from_numpy(a.view(np.int64) if a.dtype == np.uint64 else a).to(dev)


def committed_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (not measured in this run), or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


# ---------------------------------------------------------------- headline

def batch_requests(quick):
    from qmctoolscl_amd import batch as B
    meth, m, d, r, seeds = B.sample_requests(512, 2024)
    if quick:
        m = np.minimum(m, 16).astype(np.int32)
    return meth, m, d, r, seeds


def section_batch(torch, dev, timed, args):
    """The config-3 batch: autotuned persistent launch under a hipGraph, vs one
    launch per request; per-slab digests of both at full size."""
    from qmctoolscl_amd import batch as B
    meth, m, d, r, seeds = batch_requests(args.quick_batch)
    ring_elems = 4 << 30  # 32 GiB of doubles
    ring = torch.empty(ring_elems, dtype=torch.float64, device=dev)
    # tune on the same requests at a quarter of the points per replication
    m_tune = np.maximum(m - 2, 8).astype(np.int32)
    t0 = time.perf_counter()
    if args.tune:
        p, g, gp, gc = args.tune.split(",")
        p = [int(v) for v in p.split("/")] if "/" in p else int(p)
        g, gp, gc = int(g), int(gp), int(gc)
        tune = {"points_per_tile": p, "grid": g, "gdn_points_per_tile": gp, "gdn_cus": gc, "ms": None,
                "table": "given on the command line"}
    else:
        tune = B.autotune(meth, m_tune, d, r, seeds, ring, reps=1, cache=False)
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan = B.Batch(meth, m, d, r, seeds, points_per_tile=tune["points_per_tile"], grid=tune["grid"],
                   ring_elems=ring_elems, gdn_points_per_tile=tune["gdn_points_per_tile"], gdn_cus=tune["gdn_cus"])
    plan_s = time.perf_counter() - t0
    pts = plan.points
    log("batch: %d requests, %.3g points (%.2f TB written); tuned %s" % (len(m), pts, 8 * pts / 1e12,
                                                                         {k: tune[k] for k in tune if k != "table"}))
    ms_b, wall = timed(lambda: plan.run(ring, graph=True), args.steps, args.warmup)
    if args.batch_only:
        del ring
        torch.cuda.empty_cache()
        gbs = 8 * pts / ms_b / 1e6
        return {"points": pts, "ms": ms_b, "wall_s": wall, "GBs": gbs, "per_call_ms": float("nan"),
                "detail": {"autotune": {k: tune[k] for k in tune if k != "table"}, "batched_ms": round(ms_b, 2)}}
    ms_p, _ = timed(lambda: plan.run_per_call(ring), 1, 0)
    dg_b = plan.digests(ring)
    dg_p = plan.digests(ring, per_call=True)
    per_method = {}
    names = ("lattice", "dnb2_lms_shift", "dnb2_nus_hash", "gdn_lms_perm")
    for k, name in enumerate(names):
        sel = meth == k
        per_method[name] = int(np.sum((np.int64(1) << m[sel].astype(np.int64)) * d[sel] * r[sel]))
    # each method's requests alone (m capped at 16 to bound the time), tuned shapes
    by_method = {}
    mq = np.minimum(m, 16).astype(np.int32)
    for k, name in enumerate(names):
        sel = meth == k
        sub = B.Batch(meth[sel], mq[sel], d[sel], r[sel], seeds[sel], points_per_tile=tune["points_per_tile"],
                      grid=tune["grid"], ring_elems=ring_elems, gdn_points_per_tile=tune["gdn_points_per_tile"])
        ms_k, _ = timed(lambda: sub.run(ring, graph=True), 2, 1)
        by_method[name] = {"points": sub.points, "ms": round(ms_k, 3), "Gpts_s": round(sub.points / ms_k / 1e6, 1),
                           "frac_hbm": round(8 * sub.points / ms_k / 1e6 / HBM_PEAK_GBS, 4)}
        del sub
    # the batch's dnb2 LMS+shift requests at their full sizes, as a batch of their
    # own: batched LMS+shift digital-net generation against one launch per request
    sel = meth == 1
    lp = B.Batch(meth[sel], m[sel], d[sel], r[sel], seeds[sel], points_per_tile=tune["points_per_tile"],
                 grid=tune["grid"], ring_elems=ring_elems)
    ms_l, _ = timed(lambda: lp.run(ring, graph=True), max(args.steps, 2), 1)
    ms_lp, _ = timed(lambda: lp.run_per_call(ring), 1, 0)
    dl_b, dl_p = lp.digests(ring), lp.digests(ring, per_call=True)
    lms_shift = {"requests": int(sel.sum()), "points": lp.points,
                 "m": [int(m[sel].min()), int(m[sel].max())], "batched_ms": round(ms_l, 3),
                 "batched_Gpts_s": round(lp.points / ms_l / 1e6, 1),
                 "batched_frac_hbm": round(8 * lp.points / ms_l / 1e6 / HBM_PEAK_GBS, 4),
                 "per_call_ms": round(ms_lp, 3), "speedup_vs_per_call": round(ms_lp / ms_l, 3),
                 "bitexact_vs_per_call": bool(np.array_equal(dl_b, dl_p)),
                 "form": "the dnb2 LMS+shift requests of the 512-request batch (full m), batched in one "
                         "persistent graph launch; per call = one launch of the same fused kernel per request"}
    del lp
    del ring
    torch.cuda.empty_cache()
    gbs = 8 * pts / ms_b / 1e6
    return {"points": pts, "ms": ms_b, "wall_s": wall, "GBs": gbs, "per_call_ms": ms_p,
            "detail": {"workload": "512 ragged requests, m in [%d,%d], d in [1,256], r in [1,64], counter-based "
                                   "inputs" % (int(m.min()), int(m.max())),
                       "points_by_method": per_method, "tiles": plan.total_tiles,
                       "bitexact_vs_per_call": bool(np.array_equal(dg_b, dg_p)), "slabs_checked": int(dg_b.size),
                       "autotune": {k: tune[k] for k in tune if k != "table"}, "autotune_table": tune["table"],
                       "autotune_s": round(tune_s, 1), "plan_setup_s": round(plan_s, 2),
                       "batched_ms": round(ms_b, 2), "batched_Gpts_s": round(pts / ms_b / 1e6, 2),
                       "batched_frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                       "per_call_ms": round(ms_p, 2), "per_call_Gpts_s": round(pts / ms_p / 1e6, 2),
                       "speedup_vs_per_call": round(ms_p / ms_b, 3),
                       "by_method_m_le_16": by_method, "lms_shift": lms_shift}}


def section_lattice(torch, q, dev, timed, steps, warmup):
    """Config 1: randomized rank-1 lattice, n=2^20, d=64, r=16, natural order:
    the fused generate+shift kernel vs lat_gen_natural + lat_shift_mod_1."""
    from qmctoolscl_amd import fused
    n_log2, D, R = 20, 64, 16
    n = 1 << n_log2
    rng = np.random.Generator(np.random.SFC64(0))
    g = (2 * rng.integers(0, 2**19, size=(1, D), dtype=np.uint64) + 1)
    shifts = q.lat_get_shifts(rng, R, D)
    g_t, s_t = t64(torch, g, dev), t64(torch, shifts, dev)
    xr = torch.empty((R, n, D), dtype=torch.float64, device=dev)
    x1 = torch.empty((1, n, D), dtype=torch.float64, device=dev)
    pts = R * n * D
    U = np.uint64
    ms_f, _ = timed(lambda: fused.lat_gen_shift(1, R, n, D, 1, 0, g_t, s_t, xr), steps, warmup)
    ms_c, _ = timed(lambda: (q.lat_gen_natural(U(1), U(n), U(D), U(0), g_t, x1, wait=False),
                             q.lat_shift_mod_1(U(R), U(n), U(D), U(1), x1, s_t, xr, wait=False)),
                    max(3, steps // 4), 1)
    del xr, x1
    torch.cuda.empty_cache()
    ach = 8.0 * pts / (ms_f * 1e-3) / 1e9
    return {"workload": "randomized rank-1 lattice: natural order, n=2^20, d=64, r=16 shifts mod 1 (fused "
                        "generate+shift, float64 r x n x d)", "points": pts,
            "fused_ms": round(ms_f, 4), "fused_Gpts_s": round(pts / ms_f / 1e6, 2),
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": committed_traffic(LATTICE_TRAFFIC),
                         "traffic_source": "committed rocprofv3 summary (profiles/), not this run"},
            "dropin_chain_ms": round(ms_c, 4), "dropin_chain_Gpts_s": round(pts / ms_c / 1e6, 2),
            "dropin_chain": "lat_gen_natural + lat_shift_mod_1"}


# ------------------------------------------------------------ CPU baselines

def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class RefC:
    """The reference C backend (oracle/_ref, built from the reference's own
    sources) called as its Python wrapper calls it, batch sizes equal to the
    sizes (qmctoolscl/util.py:107); or this project's C restatement of it
    (oracle/liboracle.so) where the reference build is absent."""

    def __init__(self):
        from qmctoolscl_amd._api import SIGNATURES
        self.sig = SIGNATURES
        ref = os.path.join(HERE, "oracle", "_ref", "libqmcref.so")
        port = os.path.join(HERE, "oracle", "liboracle.so")
        if os.path.exists(ref):
            self.L, self.kind, self.prefix = ctypes.CDLL(ref), "reference", ""
        elif os.path.exists(port):
            self.L, self.kind, self.prefix = ctypes.CDLL(port), "port", "oracle_"
        else:
            self.L = None

    def __call__(self, name, *args):
        spec = self.sig[name]
        argv = []
        for s, a in zip(spec, args):
            argv.append(ctypes.c_uint64(int(a)) if s[1] == "u64" else ctypes.c_void_p(a.ctypes.data))
        if self.prefix == "":
            argv = argv[:3] + [ctypes.c_uint64(int(a)) for a in args[:3]] + argv[3:]
        fn = getattr(self.L, self.prefix + name)
        fn.restype = None
        fn(*argv)


def _timeit(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def cpu_batch(C, quick):
    """Config 3 on the CPU: the first requests of each method with a reference
    chain (lattice, dnb2 LMS + shift, gdn LMS + permutation), the reference's
    calls per request on the request's exported inputs, m capped at 14 and r
    at 2 (gdn: 11 and 1); the rate of each method weighted by the full batch's
    points (the hash-NUS method, which has no reference counterpart, at the
    dnb2 rate)."""
    from qmctoolscl_amd import batch as B
    meth, m, d, r, seeds = batch_requests(quick)
    rates, spent, sample_pts = {}, 0.0, 0
    for k in (0, 1, 3):
        pts, sec = 0, 0.0
        for q in np.flatnonzero(meth == k)[:6]:
            mm, dd, rr, sd = min(int(m[q]), 14 if k != 3 else 11), int(d[q]), min(int(r[q]), 2 if k != 3 else 1), \
                int(seeds[q])
            n = 1 << mm
            if k == 0:
                g, sh = B.lattice_inputs(sd, mm, dd, rr)
                x = np.empty((1, n, dd))
                xr = np.empty((rr, n, dd))
                sec += _timeit(lambda: (C("lat_gen_natural", 1, n, dd, 0, g, x),
                                        C("lat_shift_mod_1", rr, n, dd, 1, x, sh, xr)))
            elif k == 1:
                Cm, S, sh = B.dnb2_inputs(sd, 1, dd, rr)
                Cl = np.empty((rr, dd, 32), dtype=np.uint64)
                xb = np.empty((rr, n, dd), dtype=np.uint64)
                xrb = np.empty_like(xb)
                x = np.empty((rr, n, dd))
                ls, tm = np.full(rr, 32, np.uint64), np.full(rr, 64, np.uint64)
                sec += _timeit(lambda: (C("dnb2_linear_matrix_scramble", rr, dd, 32, 1, 32, S, Cm, Cl),
                                        C("dnb2_gen_natural", rr, n, dd, 0, 32, Cl, xb),
                                        C("dnb2_digital_shift", rr, n, dd, rr, ls, xb, sh, xrb),
                                        C("dnb2_integer_to_float", rr, n, dd, tm, xrb, x)))
            else:
                bases, S, perms = B.gdn_inputs(sd, mm, rr, j0=0, dj=dd)
                bmax = int(bases.max())
                Cm = np.ascontiguousarray(np.tile(np.eye(mm, dtype=np.uint64)[None, None], (1, dd, 1, 1)))
                Cl = np.empty((rr, dd, mm, 52), dtype=np.uint64)
                xd = np.empty((rr, n, dd, 52), dtype=np.uint64)
                xp = np.empty_like(xd)
                x = np.empty((rr, n, dd))
                sec += _timeit(lambda: (C("gdn_linear_matrix_scramble", rr, dd, mm, 1, 1, mm, 52, bases, S, Cm, Cl),
                                        C("gdn_gen_natural", rr, n, dd, 1, mm, 52, 0, bases, Cl, xd),
                                        C("gdn_digital_permutation", rr, n, dd, rr, 1, 52, 52, bmax, perms, xd, xp),
                                        C("gdn_integer_to_float", rr, n, dd, 1, 52, bases, xp, x)))
            pts += rr * n * dd
        rates[k] = pts / sec
        spent += sec
        sample_pts += pts
    full = {k: float(np.sum((np.int64(1) << m[meth == k].astype(np.int64)) * d[meth == k] * r[meth == k]))
            for k in range(4)}
    rates[2] = rates[1]
    t_full = sum(full[k] / rates[k] for k in range(4))
    return {"value": round(sum(full.values()) / t_full / 1e9, 5), "unit": "Gpts/s",
            "sample": "first 6 requests of the lattice, dnb2 LMS+shift and gdn LMS+permutation methods, m capped at "
                      "14, r at 2 (gdn 11, 1): the reference's call chain per request on the exported inputs (%.3g "
                      "points in "
                      "%.1f s); per-method rates weighted by the batch's points (hash NUS at the dnb2 rate: no "
                      "reference counterpart)" % (sample_pts, spent),
            "per_method_Mpts_s": {n_: round(rates[k] / 1e6, 3) for k, n_ in
                                  ((0, "lattice"), (1, "dnb2_lms_shift"), (3, "gdn_lms_perm"))},
            "projected_full_batch_s": round(t_full, 1)}


def cpu_legs(C):
    """Configs 0, 1, 2, 4 on one core, bounded samples, scale stated."""
    from qmctoolscl_amd import rand_funcs as rf
    legs = {}
    U = np.uint64
    # config 0: dnb2, n=2^16, d=8, r=4, full size
    rng = np.random.Generator(np.random.SFC64(0))
    n, d, r, t = 1 << 16, 8, 4, 32
    Cm = np.empty((1, d, t), dtype=np.uint64)
    for j in range(d):
        for k in range(t):
            above = rng.integers(0, 2, size=k, dtype=np.uint64)
            Cm[0, j, k] = (1 << (t - 1 - k)) + sum(int(b) << (t - 1 - c) for c, b in enumerate(above))
    S = rf.dnb2_get_linear_scramble_matrix(rng, r, d, t, t, False)
    sh = rf.dnb2_get_digital_shifts(rng, r, d, 64)
    Cl = np.empty((r, d, t), dtype=np.uint64)
    xb = np.empty((r, n, d), dtype=np.uint64)
    xrb = np.empty_like(xb)
    x = np.empty((r, n, d))
    sec = _timeit(lambda: (C("dnb2_linear_matrix_scramble", r, d, t, 1, t, S, Cm, Cl),
                           C("dnb2_gen_natural", r, n, d, 0, t, Cl, xb),
                           C("dnb2_digital_shift", r, n, d, r, np.full(r, 32, U), xb, sh, xrb),
                           C("dnb2_integer_to_float", r, n, d, np.full(r, 64, U), xrb, x)))
    legs["config0_dnb2"] = {"value": round(r * n * d / sec / 1e9, 5), "unit": "Gpts/s",
                            "sample": "full size (n=2^16, d=8, r=4): LMS, natural order, 64-bit shift, float in %.2f s"
                                      % sec}
    # config 1: lattice, first 2^15 of 2^20 rows, d=64, r=16
    rows, D, R = 1 << 15, 64, 16
    g = (2 * rng.integers(0, 2**19, size=(1, D), dtype=np.uint64) + 1)
    shifts = rf.lat_get_shifts(rng, R, D)
    x1 = np.empty((1, rows, D))
    xr = np.empty((R, rows, D))
    sec = _timeit(lambda: (C("lat_gen_natural", 1, rows, D, 0, g, x1), C("lat_shift_mod_1", R, rows, D, 1, x1, shifts,
                                                                         xr)))
    legs["config1_lattice"] = {"value": round(R * rows * D / sec / 1e9, 5), "unit": "Gpts/s",
                               "sample": "lat_gen_natural(r=1) + lat_shift_mod_1(r=16), first 2^15 of the 2^20 rows "
                                         "(1/32 of the points), d=64, in %.2f s" % sec}
    # config 2: gdn chain on the first 3^7 of 3^12 points, r=1 of 8
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101,
              103, 107, 109, 113, 127, 131]
    n2, d2, mmax, tnew = 3 ** 7, 32, 20, 52
    bases = np.array([primes], dtype=np.uint64)
    Cg = rf.gdn_get_halton_generating_matrix(1, d2, mmax)
    S2 = rf.gdn_get_linear_scramble_matrix(rng, 1, d2, mmax, tnew, 1, bases)
    perms = rf.gdn_get_digital_permutations(rng, 1, d2, tnew, 1, bases)
    Cl2 = np.empty((1, d2, mmax, tnew), dtype=np.uint64)
    xd = np.empty((1, n2, d2, tnew), dtype=np.uint64)
    xp = np.empty_like(xd)
    x2 = np.empty((1, n2, d2))
    sec = _timeit(lambda: (C("gdn_linear_matrix_scramble", 1, d2, mmax, 1, 1, mmax, tnew, bases, S2, Cg, Cl2),
                           C("gdn_gen_natural", 1, n2, d2, 1, mmax, tnew, 0, bases, Cl2, xd),
                           C("gdn_digital_permutation", 1, n2, d2, 1, 1, tnew, tnew, 131, perms, xd, xp),
                           C("gdn_integer_to_float", 1, n2, d2, 1, tnew, bases, xp, x2)))
    legs["config2_gdn"] = {"value": round(n2 * d2 / sec / 1e9, 6), "unit": "Gpts/s",
                           "sample": "LMS -> gen_natural -> digital_permutation -> integer_to_float, first 3^7 of the "
                                     "3^12 points, r=1 of 8 (1/648 of the points), d=32, 52 digits, in %.2f s" % sec}
    # config 4: transforms, one vector each at 2^20, 256 vectors at 2^12
    for m, V in ((20, 1), (12, 256)):
        nn = 1 << m
        xw = rng.standard_normal((V, nn))
        sec = _timeit(lambda: C("fwht_1d_radix2", 1, V, nn // 2, xw))
        yr, yi = rng.standard_normal((V, nn)), rng.standard_normal((V, nn))
        twr, twi = np.empty(nn), np.empty(nn)
        sec_f = _timeit(lambda: C("fft_bro_1d_radix2", 1, V, nn // 2, twr, twi, yr, yi))
        legs["config4_2^%d" % m] = {"fwht_GBs": round(16 * V * nn / sec / 1e9, 3),
                                    "fft_GBs": round(32 * V * nn / sec_f / 1e9, 3),
                                    "sample": "%d vector(s) of 2^%d (config 4 batches 64 at 2^20, 4096 at 2^12): fwht "
                                              "%.3f s, fft %.3f s" % (V, m, sec, sec_f)}
    return legs


def cpu_baseline(quick):
    C = RefC()
    if C.L is None:
        return None
    out = cpu_batch(C, quick)
    out.update({"cores": 1, "kind": C.kind, "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
                "opencl_cpu": "not measured: no pyopencl / CPU OpenCL runtime in this image"})
    out["legs"] = cpu_legs(C)
    return out


# ------------------------------------------------------------ sections

def section_batch_vs_chain(torch, q, dev, timed):
    """A bounded ragged sample drawn like the batch (lattice, dnb2 LMS+shift,
    gdn LMS+permutation requests) three ways: the persistent batch, one fused
    launch per request, and the reference's call sequence per request through
    the drop-in kernels on the requests' own inputs (lat_gen_natural ->
    lat_shift_mod_1; dnb2_linear_matrix_scramble -> dnb2_gen_natural ->
    dnb2_digital_shift -> dnb2_integer_to_float; gdn_linear_matrix_scramble ->
    gdn_gen_natural -> gdn_digital_permutation -> gdn_integer_to_float on
    52-digit arrays), all on device tensors."""
    from qmctoolscl_amd import batch as B
    U = np.uint64
    meth, m, d, r, seeds = B.sample_requests(48, 7, m_range=(8, 14), d_range=(1, 64), r_range=(1, 8),
                                             methods=(0, 1, 3))
    plan = B.Batch(meth, m, d, r, seeds, points_per_tile=1 << 15)
    ring = torch.empty(plan.ring_elems, dtype=torch.float64, device=dev)
    pts = plan.points
    ms_b, _ = timed(lambda: plan.run(ring, graph=True), 5, 2)
    ms_f, _ = timed(lambda: plan.run_per_call(ring), 5, 2)
    del ring
    # inputs per request, device-resident before timing; scratch shared across requests
    n_max = int((1 << m.astype(np.int64)).max())
    big = int(max((1 << int(mm)) * int(dd) * int(rr) for mm, dd, rr in zip(m, d, r)))
    xdig = torch.empty(big * 52, dtype=torch.int64, device=dev)
    xdig2 = torch.empty(big * 52, dtype=torch.int64, device=dev)
    xb = torch.empty(big, dtype=torch.int64, device=dev)
    xrb = torch.empty(big, dtype=torch.int64, device=dev)
    out = torch.empty(big, dtype=torch.float64, device=dev)
    x1 = torch.empty(n_max * 64, dtype=torch.float64, device=dev)
    calls = []
    for i in range(len(m)):
        n, dd, rr, mm = 1 << int(m[i]), int(d[i]), int(r[i]), int(m[i])
        if meth[i] == 0:
            g, sh = B.lattice_inputs(seeds[i], mm, dd, rr)
            g_t, sh_t = t64(torch, g.reshape(1, dd), dev), t64(torch, sh, dev)
            calls.append(lambda n=n, dd=dd, rr=rr, g_t=g_t, sh_t=sh_t: (
                q.lat_gen_natural(U(1), U(n), U(dd), U(0), g_t, x1, wait=False),
                q.lat_shift_mod_1(U(rr), U(n), U(dd), U(1), x1, sh_t, out, wait=False)))
        elif meth[i] == 1:
            C, S, sh = B.dnb2_inputs(seeds[i], 1, dd, rr)
            C_t, S_t, sh_t = t64(torch, C, dev), t64(torch, S, dev), t64(torch, sh, dev)
            Cl = torch.empty((rr, dd, 32), dtype=torch.int64, device=dev)
            ls = t64(torch, np.full(rr, 32, np.uint64), dev)
            tm = t64(torch, np.full(rr, 64, np.uint64), dev)
            calls.append(lambda n=n, dd=dd, rr=rr, C_t=C_t, S_t=S_t, sh_t=sh_t, Cl=Cl, ls=ls, tm=tm: (
                q.dnb2_linear_matrix_scramble(U(rr), U(dd), U(32), U(1), U(32), S_t, C_t, Cl, wait=False),
                q.dnb2_gen_natural(U(rr), U(n), U(dd), U(0), U(32), Cl, xb, wait=False),
                q.dnb2_digital_shift(U(rr), U(n), U(dd), U(rr), ls, xb, sh_t, xrb, wait=False),
                q.dnb2_integer_to_float(U(rr), U(n), U(dd), tm, xrb, out, wait=False)))
        else:
            # the request's own inputs: identity C (m x m), LMS to 52 digits, permutations
            bases, S, perms = B.gdn_inputs(seeds[i], mm, rr, j0=0, dj=dd)
            bmax = int(bases.max())
            Cg = np.ascontiguousarray(np.tile(np.eye(mm, dtype=np.uint64)[None, None], (1, dd, 1, 1)))
            b_t, S_t, C_t, p_t = (t64(torch, a, dev) for a in (bases, S, Cg, perms))
            Cl = torch.empty((rr, dd, mm, 52), dtype=torch.int64, device=dev)
            calls.append(lambda n=n, dd=dd, rr=rr, mm=mm, bmax=bmax, b_t=b_t, S_t=S_t, C_t=C_t, p_t=p_t, Cl=Cl: (
                q.gdn_linear_matrix_scramble(U(rr), U(dd), U(mm), U(1), U(1), U(mm), U(52), b_t, S_t, C_t, Cl,
                                             wait=False),
                q.gdn_gen_natural(U(rr), U(n), U(dd), U(1), U(mm), U(52), U(0), b_t, Cl, xdig, wait=False),
                q.gdn_digital_permutation(U(rr), U(n), U(dd), U(rr), U(1), U(52), U(52), U(bmax), p_t, xdig,
                                          xdig2, wait=False),
                q.gdn_integer_to_float(U(rr), U(n), U(dd), U(1), U(52), b_t, xdig2, out, wait=False)))

    def chain():
        for f in calls:
            f()

    ms_c, _ = timed(chain, 3, 1)
    # the dnb2 LMS+shift requests alone, both ways
    sel = meth == 1
    lplan = B.Batch(meth[sel], m[sel], d[sel], r[sel], seeds[sel], points_per_tile=1 << 15)
    lring = torch.empty(lplan.ring_elems, dtype=torch.float64, device=dev)
    ms_lb, _ = timed(lambda: lplan.run(lring, graph=True), 5, 2)
    del lring
    lcalls = [f for f, k in zip(calls, meth) if k == 1]

    def lchain():
        for f in lcalls:
            f()

    ms_lc, _ = timed(lchain, 3, 1)
    by = {}
    for k, name in enumerate(("lattice", "dnb2_lms_shift", "gdn_lms_perm")):
        sel = meth == (0, 1, 3)[k]
        by[name] = int(np.sum((np.int64(1) << m[sel].astype(np.int64)) * d[sel] * r[sel]))
    return {"workload": "48 ragged requests (lattice, dnb2 LMS+shift, gdn LMS+permutation), m in [8,14], "
                        "d in [1,64], r in [1,8]", "points": pts, "points_by_method": by,
            "batched_ms": round(ms_b, 4), "batched_Gpts_s": round(pts / ms_b / 1e6, 2),
            "fused_per_request_ms": round(ms_f, 4), "fused_per_request_Gpts_s": round(pts / ms_f / 1e6, 2),
            "dropin_chain_per_request_ms": round(ms_c, 4),
            "dropin_chain_per_request_Gpts_s": round(pts / ms_c / 1e6, 2),
            "speedup_vs_fused_per_request": round(ms_f / ms_b, 3),
            "speedup_vs_dropin_chain": round(ms_c / ms_b, 3),
            "lms_shift_only": {"requests": int(sel.sum()), "points": lplan.points,
                               "batched_ms": round(ms_lb, 4), "dropin_chain_ms": round(ms_lc, 4),
                               "speedup_vs_dropin_chain": round(ms_lc / ms_lb, 3)}}


def section_dnb2(torch, q, fused, timed, dev):
    rng = np.random.Generator(np.random.SFC64(0))
    n, d, r, t = 1 << 16, 8, 4, 32
    # upper-triangular unit-diagonal generating matrices, MSB-first columns
    C = np.empty((1, d, t), dtype=np.uint64)
    for j in range(d):
        for k in range(t):
            above = rng.integers(0, 2, size=k, dtype=np.uint64)
            C[0, j, k] = (1 << (t - 1 - k)) + sum(int(b) << (t - 1 - c) for c, b in enumerate(above))
    S = q.dnb2_get_linear_scramble_matrix(rng, r, d, t, t, False)
    shiftsb = q.dnb2_get_digital_shifts(rng, r, d, 64)
    Cd, Sd, shd = t64(torch, C, dev), t64(torch, S, dev), t64(torch, shiftsb, dev)
    lsh = t64(torch, np.full(r, 32, np.uint64), dev)
    tm = t64(torch, np.full(r, 64, np.uint64), dev)
    C_lms = torch.empty((r, d, t), dtype=torch.int64, device=dev)
    xb = torch.empty((r, n, d), dtype=torch.int64, device=dev)
    xrb = torch.empty((r, n, d), dtype=torch.int64, device=dev)
    x = torch.empty((r, n, d), dtype=torch.float64, device=dev)
    U = np.uint64

    def fused_step():
        q.dnb2_linear_matrix_scramble(U(r), U(d), U(t), U(1), U(t), Sd, Cd, C_lms, wait=False)
        fused.dnb2_gen_shift_float(1, r, n, d, r, 0, t, C_lms, lsh, shd, tm, xrb, x)

    def chain_step():
        q.dnb2_linear_matrix_scramble(U(r), U(d), U(t), U(1), U(t), Sd, Cd, C_lms, wait=False)
        q.dnb2_gen_natural(U(r), U(n), U(d), U(0), U(t), C_lms, xb, wait=False)
        q.dnb2_digital_shift(U(r), U(n), U(d), U(r), lsh, xb, shd, xrb, wait=False)
        q.dnb2_integer_to_float(U(r), U(n), U(d), tm, xrb, x, wait=False)

    def one_launch_step():
        fused.dnb2_lms_gen_shift_float(1, r, n, d, 1, 0, t, t, Sd, Cd, lsh, shd, tm, xrb, x)

    ms_1, _ = timed(one_launch_step, 200, 10)
    ms_f, _ = timed(fused_step, 200, 10)
    ms_c, _ = timed(chain_step, 200, 10)
    pts = r * n * d
    out = {"workload": "dnb2 LMS + natural order + 64-bit digital shift + float, n=2^16, d=8, r=4, tmax=32 "
                       "(writes the shifted integers and their doubles, 32 MiB)",
           "points": pts, "one_launch_ms": round(ms_1, 4), "one_launch_Gpts_s": round(pts / ms_1 / 1e6, 2),
           "one_launch_GBs_written": round(16 * pts / ms_1 / 1e6, 1),
           "fused_ms": round(ms_f, 4), "fused_Gpts_s": round(pts / ms_f / 1e6, 2),
           "fused_GBs_written": round(16 * pts / ms_f / 1e6, 1), "fused": "LMS launch + fused generation launch",
           "per_call_ms": round(ms_c, 4), "per_call_Gpts_s": round(pts / ms_c / 1e6, 2)}
    # nested uniform scrambling of the same unscrambled net (host trees with numpy draws in the
    # reference's order, GPU walk), 32 -> 64 bits; a repeat call on the grown trees draws nothing
    from qmctoolscl_amd import nus
    q.dnb2_gen_natural(U(1), U(n), U(d), U(0), U(t), Cd, xb[:1], wait=False)
    torch.cuda.synchronize()
    seqs = np.random.SeedSequence(7).spawn(r * d)
    rngs = np.empty((r, d), dtype=object)
    for k in range(r * d):
        rngs[k // d, k % d] = np.random.Generator(np.random.SFC64(seqs[k]))
    roots = np.array([[nus.NUSNode_dnb2() for _ in range(d)] for _ in range(r)], dtype=object)
    xb1 = xb[:1].contiguous()
    t0 = time.perf_counter()
    nus.dnb2_nested_uniform_scramble(r, n, d, 1, t, 64, rngs, roots, xb1, xrb)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    nus.dnb2_nested_uniform_scramble(r, n, d, 1, t, 64, rngs, roots, xb1, xrb)
    torch.cuda.synchronize()
    t_again = time.perf_counter() - t0
    out["nus"] = {"workload": "dnb2 nested uniform scramble, n=2^16, d=8, r=4, 32 -> 64 bits",
                  "points": pts, "first_call_s": round(t_first, 3), "repeat_call_s": round(t_again, 4),
                  "first_call_Mpts_s": round(pts / t_first / 1e6, 3)}
    return out


def section_gdn(torch, q, dev, timed, r=8):
    """Halton-type generalized net: bases = first 32 primes, n = 3^12, d = 32,
    Halton identity C (mmax = 20), LMS to tmax_new = 52 digits, per-level digit
    permutations, then float.  Digits are materialized (r*n*d*52 uint64 per copy)."""
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71,
              73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131]
    n, d, mmax, tnew = 3 ** 12, 32, 20, 52
    rng = np.random.Generator(np.random.SFC64(2))
    bases = np.array([primes], dtype=np.uint64)
    C = q.gdn_get_halton_generating_matrix(1, d, mmax)
    S = q.gdn_get_linear_scramble_matrix(rng, r, d, mmax, tnew, 1, bases)
    perms = q.gdn_get_digital_permutations(rng, r, d, tnew, 1, bases)
    bmax = int(bases.max())
    Bd, Cd, Sd, Pd = (t64(torch, a, dev) for a in (bases, C, S, perms))
    C_lms = torch.empty((r, d, mmax, tnew), dtype=torch.int64, device=dev)
    xdig = torch.empty((r, n, d, tnew), dtype=torch.int64, device=dev)
    xperm = torch.empty_like(xdig)
    x = torch.empty((r, n, d), dtype=torch.float64, device=dev)
    U = np.uint64
    q.gdn_linear_matrix_scramble(U(r), U(d), U(mmax), U(1), U(1), U(mmax), U(tnew), Bd, Sd, Cd, C_lms)
    ms_gen, _ = timed(lambda: q.gdn_gen_natural(U(r), U(n), U(d), U(1), U(mmax), U(tnew), U(0), Bd, C_lms, xdig,
                                                wait=False), 3, 1)
    ms_perm, _ = timed(lambda: q.gdn_digital_permutation(U(r), U(n), U(d), U(r), U(1), U(tnew), U(tnew), U(bmax),
                                                         Pd, xdig, xperm, wait=False), 3, 1)
    ms_flt, _ = timed(lambda: q.gdn_integer_to_float(U(r), U(n), U(d), U(1), U(tnew), Bd, xperm, x, wait=False), 3, 1)
    pts = r * n * d
    dig_bytes = 8 * pts * tnew
    # the digital shift of the same digits (the other gdn scramble)
    shifts = q.gdn_get_digital_shifts(rng, r, d, tnew, 1, bases)
    Shd = t64(torch, shifts, dev)
    ms_shift, _ = timed(lambda: q.gdn_digital_shift(U(r), U(n), U(d), U(r), U(1), U(tnew), U(tnew), Bd, Shd, xdig,
                                                    xperm, wait=False), 3, 1)
    xfs = torch.empty_like(x)
    # the same chain fused: digits walk in registers, only the doubles are written
    from qmctoolscl_amd import fused
    xf = torch.empty_like(x)
    ms_fused, _ = timed(lambda: fused.gdn_gen_scramble_float(1, r, n, d, 1, mmax, tnew, tnew, 0, bmax, Bd, C_lms, Pd,
                                                             xf), 3, 1)
    fused_exact = bool(torch.equal(xf.view(torch.int64), x.view(torch.int64)))
    ms_fused_shift, _ = timed(lambda: fused.gdn_gen_scramble_float(2, r, n, d, 1, mmax, tnew, tnew, 0, bmax, Bd, C_lms,
                                                                   Shd, xfs), 3, 1)
    del xf, xfs, Shd
    out = {"workload": "Halton-type net, bases = first 32 primes, n=3^12, d=32, r=%d, LMS to 52 digits, "
                       "digit permutations, float" % r,
           "points": pts, "digit_bytes_per_copy": dig_bytes,
           "gen_ms": round(ms_gen, 3), "gen_GBs_written": round(dig_bytes / ms_gen / 1e6, 1),
           "perm_ms": round(ms_perm, 3), "perm_GBs": round(2 * dig_bytes / ms_perm / 1e6, 1),
           "to_float_ms": round(ms_flt, 3), "to_float_GBs": round((dig_bytes + 8 * pts) / ms_flt / 1e6, 1),
           "chain_Gpts_s": round(pts / (ms_gen + ms_perm + ms_flt) / 1e6, 2),
           "shift_ms": round(ms_shift, 3), "shift_GBs": round(2 * dig_bytes / ms_shift / 1e6, 1),
           "fused_shift_ms": round(ms_fused_shift, 3),
           "fused_ms": round(ms_fused, 3), "fused_Gpts_s": round(pts / ms_fused / 1e6, 2),
           "fused_GBs_written": round(8 * pts / ms_fused / 1e6, 1),
           "fused_speedup_vs_chain": round((ms_gen + ms_perm + ms_flt) / ms_fused, 1),
           "fused_bit_exact_vs_chain": fused_exact}
    del xperm, x, C_lms
    torch.cuda.empty_cache()
    # nested uniform scrambling of the same net at full size for one of its r = 8
    # replications (the 32 trees of one replication; the other seven are
    # independent trees of the same sizes, so r = 8 costs 8x the time and the
    # memory), 20 -> 52 digits.  The reference's leaves would hold
    # tmax_new - t - 1 permutations of b values each (rand_funcs.py:328-351);
    # here the host replays each tree's SFC64 stream on its own thread and keeps
    # one generator state per leaf, and the GPU regenerates the leaves' rows.
    from qmctoolscl_amd import nus
    n8, tmax, r = n, mmax, 1
    xd = xdig[:1, :n8, :, :tmax].contiguous()  # the unscrambled Halton digits
    q.gdn_gen_natural(U(1), U(n8), U(d), U(1), U(mmax), U(tmax), U(0), Bd, Cd, xd)
    xr_nus = torch.empty((r, n8, d, tnew), dtype=torch.int64, device=dev)
    seqs = np.random.SeedSequence(11).spawn(r * d)
    rngs = np.empty((r, d), dtype=object)
    for k in range(r * d):
        rngs[k // d, k % d] = np.random.Generator(np.random.SFC64(seqs[k]))
    roots = np.array([[nus.NUSNode_gdn() for _ in range(d)] for _ in range(r)], dtype=object)
    t0 = time.perf_counter()
    nus.gdn_nested_uniform_scramble(r, n8, d, 1, 1, tmax, tnew, rngs, roots, Bd, xd, xr_nus)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    nus.gdn_nested_uniform_scramble(r, n8, d, 1, 1, tmax, tnew, rngs, roots, Bd, xd, xr_nus)
    torch.cuda.synchronize()
    t_again = time.perf_counter() - t0
    L = nus._lib_nus()
    hs = [roots[l, j]._tree.handle for l in range(r) for j in range(d)]
    tree_bytes = sum(int(L.qmc_nusb_tree_bytes(h)) for h in hs)
    leaves = sum(int(L.qmc_nusb_tree_states(h)) for h in hs)
    ref_rows = sum((tnew - 1) * int(bases[0, j]) for j in range(d)) * n8  # upper bound: one block per point
    out["nus"] = {"workload": "gdn nested uniform scramble, n=3^12, d=32 (first 32 primes), r=1 of the config's 8, "
                              "20 -> 52 digits, SFC64 generators",
                  "points": r * n8 * d, "first_call_s": round(t_first, 2), "repeat_call_s": round(t_again, 3),
                  "first_call_Mpts_s": round(r * n8 * d / t_first / 1e6, 3),
                  "host_threads": nus._host_threads(), "tree_host_GB": round(tree_bytes / 1e9, 2),
                  "leaf_states": leaves,
                  "projected_r8_first_call_s": round(8 * t_first, 1),
                  "projected_r8_tree_host_GB": round(8 * tree_bytes / 1e9, 1),
                  "reference_leaf_rows_GB_upper_bound_r8": round(8 * ref_rows * 8 / 1e9, 0)}
    del xdig, xd, xr_nus
    torch.cuda.empty_cache()
    # base-2 interlacing of a d = 16 net at alpha = 2 into 8 columns of 64 digits
    Cn = rng.integers(0, 2 ** 32, size=(r, 16, 32), dtype=np.uint64)
    Cn_t = t64(torch, Cn, dev)
    Ca = torch.empty((r, 8, 32), dtype=torch.int64, device=dev)
    ms_il, _ = timed(lambda: q.dnb2_interlace(U(r), U(8), U(32), U(16), U(32), U(64), U(2), Cn_t, Ca, wait=False),
                     20, 3)
    out["dnb2_interlace_d16_alpha2_ms"] = round(ms_il, 4)
    return out


def section_transforms(torch, q, dev, timed0):
    from qmctoolscl_amd.mfma import fft_bro_mfma, fwht_mfma, ifft_bro_mfma, relative_errors

    def timed(fn, steps, warmup):
        # the best of three rounds for every path alike (the FP64 and the matrix-core
        # transforms run back to back here, and single rounds vary by up to 15%)
        return min(timed0(fn, steps, warmup if k == 0 else 1) for k in range(3))
    # cold-cache timing: a 1 GiB buffer (4x the 256 MB Infinity Cache) is rewritten before
    # every call, so the call reads its vectors from HBM.  Back to back, a working set that
    # fits the cache (2^12 x 4096: 134 MB for the FWHT, 268 MB for the FFT's two parts) is
    # partly served from it, which helps the paths with ordinary loads (the FP64 ones)
    flush = torch.empty(1 << 28, dtype=torch.float32, device=dev)

    def cold(fn, steps=20):
        ts = []
        for _ in range(steps):
            flush.add_(1.0)
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        return float(np.median(ts))

    out = []
    U = np.uint64
    for m, V in ((12, 4096), (20, 64), (24, 4)):
        n = 1 << m
        rng = np.random.Generator(np.random.SFC64(m))
        x0 = torch.from_numpy(rng.standard_normal((V, n))).to(dev)
        x = x0.clone()
        ms64, _ = timed(lambda: q.fwht_1d_radix2(U(1), U(V), U(n // 2), x, wait=False), 20, 3)
        # accuracy: one transform of the same input each way
        ref = x0.clone()
        q.fwht_1d_radix2(U(1), U(V), U(n // 2), ref)
        xm = x0.clone()
        fwht_mfma(1, V, n // 2, xm)
        torch.cuda.synchronize()
        l2, mx = relative_errors(xm.cpu().numpy(), ref.cpu().numpy())
        work = torch.empty(V * n, dtype=torch.float32, device=dev)
        msm, _ = timed(lambda: fwht_mfma(1, V, n // 2, x, work), 20, 3)
        xr, xi = x0.clone(), x0.clone()
        twr = torch.empty(n, dtype=torch.float64, device=dev)
        twi = torch.empty(n, dtype=torch.float64, device=dev)
        msf, _ = timed(lambda: q.fft_bro_1d_radix2(U(1), U(V), U(n // 2), twr, twi, xr, xi, wait=False), 20, 3)
        # FFT accuracy and time on the matrix cores
        yr, yi = x0.clone(), x0.flip(-1).contiguous()
        q.fft_bro_1d_radix2(U(1), U(V), U(n // 2), twr, twi, yr, yi)
        zr, zi = x0.clone(), x0.flip(-1).contiguous()
        fft_bro_mfma(1, V, n // 2, zr, zi)
        torch.cuda.synchronize()
        fl2, fmx = relative_errors(np.stack([zr.cpu().numpy(), zi.cpu().numpy()]),
                                   np.stack([yr.cpu().numpy(), yi.cpu().numpy()]))
        del yr, yi, zr, zi
        fwork = torch.empty(2 * V * n, dtype=torch.float32, device=dev)
        msfm, _ = timed(lambda: fft_bro_mfma(1, V, n // 2, xr, xi, fwork), 20, 3)
        # inverse: float64 drop-in and matrix cores, accuracy on the same input
        msi, _ = timed(lambda: q.ifft_bro_1d_radix2(U(1), U(V), U(n // 2), twr, twi, xr, xi, wait=False), 20, 3)
        msim, _ = timed(lambda: ifft_bro_mfma(1, V, n // 2, xr, xi, fwork), 20, 3)
        yr, yi = x0.clone(), x0.flip(-1).contiguous()
        q.ifft_bro_1d_radix2(U(1), U(V), U(n // 2), twr, twi, yr, yi)
        zr, zi = x0.clone(), x0.flip(-1).contiguous()
        ifft_bro_mfma(1, V, n // 2, zr, zi)
        torch.cuda.synchronize()
        il2, imx = relative_errors(np.stack([zr.cpu().numpy(), zi.cpu().numpy()]),
                                   np.stack([yr.cpu().numpy(), yi.cpu().numpy()]))
        del yr, yi, zr, zi
        c64 = cold(lambda: q.fwht_1d_radix2(U(1), U(V), U(n // 2), x, wait=False))
        cm = cold(lambda: fwht_mfma(1, V, n // 2, x, work))
        cf = cold(lambda: q.fft_bro_1d_radix2(U(1), U(V), U(n // 2), twr, twi, xr, xi, wait=False))
        cfm = cold(lambda: fft_bro_mfma(1, V, n // 2, xr, xi, fwork))
        ci = cold(lambda: q.ifft_bro_1d_radix2(U(1), U(V), U(n // 2), twr, twi, xr, xi, wait=False))
        cim = cold(lambda: ifft_bro_mfma(1, V, n // 2, xr, xi, fwork))
        del fwork
        pts = V * n
        # matrix-core work of the FWHT: per radix-16 axis, four 16x16x4 fp32
        # products per 16x16 block = 32 flops per value, on the fp32 MFMA peak
        axes = m // 4
        fwht_mfma_flops = pts * axes * 4 * (2 * 16 * 16 * 4) / 256
        out.append({"n": "2^%d" % m, "vectors": V,
                    "fwht_fp64_ms": round(ms64, 4), "fwht_fp64_GBs": round(16 * pts / ms64 / 1e6, 1),
                    "fwht_mfma_ms": round(msm, 4), "fwht_mfma_GBs": round(16 * pts / msm / 1e6, 1),
                    "mfma_speedup": round(ms64 / msm, 3),
                    "mfma_rel_l2": float("%.3g" % l2), "mfma_rel_max": float("%.3g" % mx),
                    "mfma_util": float("%.3g" % (fwht_mfma_flops / (msm * 1e-3) / (FP32_MFMA_PEAK_TFLOPS * 1e12))),
                    "mfma_util_peak": "fp32 MFMA, %.1f TFLOP/s" % FP32_MFMA_PEAK_TFLOPS,
                    "fft_fp64_ms": round(msf, 4), "fft_fp64_GBs": round(32 * pts / msf / 1e6, 1),
                    "fft_mfma_ms": round(msfm, 4), "fft_mfma_GBs": round(32 * pts / msfm / 1e6, 1),
                    "fft_mfma_speedup": round(msf / msfm, 3),
                    "fft_mfma_rel_l2": float("%.3g" % fl2), "fft_mfma_rel_max": float("%.3g" % fmx),
                    "ifft_fp64_ms": round(msi, 4), "ifft_mfma_ms": round(msim, 4),
                    "ifft_mfma_speedup": round(msi / msim, 3),
                    "ifft_mfma_rel_l2": float("%.3g" % il2), "ifft_mfma_rel_max": float("%.3g" % imx),
                    "cold_cache_ms": {"fwht_fp64": round(c64, 4), "fwht_mfma": round(cm, 4),
                                      "fft_fp64": round(cf, 4), "fft_mfma": round(cfm, 4),
                                      "ifft_fp64": round(ci, 4), "ifft_mfma": round(cim, 4)},
                    "cold_cache_speedup": {"fwht": round(c64 / cm, 3), "fft": round(cf / cfm, 3),
                                           "ifft": round(ci / cim, 3)},
                    # 12 products of 16x16x32 bf16 per 256 complex values per radix-16 level
                    "fft_mfma_util": float("%.3g" % (pts * (m // 4) * 768 / (msfm * 1e-3) /
                                                     (BF16_PEAK_TFLOPS * 1e12))),
                    "fft_mfma_util_peak": "bf16 MFMA, %.0f TFLOP/s" % BF16_PEAK_TFLOPS})
        del x0, x, ref, xm, work, xr, xi
        torch.cuda.empty_cache()
    del flush
    torch.cuda.empty_cache()
    return out


def section_kernel_methods(torch, q, dev, timed):
    """Gram-matrix transforms for fast QMC kernel methods: the first column of
    a smoothness-1 shift-invariant kernel on lattice points (n = 2^20, d = 64,
    4 generating vectors) evaluated from g in registers, then fft_bro; beside
    it the chain that materializes the points (lat_gen_natural, the kernel in
    torch, fft_bro).  Same for the order-1 digital kernel on a base-2 net
    (32-bit, d = 64) with fwht."""
    from qmctoolscl_amd import kernel_methods as km
    U = np.uint64
    r, m, d = 4, 20, 64
    n = 1 << m
    rng = np.random.Generator(np.random.SFC64(5))
    g = 2 * rng.integers(0, 2 ** 19, size=(r, d), dtype=np.uint64) + U(1)
    coef_row = np.array(km.lat_kernel_coef(1.0 / (1.0 + np.arange(d)) ** 2, 1))
    gt = torch.from_numpy(g.view(np.int64)).to(dev)
    ct = torch.from_numpy(np.tile(coef_row, (r, 1))).to(dev)
    ms_f, _ = timed(lambda: km.lat_kernel_fft(1, gt, ct, n), 10, 2)
    x = torch.empty((r, n, d), dtype=torch.float64, device=dev)
    kr = torch.empty((r, n), dtype=torch.float64, device=dev)
    ki = torch.zeros_like(kr)
    twr = torch.empty(n, dtype=torch.float64, device=dev)
    twi = torch.empty_like(twr)

    def chain():
        q.lat_gen_natural(U(r), U(n), U(d), U(0), gt, x, wait=False)
        b = (x - 1.0) * x + 1.0 / 6.0
        kr.copy_(torch.prod(1.0 + ct[:, None, :] * b, dim=-1))
        ki.zero_()
        q.fft_bro_1d_radix2(U(1), U(r), U(n // 2), twr, twi, kr, ki, wait=False)
    ms_c, _ = timed(chain, 3, 1)
    del x
    torch.cuda.empty_cache()
    # base-2 net, order-1 digital kernel, fwht
    t = 32
    C = np.empty((r, d, t), dtype=np.uint64)
    for l in range(r):
        for j in range(d):
            above = rng.integers(0, 2 ** 31, size=t, dtype=np.uint64)
            for k in range(t):
                C[l, j, k] = (1 << (t - 1 - k)) | ((int(above[k]) & ((1 << k) - 1)) << (t - k))
    Ct = torch.from_numpy(C.view(np.int64)).to(dev)
    cn = torch.from_numpy(np.tile(0.9 / (1.0 + np.arange(d)), (r, 1))).to(dev)
    ms_n, _ = timed(lambda: km.dnb2_kernel_fwht(Ct, cn, n, t), 10, 2)
    ms_col, _ = timed(lambda: km.dnb2_kernel_column(Ct, cn, n, t), 10, 2)
    ms_lcol, _ = timed(lambda: km.lat_kernel_column(1, gt, ct, n), 10, 2)
    pts = r * n * d
    return {"workload": "kernel-matrix transforms, r=4 point sets, n=2^20, d=64: lattice (smoothness-1 "
                        "shift-invariant kernel, fft_bro) and 32-bit base-2 net (order-1 digital kernel, fwht)",
            "lattice_fused_ms": round(ms_f, 4), "lattice_column_ms": round(ms_lcol, 4),
            "lattice_chain_ms": round(ms_c, 4), "lattice_speedup_vs_chain": round(ms_c / ms_f, 2),
            "lattice_Gevals_s": round(pts / ms_lcol / 1e6, 1),
            "net_fused_ms": round(ms_n, 4), "net_column_ms": round(ms_col, 4),
            "net_Gevals_s": round(pts / ms_col / 1e6, 1)}


def section_halton(torch, dev, timed):
    """Generalized Halton (halton_qrng): n = 2^20 points in the first 32 prime
    bases, random digit shift, generalized factors; res is d x n float64."""
    from qmctoolscl_amd.halton import halton_qrng
    n, d = 1 << 20, 32
    rng = np.random.Generator(np.random.SFC64(3))
    randu = torch.from_numpy(rng.random((d, 32))).to(dev)
    dvec = torch.arange(d, dtype=torch.int32, device=dev)
    res = torch.empty((d, n), dtype=torch.float64, device=dev)
    ms, _ = timed(lambda: halton_qrng(n, d, 0, 1, res, randu, dvec), 20, 3)
    pts = n * d
    del res
    return {"workload": "generalized Halton, n=2^20, d=32 (first 32 primes), random digit shift, "
                        "generalized factors", "points": pts, "ms": round(ms, 4),
            "Gpts_s": round(pts / ms / 1e6, 2), "GBs_written": round(8 * pts / ms / 1e6, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed batch steps (one step: the whole 512-request batch)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sections", default="lattice,dnb2,transforms,gdn,halton,kernels,chain",
                    help="further sections after the headline batch (also: none)")
    ap.add_argument("--quick-batch", action="store_true", help="cap request sizes at n=2^16")
    ap.add_argument("--tune", default="", help="points_per_tile,grid,gdn_points_per_tile,gdn_cus (skip the autotuner; "
                                               "points_per_tile may be lattice/dnb2/nus)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 2 GiB fill and a 2 GiB device-to-device copy (known-size kernels that scale "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--batch-only", action="store_true",
                    help="time the batched launches only (no per-call run, digests or per-method runs): for profiles")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (generation shards over "
              "replications or n_start chunks, transforms over the batch, with no exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    sections = set(args.sections.split(","))

    import torch
    import qmctoolscl_amd as q

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    timed = Timer(torch, stream)
    if args.calibrate_copy:
        a = torch.ones(1 << 28, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        torch.cuda.empty_cache()
    log("headline: 512-request ragged batch")
    b = section_batch(torch, dev, timed, args)
    pts = b["points"]
    out = {
        "metric": "batched vs per-call Gpts/s (% HBM peak); MFMA vs FP64 FWHT/FFT GB/s @2e-6",
        "value": round(pts / (b["ms"] * 1e-3) / 1e9, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "ragged batch of 512 generation requests (method uniform over lattice+shift, "
                               "dnb2 LMS+shift, dnb2 hash-NUS, gdn LMS+permutation; n=2^m, m in [8,22], d in "
                               "[1,256], r in [1,64]) in persistent hipGraph launches" +
                               (" (m capped at 16)" if args.quick_batch else ""),
                   "requests": 512, "points_per_step": pts, "bytes_written_per_step": 8 * pts},
        "roofline": {"bound": "hbm", "achieved": round(b["GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(b["GBs"] / HBM_PEAK_GBS, 4),
                     "traffic": None if args.quick_batch else committed_traffic(BATCH_TRAFFIC),
                     "traffic_source": "committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/), not this run",
                     "kernel": "batch_persistent_kernel (store-bound and gdn classes side by side), one graph replay"},
        "per_call": {"same_kernels_per_request": {
            "ms_per_step": round(b["per_call_ms"], 3), "value": round(pts / b["per_call_ms"] / 1e6, 3),
            "unit": "Gpts/s", "speedup_of_batched": round(b["per_call_ms"] / b["ms"], 3),
            "form": "the full batch's requests, one launch of the batch's own fused kernels per request"}},
        "wall_s_timed": round(b["wall_s"], 3),
        "sections": {"batch": b["detail"]},
    }
    if "lms_shift" in b["detail"]:
        ls = b["detail"]["lms_shift"]
        out["lms_shift_batched"] = {"value": ls["batched_Gpts_s"], "unit": "Gpts/s",
                                    "frac_hbm": ls["batched_frac_hbm"],
                                    "speedup_vs_per_call": ls["speedup_vs_per_call"],
                                    "workload": "the batch's %d dnb2 LMS+shift requests at full size" % ls["requests"]}
    if "chain" in sections:
        log("section: batch vs per-request drop-in chains")
        out["sections"]["batch_vs_dropin_chain"] = section_batch_vs_chain(torch, q, dev, timed)
        c = out["sections"]["batch_vs_dropin_chain"]
        # the per-call GPU path: the reference's call surface (its drop-in
        # kernels, one chain of calls per request) on this library's GPU
        out["per_call"].update({
            "form": "the per-call GPU path: each request as the reference's chain of drop-in calls "
                    "(lat_gen_natural + lat_shift_mod_1; dnb2 LMS, gen_natural, digital_shift, integer_to_float; "
                    "gdn LMS, gen_natural, digital_permutation, integer_to_float) on the GPU, against one batched "
                    "launch of the same requests",
            "workload": c["workload"], "value": c["dropin_chain_per_request_Gpts_s"], "unit": "Gpts/s",
            "batched_Gpts_s": c["batched_Gpts_s"], "speedup_of_batched": c["speedup_vs_dropin_chain"]})
        if "lms_shift_batched" in out:
            out["lms_shift_batched"]["speedup_vs_dropin_chain"] = c["lms_shift_only"]["speedup_vs_dropin_chain"]
    if "lattice" in sections:
        log("section: lattice (config 1)")
        out["sections"]["lattice"] = section_lattice(torch, q, dev, timed, 50, 20)
    if "dnb2" in sections:
        log("section: dnb2")
        from qmctoolscl_amd import fused
        out["sections"]["dnb2"] = section_dnb2(torch, q, fused, timed, dev)
    if "transforms" in sections:
        log("section: transforms")
        out["sections"]["transforms"] = section_transforms(torch, q, dev, timed)
    if "gdn" in sections:
        log("section: gdn")
        out["sections"]["gdn"] = section_gdn(torch, q, dev, timed)
    if "kernels" in sections:
        log("section: kernel methods")
        out["sections"]["kernel_methods"] = section_kernel_methods(torch, q, dev, timed)
    if "halton" in sections:
        log("section: halton")
        out["sections"]["halton"] = section_halton(torch, dev, timed)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.quick_batch)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
