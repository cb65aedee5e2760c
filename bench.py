"""Benchmark of the QMC point-generation hot path on one MI355X.

Workload (one step): a randomized rank-1 lattice -- n = 2^20 points in
natural order, d = 64, one generating vector of odd integers in [1, 2^20),
r = 16 uniform random shifts mod 1 -- produced as the r x n x d float64 array
(1.07e9 points, 8 GiB) by the fused generate+shift kernel; inputs resident in
HBM before timing.  The same points through the two drop-in kernels
(lat_gen_natural, then lat_shift_mod_1) are timed beside it ("per_call").

Printed: one JSON line (rank 0) with value = points/s in Gpts/s, the
dominant kernel's HBM roofline (algorithmic bytes: 8 B written per point;
duration from HIP events around the launches on the launching stream), and a
CPU baseline: the reference C backend compiled from its own sources
(oracle/_ref), or this project's C restatement of it where that build is
absent, on a bounded sample of the same workload, single-threaded as the
reference C backend is.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
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
N_LOG2, D, R = 20, 64, 16
PROFILE_TRAFFIC = os.path.join(HERE, "profiles", "lat_gen_shift_traffic.json")


def cpu_baseline(g, shifts, rows):
    """Time lat_gen_natural(r=1) + lat_shift_mod_1(r=16) on `rows` rows."""
    ref = os.path.join(HERE, "oracle", "_ref", "libqmcref.so")
    port = os.path.join(HERE, "oracle", "liboracle.so")
    u = ctypes.c_uint64
    vp = ctypes.c_void_p
    x = np.empty((1, rows, D))
    xr = np.empty((R, rows, D))
    if os.path.exists(ref):
        L = ctypes.CDLL(ref)
        kind = "reference"
        gen = lambda: L.lat_gen_natural(u(1), u(rows), u(D), u(1), u(rows), u(D), u(0), vp(g.ctypes.data),
                                        vp(x.ctypes.data))
        sh = lambda: L.lat_shift_mod_1(u(R), u(rows), u(D), u(R), u(rows), u(D), u(1), vp(x.ctypes.data),
                                       vp(shifts.ctypes.data), vp(xr.ctypes.data))
    elif os.path.exists(port):
        L = ctypes.CDLL(port)
        kind = "port"
        gen = lambda: L.oracle_lat_gen_natural(u(1), u(rows), u(D), u(0), vp(g.ctypes.data), vp(x.ctypes.data))
        sh = lambda: L.oracle_lat_shift_mod_1(u(R), u(rows), u(D), u(1), vp(x.ctypes.data),
                                              vp(shifts.ctypes.data), vp(xr.ctypes.data))
    else:
        return None
    t0 = time.perf_counter()
    gen()
    sh()
    dt = time.perf_counter() - t0
    pts = R * rows * D
    return {"value": round(pts / dt / 1e9, 4), "unit": "Gpts/s", "cores": 1, "kind": kind,
            "sample": "lat_gen_natural(r=1) + lat_shift_mod_1(r=16) over the first %d of 2^%d rows "
                      "(n_start=0), d=%d: %.3g points in %.2f s" % (rows, N_LOG2, D, pts, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-rows", type=int, default=1 << 19)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the lattice workload would "
              "shard over replications or n_start chunks with no exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from qmctoolscl_amd import fused
    import qmctoolscl_amd as q

    dev = torch.device("cuda:0")
    n = 1 << N_LOG2
    rng = np.random.Generator(np.random.SFC64(0))
    g = (2 * rng.integers(0, 2**19, size=(1, D), dtype=np.uint64) + 1)
    shifts = q.lat_get_shifts(rng, R, D)
    g_t = torch.from_numpy(g.view(np.int64)).to(dev)
    s_t = torch.from_numpy(shifts).to(dev)
    xr = torch.empty((R, n, D), dtype=torch.float64, device=dev)
    x1 = torch.empty((1, n, D), dtype=torch.float64, device=dev)
    pts = R * n * D
    stream = torch.cuda.current_stream(dev)

    def step_fused():
        fused.lat_gen_shift(1, R, n, D, 1, 0, g_t, s_t, xr)

    def step_per_call():
        q.lat_gen_natural(np.uint64(1), np.uint64(n), np.uint64(D), np.uint64(0), g_t, x1, wait=False)
        q.lat_shift_mod_1(np.uint64(R), np.uint64(n), np.uint64(D), np.uint64(1), x1, s_t, xr, wait=False)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        start.record(stream)
        for _ in range(steps):
            fn()
        end.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return start.elapsed_time(end) / steps, wall

    ms_fused, wall = timed(step_fused, args.steps, args.warmup)
    ms_pc, _ = timed(step_per_call, max(3, args.steps // 4), 1)

    achieved = 8.0 * pts / (ms_fused * 1e-3) / 1e9  # GB/s written, algorithmic
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    out = {
        "metric": "batched vs per-call Gpts/s (% HBM peak); MFMA vs FP64 FWHT/FFT GB/s @2e-6",
        "value": round(pts / (ms_fused * 1e-3) / 1e9, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_fused, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "randomized rank-1 lattice: natural order, n=2^20, d=64, r=16 shifts mod 1 "
                               "(fused generate+shift, float64 r x n x d output)",
                   "points_per_step": pts, "bytes_written_per_step": 8 * pts},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "per_call": {"ms_per_step": round(ms_pc, 4), "value": round(pts / (ms_pc * 1e-3) / 1e9, 3),
                     "unit": "Gpts/s", "kernels": "lat_gen_natural + lat_shift_mod_1"},
        "wall_s_timed": round(wall, 3),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g, shifts, args.cpu_rows)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
