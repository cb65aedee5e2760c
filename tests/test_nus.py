"""Nested uniform scrambling: the product's native tree builder must request
random draws in exactly the reference's order (checked on the CPU against
the oracle, which follows qmctoolscl/rand_funcs.py:155-356 call for call),
and the GPU tree walk must reproduce the oracle's scrambled points bit for
bit, including across repeated calls that extend the same trees."""
import ctypes
import os
import sys

import numpy as np
import pytest

import _clib
import qmctoolscl_amd as q
from qmctoolscl_amd import _lib, nus

sys.path.insert(0, os.path.join(_clib.REPO, "oracle"))
import nus_oracle  # noqa: E402

needs_lib = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")


class Recorder:
    """numpy Generator proxy that records the sizes of integers()/permutation() calls."""

    def __init__(self, rng):
        self.rng = rng
        self.calls = []

    def integers(self, low, high=None, size=None, dtype=np.int64):
        if size is None:
            self.calls.append(("int", int(high) - int(low)))
        return self.rng.integers(low, high, size, dtype=dtype)

    def permutation(self, b):
        self.calls.append(("perm", int(b)))
        return self.rng.permutation(b)


def _points2(rng, n, tmax, dup=True):
    v = rng.integers(0, 2**tmax, size=n, dtype=np.uint64)
    if dup and n > 4:
        v[3] = v[1]
        v[n - 1] = 0
    return v


@needs_lib
def test_dnb2_draw_schedule_matches_reference_order():
    L = nus._lib_nus()
    rng = np.random.default_rng(0)
    for tmax, tmax_new in [(4, 8), (10, 10), (16, 40), (30, 64)]:
        pts_calls = [_points2(rng, 50, tmax), _points2(rng, 30, tmax)]
        # oracle, recording widths of the leaf draws (root draw excluded)
        rec = Recorder(np.random.Generator(np.random.SFC64(1)))
        roots = np.array([[nus_oracle.Node2()]], dtype=object)
        rngs = np.array([[rec]], dtype=object)
        for pts in pts_calls:
            out = np.zeros((1, len(pts), 1), dtype=np.uint64)
            nus_oracle.dnb2_nus(1, len(pts), 1, 1, tmax, tmax_new, rngs, roots, pts.reshape(1, -1, 1), out)
        want = [w for kind, w in rec.calls]
        # native builder
        h = L.qmc_nus2_tree_new(tmax_new)
        try:
            got = []
            widths = np.empty(64, dtype=np.uint32)
            for pts in pts_calls:
                k = L.qmc_nus2_insert(h, pts.ctypes.data, len(pts), 1, tmax_new - tmax, widths.ctypes.data)
                got += [1 << int(w) for w in widths[:k]]
        finally:
            L.qmc_nus2_tree_free(h)
        assert got == want[:len(got)] and len(got) == len(want), (tmax, tmax_new)


@needs_lib
def test_gdn_draw_schedule_matches_reference_order():
    L = nus._lib_nus()
    rng = np.random.default_rng(1)
    # tmax == tmax_new is left out: there the reference raises IndexError when a
    # repeated point reaches a leaf below the last digit (rand_funcs.py:323-325)
    for b, tmax, tmax_new in [(2, 3, 8), (3, 4, 6), (5, 2, 5), (7, 3, 4)]:
        calls = []
        for n in (20, 15):
            dig = rng.integers(0, b, size=(n, tmax), dtype=np.uint64)
            dig[2] = dig[0]
            calls.append(dig)
        rec = Recorder(np.random.Generator(np.random.SFC64(2)))
        roots = np.array([[nus_oracle.NodeB()]], dtype=object)
        rngs = np.array([[rec]], dtype=object)
        bases = np.array([[b]], dtype=np.uint64)
        for dig in calls:
            out = np.zeros((1, len(dig), 1, tmax_new), dtype=np.uint64)
            nus_oracle.gdn_nus(1, len(dig), 1, 1, 1, tmax, tmax_new, rngs, roots, bases,
                               dig.reshape(1, -1, 1, tmax), out)
        want = len([c for c in rec.calls if c[0] == "perm"]) - tmax_new  # rows after the root's
        h = L.qmc_nusb_tree_new(tmax, tmax_new, b)
        try:
            got = 0
            counts = np.empty(64, dtype=np.uint32)
            for dig in calls:
                k = L.qmc_nusb_insert(h, dig.ctypes.data, len(dig), tmax, 1, counts.ctypes.data)
                got += int(counts[:k].sum())
        finally:
            L.qmc_nusb_tree_free(h)
        assert got == want, (b, tmax, tmax_new)


@needs_lib
def test_sfc64_permutation_replay_matches_numpy():
    """The builder's SFC64 replay of Generator.permutation(b) equals numpy's own
    draws, row for row, and leaves the generator in numpy's state, from states
    with and without a buffered 32-bit half."""
    for seed in range(4):
        for b in (2, 3, 5, 7, 131, 1619, 65536):
            a = np.random.Generator(np.random.SFC64(seed))
            c = np.random.Generator(np.random.SFC64(seed))
            if seed % 2:  # leave half of a 64-bit output buffered
                a.integers(0, 2**31, dtype=np.uint32)
                c.integers(0, 2**31, dtype=np.uint32)
            count = 3 if b > 2000 else 12
            want = np.stack([a.permutation(b) for _ in range(count)])
            got = nus.sfc64_permutations(c, count, b)
            assert np.array_equal(got.astype(np.int64), want), (seed, b)
            sa, sc = a.bit_generator.state, c.bit_generator.state
            assert list(sa["state"]["state"]) == list(sc["state"]["state"]), (seed, b)
            assert (sa["has_uint32"], sa["uinteger"]) == (sc["has_uint32"], sc["uinteger"]), (seed, b)


@needs_lib
@pytest.mark.parametrize("bases,tmax,tmax_new", [([2, 3, 5], 3, 8), ([3, 7], 4, 6), ([131], 2, 5), ([3], 12, 52)])
def test_gdn_replay_consumes_the_stream_as_the_reference(bases, tmax, tmax_new):
    """After two calls that grow the same trees, each tree's generator is where
    the reference's draws leave it (rand_funcs.py:312-352, through the oracle),
    and the root rows equal the reference's root_node.perm.  Host side only:
    the builder's insertion and replay run without a GPU."""
    L = nus._lib_nus()
    rng = np.random.default_rng(len(bases) * 7 + tmax)
    d, r = len(bases), 2
    n = 200 if tmax > 4 else 60
    b = np.array([bases], dtype=np.uint64)
    xdig = np.ascontiguousarray(np.stack([rng.integers(0, bb, size=(1, n, tmax), dtype=np.uint64)
                                          for bb in bases], axis=2))
    xdig[:, 7] = xdig[:, 3]
    roots_o = np.array([nus_oracle.NodeB() for _ in range(r * d)]).reshape(r, d)
    rngs_o, rngs_p = _rngs(r, d, 11), _rngs(r, d, 11)
    trees = [L.qmc_nusb_tree_new(tmax, tmax_new, int(bb)) for _ in range(r) for bb in bases]
    handles = (ctypes.c_void_p * (r * d))(*trees)
    try:
        for part in (xdig[:, :n // 2], xdig[:, n // 2:]):
            part = np.ascontiguousarray(part)
            m = part.shape[1]
            want = np.zeros((r, m, d, tmax_new), dtype=np.uint64)
            nus_oracle.gdn_nus(r, m, d, 1, 1, tmax, tmax_new, rngs_o, roots_o, b, part, want)
            states = np.stack([nus._sfc64_state(g) for g in rngs_p.ravel()])
            replay = np.ones(r * d, dtype=np.uint8)
            _lib.check(L.qmc_nusb_grow(handles, r * d, part.ctypes.data, r, m, d, 1, states.ctypes.data,
                                       replay.ctypes.data, 4))
            for k, g in enumerate(rngs_p.ravel()):
                nus._set_sfc64_state(g, states[k])
            for go, gp in zip(rngs_o.ravel(), rngs_p.ravel()):
                so, sp = go.bit_generator.state, gp.bit_generator.state
                assert list(so["state"]["state"]) == list(sp["state"]["state"])
                assert (so["has_uint32"], so["uinteger"]) == (sp["has_uint32"], sp["uinteger"])
        for k in range(r * d):
            bb = bases[k % d]
            rows = np.empty((tmax_new, bb), dtype=np.uint16)
            _lib.check(L.qmc_nusb_root_rows(trees[k], rows.ctypes.data))
            root = roots_o.ravel()[k]
            first = root.perm if root.perm.ndim == 2 else None
            if first is not None:  # the root is still a leaf in the oracle: all its rows
                assert np.array_equal(rows.astype(np.uint64), first)
            else:  # split: its own row is the first one drawn
                assert np.array_equal(rows[0].astype(np.uint64), root.perm)
            # leaves keep a state, not rows: materialized rows stay far below the reference's
            assert L.qmc_nusb_tree_rows(trees[k]) <= tmax_new + L.qmc_nusb_tree_nodes(trees[k])
    finally:
        for h in trees:
            L.qmc_nusb_tree_free(h)


def _rngs(r, d, seed):
    seeds = np.random.SeedSequence(seed).spawn(r * d)
    return np.array([np.random.Generator(np.random.SFC64(s)) for s in seeds]).reshape(r, d)


@pytest.mark.gpu
@pytest.mark.parametrize("tmax,tmax_new", [(4, 8), (12, 12), (20, 52), (32, 64)])
def test_dnb2_nus_matches_oracle_gpu(tmax, tmax_new):
    rng = np.random.default_rng(tmax)
    r_x, r, n, d = 2, 3, 200, 3
    xb = rng.integers(0, 2**tmax, size=(r_x, n, d), dtype=np.uint64)
    xb[:, 5] = xb[:, 2]  # duplicates
    roots_o = np.array([nus_oracle.Node2() for _ in range(r * d)]).reshape(r, d)
    roots_p = np.array([q.NUSNode_dnb2() for _ in range(r * d)]).reshape(r, d)
    rngs_o, rngs_p = _rngs(r, d, 9), _rngs(r, d, 9)
    for half in (xb[:, :120], xb[:, 120:]):  # two calls extend the same trees
        half = np.ascontiguousarray(half)
        want = np.zeros((r, half.shape[1], d), dtype=np.uint64)
        got = np.zeros_like(want)
        nus_oracle.dnb2_nus(r, half.shape[1], d, r_x, tmax, tmax_new, rngs_o, roots_o, half, want)
        q.dnb2_nested_uniform_scramble(r, half.shape[1], d, r_x, tmax, tmax_new, rngs_p, roots_p, half, got)
        assert np.array_equal(got, want)
    # generators stay in step with the reference's consumption
    assert all(a.integers(0, 2**40) == b.integers(0, 2**40) for a, b in zip(rngs_o.ravel(), rngs_p.ravel()))


@pytest.mark.gpu
@pytest.mark.parametrize("bases,tmax,tmax_new", [([2, 3, 5], 3, 8), ([3, 7], 4, 6), ([2, 131], 2, 5)])
def test_gdn_nus_matches_oracle_gpu(bases, tmax, tmax_new):
    rng = np.random.default_rng(len(bases) + tmax)
    d = len(bases)
    r_x, r, n = 1, 2, 60
    b = np.array([bases], dtype=np.uint64)
    xdig = np.stack([rng.integers(0, bb, size=(r_x, n, tmax), dtype=np.uint64) for bb in bases], axis=2)
    xdig = np.ascontiguousarray(xdig)
    roots_o = np.array([nus_oracle.NodeB() for _ in range(r * d)]).reshape(r, d)
    roots_p = np.array([q.NUSNode_gdn() for _ in range(r * d)]).reshape(r, d)
    rngs_o, rngs_p = _rngs(r, d, 5), _rngs(r, d, 5)
    for part in (xdig[:, :40], xdig[:, 40:]):
        part = np.ascontiguousarray(part)
        want = np.zeros((r, part.shape[1], d, tmax_new), dtype=np.uint64)
        got = np.zeros_like(want)
        nus_oracle.gdn_nus(r, part.shape[1], d, r_x, 1, tmax, tmax_new, rngs_o, roots_o, b, part, want)
        q.gdn_nested_uniform_scramble(r, part.shape[1], d, r_x, 1, tmax, tmax_new, rngs_p, roots_p, b, part, got)
        assert np.array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("gen", ["sfc64", "pcg64"])
def test_gdn_nus_large_bases_and_other_generators_gpu(gen):
    """Leaf rows regenerated on the device for a base whose per-lane draw
    buffers exceed the LDS budget (1619: device scratch), beside small bases;
    with PCG64 generators every row is drawn by numpy and kept instead."""
    rng = np.random.default_rng(40)
    bases = [3, 1619, 131]
    d, r, n, tmax, tmax_new = len(bases), 2, 90, 3, 9
    b = np.array([bases], dtype=np.uint64)
    xdig = np.ascontiguousarray(np.stack([rng.integers(0, bb, size=(1, n, tmax), dtype=np.uint64)
                                          for bb in bases], axis=2))
    xdig[:, 11] = xdig[:, 5]

    def gens(seed):
        ss = np.random.SeedSequence(seed).spawn(r * d)
        bg = np.random.SFC64 if gen == "sfc64" else np.random.PCG64
        return np.array([np.random.Generator(bg(s)) for s in ss]).reshape(r, d)

    roots_o = np.array([nus_oracle.NodeB() for _ in range(r * d)]).reshape(r, d)
    roots_p = np.array([q.NUSNode_gdn() for _ in range(r * d)]).reshape(r, d)
    rngs_o, rngs_p = gens(8), gens(8)
    for part in (xdig[:, :50], xdig[:, 50:]):
        part = np.ascontiguousarray(part)
        want = np.zeros((r, part.shape[1], d, tmax_new), dtype=np.uint64)
        got = np.zeros_like(want)
        nus_oracle.gdn_nus(r, part.shape[1], d, 1, 1, tmax, tmax_new, rngs_o, roots_o, b, part, want)
        q.gdn_nested_uniform_scramble(r, part.shape[1], d, 1, 1, tmax, tmax_new, rngs_p, roots_p, b, part, got)
        assert np.array_equal(got, want)
    assert all(a.integers(0, 2**40) == c.integers(0, 2**40) for a, c in zip(rngs_o.ravel(), rngs_p.ravel()))


@pytest.mark.gpu
def test_gdn_nus_first_32_primes_many_points_gpu():
    """The config-2 shape in miniature: the first 32 primes, 20 -> 52 digits,
    3^6 points in two calls, against the oracle (which materializes every
    row); most leaves' rows come from stored generator states here."""
    from qmctoolscl_amd.batch import first_primes
    rng = np.random.default_rng(41)
    bases = first_primes(32)
    d, r, n, tmax, tmax_new = 32, 1, 3 ** 6, 20, 52
    b = np.array([bases], dtype=np.uint64)
    xdig = np.ascontiguousarray(np.stack([rng.integers(0, bb, size=(1, n, tmax), dtype=np.uint64)
                                          for bb in bases], axis=2))
    roots_o = np.array([nus_oracle.NodeB() for _ in range(r * d)]).reshape(r, d)
    roots_p = np.array([q.NUSNode_gdn() for _ in range(r * d)]).reshape(r, d)
    rngs_o, rngs_p = _rngs(r, d, 12), _rngs(r, d, 12)
    for part in (xdig[:, :300], xdig[:, 300:]):
        part = np.ascontiguousarray(part)
        want = np.zeros((r, part.shape[1], d, tmax_new), dtype=np.uint64)
        got = np.zeros_like(want)
        nus_oracle.gdn_nus(r, part.shape[1], d, 1, 1, tmax, tmax_new, rngs_o, roots_o, b, part, want)
        q.gdn_nested_uniform_scramble(r, part.shape[1], d, 1, 1, tmax, tmax_new, rngs_p, roots_p, b, part, got)
        assert np.array_equal(got, want)


@pytest.mark.gpu
def test_nus_device_copies_follow_the_trees_passed():
    """The apply calls keep the packed device copy of the last trees they were
    given; calls alternating between two sets of trees (A, B, A grown, B
    unchanged) must each use their own trees."""
    rng = np.random.default_rng(21)
    r, d, tmax, tnew = 2, 3, 6, 10
    # base 2
    xb = rng.integers(0, 2**tmax, size=(1, 80, d), dtype=np.uint64)
    sets = []
    for seed in (3, 4):
        sets.append(dict(ro=np.array([nus_oracle.Node2() for _ in range(r * d)]).reshape(r, d),
                         rp=np.array([q.NUSNode_dnb2() for _ in range(r * d)]).reshape(r, d),
                         go=_rngs(r, d, seed), gp=_rngs(r, d, seed)))
    for k, part in ((0, xb[:, :40]), (1, xb[:, :40]), (0, xb[:, 40:]), (1, xb[:, :40])):
        part = np.ascontiguousarray(part)
        S = sets[k]
        want = np.zeros((r, part.shape[1], d), dtype=np.uint64)
        got = np.zeros_like(want)
        nus_oracle.dnb2_nus(r, part.shape[1], d, 1, tmax, tnew, S["go"], S["ro"], part, want)
        q.dnb2_nested_uniform_scramble(r, part.shape[1], d, 1, tmax, tnew, S["gp"], S["rp"], part, got)
        assert np.array_equal(got, want), k
    # base b
    bases = [2, 3, 7]
    b = np.array([bases], dtype=np.uint64)
    xdig = np.stack([rng.integers(0, bb, size=(1, 50, tmax), dtype=np.uint64) for bb in bases], axis=2)
    sets = []
    for seed in (5, 6):
        sets.append(dict(ro=np.array([nus_oracle.NodeB() for _ in range(r * d)]).reshape(r, d),
                         rp=np.array([q.NUSNode_gdn() for _ in range(r * d)]).reshape(r, d),
                         go=_rngs(r, d, seed), gp=_rngs(r, d, seed)))
    for k, part in ((0, xdig[:, :25]), (1, xdig[:, :25]), (0, xdig[:, 25:]), (1, xdig[:, :25])):
        part = np.ascontiguousarray(part)
        S = sets[k]
        want = np.zeros((r, part.shape[1], d, tnew), dtype=np.uint64)
        got = np.zeros_like(want)
        nus_oracle.gdn_nus(r, part.shape[1], d, 1, 1, tmax, tnew, S["go"], S["ro"], b, part, want)
        q.gdn_nested_uniform_scramble(r, part.shape[1], d, 1, 1, tmax, tnew, S["gp"], S["rp"], b, part, got)
        assert np.array_equal(got, want), k
