"""Pure-Python oracle for nested uniform scrambling (NUS).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker, never by the product.
Small cases only (it walks one point at a time, like the reference).

Restates qmctoolscl/rand_funcs.py:155-266 (base 2) and :268-356 (base b): a
lazily grown tree per (replication l, dimension j).  A leaf stores the rest of
a point's digits together with the random material for every level below it;
when a later point shares the leaf but differs, the leaf is pushed down one
level at a time until the two separate.  Random draws happen only when a new
leaf is created, in point order, with the same numpy calls as the reference,
so the numpy stream is consumed identically:
  base 2: root      -> random_tbit_uint64s(rng, tmax_new, 1)[0]
          new leaf  -> rng.integers(0, 2**(t-1))        (t = levels left)
  base b: root      -> tmax_new rows of rng.permutation(b)
          new leaf  -> (tmax_new - t - 1) rows of rng.permutation(b)

Pinned by the reference's doctest outputs (docs/index.md:719-736 for base 2,
:1457-1552 for base b), transcribed into tests/golden/doctest_vectors.json.
"""
import numpy as np


def _tbit_draw(rng, t, shape):
    # rand_funcs.py:4-20
    if t < 64:
        return rng.integers(0, 1 << int(t), shape, dtype=np.uint64)
    s = rng.integers(-(1 << 63), 1 << 63, shape, dtype=np.int64)
    return np.asarray(s).view(np.uint64) ^ np.uint64(1 << 63)


def _perm_rows(rng, k, b):
    # rand_funcs.py:22-32
    out = np.empty((int(k), int(b)), dtype=np.uint64)
    for row in range(int(k)):
        out[row] = rng.permutation(b)
    return out


class Node2:
    """Base-2 tree node.  Leaf: `rest` holds the remaining bits of the one
    point seen below here and `bits` the random bits for all levels left.
    Internal: `rest` is None and `bits` is this level's single random bit."""
    __slots__ = ("bits", "rest", "kids")

    def __init__(self, bits=None, rest=None):
        self.bits = bits
        self.rest = rest
        self.kids = [None, None]


def dnb2_nus(r, n, d, r_x, tmax, tmax_new, rngs, roots, xb, xrb):
    """rand_funcs.py:162-266.  roots: (r, d) object array of Node2() placeholders
    that persist between calls; xb: (r_x, n, d) uint64; xrb: (r, n, d) out."""
    r, n, d, r_x, tmax, tmax_new = (int(v) for v in (r, n, d, r_x, tmax, tmax_new))
    grow = tmax_new - tmax
    for l in range(r):
        for j in range(d):
            rng = rngs[l, j]
            root = roots[l, j]
            if root.bits is None:  # first use: the root is a leaf for the value 0
                root.rest = 0
                root.bits = int(_tbit_draw(rng, tmax_new, 1)[0])
            for i in range(n):
                v = int(xb[l % r_x, i, j]) << grow
                cur, node, t, acc = v, root, tmax_new, 0
                while t > 0:
                    half = 1 << (t - 1)
                    top = (cur >> (t - 1)) & 1
                    rest = cur & (half - 1)
                    if node.rest is None:  # internal node
                        if node.bits:
                            acc += half
                        kid = node.kids[top]
                        if kid is None:
                            draw = int(rng.integers(0, 2 ** (t - 1)))
                            node.kids[top] = Node2(draw, rest)
                            acc += draw
                            break
                        node = kid
                    elif node.rest == cur:  # the same point again
                        acc += node.bits
                        break
                    else:  # a different point reaches this leaf: push it down
                        ntop = (node.rest >> (t - 1)) & 1
                        lead = (node.bits >> (t - 1)) & 1
                        moved = Node2(node.bits & (half - 1), node.rest & (half - 1))
                        if lead:
                            acc += half
                        node.kids[ntop] = moved
                        node.rest, node.bits = None, lead
                        if ntop != top:
                            draw = int(rng.integers(0, 2 ** (t - 1)))
                            node.kids[top] = Node2(draw, rest)
                            acc += draw
                            break
                        node = moved
                    t -= 1
                    cur = rest
                xrb[l, i, j] = np.uint64(v ^ acc)


class NodeB:
    """Base-b tree node.  Leaf: `rest` holds the remaining (zero-padded) digits
    of the one point seen below and `perm` one permutation row per level left.
    Internal: `rest` is None and `perm` is this level's single row."""
    __slots__ = ("perm", "rest", "kids")

    def __init__(self, perm=None, rest=None, kids=None):
        self.perm = perm
        self.rest = rest
        self.kids = kids


def gdn_nus(r, n, d, r_x, r_b, tmax, tmax_new, rngs, roots, bases, xdig, xrdig):
    """rand_funcs.py:274-356.  bases: (r_b, d); xdig: (r_x, n, d, tmax);
    xrdig: (r, n, d, tmax_new) out."""
    r, n, d, r_x, r_b, tmax, tmax_new = (int(v) for v in (r, n, d, r_x, r_b, tmax, tmax_new))
    bases = np.atleast_2d(bases)
    for l in range(r):
        for j in range(d):
            rng = rngs[l, j]
            root = roots[l, j]
            b = bases[l % r_b, j]
            bi = int(b)
            if root.perm is None:
                root.rest = np.zeros(tmax_new, dtype=np.uint64)
                root.perm = _perm_rows(rng, tmax_new, b)
                root.kids = [None] * bi
            for i in range(n):
                node, t = root, 0
                out = np.zeros(tmax_new, dtype=np.uint64)
                while t <= tmax:
                    rest = np.zeros(tmax_new - t, dtype=np.uint64)
                    rest[:tmax - t] = xdig[l % r_x, i, j, t:]
                    dig = int(rest[0])
                    if node.rest is None:  # internal node
                        out[t] = node.perm[dig]
                        if node.kids[dig] is None:
                            rows = _perm_rows(rng, tmax_new - t - 1, b)
                            node.kids[dig] = NodeB(rows, rest[1:], [None] * bi)
                            out[t + 1:] = rows[np.arange(tmax_new - t - 1), rest[1:]]
                            break
                        node = node.kids[dig]
                    elif (node.rest == rest).all():  # the same point again
                        out[t:] = node.perm[np.arange(tmax_new - t), rest]
                        break
                    else:  # a different point reaches this leaf: push it down
                        ndig = int(node.rest[0])
                        out[t] = node.perm[0, dig]
                        node.kids[ndig] = NodeB(node.perm[1:], node.rest[1:], [None] * bi)
                        node.perm, node.rest = node.perm[0], None
                        if ndig == dig:
                            node = node.kids[dig]
                        else:
                            rows = _perm_rows(rng, tmax_new - t - 1, b)
                            node.kids[dig] = NodeB(rows, rest[1:], [None] * bi)
                            out[t + 1:] = rows[np.arange(tmax_new - t - 1), rest[1:]]
                            break
                    t += 1
                xrdig[l, i, j] = out
