// Nested uniform scrambling (NUS) for base-2 and base-b digital nets.
//
// The reference (qmctoolscl/rand_funcs.py:155-356) grows one random tree per
// (replication l, dimension j) lazily, one point at a time, drawing random
// material from a numpy Generator whenever a new leaf appears.  The *shape*
// of a tree and the order of its draws depend only on the sequence of point
// values, never on the random values themselves, so the work splits in three:
//
//  1. host (this file, C++): insert the points in order into a flattened,
//     persistent tree -- the same leaf creation / leaf push-down rules as the
//     reference -- recording for every draw its size (base 2: a bit count;
//     base b: a count of permutation rows).  Every node remembers the draw
//     that owns it: a node at depth t takes bit (tmax_new-1-t) of its owner's
//     bits (base 2) or row (t - start depth of the owner's block) (base b).
//  2. Python: make the draws with the caller's numpy Generators in the
//     recorded order (one vectorized call per tree: the stream is consumed
//     exactly as by the reference's one-call-per-leaf loop).
//  3. device (this file, HIP): every point walks its finished tree in
//     parallel and assembles its scrambled value.  Later splits move random
//     material down a path but never change it, so the finished tree gives
//     every point the value it got when it was inserted.
//
// Trees persist between calls (the reference keeps them in root_nodes), so
// extending a point set scrambles consistently.
#include <cstring>
#include <vector>

#include <atomic>
#include <mutex>

#include "qmc_common.hpp"

namespace {

// ------------------------------------------------------------------ base 2

struct Node2 {
    int32_t kid[2];
    int32_t owner;   // draw index
    int32_t leaf;
    u64 rest;        // leaf: remaining low bits of the one point below
};

struct Tree2 {
    int tmax_new;
    u64 id = 0, version = 0;  // identity and mutation count: the device copy is reused while they match
    std::vector<Node2> nodes;
    std::vector<u64> draws;  // draw values, index = owner id
    u64 ndraws_scheduled;    // draws recorded so far (root included)
};

int32_t new_leaf2(Tree2 *T, u64 rest, int32_t owner) {
    Node2 nd;
    nd.kid[0] = nd.kid[1] = -1;
    nd.owner = owner;
    nd.leaf = 1;
    nd.rest = rest;
    T->nodes.push_back(nd);
    return (int32_t)T->nodes.size() - 1;
}

// rand_funcs.py:200-262 for one point value v (already shifted to tmax_new bits)
void insert2(Tree2 *T, u64 v, std::vector<uint32_t> &widths) {
    int32_t node = 0;
    u64 cur = v;
    for (int t = T->tmax_new; t > 0; t--) {
        const u64 half = 1ull << (t - 1);
        const int top = (int)((cur >> (t - 1)) & 1);
        const u64 rest = cur & (half - 1);
        Node2 &N = T->nodes[node];
        if (!N.leaf) {
            if (N.kid[top] < 0) {
                const int32_t id = (int32_t)T->ndraws_scheduled++;
                widths.push_back((uint32_t)(t - 1));
                const int32_t k = new_leaf2(T, rest, id);
                T->nodes[node].kid[top] = k;
                return;
            }
            node = N.kid[top];
        } else if (N.rest == cur) {
            return;  // seen before
        } else {
            const int ntop = (int)((N.rest >> (t - 1)) & 1);
            const int32_t owner = N.owner;
            const u64 moved_rest = N.rest & (half - 1);
            const int32_t moved = new_leaf2(T, moved_rest, owner);  // may reallocate: no N below
            Node2 &M = T->nodes[node];
            M.leaf = 0;
            M.kid[0] = M.kid[1] = -1;
            M.kid[ntop] = moved;
            if (ntop != top) {
                const int32_t id = (int32_t)T->ndraws_scheduled++;
                widths.push_back((uint32_t)(t - 1));
                const int32_t k = new_leaf2(T, rest, id);
                T->nodes[node].kid[top] = k;
                return;
            }
            node = moved;
        }
        cur = rest;
    }
}

// ------------------------------------------------------------------ base b

struct NodeB {
    int32_t kids;    // offset of b child slots in Tree::kids, -1 for a leaf
    int32_t owner;   // permutation block
    int64_t rest;    // leaf: offset of its (tmax_new - depth) remaining digits in Tree::digits
};

struct Block {
    int64_t row0;    // first row of the block in the row pool (filled when the draws arrive)
    int32_t depth0;  // depth of the block's first row
    int32_t rows;
};

struct TreeB {
    int tmax, tmax_new;
    u64 b;                    // <= 65536, so a permuted digit fits 16 bits
    u64 id, version;          // identity and mutation count: the device copy is reused while both match
    std::vector<NodeB> nodes;
    std::vector<int32_t> kids;
    std::vector<u64> digits;  // leaf digit pool
    std::vector<Block> blocks;
    std::vector<uint16_t> rows;  // permutation rows, b values each, in draw order
};

std::atomic<u64> g_treeb_ids{1};

int32_t new_leafB(TreeB *T, const u64 *rest, int len, int32_t owner) {
    NodeB nd;
    nd.kids = -1;
    nd.owner = owner;
    nd.rest = (int64_t)T->digits.size();
    T->digits.insert(T->digits.end(), rest, rest + len);
    T->nodes.push_back(nd);
    return (int32_t)T->nodes.size() - 1;
}

int32_t new_blockB(TreeB *T, int depth0, int rows, std::vector<uint32_t> &counts) {
    Block bl;
    bl.row0 = -1;
    bl.depth0 = depth0;
    bl.rows = rows;
    T->blocks.push_back(bl);
    counts.push_back((uint32_t)rows);
    return (int32_t)T->blocks.size() - 1;
}

// rand_funcs.py:318-353 for one point's digits (tmax of them)
void insertB(TreeB *T, const u64 *dig, u64 stride, std::vector<uint32_t> &counts) {
    const int tn = T->tmax_new, tm = T->tmax;
    std::vector<u64> rest(tn);
    int32_t node = 0;
    for (int t = 0; t <= tm; t++) {
        const int len = tn - t;
        for (int k = 0; k < len; k++) rest[k] = (k < tm - t) ? dig[(u64)(t + k) * stride] : 0;
        const u64 dg = rest[0];
        if (dg >= T->b) return;  // a digit outside [0, b): the reference would index past a permutation
        NodeB &N = T->nodes[node];
        if (N.kids >= 0) {
            const int32_t kid = T->kids[N.kids + dg];
            if (kid < 0) {
                const int32_t blk = new_blockB(T, t + 1, tn - t - 1, counts);
                const int32_t k = new_leafB(T, rest.data() + 1, len - 1, blk);
                T->kids[T->nodes[node].kids + dg] = k;
                return;
            }
            node = kid;
        } else if (std::memcmp(&T->digits[N.rest], rest.data(), sizeof(u64) * len) == 0) {
            return;  // seen before
        } else {
            const u64 ndg = T->digits[N.rest];
            const int32_t owner = N.owner;
            const int64_t old_rest = N.rest;
            // the leaf's contents move one level down, keeping its owner block
            NodeB moved;
            moved.kids = -1;
            moved.owner = owner;
            moved.rest = old_rest + 1;  // its digits shifted by one level (shares the pool)
            T->nodes.push_back(moved);
            const int32_t mid = (int32_t)T->nodes.size() - 1;
            const int32_t koff = (int32_t)T->kids.size();
            T->kids.insert(T->kids.end(), (size_t)T->b, -1);
            NodeB &M = T->nodes[node];
            M.kids = koff;
            T->kids[koff + ndg] = mid;
            if (ndg == dg) {
                node = mid;
            } else {
                const int32_t blk = new_blockB(T, t + 1, tn - t - 1, counts);
                const int32_t k = new_leafB(T, rest.data() + 1, len - 1, blk);
                T->kids[koff + dg] = k;
                return;
            }
        }
    }
}

// ------------------------------------------------------------ device walks

struct Packed2 {
    const int32_t *kid;    // 2 per node
    const int32_t *owner;
    const int32_t *leaf;
    const u64 *draws;
    const int64_t *node0;  // first node of tree (l, j)
    const int64_t *draw0;  // first draw of tree (l, j)
};

__global__ __launch_bounds__(kBlock) void nus2_apply_kernel(u64 r, u64 n, u64 d, u64 r_x, int tmax_new, int tdelta,
                                                            Packed2 P, const u64 *__restrict__ xb,
                                                            u64 *__restrict__ xrb) {
    const u64 total = r * n * d;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const u64 j = e % d;
        const u64 li = e / d;
        const u64 i = li % n;
        const u64 l = li / n;
        const u64 v = xb[((l % r_x) * n + i) * d + j] << tdelta;
        const int64_t nb = P.node0[l * d + j];
        const int64_t db = P.draw0[l * d + j];
        int64_t node = nb;
        u64 shift = 0;
        for (int t = tmax_new; t > 0; t--) {
            const u64 w = P.draws[db + P.owner[node]];
            if (P.leaf[node]) {
                shift |= (t >= 64) ? w : (w & ((1ull << t) - 1));
                break;
            }
            shift |= w & (1ull << (t - 1));
            const int top = (int)((v >> (t - 1)) & 1);
            node = nb + P.kid[2 * node + top];  // child ids are local to the tree
        }
        xrb[e] = v ^ shift;
    }
}

struct PackedB {
    const int32_t *kids_off;   // per node, -1 for a leaf
    const int32_t *owner;      // per node: block id (tree-local)
    const int32_t *kids;       // child slots (tree-local node ids)
    const int64_t *blk_row0;   // per block: first row (tree-local)
    const int32_t *blk_depth0; // per block
    const uint16_t *rows;      // b values per row, each tree's rows packed
    const int64_t *node0, *kid0, *blk0, *row0;  // per tree offsets (row0 in values)
    const uint32_t *tb;        // per tree base
};

__global__ __launch_bounds__(kBlock) void nusb_apply_kernel(u64 r, u64 n, u64 d, u64 r_x, u64 r_b, int tmax,
                                                            int tmax_new, u64 bmax, PackedB P,
                                                            const u64 *__restrict__ xdig,
                                                            u64 *__restrict__ xrdig) {
    const u64 total = r * n * d;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const u64 j = e % d;
        const u64 li = e / d;
        const u64 i = li % n;
        const u64 l = li / n;
        const u64 tree = l * d + j;
        const u64 *src = xdig + (((l % r_x) * n + i) * d + j) * tmax;
        u64 *out = xrdig + e * tmax_new;
        const int64_t nb = P.node0[tree], kb = P.kid0[tree], bb = P.blk0[tree];
        const uint16_t *rows = P.rows + P.row0[tree];
        const int64_t b = P.tb[tree];
        int64_t node = 0;  // tree-local
        int t = 0;
        for (; t < tmax_new; t++) {
            const u64 dg = t < tmax ? src[t] : 0;
            const int32_t blk = P.owner[nb + node];
            const int64_t row = P.blk_row0[bb + blk] + (t - P.blk_depth0[bb + blk]);
            out[t] = rows[row * b + (int64_t)dg];
            const int32_t ko = P.kids_off[nb + node];
            if (ko < 0) {  // leaf: its block's rows cover every deeper level
                for (int tt = t + 1; tt < tmax_new; tt++) {
                    const u64 g = tt < tmax ? src[tt] : 0;
                    const int64_t rw = P.blk_row0[bb + blk] + (tt - P.blk_depth0[bb + blk]);
                    out[tt] = rows[rw * b + (int64_t)g];
                }
                break;
            }
            node = P.kids[kb + ko + (int64_t)dg];
        }
    }
}

// The packed device copy of the last set of trees applied, reused while the
// trees (by identity and mutation count), r and d are unchanged: a repeat call
// on the same points uploads nothing.  Calls are synchronous (the copy is
// released or replaced only after the previous call's kernel finished).
struct NusbDevice {
    std::vector<u64> sig;
    int device = -1;
    char *buf = nullptr;
    PackedB P{};
};
NusbDevice g_nusb;  // guarded by g_nusb_mu
std::mutex g_nusb_mu;
// the same for base-2 trees
struct Nus2Device {
    std::vector<u64> sig;
    int device = -1;
    char *buf = nullptr;
    Packed2 P{};
};
Nus2Device g_nus2;  // guarded by g_nus2_mu
std::mutex g_nus2_mu;

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

QMC_EXPORT void *qmc_nus2_tree_new(int tmax_new) {
    if (tmax_new < 1 || tmax_new > 64) {
        qmc_fail(QMC_ERR_ARG, "nus: tmax_new must be in [1, 64]");
        return nullptr;
    }
    Tree2 *T = new Tree2();
    T->tmax_new = tmax_new;
    T->id = g_treeb_ids++;
    T->ndraws_scheduled = 1;  // draw 0: the root's tmax_new bits, drawn by the caller
    new_leaf2(T, 0, 0);       // the root starts as a leaf for the value 0 (rand_funcs.py:195-199)
    return T;
}

QMC_EXPORT void qmc_nus2_tree_free(void *h) { delete (Tree2 *)h; }

QMC_EXPORT u64 qmc_nus2_tree_nodes(void *h) { return ((Tree2 *)h)->nodes.size(); }
QMC_EXPORT u64 qmc_nus2_tree_draws(void *h) { return ((Tree2 *)h)->draws.size(); }

// Insert n values (v[k*stride] << tdelta) in order; writes the bit widths of
// the new draws to widths (capacity n) and returns how many there are.
QMC_EXPORT int64_t qmc_nus2_insert(void *h, const u64 *v, u64 n, u64 stride, int tdelta, uint32_t *widths) {
    Tree2 *T = (Tree2 *)h;
    std::vector<uint32_t> w;
    w.reserve(n);
    const size_t before = T->nodes.size();
    for (u64 k = 0; k < n; k++) insert2(T, v[k * stride] << tdelta, w);
    if (T->nodes.size() != before || !w.empty()) T->version++;
    if (!w.empty()) std::memcpy(widths, w.data(), w.size() * sizeof(uint32_t));
    return (int64_t)w.size();
}

// Append draw values (the root draw first, then in recorded order).
QMC_EXPORT int qmc_nus2_add_draws(void *h, const u64 *vals, u64 count) {
    Tree2 *T = (Tree2 *)h;
    if (T->draws.size() + count > T->ndraws_scheduled)
        return qmc_fail(QMC_ERR_ARG, "nus: more draws than the tree asked for");
    T->draws.insert(T->draws.end(), vals, vals + count);
    if (count) T->version++;
    return QMC_OK;
}

// Scramble: xrb[l,i,j] from xb[l % r_x, i, j] with trees[l*d + j] (all points
// already inserted and all draws added).  Packs the trees, copies them to the
// device on stream s, and launches the walk.
QMC_EXPORT int qmc_nus2_apply_dev(void **trees, u64 r, u64 n, u64 d, u64 r_x, int tmax, int tmax_new,
                                  const u64 *xb, u64 *xrb, qmc_stream_t s) {
    if (!(r && n && d)) return QMC_OK;
    const u64 nt = r * d;
    std::vector<u64> sig;
    sig.reserve(2 * nt + 2);
    sig.push_back(r);
    sig.push_back(d);
    for (u64 k = 0; k < nt; k++) {
        Tree2 *T = (Tree2 *)trees[k];
        if (T->draws.size() != T->ndraws_scheduled)
            return qmc_fail(QMC_ERR_ARG, "nus: tree %llu is missing draws", (unsigned long long)k);
        sig.push_back(T->id);
        sig.push_back(T->version);
    }
    int dev = 0;
    QMC_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_nus2_mu);
    if (g_nus2.buf == nullptr || g_nus2.device != dev || g_nus2.sig != sig) {
        std::vector<int64_t> node0(nt), draw0(nt);
        u64 nn = 0, ndr = 0;
        for (u64 k = 0; k < nt; k++) {
            Tree2 *T = (Tree2 *)trees[k];
            node0[k] = (int64_t)nn;
            draw0[k] = (int64_t)ndr;
            nn += T->nodes.size();
            ndr += T->draws.size();
        }
        std::vector<int32_t> kid(2 * nn), owner(nn), leaf(nn);
        std::vector<u64> draws(ndr > 0 ? ndr : 1);
        for (u64 k = 0; k < nt; k++) {
            Tree2 *T = (Tree2 *)trees[k];
            for (size_t q = 0; q < T->nodes.size(); q++) {
                const Node2 &N = T->nodes[q];
                const u64 g = node0[k] + q;
                kid[2 * g] = N.kid[0] < 0 ? 0 : N.kid[0];
                kid[2 * g + 1] = N.kid[1] < 0 ? 0 : N.kid[1];
                owner[g] = N.owner;
                leaf[g] = N.leaf;
            }
            if (!T->draws.empty())
                std::memcpy(draws.data() + draw0[k], T->draws.data(), T->draws.size() * sizeof(u64));
        }
        if (g_nus2.buf) {
            (void)hipFree(g_nus2.buf);
            g_nus2.buf = nullptr;
        }
        g_nus2.sig.clear();
        const size_t total = kid.size() * 4 + owner.size() * 4 + leaf.size() * 4 + draws.size() * 8 + 2 * nt * 8 + 64;
        char *buf = nullptr;
        if (hipMalloc((void **)&buf, total) != hipSuccess)
            return qmc_fail(QMC_ERR_NOMEM, "nus: device allocation of %zu bytes for the trees failed", total);
        size_t o = 0;
        hipError_t ce = hipSuccess;
        auto put = [&](const void *src, size_t bytes) -> void * {
            o = (o + 7) & ~(size_t)7;
            void *dst = buf + o;
            if (bytes && ce == hipSuccess) ce = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
            o += bytes;
            return dst;
        };
        Packed2 P;
        P.kid = (const int32_t *)put(kid.data(), kid.size() * 4);
        P.owner = (const int32_t *)put(owner.data(), owner.size() * 4);
        P.leaf = (const int32_t *)put(leaf.data(), leaf.size() * 4);
        P.draws = (const u64 *)put(draws.data(), draws.size() * 8);
        P.node0 = (const int64_t *)put(node0.data(), nt * 8);
        P.draw0 = (const int64_t *)put(draw0.data(), nt * 8);
        if (ce != hipSuccess) {
            (void)hipFree(buf);
            return qmc_fail_hip(ce, "nus: copying the trees");
        }
        g_nus2.buf = buf;
        g_nus2.P = P;
        g_nus2.device = dev;
        g_nus2.sig = sig;
    }
    const int tdelta = tmax_new - tmax;
    hipLaunchKernelGGL(nus2_apply_kernel, dim3(grid_for(r * n * d, kBlock, 2048 * 8)), dim3(kBlock), 0, s, r, n, d,
                       r_x, tmax_new, tdelta, g_nus2.P, xb, xrb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // the cached copy may be replaced by the next call
    if (e != hipSuccess) return qmc_fail_hip(e, "nus2_apply_kernel");
    return QMC_OK;
}

QMC_EXPORT void *qmc_nusb_tree_new(int tmax, int tmax_new, u64 b) {
    if (tmax < 0 || tmax_new < 1 || tmax > tmax_new || b < 1 || b > (1u << 16)) {
        qmc_fail(QMC_ERR_ARG, "nus: need 0 <= tmax <= tmax_new, 1 <= b <= 2^16");
        return nullptr;
    }
    TreeB *T = new TreeB();
    T->tmax = tmax;
    T->tmax_new = tmax_new;
    T->b = b;
    T->id = g_treeb_ids++;
    T->version = 0;
    std::vector<uint32_t> dummy;
    new_blockB(T, 0, tmax_new, dummy);  // block 0: the root's tmax_new rows, drawn by the caller
    std::vector<u64> zeros(tmax_new, 0);
    new_leafB(T, zeros.data(), tmax_new, 0);  // the root starts as a leaf for the all-zero digits
    return T;
}

QMC_EXPORT void qmc_nusb_tree_free(void *h) { delete (TreeB *)h; }
QMC_EXPORT u64 qmc_nusb_tree_rows(void *h) { return ((TreeB *)h)->rows.size() / ((TreeB *)h)->b; }
QMC_EXPORT u64 qmc_nusb_tree_rows_needed(void *h) {
    TreeB *T = (TreeB *)h;
    u64 k = 0;
    for (auto &bl : T->blocks) k += bl.rows;
    return k;
}

// Insert n points (digits at v[k*stride_pt + t*stride_dig], t < tmax) in
// order; writes the row counts of the new permutation blocks to counts
// (capacity n) and returns how many blocks were added (-1 on a digit >= b).
QMC_EXPORT int64_t qmc_nusb_insert(void *h, const u64 *v, u64 n, u64 stride_pt, u64 stride_dig, uint32_t *counts) {
    TreeB *T = (TreeB *)h;
    std::vector<uint32_t> c;
    for (u64 k = 0; k < n; k++) {
        const u64 *p = v + k * stride_pt;
        for (int t = 0; t < T->tmax; t++)
            if (p[(u64)t * stride_dig] >= T->b) {
                qmc_fail(QMC_ERR_RANGE, "nus: digit %llu is not below the base %llu",
                         (unsigned long long)p[(u64)t * stride_dig], (unsigned long long)T->b);
                return -1;
            }
        const size_t nodes = T->nodes.size();
        insertB(T, p, stride_dig, c);
        if (T->nodes.size() != nodes) T->version++;
    }
    if (!c.empty()) std::memcpy(counts, c.data(), c.size() * sizeof(uint32_t));
    return (int64_t)c.size();
}

// Append permutation rows (b values each, below b) in draw order.
QMC_EXPORT int qmc_nusb_add_rows(void *h, const uint16_t *rows, u64 nrows) {
    TreeB *T = (TreeB *)h;
    const u64 have = T->rows.size() / T->b;
    if (have + nrows > qmc_nusb_tree_rows_needed(h)) return qmc_fail(QMC_ERR_ARG, "nus: more rows than asked for");
    T->rows.insert(T->rows.end(), rows, rows + nrows * T->b);
    T->version++;
    // assign row offsets to the blocks now fully covered
    u64 acc = 0;
    for (auto &bl : T->blocks) {
        if (bl.row0 < 0 && acc + bl.rows <= have + nrows) bl.row0 = (int64_t)acc;
        acc += bl.rows;
    }
    return QMC_OK;
}


QMC_EXPORT int qmc_nusb_apply_dev(void **trees, u64 r, u64 n, u64 d, u64 r_x, u64 r_b, int tmax, int tmax_new,
                                  const u64 *bases, const u64 *xdig, u64 *xrdig, qmc_stream_t s) {
    (void)bases;  // each tree carries its base (checked against the bases when it was grown)
    if (!(r && n && d)) return QMC_OK;
    const u64 nt = r * d;
    std::vector<u64> sig;
    sig.reserve(2 * nt + 2);
    sig.push_back(r);
    sig.push_back(d);
    for (u64 k = 0; k < nt; k++) {
        TreeB *T = (TreeB *)trees[k];
        if (T->tmax != tmax || T->tmax_new != tmax_new)
            return qmc_fail(QMC_ERR_ARG, "nus: tree %llu was built with other tmax/tmax_new", (unsigned long long)k);
        if (T->rows.size() / T->b != qmc_nusb_tree_rows_needed(T))
            return qmc_fail(QMC_ERR_ARG, "nus: tree %llu is missing permutation rows", (unsigned long long)k);
        sig.push_back(T->id);
        sig.push_back(T->version);
    }
    int dev = 0;
    QMC_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_nusb_mu);
    if (g_nusb.buf == nullptr || g_nusb.device != dev || g_nusb.sig != sig) {
        std::vector<int64_t> node0(nt), kid0(nt), blk0(nt), row0(nt);
        std::vector<uint32_t> tb(nt);
        u64 nn = 0, nk = 0, nb = 0, nr = 0;
        for (u64 k = 0; k < nt; k++) {
            TreeB *T = (TreeB *)trees[k];
            node0[k] = (int64_t)nn;
            kid0[k] = (int64_t)nk;
            blk0[k] = (int64_t)nb;
            row0[k] = (int64_t)nr;
            tb[k] = (uint32_t)T->b;
            nn += T->nodes.size();
            nk += T->kids.size();
            nb += T->blocks.size();
            nr += T->rows.size();
        }
        std::vector<int32_t> kids_off(nn), owner(nn), kids(nk > 0 ? nk : 1), bdepth(nb);
        std::vector<int64_t> brow(nb);
        std::vector<uint16_t> rows(nr > 0 ? nr : 1);
        for (u64 k = 0; k < nt; k++) {
            TreeB *T = (TreeB *)trees[k];
            for (size_t q = 0; q < T->nodes.size(); q++) {
                kids_off[node0[k] + q] = T->nodes[q].kids;
                owner[node0[k] + q] = T->nodes[q].owner;
            }
            for (size_t q = 0; q < T->kids.size(); q++) kids[kid0[k] + q] = T->kids[q] < 0 ? 0 : T->kids[q];
            for (size_t q = 0; q < T->blocks.size(); q++) {
                brow[blk0[k] + q] = T->blocks[q].row0;
                bdepth[blk0[k] + q] = T->blocks[q].depth0;
            }
            if (!T->rows.empty()) std::memcpy(&rows[row0[k]], T->rows.data(), T->rows.size() * sizeof(uint16_t));
        }
        if (g_nusb.buf) {
            (void)hipFree(g_nusb.buf);
            g_nusb.buf = nullptr;
        }
        g_nusb.sig.clear();
        const size_t total = kids_off.size() * 4 + owner.size() * 4 + kids.size() * 4 + brow.size() * 8 +
                             bdepth.size() * 4 + rows.size() * 2 + 4 * nt * 8 + nt * 4 + 128;
        char *buf = nullptr;
        if (hipMalloc((void **)&buf, total) != hipSuccess)
            return qmc_fail(QMC_ERR_NOMEM, "nus: device allocation of %zu bytes for the trees failed", total);
        size_t o = 0;
        hipError_t ce = hipSuccess;
        auto put = [&](const void *src, size_t bytes) -> void * {
            o = (o + 7) & ~(size_t)7;
            void *dst = buf + o;
            if (bytes && ce == hipSuccess) ce = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
            o += bytes;
            return dst;
        };
        PackedB P;
        P.kids_off = (const int32_t *)put(kids_off.data(), kids_off.size() * 4);
        P.owner = (const int32_t *)put(owner.data(), owner.size() * 4);
        P.kids = (const int32_t *)put(kids.data(), kids.size() * 4);
        P.blk_row0 = (const int64_t *)put(brow.data(), brow.size() * 8);
        P.blk_depth0 = (const int32_t *)put(bdepth.data(), bdepth.size() * 4);
        P.rows = (const uint16_t *)put(rows.data(), rows.size() * 2);
        P.node0 = (const int64_t *)put(node0.data(), nt * 8);
        P.kid0 = (const int64_t *)put(kid0.data(), nt * 8);
        P.blk0 = (const int64_t *)put(blk0.data(), nt * 8);
        P.row0 = (const int64_t *)put(row0.data(), nt * 8);
        P.tb = (const uint32_t *)put(tb.data(), nt * 4);
        if (ce != hipSuccess) {
            (void)hipFree(buf);
            return qmc_fail_hip(ce, "nus: copying the trees");
        }
        g_nusb.buf = buf;
        g_nusb.P = P;
        g_nusb.device = dev;
        g_nusb.sig = sig;
    }
    hipLaunchKernelGGL(nusb_apply_kernel, dim3(grid_for(r * n * d, kBlock, 2048 * 8)), dim3(kBlock), 0, s, r, n, d,
                       r_x, r_b, tmax, tmax_new, (u64)0, g_nusb.P, xdig, xrdig);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return qmc_fail_hip(e, "nusb_apply_kernel");
    return QMC_OK;
}

}  // extern "C"
