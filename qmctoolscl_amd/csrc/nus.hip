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

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>

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
//
// A reference leaf holds tmax_new - depth permutation rows of b entries
// (random_uint64_permutations, rand_funcs.py:22-32), drawn when the leaf
// appears: at n = 3^12 points, 52 digits and bases up to 131 that is
// terabytes per point set.  Here a tree materializes only the row of each
// internal node (its own level) and keeps, per leaf, the state of the tree's
// generator where the leaf's rows begin (5 words); a point that ends at the
// leaf regenerates those rows on the device.  For numpy's SFC64 the builder
// replays the generator itself (Sfc64); for any other bit generator the caller
// draws every block's rows with numpy and all of them are materialized.

// numpy's SFC64 bit generator (C. Doty-Humphrey's Small Fast Chaotic PRNG:
// 256-bit state, one 64-bit output per step) with numpy's 32-bit buffering (a
// 64-bit output serves two 32-bit draws, low half first), and the draws of
// Generator.permutation(b): a Fisher-Yates shuffle of 0..b-1, for i = b-1 down
// to 1 swapping entries i and j = random_interval(i) (a 32-bit draw masked to
// the bit width of i, rejected while above i).  Checked on the host against
// numpy's own Generator (qmc_nus_sfc64_permutations).
struct Sfc64 {
    u64 s[4];
    u32 has, buf;
    __host__ __device__ __forceinline__ u64 next64() {
        const u64 t = s[0] + s[1] + s[3]++;
        s[0] = s[1] ^ (s[1] >> 11);
        s[1] = s[2] + (s[2] << 3);
        s[2] = ((s[2] << 24) | (s[2] >> 40)) + t;
        return t;
    }
    __host__ __device__ __forceinline__ u32 next32() {
        if (has) {
            has = 0;
            return buf;
        }
        const u64 v = next64();
        has = 1;
        buf = (u32)(v >> 32);
        return (u32)v;
    }
    // random_interval(mx) for 1 <= mx < 2^32
    __host__ __device__ __forceinline__ u32 interval(u32 mx) {
        u32 mask = mx;
        mask |= mask >> 1;
        mask |= mask >> 2;
        mask |= mask >> 4;
        mask |= mask >> 8;
        mask |= mask >> 16;
        u32 v;
        do {
            v = next32() & mask;
        } while (v > mx);
        return v;
    }
    // packed as 5 words: s[0..3], has << 32 | buf
    __host__ __device__ void load(const u64 *w) {
        for (int k = 0; k < 4; k++) s[k] = w[k];
        has = (u32)(w[4] >> 32);
        buf = (u32)w[4];
    }
    __host__ __device__ void store(u64 *w) const {
        for (int k = 0; k < 4; k++) w[k] = s[k];
        w[4] = ((u64)has << 32) | buf;
    }
};

// one Generator.permutation(b) into out[0..b)
void sfc_perm(Sfc64 &g, u32 b, uint16_t *out) {
    for (u32 k = 0; k < b; k++) out[k] = (uint16_t)k;
    for (u32 i = b - 1; i >= 1; i--) {
        const u32 j = g.interval(i);
        const uint16_t a = out[i];
        out[i] = out[j];
        out[j] = a;
    }
}
// The draws of `rows` permutations of b, not kept (most rows: the leaves'
// rows are regenerated on the device).  Written as a scan over the stream of
// 32-bit values (low half of each 64-bit output first) without data-dependent
// branches: a value is accepted when (v & mask) <= i, which moves i down by
// one (and halves mask when i falls to half of it); a row ends when i reaches
// 0.  The same values are consumed as by Sfc64::interval, about four times
// faster (a rejected draw no longer costs a mispredicted branch).
void sfc_skip(Sfc64 &g, u32 b, u64 rows) {
    if (b < 2 || rows == 0) return;
    u64 s0 = g.s[0], s1 = g.s[1], s2 = g.s[2], s3 = g.s[3];
    u32 full = b - 1;
    full |= full >> 1;
    full |= full >> 2;
    full |= full >> 4;
    full |= full >> 8;
    full |= full >> 16;
    u32 i = b - 1, mask = full;
    u64 left = rows;
    auto take = [&](u32 v) -> bool {  // true when the last row ends
        i -= (v & mask) <= i ? 1u : 0u;
        mask = i <= (mask >> 1) ? mask >> 1 : mask;
        if (i == 0) {
            if (--left == 0) return true;
            i = b - 1;
            mask = full;
        }
        return false;
    };
    if (g.has) {
        g.has = 0;
        if (take(g.buf)) return;
    }
    for (;;) {
        const u64 t = s0 + s1 + s3++;
        s0 = s1 ^ (s1 >> 11);
        s1 = s2 + (s2 << 3);
        s2 = ((s2 << 24) | (s2 >> 40)) + t;
        if (take((u32)t)) {
            g.has = 1;
            g.buf = (u32)(t >> 32);
            break;
        }
        if (take((u32)(t >> 32))) {
            g.has = 0;  // numpy leaves the used high half in its buffer word
            g.buf = (u32)(t >> 32);
            break;
        }
    }
    g.s[0] = s0;
    g.s[1] = s1;
    g.s[2] = s2;
    g.s[3] = s3;
}

struct NodeB {
    int32_t kids;  // offset of b child slots in TreeB::kids, -1 for a leaf
    int32_t owner; // block (leaf creation) whose rows the node uses: row (depth - depth0) at its level
    int64_t rest;  // its digits at levels depth .. tmax-1 in TreeB::digits (later levels are 0)
    int32_t row;   // materialized row of its own level, or -1 (leaves of blocks that are not materialized)
    int32_t st;    // leaf of a replayed block: generator state where its rows begin, or -1
};

struct Block {
    int32_t depth0, rows;  // first level and number of its rows (rand_funcs.py:327, 345: tmax_new - t - 1)
    int32_t first;         // the leaf created with it
    int32_t mat0;          // first of its rows in TreeB::rows when all of them are materialized, or -1
};

struct TreeB {
    int tmax, tmax_new;
    u64 b;                        // <= 65536, so a digit fits 16 bits
    u64 id, version;              // identity and mutation count: the device copy is reused while both match
    std::vector<NodeB> nodes;
    std::vector<int32_t> kids;
    std::vector<uint16_t> digits;  // leaf digit pool
    std::vector<Block> blocks;    // in draw order
    size_t ready = 0;             // blocks whose rows are drawn (replayed or supplied)
    u64 supplied = 0;             // rows received through qmc_nusb_add_rows (numpy mode) not yet assigned
    std::vector<uint16_t> pending;  // those rows
    std::vector<uint16_t> rows;   // materialized rows, b entries each
    std::vector<u64> states;      // leaf generator states, 5 words each
};

std::atomic<u64> g_treeb_ids{1};

int32_t push_rowB(TreeB *T, Sfc64 &g) {
    const size_t o = T->rows.size();
    T->rows.resize(o + T->b);
    sfc_perm(g, (u32)T->b, &T->rows[o]);
    return (int32_t)(o / T->b);
}

int32_t push_stateB(TreeB *T, const Sfc64 &g) {
    const size_t o = T->states.size();
    T->states.resize(o + 5);
    g.store(&T->states[o]);
    return (int32_t)(o / 5);
}

inline u32 digit_atB(const TreeB *T, const NodeB &N, int depth) {
    return depth < T->tmax ? T->digits[N.rest] : 0u;
}

// a new leaf at `depth` for the point pd (its digits at levels >= depth), with
// a new block of tmax_new - depth rows
int32_t new_leafB(TreeB *T, int depth, const uint16_t *pd, std::vector<uint32_t> *counts) {
    Block bl;
    bl.depth0 = depth;
    bl.rows = T->tmax_new - depth;
    bl.first = (int32_t)T->nodes.size();
    bl.mat0 = -1;
    T->blocks.push_back(bl);
    if (counts) counts->push_back((uint32_t)bl.rows);
    NodeB nd;
    nd.kids = -1;
    nd.owner = (int32_t)T->blocks.size() - 1;
    nd.rest = (int64_t)T->digits.size();
    nd.row = -1;
    nd.st = -1;
    if (depth < T->tmax) T->digits.insert(T->digits.end(), pd + depth, pd + T->tmax);
    T->nodes.push_back(nd);
    return bl.first;
}

// The leaf `node` at depth t becomes internal: its contents move one level
// down (rand_funcs.py:336-340), keeping the owner block.  Its row for level t
// comes from the materialized block, or is drawn now from the leaf's stored
// state (the moved leaf keeps the state after it); for a block whose rows are
// not drawn yet, the replay resolves the chain.
void split_leafB(TreeB *T, int32_t node, int t, u32 ndg) {
    NodeB moved;
    moved.kids = -1;
    moved.owner = T->nodes[node].owner;
    moved.rest = T->nodes[node].rest + 1;
    moved.row = -1;
    moved.st = -1;
    const Block &B = T->blocks[moved.owner];
    if (B.mat0 >= 0) {
        T->nodes[node].row = B.mat0 + (t - B.depth0);
        moved.row = T->nodes[node].row + 1;
    } else if (T->nodes[node].st >= 0) {
        Sfc64 g;
        g.load(&T->states[5 * (size_t)T->nodes[node].st]);
        T->nodes[node].row = push_rowB(T, g);
        T->nodes[node].st = -1;
        moved.st = push_stateB(T, g);
    }
    const int32_t mid = (int32_t)T->nodes.size();
    T->nodes.push_back(moved);
    const int32_t koff = (int32_t)T->kids.size();
    T->kids.insert(T->kids.end(), (size_t)T->b, -1);
    T->nodes[node].kids = koff;
    T->kids[koff + ndg] = mid;
}

// rand_funcs.py:318-353 for one point: pd holds its tmax digits (each < b).
// Where the reference's loop would read past the last digit (tmax ==
// tmax_new, rand_funcs.py:323-325, an IndexError there) the point's path is
// already complete and the insertion stops.
void insertB(TreeB *T, const uint16_t *pd, std::vector<uint32_t> *counts) {
    const int tn = T->tmax_new, tm = T->tmax;
    int32_t node = 0;
    for (int t = 0; t <= tm && t < tn; t++) {
        const u32 dg = t < tm ? pd[t] : 0u;
        if (T->nodes[node].kids >= 0) {
            const int32_t kid = T->kids[T->nodes[node].kids + dg];
            if (kid < 0) {
                const int32_t k = new_leafB(T, t + 1, pd, counts);
                T->kids[T->nodes[node].kids + dg] = k;
                return;
            }
            node = kid;
            continue;
        }
        const int nd = tm - t;  // digits the leaf holds (levels t .. tmax-1)
        if (nd <= 0 || std::memcmp(&T->digits[T->nodes[node].rest], pd + t, 2 * (size_t)nd) == 0)
            return;  // seen before
        const u32 ndg = T->digits[T->nodes[node].rest];
        split_leafB(T, node, t, ndg);
        if (ndg == dg) {
            node = T->kids[T->nodes[node].kids + dg];
            continue;
        }
        const int32_t k = new_leafB(T, t + 1, pd, counts);
        T->kids[T->nodes[node].kids + dg] = k;
        return;
    }
}

// All rows of block B are materialized from row m0 on: give every node of its
// chain (its first leaf, pushed down by later splits) its level's row.
void assign_chainB(TreeB *T, Block &B, int32_t m0) {
    B.mat0 = m0;
    int32_t node = B.first;
    for (int D = B.depth0;; D++) {
        NodeB &N = T->nodes[node];
        N.row = m0 + (D - B.depth0);
        if (N.kids < 0) break;
        node = T->kids[N.kids + digit_atB(T, N, D)];
    }
}

// Draw the rows of the blocks not drawn yet, in creation order, from g: the
// root's block is materialized whole (the reference draws it when the tree
// starts and keeps it as root_node.perm); in every other block the rows of
// the chain's internal nodes are materialized and the leaf where the chain
// ends keeps the state at its first row, the remaining rows only advancing g.
void replayB(TreeB *T, Sfc64 &g) {
    const u32 b = (u32)T->b;
    for (; T->ready < T->blocks.size(); T->ready++) {
        Block &B = T->blocks[T->ready];
        if (T->ready == 0) {
            int32_t m0 = -1;
            for (int k = 0; k < B.rows; k++) {
                const int32_t rw = push_rowB(T, g);
                if (k == 0) m0 = rw;
            }
            if (m0 < 0) m0 = (int32_t)(T->rows.size() / b);
            assign_chainB(T, B, m0);
            continue;
        }
        int32_t node = B.first;
        int D = B.depth0;
        for (int k = 0; k < B.rows; k++, D++) {
            NodeB &N = T->nodes[node];
            if (N.kids < 0) {
                N.st = push_stateB(T, g);
                sfc_skip(g, b, (u64)(B.rows - k));
                break;
            }
            N.row = push_rowB(T, g);
            node = T->kids[N.kids + digit_atB(T, N, D)];
        }
    }
    T->version++;
}

// numpy mode: the rows of the blocks not drawn yet arrive in draw order;
// every block they cover is materialized.
void take_rowsB(TreeB *T) {
    const u64 b = T->b;
    size_t used = 0;
    while (T->ready < T->blocks.size()) {
        Block &B = T->blocks[T->ready];
        const size_t need = (size_t)B.rows * b;
        if (T->pending.size() - used < need) break;
        const int32_t m0 = (int32_t)(T->rows.size() / b);
        T->rows.insert(T->rows.end(), T->pending.begin() + used, T->pending.begin() + used + need);
        used += need;
        assign_chainB(T, B, m0);
        T->ready++;
    }
    T->pending.erase(T->pending.begin(), T->pending.begin() + used);
    T->version++;
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
    const int32_t *kids_off;  // per node, -1 for a leaf
    const int32_t *row;       // per node: materialized row of its level (tree-local), or -1
    const int32_t *st;        // per node: leaf state (tree-local), or -1
    const int32_t *kids;      // child slots (tree-local node ids)
    const uint16_t *rows;     // b values per row, each tree's rows packed
    const u64 *states;        // 5 words per state
    const int64_t *node0, *kid0, *row0, *st0;  // per tree offsets (row0 in values, st0 in states)
    const uint32_t *tb;       // per tree base
};

constexpr int kNusbLanes = 64;

// One wave per workgroup; point e = tree * n + i, so the lanes of a wave take
// consecutive points of one tree.  A point walks its tree from the root: an
// internal node gives the permuted digit of its level from its row
// (rand_funcs.py:320-327); at its leaf the remaining levels come from the
// leaf's rows (rand_funcs.py:330-331), read from a materialized block or
// regenerated from the leaf's state, one Generator.permutation(b) per level.
// Only entry x of a regenerated permutation is needed: the shuffle's entry x
// is x sent through its swaps (i <-> j_i) in reverse order of drawing, so the
// draws j_{b-1} .. j_1 go to a per-lane buffer (jb, one u16 per lane and
// level, lanes interleaved) and are read back from j_1 up.
template <bool LBUF>
__global__ __launch_bounds__(kNusbLanes) void nusb_apply_kernel(u64 r, u64 n, u64 d, u64 r_x, int tmax, int tmax_new,
                                                              u64 bmax, PackedB P, const u64 *__restrict__ xdig,
                                                              u64 *__restrict__ xrdig, uint16_t *scratch) {
    extern __shared__ uint16_t lbuf[];
    uint16_t *jb = (LBUF ? lbuf : scratch + (u64)blockIdx.x * kNusbLanes * bmax) + threadIdx.x;
    const u64 total = r * d * n;
    for (u64 e = (u64)blockIdx.x * kNusbLanes + threadIdx.x; e < total; e += (u64)gridDim.x * kNusbLanes) {
        const u64 tree = e / n, i = e - tree * n;
        const u64 l = tree / d, j = tree - l * d;
        const u64 *src = xdig + (((l % r_x) * n + i) * d + j) * (u64)tmax;
        u64 *out = xrdig + ((l * n + i) * d + j) * (u64)tmax_new;
        const int64_t nb = P.node0[tree], kb = P.kid0[tree];
        const uint16_t *rows = P.rows + P.row0[tree];
        const u32 b = P.tb[tree];
        int64_t node = 0;
        for (int t = 0; t < tmax_new; t++) {
            const u32 dg = t < tmax ? (u32)src[t] : 0u;
            const int32_t ko = P.kids_off[nb + node];
            if (ko >= 0) {
                out[t] = rows[(u64)P.row[nb + node] * b + dg];
                node = P.kids[kb + ko + dg];
                continue;
            }
            const int32_t st = P.st[nb + node];
            if (st < 0) {  // materialized rows, consecutive from the leaf's own level
                const u64 rw = (u64)P.row[nb + node];
                for (int tt = t; tt < tmax_new; tt++) {
                    const u32 x = tt < tmax ? (u32)src[tt] : 0u;
                    out[tt] = rows[(rw + (u64)(tt - t)) * b + x];
                }
            } else {
                Sfc64 g;
                g.load(P.states + (u64)(P.st0[tree] + st) * 5);
                for (int tt = t; tt < tmax_new; tt++) {
                    for (u32 q = b - 1; q >= 1; q--) jb[(u64)q * kNusbLanes] = (uint16_t)g.interval(q);
                    u32 p = tt < tmax ? (u32)src[tt] : 0u;
                    for (u32 q = 1; q < b; q++) {
                        const u32 jq = jb[(u64)q * kNusbLanes];
                        p = p == q ? jq : (p == jq ? q : p);
                    }
                    out[tt] = p;
                }
            }
            break;
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
    // the root starts as a leaf for the all-zero digits, with block 0: its
    // tmax_new rows (rand_funcs.py:312-317)
    std::vector<uint16_t> zeros(tmax > 0 ? tmax : 1, 0);
    new_leafB(T, 0, zeros.data(), nullptr);
    return T;
}

QMC_EXPORT void qmc_nusb_tree_free(void *h) { delete (TreeB *)h; }
QMC_EXPORT u64 qmc_nusb_tree_rows(void *h) { return ((TreeB *)h)->rows.size() / ((TreeB *)h)->b; }
QMC_EXPORT u64 qmc_nusb_tree_rows_needed(void *h) {
    TreeB *T = (TreeB *)h;
    u64 k = 0;
    for (size_t q = T->ready; q < T->blocks.size(); q++) k += (u64)T->blocks[q].rows;
    return k - T->pending.size() / T->b;
}
QMC_EXPORT u64 qmc_nusb_tree_states(void *h) { return ((TreeB *)h)->states.size() / 5; }
QMC_EXPORT u64 qmc_nusb_tree_nodes(void *h) { return ((TreeB *)h)->nodes.size(); }

// Host bytes a tree holds (nodes, child slots, leaf digits, materialized rows,
// leaf states).
QMC_EXPORT u64 qmc_nusb_tree_bytes(void *h) {
    TreeB *T = (TreeB *)h;
    return T->nodes.size() * sizeof(NodeB) + T->kids.size() * 4 + T->digits.size() * 2 +
           T->blocks.size() * sizeof(Block) + T->rows.size() * 2 + T->states.size() * 8;
}

// Insert n points (digits at v[k*stride_pt + t*stride_dig], t < tmax) in
// order; writes the row counts of the new blocks to counts (capacity n) and
// returns how many blocks were added (-1 on a digit >= b).
QMC_EXPORT int64_t qmc_nusb_insert(void *h, const u64 *v, u64 n, u64 stride_pt, u64 stride_dig, uint32_t *counts) {
    TreeB *T = (TreeB *)h;
    std::vector<uint32_t> c;
    std::vector<uint16_t> pd(T->tmax > 0 ? T->tmax : 1);
    for (u64 k = 0; k < n; k++) {
        const u64 *p = v + k * stride_pt;
        for (int t = 0; t < T->tmax; t++) {
            const u64 x = p[(u64)t * stride_dig];
            if (x >= T->b) {
                qmc_fail(QMC_ERR_RANGE, "nus: digit %llu is not below the base %llu", (unsigned long long)x,
                         (unsigned long long)T->b);
                return -1;
            }
            pd[t] = (uint16_t)x;
        }
        const size_t nodes0 = T->nodes.size();
        insertB(T, pd.data(), &c);
        if (T->nodes.size() != nodes0) T->version++;
    }
    if (!c.empty()) std::memcpy(counts, c.data(), c.size() * sizeof(uint32_t));
    return (int64_t)c.size();
}

// numpy mode: append permutation rows (b values each, below b) in draw order;
// every block they complete is materialized.
QMC_EXPORT int qmc_nusb_add_rows(void *h, const uint16_t *rows, u64 nrows) {
    TreeB *T = (TreeB *)h;
    if (nrows > qmc_nusb_tree_rows_needed(h)) return qmc_fail(QMC_ERR_ARG, "nus: more rows than asked for");
    T->pending.insert(T->pending.end(), rows, rows + nrows * T->b);
    take_rowsB(T);
    return QMC_OK;
}

// The root's tmax_new rows (block 0) once drawn, b values each: the
// reference's root_node.perm (rand_funcs.py:316).
QMC_EXPORT int qmc_nusb_root_rows(void *h, uint16_t *out) {
    TreeB *T = (TreeB *)h;
    if (T->ready == 0 || T->blocks[0].mat0 < 0) return qmc_fail(QMC_ERR_ARG, "nus: the root's rows are not drawn");
    std::memcpy(out, &T->rows[(size_t)T->blocks[0].mat0 * T->b], (size_t)T->blocks[0].rows * T->b * 2);
    return QMC_OK;
}

// One call's host work for nt = r*d trees: tree l*d + j takes the points
// xdig[(l % r_x), i, j, 0 .. tmax) for i < n, in order; then, where replay[k]
// is set, the rows of its new blocks are drawn from the SFC64 state
// states[5k .. 5k+5) (numpy's s[0..3], has_uint32 << 32 | uinteger), which is
// advanced in place exactly as the reference's draws would advance it.
// Trees are independent: up to `threads` host threads take them in turn.
QMC_EXPORT int qmc_nusb_grow(void **trees, u64 nt, const u64 *xdig, u64 r, u64 n, u64 d, u64 r_x, u64 *states,
                             const uint8_t *replay, int threads) {
    if (nt != r * d) return qmc_fail(QMC_ERR_ARG, "nus: %llu trees for r*d = %llu", (unsigned long long)nt,
                                     (unsigned long long)(r * d));
    // largest bases first: a tree's replay costs about (b - 1) draws per row
    std::vector<u64> order(nt);
    for (u64 k = 0; k < nt; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](u64 a, u64 c) { return ((TreeB *)trees[a])->b > ((TreeB *)trees[c])->b; });
    std::atomic<u64> next{0};
    std::atomic<int> bad{0};
    std::atomic<u64> bad_digit{0}, bad_tree{0};
    auto work = [&]() {
        for (;;) {
            const u64 q = next++;
            if (q >= nt || bad.load()) return;
            const u64 k = order[q];
            TreeB *T = (TreeB *)trees[k];
            const u64 l = k / d, j = k % d;
            const u64 tm = (u64)T->tmax;
            const size_t nodes0 = T->nodes.size(), blocks0 = T->blocks.size();
            std::vector<uint16_t> pd(tm > 0 ? tm : 1);
            for (u64 i = 0; i < n; i++) {
                const u64 *p = xdig + (((l % r_x) * n + i) * d + j) * tm;
                for (u64 t = 0; t < tm; t++) {
                    if (p[t] >= T->b) {
                        bad_digit = p[t];
                        bad_tree = k;
                        bad = 1;
                        return;
                    }
                    pd[t] = (uint16_t)p[t];
                }
                insertB(T, pd.data(), nullptr);
            }
            // the device copy is kept while no tree changes (a repeat call on the same points)
            if (T->nodes.size() != nodes0 || T->blocks.size() != blocks0) T->version++;
            if (replay && replay[k] && T->ready < T->blocks.size()) {
                Sfc64 g;
                g.load(states + 5 * k);
                replayB(T, g);
                g.store(states + 5 * k);
            }
        }
    };
    const int nth = (int)std::max<u64>(1, std::min<u64>((u64)std::max(threads, 1), nt));
    std::vector<std::thread> pool;
    for (int q = 1; q < nth; q++) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    if (bad.load())
        return qmc_fail(QMC_ERR_RANGE, "nus: digit %llu of tree %llu is not below its base",
                        (unsigned long long)bad_digit.load(), (unsigned long long)bad_tree.load());
    return QMC_OK;
}

// count Generator.permutation(b) draws from the SFC64 state w (5 words,
// advanced in place) into out (count x b): the replay the builder uses
QMC_EXPORT void qmc_nus_sfc64_permutations(u64 *w, u64 b, u64 count, uint16_t *out) {
    Sfc64 g;
    g.load(w);
    for (u64 k = 0; k < count; k++) sfc_perm(g, (u32)b, out + k * b);
    g.store(w);
}

QMC_EXPORT int qmc_nusb_apply_dev(void **trees, u64 r, u64 n, u64 d, u64 r_x, u64 r_b, int tmax, int tmax_new,
                                  const u64 *bases, const u64 *xdig, u64 *xrdig, qmc_stream_t s) {
    (void)bases;  // each tree carries its base (checked against the bases when it was grown)
    (void)r_b;
    if (!(r && n && d)) return QMC_OK;
    const u64 nt = r * d;
    std::vector<u64> sig;
    sig.reserve(2 * nt + 2);
    sig.push_back(r);
    sig.push_back(d);
    u64 bmax = 1;
    for (u64 k = 0; k < nt; k++) {
        TreeB *T = (TreeB *)trees[k];
        if (T->tmax != tmax || T->tmax_new != tmax_new)
            return qmc_fail(QMC_ERR_ARG, "nus: tree %llu was built with other tmax/tmax_new", (unsigned long long)k);
        if (T->ready != T->blocks.size())
            return qmc_fail(QMC_ERR_ARG, "nus: tree %llu is missing permutation rows", (unsigned long long)k);
        sig.push_back(T->id);
        sig.push_back(T->version);
        bmax = std::max<u64>(bmax, T->b);
    }
    int dev = 0;
    QMC_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_nusb_mu);
    if (g_nusb.buf == nullptr || g_nusb.device != dev || g_nusb.sig != sig) {
        std::vector<int64_t> node0(nt), kid0(nt), row0(nt), st0(nt);
        std::vector<uint32_t> tb(nt);
        u64 nn = 0, nk = 0, nr = 0, ns = 0;
        for (u64 k = 0; k < nt; k++) {
            TreeB *T = (TreeB *)trees[k];
            node0[k] = (int64_t)nn;
            kid0[k] = (int64_t)nk;
            row0[k] = (int64_t)nr;
            st0[k] = (int64_t)ns;
            tb[k] = (uint32_t)T->b;
            nn += T->nodes.size();
            nk += T->kids.size();
            nr += T->rows.size();
            ns += T->states.size() / 5;
        }
        std::vector<int32_t> kids_off(nn), row(nn), st(nn), kids(nk > 0 ? nk : 1);
        std::vector<uint16_t> rows(nr > 0 ? nr : 1);
        std::vector<u64> states(ns > 0 ? 5 * ns : 5);
        for (u64 k = 0; k < nt; k++) {
            TreeB *T = (TreeB *)trees[k];
            for (size_t q = 0; q < T->nodes.size(); q++) {
                kids_off[node0[k] + q] = T->nodes[q].kids;
                row[node0[k] + q] = T->nodes[q].row;
                st[node0[k] + q] = T->nodes[q].st;
            }
            for (size_t q = 0; q < T->kids.size(); q++) kids[kid0[k] + q] = T->kids[q] < 0 ? 0 : T->kids[q];
            if (!T->rows.empty()) std::memcpy(&rows[row0[k]], T->rows.data(), T->rows.size() * sizeof(uint16_t));
            if (!T->states.empty()) std::memcpy(&states[5 * st0[k]], T->states.data(), T->states.size() * 8);
        }
        if (g_nusb.buf) {
            (void)hipFree(g_nusb.buf);
            g_nusb.buf = nullptr;
        }
        g_nusb.sig.clear();
        const size_t total = (kids_off.size() + row.size() + st.size() + kids.size()) * 4 + rows.size() * 2 +
                             states.size() * 8 + 4 * nt * 8 + nt * 4 + 128;
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
        P.row = (const int32_t *)put(row.data(), row.size() * 4);
        P.st = (const int32_t *)put(st.data(), st.size() * 4);
        P.kids = (const int32_t *)put(kids.data(), kids.size() * 4);
        P.rows = (const uint16_t *)put(rows.data(), rows.size() * 2);
        P.states = (const u64 *)put(states.data(), states.size() * 8);
        P.node0 = (const int64_t *)put(node0.data(), nt * 8);
        P.kid0 = (const int64_t *)put(kid0.data(), nt * 8);
        P.row0 = (const int64_t *)put(row0.data(), nt * 8);
        P.st0 = (const int64_t *)put(st0.data(), nt * 8);
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
    // the per-lane draw buffers: 64 lanes x bmax u16 per workgroup, in LDS up
    // to 64 KB, else in a device scratch of at most 512 MB
    const u64 total = r * n * d;
    const size_t per_wg = (size_t)kNusbLanes * bmax * 2;
    hipError_t e = hipSuccess;
    if (per_wg <= 64 * 1024) {
        const unsigned grid = grid_for(total, kNusbLanes, 4096);
        hipLaunchKernelGGL((nusb_apply_kernel<true>), dim3(grid), dim3(kNusbLanes), per_wg, s, r, n, d, r_x, tmax,
                           tmax_new, bmax, g_nusb.P, xdig, xrdig, (uint16_t *)nullptr);
        e = hipGetLastError();
    } else {
        const unsigned grid =
            (unsigned)std::max<u64>(1, std::min<u64>(grid_for(total, kNusbLanes, 4096), (512ull << 20) / per_wg));
        uint16_t *scratch = nullptr;
        e = hipMallocAsync((void **)&scratch, (size_t)grid * per_wg, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL((nusb_apply_kernel<false>), dim3(grid), dim3(kNusbLanes), 0, s, r, n, d, r_x, tmax,
                               tmax_new, bmax, g_nusb.P, xdig, xrdig, scratch);
            e = hipGetLastError();
            const hipError_t f = hipFreeAsync(scratch, s);
            if (e == hipSuccess) e = f;
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return qmc_fail_hip(e, "nusb_apply_kernel");
    return QMC_OK;
}

}  // extern "C"
