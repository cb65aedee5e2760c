// Ragged batches of point-generation requests in one persistent kernel.
//
// A request asks for r x n x d float64 points (n = 2^m, natural order) of one
// method:
//   0 LATTICE    rank-1 lattice + r random shifts mod 1
//                 (lat_gen_natural -> lat_shift_mod_1)
//   1 DNB2_LMS   base-2 net, 32x32 upper-triangular generating matrices, a
//                 random lower-triangular LMS per replication, 64-bit digital
//                 shift, float (dnb2_linear_matrix_scramble -> dnb2_gen_natural
//                 -> dnb2_digital_shift -> dnb2_integer_to_float)
//   2 DNB2_NUS   base-2 net with a hash-based nested uniform scramble
//   3 GDN_LMS    Halton-type base-b net (base = j-th prime), random
//                 lower-triangular LMS, random digit permutation per level,
//                 float
//
// Random inputs are counter-based: every generating vector, matrix, shift and
// permutation is a hash of (request seed, role, indices), so a batch needs no
// host-drawn arrays and any tile can be produced independently.  This is a
// stream of its own -- the points equal this library's one-launch-per-request
// path bit for bit, and for methods 0 and 1 they equal the reference kernels
// applied to the same inputs (qmc_batch_*_inputs export them), but they are not
// numpy's draws.
//
// Scheduling: each request is cut into tiles of whole rows (about P points
// each); a setup kernel builds every request's small tables (prefix-XOR
// columns, shifts, LMS rows, permutation parameters) in a workspace, then
// persistent grids drain tile counters: one kernel for the lattice and dnb2
// tiles (store-bound, few registers) and one for the gdn tiles (digit walks,
// ~110 registers, ALU-bound), each compiled without the other's code; when a
// batch has both, the gdn kernel runs on a second stream beside the other
// (one graph either way).  Outputs go to a ring buffer:
// every replication slab (n*d doubles) of every request is placed contiguously
// at an offset chosen on the host, wrapping to the start when the ring is full,
// so batches far larger than HBM can stream through a bounded buffer.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qmc_common.hpp"

namespace {

enum { M_LATTICE = 0, M_DNB2_LMS = 1, M_DNB2_NUS = 2, M_GDN_LMS = 3 };

constexpr int kTB = 32;     // base-2 digits (tmax) of the dnb2 methods
constexpr int kMaxChainLog = 12;  // dnb2 chains start in the tile's first 2^12 rows at most
constexpr int kMaxGroupLog = 8;   // dnb2 chains have at most 8 * 2^(8-1) = 1024 points
constexpr int kTNmax = 52;  // base-b digits kept (2^-52 resolution)

struct Req {
    int32_t method, m, d, r;
    u64 seed;
    u64 tile0;           // first global tile
    u32 rows_per_tile;   // power of two, <= n
    u32 tiles_per_l;     // n / rows_per_tile
    u64 tab;             // word offset of the request's tables in the workspace
    u64 slabs;           // word offset of r slab bases (output element offsets)
    u32 clog;            // dnb2: log2 of the row stride between a chain's points
    u32 gng;             // gdn: lane groups (ranges of consecutive columns)
    u32 gunits;          // gdn: units of work per tile (sum over groups)
    u64 ggrp;            // gdn: word offset (in the request's tables) of the group table
    u32 gtw;             // gdn: u32 words of one replication's digit tables
    u64 gtab;            // gdn: word offset (in the request's tables) of the replication tables
    u32 slab0;           // index of the request's first replication among all slabs (digests)
    DivU32 ddiv;         // division by d
};

// ------------------------------------------------------------------ hashing

__host__ __device__ __forceinline__ u64 mix64(u64 z) {
    // SplitMix64 finalizer (Steele, Lea, Flood 2014)
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ u64 key(u64 seed, u64 role, u64 a, u64 b) {
    return mix64(mix64(mix64(seed ^ (role * 0x9e3779b97f4a7c15ull)) + a) + b);
}

// Nested uniform scramble of the 64 bits of x (most significant first):
// output bit k = input bit k XOR f(bits above k).  Built, as in Burley's
// hash-based Owen scrambling (JCGT 2020), on the bit-reversed value, where an
// addition and xor-multiplies by even constants only carry information from
// lower to higher bits.
__host__ __device__ __forceinline__ u64 brev64(u64 x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __brevll(x);
#else
    x = ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0f0f0f0f0f0f0f0full) | ((x & 0x0f0f0f0f0f0f0f0full) << 4);
    x = ((x >> 8) & 0x00ff00ff00ff00ffull) | ((x & 0x00ff00ff00ff00ffull) << 8);
    x = ((x >> 16) & 0x0000ffff0000ffffull) | ((x & 0x0000ffff0000ffffull) << 16);
    return (x >> 32) | (x << 32);
#endif
}

__host__ __device__ __forceinline__ u64 owen_hash64(u64 x, u64 seed) {
    x = brev64(x);
    x += seed;
    x ^= x * 0x6c50b47cc8f83c2eull;
    x ^= x * 0xb82f1e52a6d41b3cull;
    x ^= x * 0xc7afe638e2d7f0a6ull;
    x ^= x * 0x8d22f6e6b1f5a3d4ull;
    return brev64(x);
}

// ------------------------------------------------- per-method random inputs

// lattice: generating vector (odd in [1, n)) and shifts in [0, 1)
__host__ __device__ __forceinline__ u64 lat_g(u64 seed, int m, u64 j) {
    return 2 * (key(seed, 1, j, 0) & ((1ull << (m - 1)) - 1)) + 1;
}
__host__ __device__ __forceinline__ double unit53(u64 h) { return (double)(h >> 11) * 0x1p-53; }
__host__ __device__ __forceinline__ double lat_shift(u64 seed, u64 l, u64 j) { return unit53(key(seed, 2, l, j)); }

// dnb2: column k of the 32x32 upper-triangular unit-diagonal generating
// matrix (MSB = row 0): bit 31-k set, Bernoulli(1/2) bits 31-t for t < k
__host__ __device__ __forceinline__ u64 dnb2_col(u64 seed, u64 j, int k) {
    const u64 above = k ? (((1ull << k) - 1) << (32 - k)) : 0;
    return (1ull << (31 - k)) | (key(seed, 3, j, k) & above);
}
// row t of the lower-triangular unit-diagonal LMS matrix: bit 31-t set,
// Bernoulli(1/2) bits 31-c for c < t
__host__ __device__ __forceinline__ u64 dnb2_srow(u64 seed, u64 l, u64 j, int t) {
    const u64 left = t ? (((1ull << t) - 1) << (32 - t)) : 0;
    return (1ull << (31 - t)) | (key(seed, 4, (l << 32) | j, t) & left);
}
__host__ __device__ __forceinline__ u64 dnb2_shift(u64 seed, u64 l, u64 j) { return key(seed, 5, l, j); }
__host__ __device__ __forceinline__ u64 nus_seed(u64 seed, u64 l, u64 j) { return key(seed, 6, l, j); }

// gdn: first primes as bases
__constant__ u32 c_primes[256];
u32 h_primes[256];

void fill_primes() {
    if (h_primes[0]) return;
    int cnt = 0;
    for (u32 p = 2; cnt < 256; p++) {
        bool ok = true;
        for (u32 q = 2; q * q <= p; q++)
            if (p % q == 0) { ok = false; break; }
        if (ok) h_primes[cnt++] = p;
    }
}

// digits of base b for 2^-52 resolution (b^tn >= 2^52), and for indices < 2^m
__host__ __device__ __forceinline__ int gdn_tn(u32 b) {
    int t = 0;
    double v = 1.0;
    while (v < 4503599627370496.0) { v *= b; t++; }
    return t;
}
__host__ __device__ __forceinline__ int gdn_mm(u32 b, int m) {
    int t = 0;
    u64 v = 1;
    while (v < (1ull << m)) { v *= b; t++; }
    return t < 1 ? 1 : t;
}
// LMS entry S[t][k] (k <= t): diagonal in [1, b), below in [0, b)
__host__ __device__ __forceinline__ u32 gdn_s(u64 seed, u64 l, u64 j, int t, int k, u32 b) {
    const u64 h = key(seed, 7, (l << 32) | j, ((u64)t << 8) | k);
    return t == k ? (u32)(1 + h % (b - 1)) : (u32)(h % b);
}
// digit permutation at level t: y = (a*x + c) mod b, a in [1, b)
__host__ __device__ __forceinline__ void gdn_perm(u64 seed, u64 l, u64 j, int t, u32 b, u32 &a, u32 &c) {
    const u64 h = key(seed, 8, (l << 32) | j, t);
    a = (u32)(1 + (h >> 32) % (b - 1));
    c = (u32)((h & 0xffffffffull) % b);
}

// ----------------------------------------------------------- table layouts
// lattice : g[d] | shift[r*d] (double bits)
// dnb2    : per column set (l, j) -- r*d sets for LMS, d for NUS -- 64 u32:
//           pre[32] (prefix XOR of the columns) then cols[32] | shift or NUS seed[r*d]
// gdn     : per j kGJ words of dimension info | per replication l, gtw u32 of
//           digit tables (layout below), one block per j at u32 offset qoff_j
//
// gdn digit tables.  Point k of dimension j has base-b digits dig_i(k); the
// scrambled digit t is z_t(k) = (a_t * sum_i S[t][i] dig_i(k) + c_t) mod b.
// Going from k to k+1 turns the c trailing digits equal to b-1 into 0 and adds
// 1 to digit c, so every z_t moves by Q_c[t] = a_t * sum_{i<=c} S[t][i] mod b
// (-(b-1) = 1 mod b).  A chain of consecutive points is then one table row
// added per point, digitwise mod b.
//   b >= 3: rows 0..mm of W u32 each; row 0 holds c_t, row 1+c holds Q_c.
//           Digits are u16 halves, two per u32: padded digit u = t + (tn & 1)
//           (odd tn gets a leading zero digit) sits in pair u/2, the more
//           significant digit (even u) in the low half.  Rows are right-aligned
//           to the width W of the widest expansion among the dimension's lane
//           group (np = ceil(tn/2) <= W pairs, zero pairs in front), so every
//           lane of a wave walks the same register pairs.
//   b == 2: u64 words (a == 1): cmask, then Q_0..Q_{mm-1}, digit t at bit 51-t.

// words of dimension info: b | L<<16 | qoff<<32,  tn | mm<<8 | np<<16 | W<<24 | b^2<<32,  ceil(2^64/b),  b^tn
constexpr int kGJ = 4;
constexpr int kNPmax = 17;   // u32 digit pairs of the longest b >= 3 expansion (b = 3: 33 digits + 1 pad)

__host__ __device__ __forceinline__ u64 recip_magic(u32 b) { return b == 1 ? 0 : (~0ull) / b + 1; }

// x mod b and x / b for x < 2^32 with M = ceil(2^64 / b) (exact, Lemire et al. 2019)
__device__ __forceinline__ u32 fast_div(u32 x, u64 M) { return (u32)__umul64hi(M, (u64)x); }

inline u32 gdn_np(u32 b) { return b == 2 ? 0 : ((u32)gdn_tn(b) + 1) / 2; }

// Lane groups of a gdn tile: ranges [j0, j0 + gc) of consecutive columns,
// [0, 8) (the small bases, up to 17 digit pairs for base 3) apart from the
// rest, which go in groups of 64 from column 8; a wave works on ch = 64/gc
// chains of one group, every lane walking w digit pairs (the group's widest
// expansion: its smallest odd base).  Chains are k rows long, k a power of two
// giving about 8 units of work per group.  (With one group for the first 64
// columns every one of them walked base 3's 17 pairs; splitting off [0, 8)
// made the gdn class 10% faster and a small ragged batch 29%.  Row segments of
// 8 columns end mid-line, which the class's cached stores complete in L2.)
struct GdnGroup {
    u32 j0, gc, ch, w, l, k, cb0, nu;
};

std::vector<GdnGroup> gdn_groups(int d, u32 RT) {
    std::vector<std::pair<u32, u32>> ranges;  // (j0, gc)
    if (d <= 8) {
        ranges.push_back({0u, (u32)d});
    } else {
        ranges.push_back({0u, 8u});
        for (u32 j = 8; j < (u32)d; j += 64) ranges.push_back({j, std::min<u32>(64, (u32)d - j)});
    }
    const u32 ng = (u32)ranges.size();
    double inv = 0;  // sum over groups of 1/ch: a tile has RT * inv / K units of K rows
    for (auto &rg : ranges) inv += 1.0 / (double)(64 / rg.second);
    const u64 want = (u64)((double)RT * inv / (8.0 * ng));
    u32 K = 1;
    while (K * 2 <= want && K < 256) K *= 2;
    if (K > RT) K = RT;
    std::vector<GdnGroup> g(ng);
    u32 cb0 = 0;
    for (u32 i = 0; i < ng; i++) {
        GdnGroup &x = g[i];
        x.j0 = ranges[i].first;
        x.gc = ranges[i].second;
        const u32 CH = 64 / x.gc;
        x.ch = CH;
        u32 bmax = 0;
        for (u32 j = x.j0; j < x.j0 + x.gc; j++) {
            x.w = std::max(x.w, gdn_np(h_primes[j]));
            bmax = std::max(bmax, h_primes[j]);
        }
        const double bm = (double)bmax;
        x.l = bm * bm * bm * bm * bm * bm < 0x1p53 ? 3 : 2;
        x.k = K;
        x.nu = (RT + CH * K - 1) / (CH * K);
        x.cb0 = cb0;
        cb0 += x.nu;
    }
    return g;
}

// Words of column j's block.  Base 2 (j = 0, read as u64) comes first and is
// even-sized; the u32 blocks of odd bases are odd-sized, so the lanes of a wave,
// each reading its own column's row at the same offset, fall on distinct LDS
// banks.
inline u32 gdn_words(const Req &q, int j, u32 W) {
    const u32 b = h_primes[j];
    const u32 mm = (u32)gdn_mm(b, q.m);
    return b == 2 ? 2 + 2 * mm : ((mm + 1) * W) | 1u;
}

u64 table_words(Req &q, const std::vector<GdnGroup> &grp) {
    const u64 d = q.d, r = q.r;
    switch (q.method) {
        case M_LATTICE: return d + r * d;
        case M_DNB2_LMS: return 32 * r * d + r * d;
        case M_DNB2_NUS: return 32 * d + r * d;
        default: {
            u32 w = 0;
            for (const auto &x : grp)
                for (u32 j = x.j0; j < x.j0 + x.gc; j++) w += gdn_words(q, (int)j, x.w);
            q.gtw = (w + 1) & ~1u;  // even: every replication's block starts 8-byte aligned
            q.ggrp = (u64)kGJ * d;
            q.gtab = q.ggrp + 2 * grp.size();
            return q.gtab + r * (u64)(q.gtw / 2);
        }
    }
}

// --------------------------------------------------------------- setup

__global__ __launch_bounds__(kBlock) void batch_setup_kernel(const Req *__restrict__ reqs, u64 *__restrict__ ws) {
    const Req q = reqs[blockIdx.y];
    u64 *tab = ws + q.tab;
    const u64 d = q.d, r = q.r;
    const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 step = (u64)gridDim.x * blockDim.x;
    if (q.method == M_LATTICE) {
        for (u64 e = tid; e < d + r * d; e += step) {
            if (e < d) {
                tab[e] = lat_g(q.seed, q.m, e);
            } else {
                const u64 f = e - d;
                tab[e] = (u64)__double_as_longlong(lat_shift(q.seed, f / d, f % d));
            }
        }
    } else if (q.method == M_DNB2_LMS || q.method == M_DNB2_NUS) {
        const u64 nset = (q.method == M_DNB2_LMS ? r : 1) * d;  // (l, j) column sets
        u32 *T32 = (u32 *)tab;
        for (u64 c = tid; c < nset + r * d; c += step) {
            if (c < nset) {
                const u64 l = c / d, j = c % d;
                u32 *pre = T32 + c * 64;
                u32 *cols = pre + 32;
                u32 acc = 0;
                for (int k = 0; k < kTB; k++) {
                    u64 col = dnb2_col(q.seed, j, k);
                    if (q.method == M_DNB2_LMS) {  // C_lms[k] = S * C[k] over GF(2) (dnb2_linear_matrix_scramble.c:33-48)
                        u64 out = 0;
                        for (int t = 0; t < kTB; t++)
                            out |= (u64)(__popcll(dnb2_srow(q.seed, l, j, t) & col) & 1) << (kTB - 1 - t);
                        col = out;
                    }
                    cols[k] = (u32)col;
                    acc ^= (u32)col;
                    pre[k] = acc;
                }
            } else {
                const u64 f = c - nset;
                tab[32 * nset + f] = q.method == M_DNB2_LMS ? dnb2_shift(q.seed, f / d, f % d)
                                                           : nus_seed(q.seed, f / d, f % d);
            }
        }
    } else {
        // gdn digit tables, one thread per (l, j, t); the dimension info was
        // written by the host
        u32 *T = (u32 *)(tab + q.gtab);
        for (u64 e = tid; e < r * d * kTNmax; e += step) {
            const int t = (int)(e % kTNmax);
            const u64 lj = e / kTNmax, l = lj / d, j = lj % d;
            const u64 w0 = tab[kGJ * j], w1 = tab[kGJ * j + 1];
            const u32 b = (u32)w0 & 0xffffu;
            const int tn = (int)(w1 & 0xff), mm = (int)((w1 >> 8) & 0xff), np = (int)((w1 >> 16) & 0xff);
            const int W = (int)((w1 >> 24) & 0xff);
            const u32 qoff = (u32)(w0 >> 32);
            const u32 qend = j + 1 < d ? (u32)(tab[kGJ * (j + 1)] >> 32) : q.gtw;
            const u32 need = b == 2 ? 2 * (u32)(mm + 1) : (u32)((mm + 1) * W);
            if (qoff + need > qend) continue;  // never: the host sizes every block (keeps writes in it)
            u32 *Tj = T + l * q.gtw + qoff;
            if (b == 2) {
                if (t) continue;
                u64 *Q2 = (u64 *)Tj;
                for (int i = 0; i <= mm; i++) Q2[i] = 0;
                u64 cm = 0;
                for (int tt = 0; tt < tn; tt++) {
                    u32 a, cc;
                    gdn_perm(q.seed, l, j, tt, 2, a, cc);
                    cm |= (u64)cc << (51 - tt);
                    u32 run = 0;
                    for (int i = 0; i < mm; i++) {
                        if (i <= tt) run ^= gdn_s(q.seed, l, j, tt, i, 2) & 1;
                        Q2[1 + i] |= (u64)run << (51 - tt);
                    }
                }
                Q2[0] = cm;
            } else {
                if (t >= tn) continue;
                // digit u of row w at H[w * 2W + 2(W - np) + u]; zero pairs in front
                unsigned short *H = (unsigned short *)Tj + 2 * (W - np);
                const int u = t + (tn & 1);
                if (t == 0)
                    for (int w = 0; w <= mm; w++)
                        for (int h = -2 * (W - np); h < (tn & 1); h++) H[w * 2 * W + h] = 0;
                u32 a, cc;
                gdn_perm(q.seed, l, j, t, b, a, cc);
                H[u] = (unsigned short)cc;
                u32 run = 0;
                for (int i = 0; i < mm; i++) {
                    if (i <= t) run = (run + a * gdn_s(q.seed, l, j, t, i, b)) % b;  // a, S < b <= 2^11
                    H[(1 + i) * 2 * W + u] = (unsigned short)run;
                }
            }
        }
    }
}

// ------------------------------------------------------------- one tile

__device__ __forceinline__ u32 xor_cols(const u32 *__restrict__ cols, u64 q) {
    u32 v = 0;
    while (q) {
        v ^= cols[__ffsll((long long)q) - 1];
        q &= q - 1;
    }
    return v;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Output of a tile, and with DIG a running wrap-around sum of the stored bit
// patterns (an order-independent per-slab digest).  STREAM: non-temporal
// stores, for the store-bound class whose chains cover whole 128-byte lines;
// the gdn class's rows (d doubles) share lines with the next row, which
// cached stores complete in L2 (5% faster there, 3-5% slower for the
// store-bound class).
template <bool DIG, bool STREAM>
struct Sink {
    u64 acc = 0;
    __device__ __forceinline__ void put(double *p, double v) {
        if (STREAM)
            store_stream(p, v);
        else
            *p = v;
        if (DIG) acc += (u64)__double_as_longlong(v);
    }
    // one atomic per wave: every lane of the block works on the same slab
    __device__ __forceinline__ void flush(unsigned long long *digest) {
        if (!DIG) return;
        u64 a = acc;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a += (u64)__shfl_xor((long long)a, o, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(digest, (unsigned long long)a);
    }
};

__device__ __forceinline__ u16x2 as_pair(u32 w) { return __builtin_bit_cast(u16x2, w); }
__device__ __forceinline__ u32 as_word(u16x2 p) { return __builtin_bit_cast(u32, p); }

// dnb2 output of column value v (the top 32 of 64 bits): LMS adds the 64-bit
// digital shift xs, NUS hashes; then x = y * 2^-64 (dnb2_integer_to_float.c:29).
// For LMS, y = (v ^ xs_hi) * 2^32 + xs_lo and one fma of exact terms rounds
// y * 2^-64 once, as the u64 -> f64 conversion does.
template <bool LMS>
__device__ __forceinline__ double dnb2_emit(u32 v, u64 xs, double lo) {
    if (LMS) return __fma_rn((double)(v ^ (u32)(xs >> 32)), 0x1p-32, lo);
    return ldexp((double)owen_hash64((u64)v << 32, xs), -64);
}

// Chains of a dnb2 tile: the tile's R*d (row, column) slots of its first R rows
// (R = 2^clog) each walk down the tile in steps of R rows, so that at every
// step the block writes R*d consecutive doubles.  From chain step g to g+1 the
// row index changes in bits clog..clog+c (c = ctz(g+1)), so the value changes by
// pre[clog+c] ^ pre[clog-1] (pre = prefix XOR of the columns).  Steps come in
// groups of 8 whose first 7 transitions use c = 0,1,0,2,0,1,0.
template <bool LMS, bool DIG, class SinkT>
__device__ __forceinline__ void dnb2_tile(const Req &q, u64 l, u64 row0, const u64 *__restrict__ tab,
                                          double *__restrict__ slab, u32 *v0, SinkT &out) {
    const u32 d = (u32)q.d;
    const u32 s = q.clog;
    const u32 R = 1u << s, H = q.rows_per_tile >> s;
    const u64 set = LMS ? l * d : 0;
    const u64 nset = (LMS ? (u64)q.r : 1) * d;
    const u32 *T32 = (const u32 *)tab;
    const u64 *extra = tab + 32 * nset + l * d;  // shifts or NUS seeds
    const u64 stride = (u64)R * d;
    // value of the tile's first row per column, shared by the column's chains
    // (all column loads issued at once: one memory round trip)
    for (u32 j = threadIdx.x; j < d; j += blockDim.x) {
        const u32 *cols = T32 + (set + j) * 64 + 32;
        u32 v = 0;
#pragma unroll 1
        for (int b0 = 0; b0 < kTB; b0 += 8) {
            if (!((row0 >> b0) & 0xff)) continue;
#pragma unroll
            for (int bit = b0; bit < b0 + 8; bit++)
                if ((row0 >> bit) & 1) v ^= cols[bit];
        }
        v0[j] = v;
    }
    __syncthreads();
    for (u32 sig = threadIdx.x; sig < R * d; sig += blockDim.x) {
        const u32 rr0 = div_q(sig, q.ddiv);
        const u32 j = sig - rr0 * d;
        const u32 *pre = T32 + (set + j) * 64;
        const u32 *cols = pre + 32;
        const u64 xs = extra[j];
        const double lo = (double)(u32)xs * 0x1p-64;
        const u32 base = s ? pre[s - 1] : 0u;
        // row0 + rr0 with rr0 < 2^s <= the tile: the low bits add independently
        u32 v = v0[j];
#pragma unroll
        for (int bit = 0; bit < kMaxChainLog; bit++)
            if (bit < (int)s && ((rr0 >> bit) & 1)) v ^= cols[bit];
        double *dst = slab + (row0 + rr0) * d + j;
        if (H >= 8) {
            // every transition of the chain, in registers: after step h the row
            // index flips bits s .. s+c with c = ctz(h+1) < log2(H) (the tile's
            // first row is a multiple of H 2^s), so pre[s .. s+kMaxGroupLog+2]
            // covers them all
            u32 xk[3 + kMaxGroupLog];
#pragma unroll
            for (int k = 0; k < 3 + kMaxGroupLog; k++) xk[k] = (s + k < (u32)kTB) ? pre[s + k] ^ base : 0u;
            const u32 x0 = xk[0], x1 = xk[1], x2 = xk[2];
            for (u32 h = 0; h < H; h += 8) {
                // the transition after this group: c = 3 + ctz((h + 8) / 8), uniform over the block
                const int c8 = __ffs((int)((h + 8) >> 3)) - 1;
                u32 next = xk[3];
#pragma unroll
                for (int k = 1; k < kMaxGroupLog; k++)
                    if (c8 == k) next = xk[3 + k];
                out.put(dst, dnb2_emit<LMS>(v, xs, lo));
                v ^= x0;
                out.put(dst + stride, dnb2_emit<LMS>(v, xs, lo));
                v ^= x1;
                out.put(dst + 2 * stride, dnb2_emit<LMS>(v, xs, lo));
                v ^= x0;
                out.put(dst + 3 * stride, dnb2_emit<LMS>(v, xs, lo));
                v ^= x2;
                out.put(dst + 4 * stride, dnb2_emit<LMS>(v, xs, lo));
                v ^= x0;
                out.put(dst + 5 * stride, dnb2_emit<LMS>(v, xs, lo));
                v ^= x1;
                out.put(dst + 6 * stride, dnb2_emit<LMS>(v, xs, lo));
                v ^= x0;
                out.put(dst + 7 * stride, dnb2_emit<LMS>(v, xs, lo));
                dst += 8 * stride;
                v ^= next;
            }
        } else {
            // H <= 4: transitions c = ctz(h+1) <= 1
            const u32 x0 = pre[s] ^ base, x1 = H > 1 ? pre[s + 1] ^ base : 0u;
            for (u32 h = 0;; h++) {
                out.put(dst, dnb2_emit<LMS>(v, xs, lo));
                if (h + 1 == H) break;
                v ^= (h & 1) ? x1 : x0;
                dst += stride;
            }
        }
    }
}

// z + q digitwise mod b for two u16 digits per word (both < b <= 2^11):
// min(z + q, z + q - b) with 16-bit wrap-around (v_pk_add/sub/min_u16)
__device__ __forceinline__ u32 pk_add_mod(u32 z, u32 q, u16x2 bb) {
    const u16x2 sum = as_pair(z) + as_pair(q);
    return as_word(__builtin_elementwise_min(sum, sum - bb));
}

// One gdn chain of a base-2 column: len consecutive points from k, row stride d.
template <class SinkT>
__device__ __forceinline__ void gdn_chain2(const u64 *__restrict__ Q, u32 k, u32 len, u32 d, double *dst,
                                           SinkT &out) {
    // Q = cmask, Q_0..Q_{mm-1}; point k's digits: cmask ^ XOR of columns S_i = Q_i ^ Q_{i-1}
    u64 z = Q[0];
    for (u32 kk = k; kk; kk &= kk - 1) {
        const int i = __ffs(kk) - 1;
        z ^= Q[1 + i] ^ (i ? Q[i] : 0ull);
    }
    for (u32 h = 0;; h++) {
        out.put(dst, (double)z * 0x1p-52);  // z < 2^52: exact, = z / 2^tn
        if (h + 1 == len) break;
        z ^= Q[__ffs(k + 1)];  // Q_c at Q[1 + c], c = ctz(k + 1)
        k++;
        dst += d;
    }
}

// One gdn chain of a column with odd base b: len consecutive points from k,
// row stride d.  T = the column's digit tables in LDS, W pairs per row (the
// lane group's width, as a template so every pair lives in a named register);
// L = pairs in the low part of the floating-point Horner split.
template <int W, int L, class SinkT>
__device__ __forceinline__ void gdn_chain(const u64 *__restrict__ gj, const u32 *T, u32 k, u32 len, u32 d,
                                          double *dst, SinkT &out) {
    const u64 w1 = gj[1];
    const u32 b = (u32)gj[0] & 0xffffu;
    const u32 mm = (u32)((w1 >> 8) & 0xff);
    const u32 b2 = (u32)(w1 >> 32);
    const u64 M = gj[2];
    const double scale = __longlong_as_double((long long)gj[3]);
    // start: z = c + sum_i dig_i * S'_i  with  S'_i = Q_i - Q_{i-1}
    //      = c + sum_i Q_i * e_i,  e_i = dig_i - dig_{i+1} mod b  (dig_mm = 0)
    u32 acc_lo[W], acc_hi[W];
#pragma unroll
    for (int P = 0; P < W; P++) {
        const u32 w = T[P];
        acc_lo[P] = w & 0xffffu;
        acc_hi[P] = w >> 16;
    }
    u32 qd = fast_div(k, M);
    u32 dig = k - qd * b;
    const u32 dlow0 = dig;
    for (u32 i = 0; i < mm; i++) {
        const u32 qn = fast_div(qd, M);
        const u32 dn = qd - qn * b;
        const u32 e = dig >= dn ? dig - dn : dig + b - dn;
        const u32 *Qi = T + (1 + i) * W;
#pragma unroll
        for (int P = 0; P < W; P++) {
            const u32 w = Qi[P];
            acc_lo[P] += e * (w & 0xffffu);  // < (1 + mm (b-1)) b < 2^32
            acc_hi[P] += e * (w >> 16);
        }
        qd = qn;
        dig = dn;
    }
    u32 z[W];
#pragma unroll
    for (int P = 0; P < W; P++) {
        const u32 lo = acc_lo[P] - fast_div(acc_lo[P], M) * b;
        const u32 hi = acc_hi[P] - fast_div(acc_hi[P], M) * b;
        z[P] = lo | (hi << 16);
    }
    const u16x2 bb = as_pair(b | (b << 16));
    const u16x2 b1 = as_pair(b | (1u << 16));  // pair value = lo * b + hi (v_dot2_u32_u16)
    const double bd2 = (double)b2;
    double BL = bd2;
#pragma unroll
    for (int i = 1; i < L; i++) BL *= bd2;  // (b^2)^L < 2^53: exact
    u32 dlow = dlow0;
    for (u32 h = 0;; h++) {
        // Y = sum of digit pairs in base b^2 as Yhi * (b^2)^L + Ylo, both parts
        // exact doubles (< 2^53); one fma rounds Y once, as u64 -> f64 does
        double Yhi = 0.0, Ylo = 0.0;
#pragma unroll
        for (int P = 0; P < W; P++) {
            const double pv = (double)__builtin_amdgcn_udot2(as_pair(z[P]), b1, 0u, false);
            if (P < W - L)
                Yhi = __fma_rn(Yhi, bd2, pv);
            else
                Ylo = __fma_rn(Ylo, bd2, pv);
        }
        out.put(dst, __fma_rn(Yhi, BL, Ylo) / scale);
        if (h + 1 == len) break;
        // carries of k -> k+1: c trailing digits equal to b-1
        u32 c = 0;
        if (dlow == b - 1) {
            dlow = 0;
            c = 1;
            u32 u = fast_div(k, M);
            for (;;) {
                const u32 un = fast_div(u, M);
                if (u - un * b != b - 1) break;
                c++;
                u = un;
            }
        } else {
            dlow++;
        }
        k++;
        const u32 *Qc = T + (1 + c) * W;
        u32 inc[W];
#pragma unroll
        for (int P = 0; P < W; P++) inc[P] = Qc[P];
#pragma unroll
        for (int P = 0; P < W; P++) z[P] = pk_add_mod(z[P], inc[P], bb);
        dst += d;
    }
}

// Tiles are drained by two persistent kernels: class 0 = lattice and dnb2
// (HBM-write bound, few registers, full occupancy), class 1 = gdn (register
// heavy digit walks).  Each class is compiled without the other's code.
__host__ __device__ constexpr int method_class(int method) { return method == M_GDN_LMS ? 1 : 0; }

template <int CLASS, bool DIG>
__device__ void run_tile(const Req &q, u64 t, const u64 *__restrict__ ws, double *__restrict__ ring, u32 *lds,
                         unsigned long long *digest) {
    const u64 d = q.d;
    const u64 l = t / q.tiles_per_l;
    const u64 row0 = (t - l * q.tiles_per_l) * (u64)q.rows_per_tile;
    const u64 *tab = ws + q.tab;
    double *slab = ring + ws[q.slabs + l];
    Sink<DIG, CLASS == 0> out;
    if (CLASS == 0 && q.method == M_LATTICE) {
        const u32 pts = q.rows_per_tile * (u32)d;
        const u64 *g = tab;
        const double *sh = (const double *)(tab + d) + l * d;
        for (u32 e = threadIdx.x; e < pts; e += blockDim.x) {
            const u32 rr = div_q(e, q.ddiv);
            const u32 j = e - rr * q.d;
            const u64 row = row0 + rr;
            const double phi = (double)__brevll(row) * 0x1p-64;  // phi2(row), exact
            const double x = frac_mod1((double)g[j] * phi);      // lat_gen_natural.c:45
            out.put(slab + row * d + j, frac_mod1(x + sh[j]));  // lat_shift_mod_1.c:31
        }
    } else if (CLASS == 0 && q.method == M_DNB2_LMS) {
        dnb2_tile<true, DIG>(q, l, row0, tab, slab, lds, out);
    } else if (CLASS == 0) {
        dnb2_tile<false, DIG>(q, l, row0, tab, slab, lds, out);
    } else {
        // this replication's digit tables into LDS
        const u32 *src = (const u32 *)(tab + q.gtab) + l * q.gtw;
        for (u32 e = threadIdx.x; e < q.gtw; e += blockDim.x) lds[e] = src[e];
        __syncthreads();
        // units = (lane group, block of ch chains of k rows); lanes = (chain, column)
        const u64 *grp = tab + q.ggrp;  // per group: j0 | gc<<16 | ch<<24 | w<<32 | l<<40,  k | cb0<<32 | nu<<48
        const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
        for (u32 u = wave; u < q.gunits; u += nwaves) {
            u32 g = 0;
            u64 g1 = grp[1];
            while (u >= ((g1 >> 32) & 0xffff) + (g1 >> 48)) g1 = grp[2 * (++g) + 1];
            const u64 g0 = grp[2 * g];
            const u32 j0 = __builtin_amdgcn_readfirstlane((u32)(g0 & 0xffff));
            const u32 gc = __builtin_amdgcn_readfirstlane((u32)((g0 >> 16) & 0xff));
            const u32 CH = __builtin_amdgcn_readfirstlane((u32)((g0 >> 24) & 0xff));
            const u32 Ww = __builtin_amdgcn_readfirstlane((u32)((g0 >> 32) & 0xff));
            const u32 Lw = __builtin_amdgcn_readfirstlane((u32)((g0 >> 40) & 0xff));
            const u32 K = __builtin_amdgcn_readfirstlane((u32)g1);
            const u32 cb = u - __builtin_amdgcn_readfirstlane((u32)((g1 >> 32) & 0xffff));
            const u32 ch = lane / gc, jj = lane - ch * gc;
            const u32 j = j0 + jj;
            const u32 rs = cb * CH * K + ch * K;
            if (ch >= CH || rs >= q.rows_per_tile) continue;
            const u32 len = q.rows_per_tile - rs < K ? q.rows_per_tile - rs : K;
            const u64 k = row0 + rs;
            const u64 *gj = tab + kGJ * j;
            const u32 *T = lds + (u32)(gj[0] >> 32);
            double *dst = slab + k * d + j;
            if (j == 0) {
                gdn_chain2((const u64 *)T, (u32)k, len, (u32)d, dst, out);
                continue;
            }
#define QMC_GDN_SHAPE(W_, L_) \
    if (Ww == W_ && Lw == L_) { gdn_chain<W_, L_>(gj, T, (u32)k, len, (u32)d, dst, out); continue; }
            // every (w, l) a group can have (host-checked): 17 for columns [0, 8),
            // 6 for [8, 72), 4 or 3 for the groups of larger bases
            QMC_GDN_SHAPE(17, 3) QMC_GDN_SHAPE(6, 3) QMC_GDN_SHAPE(4, 3) QMC_GDN_SHAPE(4, 2) QMC_GDN_SHAPE(3, 2)
#undef QMC_GDN_SHAPE
        }
    }
    out.flush(digest + q.slab0 + l);
}

// ------------------------------------------------------------ the kernels

template <int CLASS, bool DIG>
__global__ __launch_bounds__(kBlock) void batch_persistent_kernel(const Req *__restrict__ reqs, int nreq,
                                                                  u64 total_tiles, unsigned long long *counter,
                                                                  const u64 *__restrict__ ws,
                                                                  double *__restrict__ out,
                                                                  unsigned long long *digest) {
    __shared__ u64 s_tile;
    __shared__ int s_req;
    extern __shared__ u32 lds_tables[];  // dnb2 first-row values or gdn digit tables of the current tile
    for (;;) {
        if (threadIdx.x == 0) {
            const u64 t = atomicAdd(counter, 1ull);
            s_tile = t;
            int lo = 0, hi = nreq - 1;  // last request with tile0 <= t
            if (t < total_tiles) {
                while (lo < hi) {
                    const int mid = (lo + hi + 1) / 2;
                    if (reqs[mid].tile0 <= t) lo = mid;
                    else hi = mid - 1;
                }
            }
            s_req = lo;
        }
        __syncthreads();
        const u64 t = s_tile;
        const int qi = s_req;
        __syncthreads();
        if (t >= total_tiles) break;  // every wave of every block reaches this
        const Req q = reqs[qi];
        run_tile<CLASS, DIG>(q, t - q.tile0, ws, out, lds_tables, digest);
    }
}

__global__ void reset_counter_kernel(unsigned long long *counter) {
    counter[0] = 0;
    counter[1] = 0;
}

// one request, one tile per block (the per-call form)
template <int CLASS, bool DIG>
__global__ __launch_bounds__(kBlock) void batch_single_kernel(const Req *__restrict__ req, const u64 *__restrict__ ws,
                                                              double *__restrict__ out, unsigned long long *digest) {
    extern __shared__ u32 lds_tables[];
    const Req q = *req;
    run_tile<CLASS, DIG>(q, blockIdx.x, ws, out, lds_tables, digest);
}

// ----------------------------------------------------------- host plan

// dnb2 chain stride 2^s for a tile of RT rows x d columns: U = ceil(2^s d / 256)
// chains per thread of RT / 2^s points each, plus about 8 points' worth of
// work to start a chain.  A step writes 2^s d consecutive doubles; when that
// is a multiple of 16 (128 bytes) every step covers whole cache lines, which
// streaming stores need to avoid partial-line writes to HBM, so the cheapest
// such s is taken whenever there is one (measured on batches of one d: d = 127
// went from 2.9 to 5.7 TB/s, d = 31 from 4.8 to 5.4; a mixed-d batch from 4.2
// to 5.4 TB/s).
u32 dnb2_chain_log(u32 RT, u32 d) {
    u32 best = 0, best_al = ~0u;
    u64 best_cost = ~0ull, best_al_cost = ~0ull;
    for (u32 s = 0; (1u << s) <= RT && s < kMaxChainLog; s++) {
        if ((RT >> s) > (8u << (kMaxGroupLog - 1))) continue;  // chains of at most 1024 points
        const u64 U = (((u64)d << s) + kBlock - 1) / kBlock;
        const u64 cost = U * ((RT >> s) + 8);
        if (cost < best_cost) {
            best_cost = cost;
            best = s;
        }
        if ((((u64)d << s) & 15) == 0 && cost < best_al_cost) {
            best_al_cost = cost;
            best_al = s;
        }
    }
    return best_al != ~0u ? best_al : best;
}

struct Plan {
    std::vector<Req> reqs;
    u64 total_tiles = 0;
    u64 ws_words = 0;
    std::vector<Req> qreqs;          // requests regrouped by class, tile0 counted within the class
    int qn[2] = {0, 0};              // requests per class (class 0 first in qreqs)
    u64 qtiles[2] = {0, 0};          // tiles per class
    Req *d_qreqs = nullptr;
    u64 ring_elems = 0;
    Req *d_reqs = nullptr;
    u64 *d_ws = nullptr;
    unsigned long long *d_counter = nullptr;
    int grid = 0;
    int device = 0;
    u32 lds_bytes = 0;               // dynamic LDS of the gdn kernel (digit tables)
    hipStream_t cap = nullptr;       // private stream for graph capture
    hipStream_t side = nullptr;      // the gdn class runs here, beside the store-bound class
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipGraphExec_t gexec = nullptr;  // [counter reset, persistent kernel]
    double *gring = nullptr;         // ring the graph was captured for
    unsigned long long *digest = nullptr;   // per-slab digests, or null
    unsigned long long *gdigest = nullptr;  // digest pointer the graph was captured with
    u64 nslabs = 0;
    std::vector<u64> slab_pos;       // ring offset of every slab (host copy of the workspace's)
};

constexpr u32 kLdsClass0 = 256 * sizeof(u32);  // dnb2 first-row values

// With both classes present, the gdn kernel (the long tiles) is forked onto the
// plan's side stream and joined back through events, so a graph captured from
// s takes in both branches.  On large batches this changes nothing (each
// persistent grid holds its CUs until its queue drains); on small ragged ones
// the store-bound tiles finish under the gdn tiles (9% on a 48-request batch).
template <bool DIG>
int launch_batch_t(Plan *P, double *ring, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, P->d_counter);
    if (P->qtiles[0] && P->qtiles[1]) {
        QMC_HIP_CHECK(hipEventRecord(P->ev_fork, s));
        QMC_HIP_CHECK(hipStreamWaitEvent(P->side, P->ev_fork, 0));
        hipLaunchKernelGGL((batch_persistent_kernel<1, DIG>), dim3(P->grid), dim3(kBlock), P->lds_bytes, P->side,
                           P->d_qreqs + P->qn[0], P->qn[1], P->qtiles[1], P->d_counter + 1, P->d_ws, ring,
                           P->digest);
        QMC_HIP_CHECK(hipEventRecord(P->ev_join, P->side));
        hipLaunchKernelGGL((batch_persistent_kernel<0, DIG>), dim3(P->grid), dim3(kBlock), kLdsClass0, s, P->d_qreqs,
                           P->qn[0], P->qtiles[0], P->d_counter, P->d_ws, ring, P->digest);
        QMC_HIP_CHECK(hipStreamWaitEvent(s, P->ev_join, 0));
        QMC_HIP_CHECK(hipGetLastError());
        return QMC_OK;
    }
    if (P->qtiles[0])
        hipLaunchKernelGGL((batch_persistent_kernel<0, DIG>), dim3(P->grid), dim3(kBlock), kLdsClass0, s, P->d_qreqs,
                           P->qn[0], P->qtiles[0], P->d_counter, P->d_ws, ring, P->digest);
    if (P->qtiles[1])
        hipLaunchKernelGGL((batch_persistent_kernel<1, DIG>), dim3(P->grid), dim3(kBlock), P->lds_bytes, s,
                           P->d_qreqs + P->qn[0], P->qn[1], P->qtiles[1], P->d_counter + 1, P->d_ws, ring,
                           P->digest);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int launch_batch(Plan *P, double *ring, hipStream_t s) {
    return P->digest ? launch_batch_t<true>(P, ring, s) : launch_batch_t<false>(P, ring, s);
}

}  // namespace

extern "C" {

QMC_EXPORT void qmc_batch_free(void *h);

// Build a plan for nreq requests: methods[q] in 0..3, m[q] = log2 n in [1, 31],
// d[q] in [1, 256], r[q] >= 1, seeds[q]; `points_per_tile` (power of two) and
// `grid` (persistent workgroups; 0 = 4 per CU) are the tunables; `ring_elems`
// is the output ring size in doubles (every replication slab must fit).
QMC_EXPORT void *qmc_batch_plan(int nreq, const int32_t *methods, const int32_t *m, const int32_t *d,
                                const int32_t *r, const u64 *seeds, u64 points_per_tile, int grid, u64 ring_elems) {
    fill_primes();
    Plan *P = new Plan();
    P->ring_elems = ring_elems;
    u64 tile = 0, words = 0;
    std::vector<std::vector<GdnGroup>> groups(nreq);
    for (int i = 0; i < nreq; i++) {
        Req q{};
        q.method = methods[i];
        q.m = m[i];
        q.d = d[i];
        q.r = r[i];
        q.seed = seeds[i];
        if (q.method < 0 || q.method > 3 || q.m < 1 || q.m > 31 || q.d < 1 || q.d > 256 || q.r < 1 ||
            (q.method == M_LATTICE && q.m > 53) || ((q.method == 1 || q.method == 2) && q.m > kTB) ||
            (q.method == M_GDN_LMS && q.m > kTB)) {
            qmc_fail(QMC_ERR_ARG, "batch: request %d has unsupported (method, m, d, r) = (%d, %d, %d, %d)", i,
                     q.method, q.m, q.d, q.r);
            qmc_batch_free(P);
            return nullptr;
        }
        const u64 n = 1ull << q.m;
        u64 rows = points_per_tile / (u64)q.d;
        if (q.method == M_GDN_LMS && rows < 256) rows = 256;  // digit tables are loaded once per tile
        if ((q.method == M_DNB2_LMS || q.method == M_DNB2_NUS) && rows < 1024) rows = 1024;  // 64-step chains at 2^s = 16
        u64 p2 = 1;
        while (p2 * 2 <= rows) p2 *= 2;
        if (p2 > n) p2 = n;
        q.rows_per_tile = (u32)p2;
        q.tiles_per_l = (u32)(n / p2);
        q.tile0 = tile;
        tile += (u64)q.r * q.tiles_per_l;
        q.ddiv = make_div((u32)q.d);
        if (q.method == M_DNB2_LMS || q.method == M_DNB2_NUS) q.clog = dnb2_chain_log(q.rows_per_tile, (u32)q.d);
        if (q.method == M_GDN_LMS) {
            groups[i] = gdn_groups(q.d, q.rows_per_tile);
            q.gng = (u32)groups[i].size();
            q.gunits = groups[i].back().cb0 + groups[i].back().nu;
        }
        q.tab = words;
        words += table_words(q, groups[i]);
        if (q.method == M_GDN_LMS) {
            if ((u64)q.gtw * 4 > 65536) {
                qmc_fail(QMC_ERR_ARG, "batch: gdn request %d needs %u bytes of digit tables (max 65536)", i,
                         q.gtw * 4);
                qmc_batch_free(P);
                return nullptr;
            }
            if (q.gtw * 4 > P->lds_bytes) P->lds_bytes = q.gtw * 4;
        }
        q.slabs = words;
        words += q.r;
        q.slab0 = (u32)P->nslabs;
        P->nslabs += (u64)q.r;
        P->reqs.push_back(q);
        const u64 slab = n * (u64)q.d;
        if (slab > ring_elems) {
            qmc_fail(QMC_ERR_ARG, "batch: a replication slab of request %d (%llu doubles) exceeds the ring", i,
                     (unsigned long long)slab);
            qmc_batch_free(P);
            return nullptr;
        }
    }
    P->total_tiles = tile;
    P->ws_words = words;
    for (int c = 0; c < 2; c++)
        for (const Req &q0 : P->reqs) {
            if (method_class(q0.method) != c) continue;
            Req q = q0;
            q.tile0 = P->qtiles[c];
            P->qtiles[c] += (u64)q.r * q.tiles_per_l;
            P->qreqs.push_back(q);
            P->qn[c]++;
        }
    if (hipGetDevice(&P->device) != hipSuccess) {
        qmc_batch_free(P);
        qmc_fail(QMC_ERR_HIP, "batch: no device");
        return nullptr;
    }
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, P->device) != hipSuccess) cus = 256;
    P->grid = grid > 0 ? grid : 4 * cus;
    std::vector<u64> ws_host(words, 0);
    // gdn dimension info and lane groups (the digit tables are built on the device)
    for (size_t qi = 0; qi < P->reqs.size(); qi++) {
        const Req &q = P->reqs[qi];
        if (q.method != M_GDN_LMS) continue;
        u32 qoff = 0;
        for (size_t gi = 0; gi < groups[qi].size(); gi++) {
            const GdnGroup &x = groups[qi][gi];
            u64 *gw = ws_host.data() + q.tab + q.ggrp + 2 * gi;
            const bool shape_ok = x.w == 0 || (x.w == 17 && x.l == 3) || (x.w == 6 && x.l == 3) ||
                                  (x.w == 4 && (x.l == 3 || x.l == 2)) || (x.w == 3 && x.l == 2);
            if (!shape_ok) {
                qmc_fail(QMC_ERR_ARG, "batch: gdn lane group with no kernel shape (w=%u, l=%u)", x.w, x.l);
                qmc_batch_free(P);
                return nullptr;
            }
            gw[0] = x.j0 | (x.gc << 16) | (x.ch << 24) | ((u64)x.w << 32) | ((u64)x.l << 40);
            gw[1] = x.k | ((u64)x.cb0 << 32) | ((u64)x.nu << 48);
            for (u32 j = x.j0; j < x.j0 + x.gc; j++) {
                const u32 b = h_primes[j];
                const u64 tn = (u64)gdn_tn(b), mm = (u64)gdn_mm(b, q.m), np = gdn_np(b);
                const u64 W = x.w, L = x.l;
                // Horner halves exact: (b^2)^L and (b^2)^(np-L) below 2^53
                if (b != 2 && !((double)b * b * b * b < 0x1p53 && (L == 2 || std::pow((double)b, 6.0) < 0x1p53) &&
                                std::pow((double)b * b, (double)(np > L ? np - L : 0)) < 0x1p53)) {
                    qmc_fail(QMC_ERR_ARG, "batch: gdn base %u does not split into exact halves (W=%u, L=%u)", b,
                             (u32)W, (u32)L);
                    qmc_batch_free(P);
                    return nullptr;
                }
                double scale = 1.0;
                for (u64 t = 0; t < tn; t++) scale *= (double)b;
                u64 sbits;
                std::memcpy(&sbits, &scale, sizeof sbits);
                u64 *ji = ws_host.data() + q.tab + (u64)kGJ * j;
                ji[0] = b | (L << 16) | ((u64)qoff << 32);
                ji[1] = tn | (mm << 8) | (np << 16) | (W << 24) | ((u64)b * b << 32);
                ji[2] = recip_magic(b);
                ji[3] = sbits;
                qoff += gdn_words(q, (int)j, x.w);
            }
        }
    }
    // slab bases in the ring: consecutive, wrapping when a slab would run past the end
    u64 pos = 0;
    for (auto &q : P->reqs) {
        const u64 slab = (1ull << q.m) * (u64)q.d;
        for (int l = 0; l < q.r; l++) {
            if (pos + slab > ring_elems) pos = 0;
            ws_host[q.slabs + l] = pos;
            P->slab_pos.push_back(pos);
            pos += slab;
        }
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_primes), h_primes, sizeof(h_primes)) != hipSuccess ||
        hipMalloc((void **)&P->d_reqs, sizeof(Req) * P->reqs.size()) != hipSuccess ||
        hipMalloc((void **)&P->d_ws, sizeof(u64) * (words ? words : 1)) != hipSuccess ||
        hipMalloc((void **)&P->d_qreqs, sizeof(Req) * P->qreqs.size()) != hipSuccess ||
        hipMalloc((void **)&P->d_counter, 2 * sizeof(unsigned long long)) != hipSuccess) {
        qmc_fail(QMC_ERR_NOMEM, "batch: device allocation failed");
        qmc_batch_free(P);
        return nullptr;
    }
    if (hipMemcpy(P->d_reqs, P->reqs.data(), sizeof(Req) * P->reqs.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P->d_qreqs, P->qreqs.data(), sizeof(Req) * P->qreqs.size(), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(P->d_ws, ws_host.data(), sizeof(u64) * words, hipMemcpyHostToDevice) != hipSuccess) {
        qmc_fail(QMC_ERR_HIP, "batch: copying the plan to the device failed");
        qmc_batch_free(P);
        return nullptr;
    }
    // tables
    hipLaunchKernelGGL(batch_setup_kernel, dim3(64, (unsigned)P->reqs.size()), dim3(kBlock), 0, 0, P->d_reqs,
                       P->d_ws);
    if (hipDeviceSynchronize() != hipSuccess) {
        qmc_fail(QMC_ERR_HIP, "batch: setup kernel failed");
        qmc_batch_free(P);
        return nullptr;
    }
    // the second class's stream and the events that fork it from, and join it to, the caller's stream
    if (hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming) != hipSuccess) {
        qmc_fail(QMC_ERR_HIP, "batch: stream or event creation failed");
        qmc_batch_free(P);
        return nullptr;
    }
    return P;
}

QMC_EXPORT void qmc_batch_free(void *h) {
    Plan *P = (Plan *)h;
    if (!P) return;
    if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
    if (P->cap) (void)hipStreamDestroy(P->cap);
    if (P->side) (void)hipStreamDestroy(P->side);
    if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
    if (P->ev_join) (void)hipEventDestroy(P->ev_join);
    (void)hipFree(P->d_reqs);
    (void)hipFree(P->d_ws);
    (void)hipFree(P->d_counter);
    (void)hipFree(P->d_qreqs);
    delete P;
}

QMC_EXPORT u64 qmc_batch_total_tiles(void *h) { return ((Plan *)h)->total_tiles; }

// Number of replication slabs (sum of r over the requests): the length of a digest array.
QMC_EXPORT u64 qmc_batch_slabs(void *h) { return ((Plan *)h)->nslabs; }

// Make later runs of the plan add, for every replication slab, the wrap-around
// sum of the stored float64 bit patterns into digest[slab] (device memory of
// qmc_batch_slabs() u64, zeroed by the caller); NULL turns it off.  The sum is
// order-independent, so batched and per-call runs of a plan give equal digests
// exactly when they write equal points, whatever the ring placement.
QMC_EXPORT int qmc_batch_set_digest(void *h, unsigned long long *digest) {
    ((Plan *)h)->digest = digest;
    return QMC_OK;
}
QMC_EXPORT u64 qmc_batch_points(void *h) {
    u64 p = 0;
    for (auto &q : ((Plan *)h)->reqs) p += (u64)q.r * (1ull << q.m) * (u64)q.d;
    return p;
}
// element offset in the ring of replication l of request q
QMC_EXPORT u64 qmc_batch_slab_offset(void *h, int q, int l) {
    const Plan *P = (const Plan *)h;
    u64 s0 = 0;  // slabs are numbered request by request
    for (int k = 0; k < q; k++) s0 += (u64)P->reqs[k].r;
    return P->slab_pos[s0 + (u64)l];
}

// All requests in one persistent launch (capturable: a fill kernel resets the
// tile counter, no memset node).
QMC_EXPORT int qmc_batch_run_dev(void *h, double *ring, qmc_stream_t s) {
    Plan *P = (Plan *)h;
    if (!P->total_tiles) return QMC_OK;
    return launch_batch(P, ring, s);
}

// The same launch replayed from a hipGraph (captured once per ring on a
// private stream, then launched on s).  Calls on one plan must not overlap:
// they share the tile counter.
QMC_EXPORT int qmc_batch_graph_run_dev(void *h, double *ring, qmc_stream_t s) {
    Plan *P = (Plan *)h;
    if (!P->total_tiles) return QMC_OK;
    if (!P->gexec || P->gring != ring || P->gdigest != P->digest) {
        if (P->gexec) {
            QMC_HIP_CHECK(hipGraphExecDestroy(P->gexec));
            P->gexec = nullptr;
        }
        if (!P->cap) QMC_HIP_CHECK(hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        QMC_HIP_CHECK(hipStreamBeginCapture(P->cap, hipStreamCaptureModeThreadLocal));
        const int rc = launch_batch(P, ring, P->cap);
        const hipError_t ec = hipStreamEndCapture(P->cap, &g);
        if (rc) return rc;
        QMC_HIP_CHECK(ec);
        QMC_HIP_CHECK(hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0));
        QMC_HIP_CHECK(hipGraphDestroy(g));
        P->gring = ring;
        P->gdigest = P->digest;
    }
    QMC_HIP_CHECK(hipGraphLaunch(P->gexec, s));
    return QMC_OK;
}

// The same requests, one ordinary launch per request.
QMC_EXPORT int qmc_batch_run_per_call_dev(void *h, double *ring, qmc_stream_t s) {
    Plan *P = (Plan *)h;
    for (size_t i = 0; i < P->reqs.size(); i++) {
        const Req &q = P->reqs[i];
        const u64 tiles = (u64)q.r * q.tiles_per_l;
        const bool c0 = method_class(q.method) == 0;
        const unsigned lds = c0 ? kLdsClass0 : P->lds_bytes;
        if (P->digest) {
            if (c0)
                hipLaunchKernelGGL((batch_single_kernel<0, true>), dim3((unsigned)tiles), dim3(kBlock), lds, s,
                                   P->d_reqs + i, P->d_ws, ring, P->digest);
            else
                hipLaunchKernelGGL((batch_single_kernel<1, true>), dim3((unsigned)tiles), dim3(kBlock), lds, s,
                                   P->d_reqs + i, P->d_ws, ring, P->digest);
        } else {
            if (c0)
                hipLaunchKernelGGL((batch_single_kernel<0, false>), dim3((unsigned)tiles), dim3(kBlock), lds, s,
                                   P->d_reqs + i, P->d_ws, ring, P->digest);
            else
                hipLaunchKernelGGL((batch_single_kernel<1, false>), dim3((unsigned)tiles), dim3(kBlock), lds, s,
                                   P->d_reqs + i, P->d_ws, ring, P->digest);
        }
    }
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

// Host copies of the counter-based inputs, for checks against the reference
// kernels.  lattice: g[d], shifts[r*d].  dnb2: C[d*32] (unscrambled columns),
// S[r*d*32] (LMS rows), shifts[r*d] (64-bit digital shifts; NUS seeds for method 2).
QMC_EXPORT void qmc_batch_lattice_inputs(u64 seed, int m, int d, int r, u64 *g, double *shifts) {
    for (int j = 0; j < d; j++) g[j] = lat_g(seed, m, j);
    for (int l = 0; l < r; l++)
        for (int j = 0; j < d; j++) shifts[l * d + j] = lat_shift(seed, l, j);
}

QMC_EXPORT void qmc_batch_dnb2_inputs(u64 seed, int method, int d, int r, u64 *C, u64 *S, u64 *shifts) {
    for (int j = 0; j < d; j++)
        for (int k = 0; k < kTB; k++) C[j * kTB + k] = dnb2_col(seed, j, k);
    for (int l = 0; l < r; l++)
        for (int j = 0; j < d; j++) {
            for (int t = 0; t < kTB; t++) S[((u64)l * d + j) * kTB + t] = dnb2_srow(seed, l, j, t);
            shifts[l * d + j] = method == M_DNB2_NUS ? nus_seed(seed, l, j) : dnb2_shift(seed, l, j);
        }
}

// gdn inputs for dimension j: base, digit counts, S (tn x mm), permutation (a, c) per level
QMC_EXPORT int qmc_batch_gdn_inputs(u64 seed, int m, int l, int j, u32 *b, int *tn, int *mm, u32 *S, u32 *pa,
                                    u32 *pc) {
    fill_primes();
    *b = h_primes[j];
    *tn = gdn_tn(*b);
    *mm = gdn_mm(*b, m);
    for (int t = 0; t < *tn; t++) {
        for (int k = 0; k < *mm; k++) S[t * *mm + k] = k <= t ? gdn_s(seed, l, j, t, k, *b) : 0;
        gdn_perm(seed, l, j, t, *b, pa[t], pc[t]);
    }
    return QMC_OK;
}

QMC_EXPORT u64 qmc_owen_hash64(u64 x, u64 seed) { return owen_hash64(x, seed); }

}  // extern "C"
