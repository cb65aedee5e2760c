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
//   3 GDN_LMS    Halton-type base-b net (base = j-th prime, identity
//                 generating matrices), random lower-triangular LMS to 52
//                 digits, a uniform random permutation of the digit values per
//                 level, float (gdn_linear_matrix_scramble -> gdn_gen_natural
//                 -> gdn_digital_permutation -> gdn_integer_to_float)
//
// Random inputs are counter-based: every generating vector, matrix, shift and
// permutation is a hash of (request seed, role, indices), so a batch needs no
// host-drawn arrays.  This is a stream of its own -- the points equal this
// library's one-launch-per-request path bit for bit, and for methods 0, 1 and
// 3 they equal the reference kernels applied to the same inputs
// (qmc_batch_*_inputs export them), but they are not numpy's draws.
//
// Scheduling: each request is cut into tiles (rows of all columns for the
// lattice and dnb2 methods, rows of one column for gdn); setup kernels build
// every request's small tables (prefix-XOR columns, shifts, LMS rows; for gdn
// the permutations of the levels a point usually needs and the walk tables) in
// a workspace, then persistent grids drain tile counters: one kernel for the
// lattice and dnb2 tiles (store-bound, few registers) and one for the gdn
// tiles (digit walks and sequential sums, ALU-bound), each compiled without
// the other's code; when a batch has both, the gdn kernel runs on a second
// stream beside the other (one graph either way).  Outputs go to a ring buffer:
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
    u32 tiles_per_l;     // tiles per replication: n / rows_per_tile (times d for gdn)
    u64 tab;             // word offset of the request's tables in the workspace
    u64 slabs;           // word offset of r slab bases (output element offsets)
    u32 clog;            // dnb2: log2 of the row stride between a chain's points
    u32 gcap;            // gdn: most permutation levels a tile reads from its tables
    u32 gng;             // gdn: column groups (tiles per row range)
    u64 gblk;            // gdn: word offset of the replication blocks in the request's tables
    u64 lw;              // gdn: words of one replication's column tables
    DivU32 gdiv;         // gdn: division by gng
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

// ---------------------------------------------------------------- gdn inputs
//
// A gdn request is the reference's Halton-type chain with one tmax_new for all
// columns: bases b_j = the j-th prime; generating matrices the m x m identity
// (gdn_get_halton_generating_matrix, mmax = tmax = m); per (l, j) an LMS matrix
// S of kTNmax x m entries (gdn_get_linear_scramble_matrix: below the diagonal
// uniform in [0, b), the diagonal uniform in [1, b), zero above) and one
// uniform random permutation of the b digit values per level t < kTNmax
// (gdn_get_digital_permutations); then gdn_linear_matrix_scramble ->
// gdn_gen_natural -> gdn_digital_permutation -> gdn_integer_to_float.  With the
// identity C, C_lms[k][t] = S[t][k] (entries are below b, so the reference's
// sum without a final mod is exact), and digit t of point k is
//     z_t(k) = sum_i dig_i(k) S[t][i] mod b,   y_t = pi_t(z_t),
//     x = ((p_0 + p_1) + ...) + p_51,   p_t = RN(rec_t * y_t),
// rec_0 = RN(1/b), rec_{t+1} = RN(rec_t / b)  (gdn_integer_to_float.c:34-38).

// base-b digits of indices below 2^m (at least 1)
__host__ __device__ __forceinline__ int gdn_mm(u32 b, int m) {
    int t = 0;
    u64 v = 1;
    while (v < (1ull << m)) { v *= b; t++; }
    return t < 1 ? 1 : t;
}

// LMS entry S[t][i] of replication l, column j
__host__ __device__ __forceinline__ u32 gdn_s(u64 seed, u64 l, u64 j, u32 t, u32 i, u32 b) {
    if (i > t) return 0;
    const u64 h = key(seed, 7, (l << 32) | j, ((u64)t << 8) | i);
    return t == i ? (u32)(1 + h % (b - 1)) : (u32)(h % b);
}

// Permutation pi_t of replication l, column j at level t: pi_t(z) = the rank of
// key(z) among key(0), ..., key(b-1), key(z) = 53 hash bits above the 11 bits
// of z.  Ranking i.i.d. uniform keys gives a uniform random permutation; the
// index bits only break ties, and two keys share their 53 bits with
// probability below b^2 2^-54 < 1.5e-10 per level (b <= 1619).
__host__ __device__ __forceinline__ u64 gdn_pkey(u64 seed, u64 l, u64 j, u32 t, u32 z) {
    return (key(seed, 8, (l << 32) | j, ((u64)t << 16) | z) & ~0x7ffull) | z;
}

// pi_t(z) by counting: b keys (the tables hold the same ranks, by sorting)
__host__ __device__ inline u32 gdn_perm_at(u64 seed, u64 l, u64 j, u32 t, u32 b, u32 z) {
    const u64 kz = gdn_pkey(seed, l, j, t, z);
    u32 c = 0;
    for (u32 w = 0; w < b; w++) c += gdn_pkey(seed, l, j, t, w) < kz ? 1u : 0u;
    return c;
}

// Exact early exit of the sum.  B_t = RN(rec_t (b - 1)) bounds every later
// term: p_s = RN(rec_s y_s) <= RN(rec_t (b - 1)) for s >= t, since y_s <= b - 1,
// rec_s <= rec_t and rounding is monotone.  With thr_t = 2^(ilogb(B_t) + 54),
// a partial sum v >= thr_t has ulp(v) / 2 >= 2^(ilogb(B_t) + 1) > B_t, so
// v + p_s rounds back to v for every later term and v is final.  A wave stops
// adding once all its lanes' sums pass thr_t; lanes already past it would only
// have added terms that leave them unchanged.
//
// Levels kept in LDS per column: TB, the smallest of 8, 10, 12, 16, 24, 48 at
// or above the first t with b^t >= 2^74 (one kernel instance per count; the
// exact counts, ten instances, were no faster).  A point needs a level t >= TB only
// if its sum stays below thr_TB <= 2^55 b^-TB <= 2^-19 (a chance of about
// 2^-19 for these uniformly spread points); such levels are computed from the
// point's index and the permutation's keys (gdn_tail).  Base 2 keeps all 52
// levels as bits of one word.
__host__ __device__ inline int gdn_levels(u32 b) {
    if (b == 2) return kTNmax;
    int t = 0;
    double v = 1.0;
    while (v < 0x1p74) { v *= (double)b; t++; }
    return t <= 8 ? 8 : t <= 10 ? 10 : t <= 12 ? 12 : t <= 16 ? 16 : t <= 24 ? 24 : 48;
}

// u16 entries per LDS row of a permutation table: TB rounded up so that a row
// spans an odd number of 4-byte banks (random digits of a wave's lanes spread
// over the banks)
__host__ __device__ inline u32 gdn_row_u16(u32 TB) {
    u32 r = (TB + 1) & ~1u;
    if (((r / 2) & 1) == 0) r += 2;
    return r;
}

// Per-base constants at the start of the workspace: rec[256][kTNmax], then
// thr[256][kTNmax] (as doubles).
constexpr u64 kBaseWords = 2ull * 256 * kTNmax;

void gdn_base_tables(u32 b, double *rec, double *thr) {
    double rt = 1.0 / (double)b;
    for (int t = 0; t < kTNmax; t++) {
        rec[t] = rt;
        const double B = rt * (double)(b - 1);
        thr[t] = std::ldexp(1.0, std::ilogb(B) + 54);
        rt /= (double)b;
    }
}

// first level whose threshold is below 1 (sums stay below 1, so no earlier
// level can end one)
inline u32 gdn_tmin(const double *thr) {
    u32 t = 0;
    while (t < (u32)kTNmax && !(thr[t] < 1.0)) t++;
    return t;
}

// x mod b and x / b for x < 2^32 with M = ceil(2^64 / b) (exact, Lemire et al. 2019)
__host__ __device__ __forceinline__ u64 recip_magic(u32 b) { return b == 1 ? 0 : (~0ull) / b + 1; }
__device__ __forceinline__ u32 fast_div(u32 x, u64 M) { return (u32)__umul64hi(M, (u64)x); }

// ----------------------------------------------------------- table layouts
// lattice : g[d] | shift[r*d] (double bits)
// dnb2    : per column set (l, j) -- r*d sets for LMS, d for NUS -- 64 u32:
//           pre[32] (prefix XOR of the columns) then cols[32] | shift or NUS seed[r*d]
// gdn     : info[d] | r replication blocks of lw words, column j's block at
//           word colo_j of each:
//             odd b: u16 perm[b][TE] (entry z*TE + t = pi_t(z), TE = TB
//                    rounded up to even), then u16 Q[mm][TE] with
//                    Q_c[t] = sum_{i <= c} S[t][i] mod b
//                    (going from point k to k + 1 turns the c trailing digits
//                    equal to b - 1 into 0 and adds 1 to digit c, so every
//                    z_t moves by Q_c[t], as -(b - 1) = 1 mod b)
//             b = 2: u64 cmask (bit 51 - t = pi_t(0)), Q_0 .. Q_{m-1} (bits
//                    51 - t, mod 2)
//           info_j = colo_j | TB << 32 | mm << 40 | tmin << 48;
//           then desc_j (column j's LDS tables in its tile, gdn_plan), the
//           column groups of the tiles (j0 | ncols << 16 | logC << 24), and
//           the replication blocks

// u16 entries per row of a column's tables in the workspace
__host__ __device__ inline u32 gdn_row_even(u32 TB) { return (TB + 1) & ~1u; }

inline u64 gdn_col_words(u32 b, int m) {
    if (b == 2) return 1 + (u64)m;
    const u64 TE = gdn_row_even((u32)gdn_levels(b)), mm = (u64)gdn_mm(b, m);
    return ((b + mm) * TE * 2 + 7) / 8;
}

constexpr u32 kBlock1 = 1024;                                 // threads of a gdn workgroup
constexpr u32 kGdnStage = 2 * kBlock1 * (u32)sizeof(double);  // two stages of step values, at LDS offset 0

// Kinds of a column's LDS tables: base 2 (the walk words), permutation rows
// (u16) or value rows (doubles); see WalkOdd
enum { GK_B2 = 0, GK_PERM = 1, GK_VAL = 2 };

// levels a column walks: the stored TB, or a smaller bucket under the level cap
inline u32 gdn_walked(u32 TB, u32 gcap) {
    return gcap >= TB ? TB : gcap >= 24 ? 24 : gcap >= 16 ? 16 : gcap >= 12 ? 12 : gcap >= 10 ? 10 : 8;
}

// LDS bytes of one column's tables in a tile (16-byte multiples): the rows,
// the walk table twice (packed row offsets and plain), the early-exit
// thresholds of the walked levels, and for permutation rows the reciprocals.  Row offsets are packed in u16 halves whose sums must
// stay below 2^16: b * rowb < 2^15 (0 when that fails).
inline u32 gdn_slot_bytes(u32 b, int m, u32 TW, int kind) {
    if (kind == GK_B2) return (8 * (1 + (u32)m) + 15) & ~15u;
    const u32 mm = (u32)gdn_mm(b, m);
    const u32 rowb = kind == GK_VAL ? 8 * (TW + 1) : 2 * gdn_row_u16(TW);
    if ((u64)b * rowb >= (1u << 15)) return 0;
    const u32 rows = (b * rowb + 15) & ~15u, walk = (12 * mm * ((TW + 1) / 2) + 15) & ~15u;
    const u32 lv = (8 * TW + 15) & ~15u;
    return rows + walk + lv + (kind == GK_VAL ? 0u : lv);
}

// The column groups of a gdn request and every column's LDS tables: the 16
// columns of each 128-byte line of a row in one group if their permutation
// rows fit in the LDS left beside the stages, else halves of it, recursively;
// then, in each group, value rows instead of permutation rows for the columns
// with the most levels first, while the group still fits.  desc_j = LDS byte
// offset | kind << 32 | walked levels << 40.
constexpr u32 kGdnTablesMax = 160 * 1024 - 16 - kGdnStage;  // LDS beside the stages and the queue slots
struct GdnPlan {
    std::vector<u64> groups, desc;
    u32 lds = 0;     // LDS bytes of the largest tile, stages included
    int bad = -1;    // a column whose tables fit nowhere
};

inline void gdn_plan_cols(GdnPlan &p, int s, int e, int m, u32 gcap) {
    const int n = e - s;
    u32 C = 1, logC = 0;
    while ((int)C < n) C *= 2, logC++;
    std::vector<int> kind(n);
    std::vector<u32> TW(n), bytes(n);
    u64 total = 0;
    for (int c = 0; c < n; c++) {
        const u32 b = h_primes[s + c];
        const u32 TB = (u32)gdn_levels(b);
        TW[c] = b == 2 ? TB : gdn_walked(TB, gcap);
        // 24 or 48 levels only for b <= 23: value rows of at most 4.6 KB
        kind[c] = b == 2 ? GK_B2 : TW[c] >= 24 ? GK_VAL : GK_PERM;
        bytes[c] = gdn_slot_bytes(b, m, TW[c], kind[c]);
        if (!bytes[c]) {
            p.bad = s + c;
            return;
        }
        total += bytes[c];
    }
    if (total > kGdnTablesMax) {
        if (n == 1) {
            p.bad = s;
            return;
        }
        gdn_plan_cols(p, s, s + (int)C / 2, m, gcap);
        gdn_plan_cols(p, s + (int)C / 2, e, m, gcap);
        return;
    }
    std::vector<int> order(n);
    for (int c = 0; c < n; c++) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return TW[a] > TW[b]; });
    for (int c : order) {
        if (kind[c] != GK_PERM) continue;
        const u32 vb = gdn_slot_bytes(h_primes[s + c], m, TW[c], GK_VAL);
        if (vb && total - bytes[c] + vb <= kGdnTablesMax) {
            total = total - bytes[c] + vb;
            bytes[c] = vb;
            kind[c] = GK_VAL;
        }
    }
    u32 off = kGdnStage;
    for (int c = 0; c < n; c++) {
        p.desc[s + c] = off | ((u64)kind[c] << 32) | ((u64)TW[c] << 40);
        off += bytes[c];
    }
    p.groups.push_back((u64)s | ((u64)n << 16) | ((u64)logC << 24));
    p.lds = std::max(p.lds, off);
}

inline GdnPlan gdn_plan(int d, int m, u32 gcap) {
    GdnPlan p;
    p.desc.assign(d, 0);
    for (int a = 0; a < d && p.bad < 0; a += 16) gdn_plan_cols(p, a, std::min(d, a + 16), m, gcap);
    return p;
}

u64 table_words(Req &q) {
    const u64 d = q.d, r = q.r;
    switch (q.method) {
        case M_LATTICE: return d + r * d;
        case M_DNB2_LMS: return 32 * r * d + r * d;
        case M_DNB2_NUS: return 32 * d + r * d;
        default: {
            u64 w = 0;
            for (u64 j = 0; j < d; j++) w += gdn_col_words(h_primes[j], q.m);
            q.lw = w;
            q.gng = (u32)gdn_plan((int)d, q.m, q.gcap).groups.size();
            q.gblk = 2 * d + q.gng;
            return q.gblk + r * w;
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
    }
    // gdn: batch_gdn_setup_kernel
}

// gdn tables, one block per column (request, l, j) of every gdn request: the
// permutations of levels t < TB by sorting the b keys in LDS (bitonic, padded
// to a power of two with keys above every real one), and the walk table.
__global__ __launch_bounds__(kBlock) void batch_gdn_setup_kernel(const Req *__restrict__ reqs,
                                                                 const u32 *__restrict__ greq,
                                                                 const u64 *__restrict__ gcol0, int ng,
                                                                 u64 *__restrict__ ws) {
    __shared__ u64 keys[2048];
    const u64 col = blockIdx.x;
    int lo = 0, hi = ng - 1;  // last gdn request with gcol0 <= col
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (gcol0[mid] <= col) lo = mid;
        else hi = mid - 1;
    }
    const Req q = reqs[greq[lo]];
    const u64 c = col - gcol0[lo];
    const u64 d = q.d, l = c / d, j = c - (c / d) * d;
    u64 *tab = ws + q.tab;
    const u64 info = tab[j];
    const u32 TB = (u32)((info >> 32) & 0xff), mm = (u32)((info >> 40) & 0xff);
    u64 *blk = tab + q.gblk + l * q.lw + (u32)info;
    const u32 b = c_primes[j];
    if (b == 2) {
        if (threadIdx.x == 0) {
            u64 cm = 0;  // pi_t(0) = 1 iff key(1) < key(0)
            for (u32 t = 0; t < (u32)kTNmax; t++)
                cm |= (u64)(gdn_pkey(q.seed, l, j, t, 1) < gdn_pkey(q.seed, l, j, t, 0)) << (kTNmax - 1 - t);
            blk[0] = cm;
        }
        for (u32 cc = threadIdx.x; cc < (u32)q.m; cc += blockDim.x) {
            u64 w = 0;
            for (u32 t = 0; t < (u32)kTNmax; t++) {
                u32 run = 0;
                for (u32 i = 0; i <= cc && i <= t; i++) run ^= gdn_s(q.seed, l, j, t, i, 2) & 1u;
                w |= (u64)run << (kTNmax - 1 - t);
            }
            blk[1 + cc] = w;
        }
        return;
    }
    const u32 TE = gdn_row_even(TB);
    unsigned short *P = (unsigned short *)blk;
    unsigned short *Qt = P + (u64)b * TE;
    u32 N = 1;
    while (N < b) N *= 2;
    for (u32 t = 0; t < TB; t++) {
        for (u32 z = threadIdx.x; z < N; z += blockDim.x) keys[z] = z < b ? gdn_pkey(q.seed, l, j, t, z) : ~0ull;
        __syncthreads();
        for (u32 size = 2; size <= N; size <<= 1)
            for (u32 stride = size >> 1; stride > 0; stride >>= 1) {
                for (u32 i = threadIdx.x; i < N / 2; i += blockDim.x) {
                    const u32 a = 2 * i - (i & (stride - 1)), e = a + stride;
                    const u64 ka = keys[a], ke = keys[e];
                    if ((ka > ke) == ((a & size) == 0)) {
                        keys[a] = ke;
                        keys[e] = ka;
                    }
                }
                __syncthreads();
            }
        for (u32 pos = threadIdx.x; pos < b; pos += blockDim.x)
            P[(u64)(keys[pos] & 0x7ff) * TE + t] = (unsigned short)pos;
        __syncthreads();
    }
    for (u32 e = threadIdx.x; e < mm * TE; e += blockDim.x) {
        const u32 cc = e / TE, t = e - cc * TE;
        u32 run = 0;
        if (t < TB)
            for (u32 i = 0; i <= cc && i <= t; i++) run = (run + gdn_s(q.seed, l, j, t, i, b)) % b;
        Qt[e] = (unsigned short)run;
    }
    if (TE > TB)  // the padding level of the permutation rows
        for (u32 z = threadIdx.x; z < b; z += blockDim.x) P[(u64)z * TE + TB] = 0;
}

// ------------------------------------------------------------- one tile

// Output of a tile, and with DIG a running wrap-around sum of the stored bit
// patterns (an order-independent per-slab digest).  STREAM: non-temporal
// stores, for the store-bound class whose chains cover whole 128-byte lines;
// the gdn class's rows (d doubles) share lines with the next row, which
// cached stores complete in L2 (5% faster there, 3-5% slower for the
// store-bound class).
template <bool DIG, bool STREAM>
struct Sink {
    static constexpr bool kDig = DIG;
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
//
// The chains read their column constants from LDS, staged once per tile:
// one counter (vmcnt) tracks a wave's loads and stores in issue order, so a
// global load issued after a chain's stores would wait for all of them to
// complete, and every chain would start on an empty store queue.  Per column
// (kChainWords(s) words): cols[0 .. s), the transitions
// xk[k] = pre[s + k] ^ pre[s - 1] for k < 3 + kMaxGroupLog, the 64-bit shift
// or seed (low word first).
__host__ __device__ constexpr u32 kChainWords(u32 s) { return s + 3 + kMaxGroupLog + 2; }

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
    const u32 W = kChainWords(s);
    u32 *cst = v0 + 256;  // column constants, W words per column
    for (u32 e = threadIdx.x; e < d * W; e += blockDim.x) {
        const u32 j = e / W, w = e - j * W;
        const u32 *pre = T32 + (set + j) * 64;
        u32 val;
        if (w < s) {
            val = pre[32 + w];
        } else if (w < s + 3 + kMaxGroupLog) {
            const u32 k = w - s;
            val = s + k < (u32)kTB ? pre[s + k] ^ (s ? pre[s - 1] : 0u) : 0u;
        } else {
            const u64 xs = extra[j];
            val = w == W - 2 ? (u32)xs : (u32)(xs >> 32);
        }
        cst[e] = val;
    }
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
        const u32 *cj = cst + j * W;
        const u32 *cols = cj;
        const u32 *xkj = cj + s;
        const u64 xs = (u64)cj[W - 2] | ((u64)cj[W - 1] << 32);
        const double lo = (double)(u32)xs * 0x1p-64;
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
            for (int k = 0; k < 3 + kMaxGroupLog; k++) xk[k] = xkj[k];
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
            const u32 x0 = xkj[0], x1 = H > 1 ? xkj[1] : 0u;
            for (u32 h = 0;; h++) {
                out.put(dst, dnb2_emit<LMS>(v, xs, lo));
                if (h + 1 == H) break;
                v ^= (h & 1) ? x1 : x0;
                dst += stride;
            }
        }
    }
}

// A gdn tile is a group of C = 2^logC adjacent columns (C <= 16) times RT rows
// of one replication.  Column c of the group has L = 1024 / C lanes of the
// workgroup (threads [c L, (c + 1) L): whole waves), lane i walking the H = RT / L
// consecutive points row0 + i H .. row0 + i H + H - 1.  Groups follow the
// 128-byte lines of the rows: the 16 columns of a line share a group whenever
// their tables fit in LDS together, else its halves or quarters do.  After
// every step the group's values meet in LDS and are stored row by row, so one
// store instruction writes whole lines (C = 16) or 64- or 32-byte pieces of
// them.  A piece that other workgroups complete reaches HBM as a partial line
// whenever the L2 evicts the line first, and HBM writes 32-byte pieces as
// read-modify-writes (685 GB/s against 4.97 TB/s for whole lines, measured on
// a plain kernel): with four columns per group those writes bounded the class.

// Base 2 (j = 0).  Q = cmask, Q_0 .. Q_{m-1}.  The permuted digits of point k
// are cmask ^ (XOR of S_i = Q_i ^ Q_{i-1} over the set bits i of k), and a
// step k -> k + 1 adds Q_c, c = ctz(k + 1).  In base 2 every term rec_t y_t
// is 0 or 2^(-t-1) and every partial sum of the reference's loop is a
// multiple of 2^-52 below 1, so all of them are exact and x = Y 2^-52 for the
// 52-bit word Y.
struct Walk2 {
    const u64 *Q;
    u64 z;
    u32 k;
    __device__ __forceinline__ void init(const u64 *Q_, u32 k0) {
        Q = Q_;
        k = k0;
        z = Q[0];
        for (u32 kk = k0; kk; kk &= kk - 1) {
            const int i = __ffs(kk) - 1;
            z ^= Q[1 + i] ^ (i ? Q[i] : 0ull);
        }
    }
    __device__ __forceinline__ double value() const { return (double)z * 0x1p-52; }
    __device__ __forceinline__ void step() {
        z ^= Q[__ffs(k + 1)];  // Q_c at Q[1 + c]
        k++;
    }
};

// A column of odd base b, as its lanes walk it.
struct GdnCol {
    const double *rec, *thr;  // the base's constants (kTNmax each)
    const u32 *Qs, *Qu;       // walk table in LDS: packed pairs of row offsets, and plain
    const double *srec;       // permutation rows: the reciprocals of the walked levels in LDS
    const double *sthr;       // the early-exit thresholds of the walked levels in LDS
    u64 seed, l, j, M;
    u32 b, mm, rowb, base;    // rowb: bytes per table row; base: LDS byte address of the rows
};

// u16 and double at byte address a of the dynamic LDS, which starts at address
// 0 in the kernels that use these (they declare no static LDS)
__device__ __forceinline__ u32 lds_u16(u32 a) {
    return *(__attribute__((address_space(3))) const unsigned short *)(uintptr_t)a;
}
__device__ __forceinline__ double lds_f64(u32 a) {
    return *(__attribute__((address_space(3))) const double *)(uintptr_t)a;
}

__device__ __forceinline__ double uniform_f64(const double *p) {
    const u64 w = (u64) * (const long long *)p;
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)w), hi = __builtin_amdgcn_readfirstlane((u32)(w >> 32));
    return __longlong_as_double((long long)(((u64)hi << 32) | lo));
}

// Levels t .. kTNmax-1 of the sum for the lanes still below thr_t: digit t from
// the point's index digits, pi_t by counting keys.  Rare (see gdn_levels).
__device__ __noinline__ double gdn_tail(GdnCol c, u32 k, u32 t, double v) {
    for (; t < (u32)kTNmax; t++) {
        const bool act = v < c.thr[t];
        if (__builtin_amdgcn_ballot_w64(act) == 0) break;
        if (act) {
            u32 z = 0, kk = k;
            for (u32 i = 0; i < c.mm && i <= t; i++) {
                const u32 qn = fast_div(kk, c.M);
                z = (z + (kk - qn * c.b) * gdn_s(c.seed, c.l, c.j, t, i, c.b)) % c.b;  // < b^2 + b
                kk = qn;
            }
            v += c.rec[t] * (double)gdn_perm_at(c.seed, c.l, c.j, t, c.b, z);
        }
    }
    return v;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as_pair(u32 w) { return __builtin_bit_cast(u16x2, w); }
__device__ __forceinline__ u32 as_word(u16x2 p) { return __builtin_bit_cast(u32, p); }

// Consecutive points of a column of odd base b, the first TW levels from LDS.
// A lane keeps the digits z_t of its point as byte offsets z_t * rowb into the
// column's table rows, two levels per register (u16 halves: b * rowb < 2^15,
// host-checked), so a level's entry is one load with t in the immediate
// offset, and a step adds a table row of packed offsets modulo b * rowb.  The
// rows hold either the permutations pi_t(z) (u16; the term is then
// rec_t * pi_t(z), rec_t read from LDS as a broadcast) or, for VT, the terms
// themselves, V[z][t] = RN(rec_t * pi_t(z)) as doubles: the reference's
// product, rounded once as there (gdn_integer_to_float.c:36-37), so a level
// costs one 8-byte load and the add.  From level tchk on, a wave stops as soon
// as every one of its sums has reached the level's threshold (a lane whose sum
// has passed it would only add terms that leave it unchanged, gdn_levels);
// after TW levels one check decides whether the rare tail runs.
// Levels at which a wave of a TW-level column checks whether all its sums are
// final: for the bases that walk TW levels, the levels at which the 64 sums of
// a wave (spread uniformly over [0, 1)) have all passed the threshold with a
// good chance.  A check costs a compare and exposes the latency of the next
// LDS reads (they cannot be issued ahead of the branch), so there are few of
// them; at a level below the column's own first useful one (gdn_tmin) a check
// never ends the sum.  Fixed per TW: a check level read at run time made the
// compiler spill hundreds of registers.
__host__ __device__ constexpr bool gdn_check_level(int TW, int t) {
    return TW == 8    ? t == 7
           : TW == 10 ? t == 8
           : TW == 12 ? t == 10
           : TW == 16 ? (t == 12 || t == 14)
           : TW == 24 ? (t == 15 || t == 18)
                      : (t == 26 || t == 38 || t == 42);
}

template <int TW, bool VT>
struct WalkOdd {
    static constexpr int TP = (TW + 1) / 2;  // level pairs (the last half is padding when TW is odd)
    u32 z[TP];
    u32 k, dlow;
    __device__ __forceinline__ void init(const GdnCol &c, u32 k0) {
        const u32 b = c.b;
        const u64 M = c.M;
        k = k0;
        // z_t = sum_i (dig_i - dig_{i+1}) Q_i[t] mod b (dig_mm = 0; Abel
        // summation of S[t][i] = Q_i[t] - Q_{i-1}[t])
        u32 acc[2 * TP];
#pragma unroll
        for (int t = 0; t < 2 * TP; t++) acc[t] = 0;
        u32 qd = fast_div(k0, M);
        u32 dig = k0 - qd * b;
        for (u32 i = 0; i < c.mm; i++) {
            const u32 qn = fast_div(qd, M);
            const u32 dn = qd - qn * b;
            const u32 e = dig >= dn ? dig - dn : dig + b - dn;
            const u32 *Qi = c.Qu + i * 2 * TP;
#pragma unroll
            for (int t = 0; t < 2 * TP; t++) acc[t] += e * Qi[t];  // < mm b^2 < 2^32
            qd = qn;
            dig = dn;
        }
#pragma unroll
        for (int p = 0; p < TP; p++)
            z[p] = (acc[2 * p] - fast_div(acc[2 * p], M) * b) * c.rowb |
                   ((acc[2 * p + 1] - fast_div(acc[2 * p + 1], M) * b) * c.rowb) << 16;
        dlow = k0 - fast_div(k0, M) * b;
    }
    // the LDS address of level t's entry (t a constant after unrolling)
    __device__ __forceinline__ u32 addr(const GdnCol &c, int t) const {
        const u32 o = (t & 1) ? (z[t / 2] >> 16) : (z[t / 2] & 0xffffu);
        return o + c.base + (VT ? 8u : 2u) * (u32)t;
    }
    // gdn_integer_to_float.c:36-37; v starts as the first term (0.0 + p = p)
    __device__ __forceinline__ double value(const GdnCol &c, double thw) const {
        // the reciprocals of the permutation rows are read from LDS on every
        // point (as broadcasts) rather than held in registers: the pointer
        // passes through opaque()
        const double *sr = c.srec + (VT ? 0 : opaque(0));
        const double *sh = c.sthr + opaque(0);
        double v = 0.0;
#pragma unroll
        for (int t = 0; t < TW; t++) {
            // at the check levels, a wave whose sums have all reached thr_t is
            // done: every later term leaves them unchanged (gdn_levels)
            if (gdn_check_level(TW, t) && __builtin_amdgcn_ballot_w64(v < sh[t]) == 0) return v;
            const double p = VT ? lds_f64(addr(c, t)) : sr[t] * (double)lds_u16(addr(c, t));
            v = t == 0 ? p : v + p;
        }
        if (TW < kTNmax && __builtin_amdgcn_ballot_w64(v < thw) != 0) v = gdn_tail(c, k, TW, v);
        return v;
    }
    __device__ __forceinline__ void step(const GdnCol &c) {
        const u32 b = c.b;
        // carries of k -> k+1: cc trailing digits equal to b-1
        u32 cc = 0;
        if (dlow == b - 1) {
            dlow = 0;
            cc = 1;
            u32 u = fast_div(k, c.M);
            for (;;) {
                const u32 un = fast_div(u, c.M);
                if (u - un * b != b - 1) break;
                cc++;
                u = un;
            }
        } else {
            dlow++;
        }
        k++;
        const u32 *Qc = c.Qs + cc * TP;
        u32 inc[TP];
#pragma unroll
        for (int p = 0; p < TP; p++) inc[p] = Qc[p];
        const u16x2 bR = as_pair((b * c.rowb) * 0x10001u);
#pragma unroll
        for (int p = 0; p < TP; p++) {
            const u16x2 s2 = as_pair(z[p]) + as_pair(inc[p]);  // < 2 b rowb < 2^16
            z[p] = as_word(__builtin_elementwise_min(s2, s2 - bR));
        }
    }
};

// Byte offset of the walk tables in an odd-base column's LDS slot (after the
// rows), and of the reciprocals (permutation rows only, after the walk tables)
__host__ __device__ inline u32 gdn_qoff(u32 b, u32 TW, bool vt) {
    return vt ? ((b * (TW + 1) * 8 + 15) & ~15u) : ((b * gdn_row_u16(TW) * 2 + 15) & ~15u);
}
__host__ __device__ inline u32 gdn_walk_bytes(u32 mm, u32 TW) { return (12 * mm * ((TW + 1) / 2) + 15) & ~15u; }

// LDS fill of an odd-base column for TW walked levels out of TB stored, by the
// column's L threads (t): the rows -- permutations (level pairs as u32 words;
// TW, TB and the row length are even) or the terms V[z][t] (rows of TW + 1
// doubles, an odd count, so that random rows spread over the banks) -- then the
// walk table as packed pairs of row offsets and plain, and for permutation rows
// the reciprocals.  Eight loads in flight per thread before their LDS stores.
template <int TW, bool VT>
__device__ __forceinline__ void gdn_fill_odd(u32 t, u32 L, const u64 *__restrict__ blk, u32 b, u32 mm, u32 TB,
                                             unsigned char *slot, u32 *Qs, u32 *Qu, const double *__restrict__ rec,
                                             double *srec, const double *__restrict__ thr, double *sthr, u32 rowb) {
    constexpr u32 TP = (TW + 1) / 2;
    constexpr u32 kF = 8;
    const u32 TE = gdn_row_even(TB);
    if (VT) {
        const unsigned short *Pg = (const unsigned short *)blk;
        double *V = (double *)slot;
        const u32 nv = b * TW;
        for (u32 e0 = 0; e0 < nv; e0 += kF * L) {
            u32 w[kF];
#pragma unroll
            for (u32 f = 0; f < kF; f++) {
                const u32 e = e0 + f * L + t;
                const u32 z = e / TW, tt = e - z * TW;
                w[f] = e < nv ? Pg[z * TE + tt] : 0u;
            }
#pragma unroll
            for (u32 f = 0; f < kF; f++) {
                const u32 e = e0 + f * L + t;
                const u32 z = e / TW, tt = e - z * TW;
                if (e < nv) V[z * (TW + 1) + tt] = rec[tt] * (double)w[f];
            }
        }
    } else {
        const u32 *Pg = (const u32 *)blk;
        u32 *Pl = (u32 *)slot;
        const u32 TBp = TE / 2, rw = rowb / 4;  // u32 words per global and per LDS row
        const u32 nw = b * TP;
        for (u32 e0 = 0; e0 < nw; e0 += kF * L) {
            u32 w[kF];
#pragma unroll
            for (u32 f = 0; f < kF; f++) {
                const u32 e = e0 + f * L + t;
                const u32 z = e / TP, tp = e - z * TP;
                w[f] = e < nw ? Pg[z * TBp + tp] : 0u;
            }
#pragma unroll
            for (u32 f = 0; f < kF; f++) {
                const u32 e = e0 + f * L + t;
                const u32 z = e / TP, tp = e - z * TP;
                if (e < nw) Pl[z * rw + tp] = w[f];
            }
        }
        if (t < TW) srec[t] = rec[t];
    }
    if (t < TW) sthr[t] = thr[t];
    const unsigned short *Qg = (const unsigned short *)blk + (u64)b * TE;
    for (u32 e = t; e < mm * TP; e += L) {
        const u32 cc = e / TP, tp = e - cc * TP;
        const u32 lo = Qg[cc * TE + 2 * tp], hi = 2 * tp + 1 < TW ? Qg[cc * TE + 2 * tp + 1] : 0u;
        Qs[e] = (lo * rowb) | ((hi * rowb) << 16);
        Qu[2 * e] = lo;
        Qu[2 * e + 1] = hi;
    }
}

// Tiles are drained by two persistent kernels: class 0 = lattice and dnb2
// (HBM-write bound, few registers, full occupancy), class 1 = gdn (digit walks
// and the sequential sums).  Each class is compiled without the other's code.
__host__ __device__ constexpr int method_class(int method) { return method == M_GDN_LMS ? 1 : 0; }

// The step loop of a gdn tile: every lane of an active column computes its
// point, the values meet in LDS and the workgroup stores them row by row
// (thread e: row e / C, column e % C of the step).  The stage slot of (row i,
// column c) is i C + (c ^ ((i >> (4 - logC)) mod C)): 16 consecutive lanes of
// a column, and 16 consecutive threads of the store, touch 16 different
// 8-byte bank pairs.  H barriers for every wave of the workgroup, whatever its
// column's kind (the waves of one group run different instantiations of this
// loop, with equal barrier counts).
template <class W, class SinkT, class Value, class Step>
__device__ __forceinline__ void gdn_steps(W &w, bool act, u32 H, u32 i, u32 cg, u32 logC, u32 ncols, u32 rows,
                                          u64 row0, u32 d, u32 j0, double *__restrict__ slab, double *stage,
                                          SinkT &out, Value value, Step step) {
    const u32 C = 1u << logC, sh = 4 - logC;
    const u32 rr = threadIdx.x >> logC, cc = threadIdx.x & (C - 1);
    const bool st = cc < ncols && rr < rows;
    const u32 wslot = (i << logC) + (cg ^ ((i >> sh) & (C - 1)));
    const u32 rslot = (rr << logC) + (cc ^ ((rr >> sh) & (C - 1)));
    double *dst = slab + (row0 + (u64)rr * H) * d + j0 + cc;
    for (u32 h = 0; h < H; h++) {
        double *sb = stage + (h & 1) * kBlock1;
        if (act) sb[wslot] = value(w);
        __syncthreads();
        if (st) out.put(dst, sb[rslot]);
        dst += d;
        if (act && h + 1 < H) step(w);
    }
}

// A gdn tile: rows [chunk RT, (chunk + 1) RT) of the column group g of
// replication l, rem = chunk * G + g (the column groups of one row range are
// consecutive tiles).  Group word: j0 | ncols << 16 | logC << 24; column word
// (desc): LDS byte offset of its tables | kind << 32 | walked levels << 40.
template <class SinkT>
__device__ __forceinline__ void gdn_tile(const Req &q, u64 l, u32 rem, const u64 *__restrict__ ws,
                                         const u64 *__restrict__ tab, double *__restrict__ slab,
                                         unsigned char *lds, SinkT &out) {
    const u32 d = (u32)q.d;
    const u32 chunk = div_q(rem, q.gdiv);
    const u32 g = rem - chunk * q.gng;
    const u64 gw = tab[2 * (u64)d + g];
    const u32 j0 = __builtin_amdgcn_readfirstlane((u32)(gw & 0xffff));
    const u32 ncols = __builtin_amdgcn_readfirstlane((u32)((gw >> 16) & 0xff));
    const u32 logC = __builtin_amdgcn_readfirstlane((u32)((gw >> 24) & 0x7));
    const u32 logL = 10 - logC, L = 1u << logL;  // lanes per column
    const u32 cg = __builtin_amdgcn_readfirstlane(threadIdx.x >> logL);
    const u32 i = threadIdx.x & (L - 1);
    const bool colact = cg < ncols;
    const u32 j = j0 + (colact ? cg : 0u);
    const u32 RT = q.rows_per_tile;
    const u64 info = tab[j], desc = tab[d + j];
    const u32 colo = __builtin_amdgcn_readfirstlane((u32)info);
    const u32 TB = __builtin_amdgcn_readfirstlane((u32)((info >> 32) & 0xff));
    const u32 mm = __builtin_amdgcn_readfirstlane((u32)((info >> 40) & 0xff));
    const u32 soff = __builtin_amdgcn_readfirstlane((u32)desc);
    const u32 kind = __builtin_amdgcn_readfirstlane((u32)((desc >> 32) & 0xff));
    const u32 TW = __builtin_amdgcn_readfirstlane((u32)((desc >> 40) & 0xff));
    const u64 *blk = tab + q.gblk + l * q.lw + colo;
    const u32 b = c_primes[j];
    const u32 H = RT >= L ? RT >> logL : 1u;
    const u32 rows = RT >= L ? L : RT;  // lanes with points
    const bool act = colact && i < rows;
    const u64 row0 = (u64)chunk * RT;
    const u32 k = (u32)row0 + i * H;
    unsigned char *slot = lds + soff;
    double *stage = (double *)lds;
    if (kind == GK_B2) {  // column 0
        u64 *Q = (u64 *)slot;
        if (colact)
            for (u32 e = i; e <= (u32)q.m; e += L) Q[e] = blk[e];
        __syncthreads();
        Walk2 w;
        if (act) w.init(Q, k);
        gdn_steps(w, act, H, i, cg, logC, ncols, rows, row0, d, j0, slab, stage, out,
                  [](Walk2 &x) { return x.value(); }, [](Walk2 &x) { x.step(); });
        return;
    }
    const bool vt = kind == GK_VAL;
    const u32 TP = (TW + 1) / 2;
    // byte offsets in the slot (pointer arithmetic keeps the LDS address space)
    const u32 qoff = gdn_qoff(b, TW, vt);
    u32 *Qs = (u32 *)(slot + qoff);
    u32 *Qu = Qs + mm * TP;
    double *sthr = (double *)(slot + qoff + gdn_walk_bytes(mm, TW));
    double *srec = sthr + (((8 * TW + 15) & ~15u) / 8);
    GdnCol c;
    c.rec = (const double *)ws + (u64)j * kTNmax;
    c.thr = (const double *)ws + (u64)(256 + j) * kTNmax;
    c.Qs = Qs;
    c.Qu = Qu;
    c.srec = srec;
    c.sthr = sthr;
    c.seed = q.seed;
    c.l = l;
    c.j = j;
    c.M = recip_magic(b);
    c.b = b;
    c.mm = mm;
    c.rowb = vt ? 8 * (TW + 1) : 2 * gdn_row_u16(TW);
    c.base = (u32)(uintptr_t)(__attribute__((address_space(3))) unsigned char *)slot;
    const double thw = uniform_f64(c.thr + (TW < (u32)kTNmax ? TW : 0u));
#define QMC_GDN_TW(T_, VT_)                                                                                    \
    case (T_) | ((VT_) ? 0x100 : 0): {                                                                         \
        if (colact) gdn_fill_odd<T_, VT_>(i, L, blk, b, mm, TB, slot, Qs, Qu, c.rec, srec, c.thr, sthr, c.rowb); \
        __syncthreads();                                                                                       \
        WalkOdd<T_, VT_> w;                                                                                    \
        if (act) w.init(c, k);                                                                                 \
        gdn_steps(w, act, H, i, cg, logC, ncols, rows, row0, d, j0, slab, stage, out,                         \
                  [&](WalkOdd<T_, VT_> &x) { return x.value(c, thw); }, [&](WalkOdd<T_, VT_> &x) { x.step(c); }); \
    } break;
    switch (TW | (vt ? 0x100u : 0u)) {
        QMC_GDN_TW(8, false) QMC_GDN_TW(10, false) QMC_GDN_TW(12, false) QMC_GDN_TW(16, false)
        QMC_GDN_TW(8, true) QMC_GDN_TW(10, true) QMC_GDN_TW(12, true) QMC_GDN_TW(16, true)
        QMC_GDN_TW(24, true) QMC_GDN_TW(48, true)
        default: break;  // never: the planner gives one of these
    }
#undef QMC_GDN_TW
}

template <int CLASS, bool DIG>
__device__ void run_tile(const Req &q, u64 t, const u64 *__restrict__ ws, double *__restrict__ ring,
                         unsigned char *lds, unsigned long long *digest) {
    const u64 d = q.d;
    const u64 l = t / q.tiles_per_l;
    const u64 rem = t - l * q.tiles_per_l;
    const u64 *tab = ws + q.tab;
    double *slab = ring + ws[q.slabs + l];
    Sink<DIG, CLASS == 0> out;
    if (CLASS == 0) {
        const u64 row0 = rem * (u64)q.rows_per_tile;
        if (q.method == M_LATTICE) {
            // g (as exact doubles) and the replication's shifts staged in LDS:
            // a global load after the stores would wait for them (dnb2_tile)
            const u32 pts = q.rows_per_tile * (u32)d;
            double *sg = (double *)lds, *ssh = sg + d;
            for (u32 j = threadIdx.x; j < d; j += blockDim.x) {
                sg[j] = (double)tab[j];
                ssh[j] = ((const double *)(tab + d))[l * d + j];
            }
            __syncthreads();
            for (u32 e = threadIdx.x; e < pts; e += blockDim.x) {
                const u32 rr = div_q(e, q.ddiv);
                const u32 j = e - rr * q.d;
                const u64 row = row0 + rr;
                const double phi = (double)__brevll(row) * 0x1p-64;  // phi2(row), exact
                const double x = frac_mod1(sg[j] * phi);             // lat_gen_natural.c:45
                out.put(slab + row * d + j, frac_mod1(x + ssh[j]));  // lat_shift_mod_1.c:31
            }
        } else if (q.method == M_DNB2_LMS) {
            dnb2_tile<true, DIG>(q, l, row0, tab, slab, (u32 *)lds, out);
        } else {
            dnb2_tile<false, DIG>(q, l, row0, tab, slab, (u32 *)lds, out);
        }
    } else {
        gdn_tile(q, l, (u32)rem, ws, tab, slab, lds, out);
    }
    out.flush(digest + q.slab0 + l);
}

// ------------------------------------------------------------ the kernels

// Persistent workgroups take tiles from a counter until it passes the class's
// tile count.  The queue slots sit at byte `slot` of the dynamic LDS, after the
// tile's tables (the kernels declare no static LDS).
//
// The gdn class (CLASS 1) takes its tiles from eight queues, one per XCD: its
// tiles come in groups (request, l, row range) of the d column tiles that
// write the same rows, each group on one XCD's queue, columns in order.  A
// workgroup drains the queue of the XCD it runs on (the hardware XCC_ID
// register), so the column tiles that write the 8-byte pieces of a 128-byte
// line run at the same time in the same L2 and the line leaves it whole; then
// it helps with the other queues.  xq = [group offsets per XCD (9) | tiles per
// XCD (8) | per group: tiles before it in its XCD's queue | per group: its
// first tile].
template <int CLASS, bool DIG>
__global__ __launch_bounds__(CLASS == 1 ? kBlock1 : kBlock) void batch_persistent_kernel(const Req *__restrict__ reqs, int nreq,
                                                                  u64 total_tiles, unsigned long long *counter,
                                                                  const u64 *__restrict__ xq,
                                                                  const u64 *__restrict__ ws,
                                                                  double *__restrict__ out,
                                                                  unsigned long long *digest, u32 slot) {
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    u64 *s_tile = (u64 *)(lds_dyn + slot);
    int *s_req = (int *)(lds_dyn + slot + 8);
    u32 qx = 0;  // queues this workgroup found empty (thread 0)
    const u32 x0 = CLASS == 1 ? (__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 7u) : 0u;  // HW_REG_XCC_ID
    for (;;) {
        if (threadIdx.x == 0) {
            u64 t = total_tiles;
            if (CLASS == 0) {
                t = atomicAdd(counter, 1ull);
            } else {
                const u64 G = xq[8];
                while (qx < 8) {
                    const u32 x = (x0 + qx) & 7u;
                    const u64 u = atomicAdd(counter + x, 1ull);
                    if (u < xq[9 + x]) {
                        u64 lo = xq[x], hi = xq[x + 1] - 1;  // last group of queue x with tiles-before <= u
                        while (lo < hi) {
                            const u64 mid = (lo + hi + 1) / 2;
                            if (xq[17 + mid] <= u) lo = mid;
                            else hi = mid - 1;
                        }
                        t = xq[17 + G + lo] + (u - xq[17 + lo]);
                        break;
                    }
                    qx++;
                }
            }
            *s_tile = t;
            int lo = 0, hi = nreq - 1;  // last request with tile0 <= t
            if (t < total_tiles) {
                while (lo < hi) {
                    const int mid = (lo + hi + 1) / 2;
                    if (reqs[mid].tile0 <= t) lo = mid;
                    else hi = mid - 1;
                }
            }
            *s_req = lo;
        }
        __syncthreads();
        const u64 t = *s_tile;
        const int qi = *s_req;
        __syncthreads();
        if (t >= total_tiles) break;  // every wave of every block reaches this
        const Req q = reqs[qi];
        run_tile<CLASS, DIG>(q, t - q.tile0, ws, out, lds_dyn, digest);
    }
}

// the tile counters of both classes (1 + 8 queues)
constexpr int kCounters = 9;
__global__ void reset_counter_kernel(unsigned long long *counter) {
    if (threadIdx.x < kCounters) counter[threadIdx.x] = 0;
}

// one request, one tile per block (the per-call form)
template <int CLASS, bool DIG>
__global__ __launch_bounds__(CLASS == 1 ? kBlock1 : kBlock) void batch_single_kernel(const Req *__restrict__ req, const u64 *__restrict__ ws,
                                                              double *__restrict__ out, unsigned long long *digest) {
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    const Req q = *req;
    run_tile<CLASS, DIG>(q, blockIdx.x, ws, out, lds_dyn, digest);
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

// LDS of a class-0 tile: dnb2 first-row values (256 words) and chain
// constants (dnb2_tile), or the lattice's g and shifts
constexpr u32 kLdsClass0 = 256 * sizeof(u32);
inline u32 class0_lds(const Req &q) {
    if (q.method == M_LATTICE) return 16u * (u32)q.d;
    return 4u * (256u + (u32)q.d * kChainWords(q.clog));
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
    u64 *d_xq = nullptr;             // the gdn class's per-XCD queues (see batch_persistent_kernel)
    int grid = 0;                    // persistent workgroups of the lattice and dnb2 class
    int grid1 = 0;                   // persistent workgroups of the gdn class (kBlock1 threads each)
    int cus = 0;
    int device = 0;
    u32 lds_bytes[2] = {kLdsClass0, 0};  // tile tables per class; the queue slots follow them
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

inline u32 lds_total(u32 bytes) { return ((bytes + 15) & ~15u) + 16; }

template <int CLASS, bool DIG>
void launch_persistent(Plan *P, double *ring, hipStream_t s) {
    const u32 slot = (P->lds_bytes[CLASS] + 15) & ~15u;
    hipLaunchKernelGGL((batch_persistent_kernel<CLASS, DIG>), dim3(CLASS ? P->grid1 : P->grid),
                       dim3(CLASS ? kBlock1 : kBlock), lds_total(P->lds_bytes[CLASS]),
                       s, P->d_qreqs + (CLASS ? P->qn[0] : 0), P->qn[CLASS], P->qtiles[CLASS], P->d_counter + CLASS,
                       P->d_xq, P->d_ws, ring, P->digest, slot);
}

// With both classes present, the gdn kernel (the long tiles) is forked onto the
// plan's side stream and joined back through events, so a graph captured from
// s takes in both branches.  On large batches this changes nothing (each
// persistent grid holds its CUs until its queue drains); on small ragged ones
// the store-bound tiles finish under the gdn tiles (9% on a 48-request batch).
template <bool DIG>
int launch_batch_t(Plan *P, double *ring, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, P->d_counter);
    if (P->qtiles[0] && P->qtiles[1]) {
        QMC_HIP_CHECK(hipEventRecord(P->ev_fork, s));
        QMC_HIP_CHECK(hipStreamWaitEvent(P->side, P->ev_fork, 0));
        launch_persistent<1, DIG>(P, ring, P->side);
        QMC_HIP_CHECK(hipEventRecord(P->ev_join, P->side));
        launch_persistent<0, DIG>(P, ring, s);
        QMC_HIP_CHECK(hipStreamWaitEvent(s, P->ev_join, 0));
        QMC_HIP_CHECK(hipGetLastError());
        return QMC_OK;
    }
    if (P->qtiles[0]) launch_persistent<0, DIG>(P, ring, s);
    if (P->qtiles[1]) launch_persistent<1, DIG>(P, ring, s);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int launch_batch(Plan *P, double *ring, hipStream_t s) {
    return P->digest ? launch_batch_t<true>(P, ring, s) : launch_batch_t<false>(P, ring, s);
}

}  // namespace

extern "C" {

QMC_EXPORT void qmc_batch_free(void *h);

// Build a plan for nreq requests: methods[q] in 0..3, m[q] = log2 n in [1, 32]
// (1..31 for lattices), d[q] in [1, 256], r[q] >= 1, seeds[q]; the tunables are
// tile_points[method] (the points a tile of that method holds: rows of all d
// columns for methods 0..2, rows of a group of columns for gdn, rounded down
// to a power of two of rows) and `grid` (persistent workgroups of the
// lattice/dnb2 class; 0 = 4 per CU); `ring_elems` is the output ring size in
// doubles (every replication slab must fit); `gdn_level_cap` (0 = none) limits
// the permutation levels a gdn tile reads from its tables, the rest taking the
// index-and-key path (a test hook for that path; the output does not change).
QMC_EXPORT void *qmc_batch_plan_tiles(int nreq, const int32_t *methods, const int32_t *m, const int32_t *d,
                                      const int32_t *r, const u64 *seeds, const u64 *tile_points, int grid,
                                      u64 ring_elems, int gdn_level_cap) {
    fill_primes();
    Plan *P = new Plan();
    P->ring_elems = ring_elems;
    u64 tile = 0, words = kBaseWords;
    for (int i = 0; i < nreq; i++) {
        Req q{};
        q.method = methods[i];
        q.m = m[i];
        q.d = d[i];
        q.r = r[i];
        q.seed = seeds[i];
        if (q.method < 0 || q.method > 3 || q.m < 1 || q.m > 32 || q.d < 1 || q.d > 256 || q.r < 1 ||
            (q.method == M_LATTICE && q.m > 31)) {
            qmc_fail(QMC_ERR_ARG, "batch: request %d has unsupported (method, m, d, r) = (%d, %d, %d, %d)", i,
                     q.method, q.m, q.d, q.r);
            qmc_batch_free(P);
            return nullptr;
        }
        const u64 n = 1ull << q.m;
        u64 rows;
        if (q.method == M_GDN_LMS) {
            rows = tile_points[3] / 4;  // a tile of 4 columns
            if (rows < 256) rows = 256;
        } else {
            rows = tile_points[q.method] / (u64)q.d;
            if ((q.method == M_DNB2_LMS || q.method == M_DNB2_NUS) && rows < 1024) rows = 1024;  // 64-step chains at 2^s = 16
        }
        u64 p2 = 1;
        while (p2 * 2 <= rows) p2 *= 2;
        if (p2 > n) p2 = n;
        q.rows_per_tile = (u32)p2;
        q.gcap = gdn_level_cap > 0 ? (u32)gdn_level_cap : 64u;
        q.tab = words;
        words += table_words(q);
        q.tiles_per_l = (u32)(n / p2) * (q.method == M_GDN_LMS ? q.gng : 1u);
        q.tile0 = tile;
        tile += (u64)q.r * q.tiles_per_l;
        q.ddiv = make_div((u32)q.d);
        q.gdiv = make_div(q.method == M_GDN_LMS ? q.gng : 1u);
        if (q.method == M_DNB2_LMS || q.method == M_DNB2_NUS) q.clog = dnb2_chain_log(q.rows_per_tile, (u32)q.d);
        if (method_class(q.method) == 0) P->lds_bytes[0] = std::max(P->lds_bytes[0], class0_lds(q));
        if (q.method == M_GDN_LMS) {
            const GdnPlan gp = gdn_plan(q.d, q.m, q.gcap);
            if (gp.bad >= 0) {
                qmc_fail(QMC_ERR_ARG, "batch: gdn column %d has too large a permutation table", gp.bad);
                qmc_batch_free(P);
                return nullptr;
            }
            P->lds_bytes[1] = std::max(P->lds_bytes[1], gp.lds);
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
    if (P->qtiles[1]) {
        // gdn workgroups can hold up to 160 KiB of LDS (more than the default 64 KiB)
        static PerDeviceOnce attrs;
        const int rc = attrs.run([&]() -> int {
            const int most = 160 * 1024;
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)batch_persistent_kernel<1, false>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)batch_persistent_kernel<1, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)batch_single_kernel<1, false>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)batch_single_kernel<1, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            return QMC_OK;
        });
        if (rc != QMC_OK || lds_total(P->lds_bytes[1]) > 160u * 1024) {
            if (rc == QMC_OK) qmc_fail(QMC_ERR_ARG, "batch: gdn tiles need %u bytes of LDS", lds_total(P->lds_bytes[1]));
            qmc_batch_free(P);
            return nullptr;
        }
    }
    // persistent gdn workgroups: one per CU (their registers fill it); with
    // both classes in a batch, the store-bound kernel runs on the CUs left
    // (qmc_batch_set_gdn_cus), by default a quarter of them
    P->cus = cus;
    P->grid1 = P->qtiles[0] ? (3 * cus) / 4 : cus;  // 192 of 256: 182 ms against 192 ms at 256 (512 requests, m <= 16)
    std::vector<u64> ws_host(words, 0);
    // per-base constants, then each gdn request's column info (its tables are built on the device)
    for (int j = 0; j < 256; j++) {
        gdn_base_tables(h_primes[j], (double *)ws_host.data() + (u64)j * kTNmax,
                        (double *)ws_host.data() + (u64)(256 + j) * kTNmax);
    }
    std::vector<u32> greq;
    std::vector<u64> gcol0;
    u64 gcols = 0;
    for (size_t qi = 0; qi < P->reqs.size(); qi++) {
        const Req &q = P->reqs[qi];
        if (q.method != M_GDN_LMS) continue;
        u64 colo = 0;
        for (int j = 0; j < q.d; j++) {
            const u32 b = h_primes[j];
            const u64 tmin = gdn_tmin((const double *)ws_host.data() + (u64)(256 + j) * kTNmax);
            ws_host[q.tab + j] = colo | ((u64)gdn_levels(b) << 32) | ((u64)gdn_mm(b, q.m) << 40) | (tmin << 48);
            colo += gdn_col_words(b, q.m);
        }
        const GdnPlan gp = gdn_plan(q.d, q.m, q.gcap);
        for (int j = 0; j < q.d; j++) ws_host[q.tab + q.d + j] = gp.desc[j];
        for (size_t g = 0; g < gp.groups.size(); g++) ws_host[q.tab + 2 * (u64)q.d + g] = gp.groups[g];
        greq.push_back((u32)qi);
        gcol0.push_back(gcols);
        gcols += (u64)q.r * (u64)q.d;
    }
    if (gcols > 0x7fffffffull) {
        qmc_fail(QMC_ERR_ARG, "batch: too many gdn columns (%llu)", (unsigned long long)gcols);
        qmc_batch_free(P);
        return nullptr;
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
    // the gdn class's tile groups (request, l, row range), dealt to the XCD queues in turn
    std::vector<u64> xq(17, 0);
    {
        std::vector<std::vector<std::pair<u64, u64>>> lists(8);  // (first tile, tiles) per XCD
        u64 gi = 0;
        for (int i = P->qn[0]; i < P->qn[0] + P->qn[1]; i++) {
            const Req &q = P->qreqs[i];
            const u64 nch = q.tiles_per_l / (u64)q.gng;
            for (u64 l = 0; l < (u64)q.r; l++)
                for (u64 ch = 0; ch < nch; ch++, gi++)
                    lists[gi & 7].push_back({q.tile0 + l * q.tiles_per_l + ch * (u64)q.gng, (u64)q.gng});
        }
        std::vector<u64> pre, first;
        for (int x = 0; x < 8; x++) {
            xq[x] = pre.size();
            u64 acc = 0;
            for (auto &g : lists[x]) {
                pre.push_back(acc);
                first.push_back(g.first);
                acc += g.second;
            }
            xq[9 + x] = acc;
        }
        xq[8] = pre.size();
        xq.insert(xq.end(), pre.begin(), pre.end());
        xq.insert(xq.end(), first.begin(), first.end());
    }
    u32 *d_greq = nullptr;
    u64 *d_gcol0 = nullptr;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_primes), h_primes, sizeof(h_primes)) != hipSuccess ||
        hipMalloc((void **)&P->d_reqs, sizeof(Req) * P->reqs.size()) != hipSuccess ||
        hipMalloc((void **)&P->d_ws, sizeof(u64) * words) != hipSuccess ||
        hipMalloc((void **)&P->d_qreqs, sizeof(Req) * P->qreqs.size()) != hipSuccess ||
        hipMalloc((void **)&P->d_counter, kCounters * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void **)&P->d_xq, sizeof(u64) * xq.size()) != hipSuccess ||
        (!greq.empty() && (hipMalloc((void **)&d_greq, sizeof(u32) * greq.size()) != hipSuccess ||
                           hipMalloc((void **)&d_gcol0, sizeof(u64) * gcol0.size()) != hipSuccess))) {
        (void)hipFree(d_greq);
        qmc_fail(QMC_ERR_NOMEM, "batch: device allocation failed");
        qmc_batch_free(P);
        return nullptr;
    }
    bool ok = hipMemcpy(P->d_reqs, P->reqs.data(), sizeof(Req) * P->reqs.size(), hipMemcpyHostToDevice) ==
                  hipSuccess &&
              hipMemcpy(P->d_qreqs, P->qreqs.data(), sizeof(Req) * P->qreqs.size(), hipMemcpyHostToDevice) ==
                  hipSuccess &&
              hipMemcpy(P->d_ws, ws_host.data(), sizeof(u64) * words, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(P->d_xq, xq.data(), sizeof(u64) * xq.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !greq.empty())
        ok = hipMemcpy(d_greq, greq.data(), sizeof(u32) * greq.size(), hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d_gcol0, gcol0.data(), sizeof(u64) * gcol0.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        (void)hipFree(d_greq);
        (void)hipFree(d_gcol0);
        qmc_fail(QMC_ERR_HIP, "batch: copying the plan to the device failed");
        qmc_batch_free(P);
        return nullptr;
    }
    // tables
    hipLaunchKernelGGL(batch_setup_kernel, dim3(64, (unsigned)P->reqs.size()), dim3(kBlock), 0, 0, P->d_reqs,
                       P->d_ws);
    if (gcols)
        hipLaunchKernelGGL(batch_gdn_setup_kernel, dim3((unsigned)gcols), dim3(kBlock), 0, 0, P->d_reqs, d_greq,
                           d_gcol0, (int)greq.size(), P->d_ws);
    const hipError_t se = hipDeviceSynchronize();
    (void)hipFree(d_greq);
    (void)hipFree(d_gcol0);
    if (se != hipSuccess) {
        qmc_fail(QMC_ERR_HIP, "batch: setup kernels failed");
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

// The same plan with one tile size for the lattice and dnb2 methods and
// `gdn_points_per_tile` (0 = points_per_tile) for gdn.
QMC_EXPORT void *qmc_batch_plan(int nreq, const int32_t *methods, const int32_t *m, const int32_t *d,
                                const int32_t *r, const u64 *seeds, u64 points_per_tile, int grid, u64 ring_elems,
                                u64 gdn_points_per_tile, int gdn_level_cap) {
    const u64 tp[4] = {points_per_tile, points_per_tile, points_per_tile,
                       gdn_points_per_tile ? gdn_points_per_tile : points_per_tile};
    return qmc_batch_plan_tiles(nreq, methods, m, d, r, seeds, tp, grid, ring_elems, gdn_level_cap);
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
    (void)hipFree(P->d_xq);
    (void)hipFree(P->d_qreqs);
    delete P;
}

QMC_EXPORT u64 qmc_batch_total_tiles(void *h) { return ((Plan *)h)->total_tiles; }

// Number of replication slabs (sum of r over the requests): the length of a digest array.
QMC_EXPORT u64 qmc_batch_slabs(void *h) { return ((Plan *)h)->nslabs; }

// CUs (persistent workgroups) of the gdn class; the lattice and dnb2 class
// takes the others while both run, then the freed ones.  A tunable: 1 ..
// the device's CU count.
QMC_EXPORT int qmc_batch_set_gdn_cus(void *h, int n) {
    Plan *P = (Plan *)h;
    if (n < 1 || n > P->cus) return qmc_fail(QMC_ERR_ARG, "batch: gdn CUs must be in [1, %d]", P->cus);
    if (n != P->grid1 && P->gexec) {
        QMC_HIP_CHECK(hipGraphExecDestroy(P->gexec));
        P->gexec = nullptr;
    }
    P->grid1 = n;
    return QMC_OK;
}

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
        const unsigned lds = lds_total(P->lds_bytes[c0 ? 0 : 1]);
        if (P->digest) {
            if (c0)
                hipLaunchKernelGGL((batch_single_kernel<0, true>), dim3((unsigned)tiles), dim3(kBlock), lds, s,
                                   P->d_reqs + i, P->d_ws, ring, P->digest);
            else
                hipLaunchKernelGGL((batch_single_kernel<1, true>), dim3((unsigned)tiles), dim3(kBlock1), lds, s,
                                   P->d_reqs + i, P->d_ws, ring, P->digest);
        } else {
            if (c0)
                hipLaunchKernelGGL((batch_single_kernel<0, false>), dim3((unsigned)tiles), dim3(kBlock), lds, s,
                                   P->d_reqs + i, P->d_ws, ring, P->digest);
            else
                hipLaunchKernelGGL((batch_single_kernel<1, false>), dim3((unsigned)tiles), dim3(kBlock1), lds, s,
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

// gdn inputs of a request (seed, m, d, r) in the reference's layouts, for
// columns j in [j0, j0 + dj): bases[dj] (the primes), S[r][dj][kTNmax][m]
// (gdn_get_linear_scramble_matrix with tmax = m, tmax_new = kTNmax) and
// perms[r][dj][kTNmax][bmax] (gdn_get_digital_permutations, zero past b),
// bmax = the largest of the bases.  The generating matrices are identities.
QMC_EXPORT int qmc_batch_gdn_inputs(u64 seed, int m, int j0, int dj, int r, u64 *bases, u64 *S, u64 *perms) {
    fill_primes();
    if (m < 1 || m > 32 || j0 < 0 || dj < 1 || j0 + dj > 256 || r < 1)
        return qmc_fail(QMC_ERR_ARG, "qmc_batch_gdn_inputs: bad (m, j0, dj, r) = (%d, %d, %d, %d)", m, j0, dj, r);
    const u64 bmax = h_primes[j0 + dj - 1];
    std::vector<u64> keys;
    for (int jj = 0; jj < dj; jj++) bases[jj] = h_primes[j0 + jj];
    for (int l = 0; l < r; l++)
        for (int jj = 0; jj < dj; jj++) {
            const u64 j = (u64)(j0 + jj);
            const u32 b = h_primes[j];
            u64 *Slj = S + ((u64)l * dj + jj) * kTNmax * m;
            u64 *Plj = perms + ((u64)l * dj + jj) * kTNmax * bmax;
            for (u32 t = 0; t < (u32)kTNmax; t++) {
                for (u32 i = 0; i < (u32)m; i++) Slj[t * m + i] = gdn_s(seed, l, j, t, i, b);
                keys.resize(b);
                for (u32 z = 0; z < b; z++) keys[z] = gdn_pkey(seed, l, j, t, z);
                std::sort(keys.begin(), keys.end());
                for (u64 z = 0; z < bmax; z++) Plj[t * bmax + z] = 0;
                for (u32 pos = 0; pos < b; pos++) Plj[t * bmax + (keys[pos] & 0x7ff)] = pos;
            }
        }
    return QMC_OK;
}

// The per-base constants of the gdn sums for column j (base = the j-th prime):
// rec[kTNmax], thr[kTNmax] (the early-exit thresholds), the levels kept in LDS
// and the first level that can end a sum.
QMC_EXPORT int qmc_batch_gdn_base(int j, double *rec, double *thr, int *levels, int *tmin) {
    fill_primes();
    if (j < 0 || j >= 256) return qmc_fail(QMC_ERR_ARG, "qmc_batch_gdn_base: j = %d outside [0, 256)", j);
    gdn_base_tables(h_primes[j], rec, thr);
    *levels = gdn_levels(h_primes[j]);
    *tmin = (int)gdn_tmin(thr);
    return QMC_OK;
}

QMC_EXPORT u64 qmc_owen_hash64(u64 x, u64 seed) { return owen_hash64(x, seed); }

}  // extern "C"
