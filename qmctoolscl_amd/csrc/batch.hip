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
// columns, shifts, LMS rows, permutation parameters) in a workspace, then a
// persistent grid drains a global tile counter.  Outputs go to a ring buffer:
// every replication slab (n*d doubles) of every request is placed contiguously
// at an offset chosen on the host, wrapping to the start when the ring is full,
// so batches far larger than HBM can stream through a bounded buffer.
#include <cmath>
#include <cstring>
#include <vector>

#include "qmc_common.hpp"

namespace {

enum { M_LATTICE = 0, M_DNB2_LMS = 1, M_DNB2_NUS = 2, M_GDN_LMS = 3 };

constexpr int kTB = 32;     // base-2 digits (tmax) of the dnb2 methods
constexpr int kTNmax = 52;  // base-b digits kept (2^-52 resolution)

struct Req {
    int32_t method, m, d, r;
    u64 seed;
    u64 tile0;           // first global tile
    u32 rows_per_tile;   // power of two, <= n
    u32 tiles_per_l;     // n / rows_per_tile
    u64 tab;             // word offset of the request's tables in the workspace
    u64 slabs;           // word offset of r slab bases (output element offsets)
    u64 sj;              // gdn: LMS entries per replication (sum over j of tn_j * mm_j)
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
// dnb2 LMS: pre[r*d*32] (prefix XOR of the LMS columns) | cols[r*d*32] | shift[r*d]
// dnb2 NUS: pre[d*32] | cols[d*32] | nseed[r*d]
// gdn     : per j: b, tn, mm, S offset, ceil(2^64/b), b^tn (double bits) (8 words)
//           | S (u32, two per word) | perm (a,c) per (l,j,t) | base-2 row masks per (l,t)

constexpr int kJInfo = 8;

struct GdnLayout {
    u64 jinfo;   // kJInfo words per j
    u64 S;       // r * SJ u32 entries, packed two per word
    u64 perm;    // r * d * kTNmax words: (a << 32) | c
    u64 mask2;   // r * kTNmax words: row t of S for base 2 as a bit mask over digits
    u64 words;
};

__host__ __device__ __forceinline__ GdnLayout gdn_layout(int d, int r, u64 SJ) {
    GdnLayout g;
    g.jinfo = 0;
    g.S = (u64)kJInfo * d;
    g.perm = g.S + (r * SJ + 1) / 2;
    g.mask2 = g.perm + (u64)r * d * kTNmax;
    g.words = g.mask2 + (u64)r * kTNmax;
    return g;
}

__host__ __device__ __forceinline__ u64 recip_magic(u32 b) { return b == 1 ? 0 : (~0ull) / b + 1; }

// x mod b and x / b for x < 2^32 with M = ceil(2^64 / b) (exact, Lemire et al. 2019)
__device__ __forceinline__ u32 fast_div(u32 x, u64 M) { return (u32)__umul64hi(M, (u64)x); }

u64 table_words(const Req &q, u64 &SJ) {
    const u64 d = q.d, r = q.r;
    SJ = 0;
    switch (q.method) {
        case M_LATTICE: return d + r * d;
        case M_DNB2_LMS: return 2 * r * d * kTB + r * d;
        case M_DNB2_NUS: return 2 * d * kTB + r * d;
        default: {
            for (int j = 0; j < q.d; j++) SJ += (u64)gdn_tn(h_primes[j]) * gdn_mm(h_primes[j], q.m);
            return gdn_layout(q.d, q.r, SJ).words;
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
        const u64 ncol = (q.method == M_DNB2_LMS ? r : 1) * d;  // (l, j) columns sets
        for (u64 c = tid; c < ncol + r * d; c += step) {
            if (c < ncol) {
                const u64 l = c / d, j = c % d;
                u64 *pre = tab + c * kTB;
                u64 *cols = tab + ncol * kTB + c * kTB;
                u64 acc = 0;
                for (int k = 0; k < kTB; k++) {
                    u64 col = dnb2_col(q.seed, j, k);
                    if (q.method == M_DNB2_LMS) {  // C_lms[k] = S * C[k] over GF(2) (dnb2_linear_matrix_scramble.c:33-48)
                        u64 out = 0;
                        for (int t = 0; t < kTB; t++)
                            out |= (u64)(__popcll(dnb2_srow(q.seed, l, j, t) & col) & 1) << (kTB - 1 - t);
                        col = out;
                    }
                    cols[k] = col;
                    acc ^= col;
                    pre[k] = acc;
                }
            } else {
                const u64 f = c - ncol;
                tab[2 * ncol * kTB + f] = q.method == M_DNB2_LMS ? dnb2_shift(q.seed, f / d, f % d)
                                                                 : nus_seed(q.seed, f / d, f % d);
            }
        }
    } else {
        // gdn: j info, S rows, permutations
        const u64 SJ = q.sj;
        const GdnLayout L = gdn_layout(q.d, q.r, SJ);
        u32 *S = (u32 *)(tab + L.S);
        // j info (offsets are cumulative; computed by every thread for its own j)
        for (u64 j = tid; j < d; j += step) {
            u64 off = 0;
            for (u64 jj = 0; jj < j; jj++) off += (u64)gdn_tn(c_primes[jj]) * gdn_mm(c_primes[jj], q.m);
            const u32 b = c_primes[j];
            const int tn = gdn_tn(b);
            double scale = 1.0;
            for (int t = 0; t < tn; t++) scale *= (double)b;
            u64 *ji = tab + L.jinfo + kJInfo * j;
            ji[0] = b;
            ji[1] = tn;
            ji[2] = gdn_mm(b, q.m);
            ji[3] = off;
            ji[4] = recip_magic(b);
            ji[5] = (u64)__double_as_longlong(scale);
        }
        for (u64 e = tid; e < r * d; e += step) {
            const u64 l = e / d, j = e % d;
            const u32 b = c_primes[j];
            const int tn = gdn_tn(b), mm = gdn_mm(b, q.m);
            u64 off = 0;
            for (u64 jj = 0; jj < j; jj++) off += (u64)gdn_tn(c_primes[jj]) * gdn_mm(c_primes[jj], q.m);
            u32 *Slj = S + l * SJ + off;
            for (int t = 0; t < tn; t++) {
                u64 mask = 0;
                for (int k = 0; k < mm; k++) {
                    const u32 s = k <= t ? gdn_s(q.seed, l, j, t, k, b) : 0;
                    Slj[t * mm + k] = s;
                    mask |= (u64)(s & 1) << k;
                }
                if (b == 2) tab[L.mask2 + l * kTNmax + t] = mask;
                u32 a, c;
                gdn_perm(q.seed, l, j, t, b, a, c);
                tab[L.perm + (l * d + j) * kTNmax + t] = ((u64)a << 32) | c;
            }
        }
    }
}

// ------------------------------------------------------------- one tile

__device__ __forceinline__ u64 xor_cols32(const u64 *__restrict__ cols, u64 q) {
    u64 v = 0;
    while (q) {
        v ^= cols[__ffsll((long long)q) - 1];
        q &= q - 1;
    }
    return v;
}

constexpr int kMaxDigB = 21;  // base-b digits of a 32-bit index, b >= 3

__device__ void run_tile(const Req &q, u64 t, const u64 *__restrict__ ws, double *__restrict__ out, u32 *sdig) {
    const u64 d = q.d;
    const u64 l = t / q.tiles_per_l;
    const u64 row0 = (t - l * q.tiles_per_l) * (u64)q.rows_per_tile;
    const u64 *tab = ws + q.tab;
    double *slab = out + ws[q.slabs + l];
    const u32 pts = q.rows_per_tile * (u32)d;
    if (q.method == M_LATTICE) {
        const u64 *g = tab;
        const double *sh = (const double *)(tab + d) + l * d;
        for (u32 e = threadIdx.x; e < pts; e += blockDim.x) {
            const u32 rr = div_q(e, q.ddiv);
            const u32 j = e - rr * q.d;
            const u64 row = row0 + rr;
            const double phi = (double)__brevll(row) * 0x1p-64;  // phi2(row), exact
            const double x = frac_mod1((double)g[j] * phi);      // lat_gen_natural.c:45
            store_stream(slab + row * d + j, frac_mod1(x + sh[j]));  // lat_shift_mod_1.c:31
        }
    } else if (q.method == M_DNB2_LMS || q.method == M_DNB2_NUS) {
        const bool lms = q.method == M_DNB2_LMS;
        const u64 ncol = (lms ? (u64)q.r : 1) * d;
        const u64 set = lms ? l * d : 0;
        // threads -> (j, run); runs of consecutive rows walk natural order with
        // one prefix-XOR per step (value(q+1) = value(q) ^ C0 ^ ... ^ C_ctz(q+1))
        // as many runs as fill the block: runs * d <= blockDim (or one run per row)
        u32 runs = blockDim.x / (u32)d;
        if (runs < 1) runs = 1;
        if (runs > q.rows_per_tile) runs = q.rows_per_tile;
        const u32 K = (q.rows_per_tile + runs - 1) / runs;
        const u64 *extra = tab + 2 * ncol * kTB + l * d;  // shifts or NUS seeds
        for (u32 w = threadIdx.x; w < runs * d; w += blockDim.x) {
            const u32 j = w % (u32)d;
            const u32 run = w / (u32)d;
            const u32 r0 = run * K;
            if (r0 >= q.rows_per_tile) continue;
            const u32 cnt = (q.rows_per_tile - r0) < K ? (q.rows_per_tile - r0) : K;
            const u64 *pre = tab + (set + j) * kTB;
            const u64 *cols = tab + ncol * kTB + (set + j) * kTB;
            const u64 xs = extra[j];
            u64 row = row0 + r0;
            u64 v = xor_cols32(cols, row);
            for (u32 s = 0;; s++) {
                const u64 xb = v << 32;  // lshift 64 - tmax (dnb2_digital_shift.c:32)
                const u64 y = lms ? (xb ^ xs) : owen_hash64(xb, xs);
                store_stream(slab + row * d + j, ldexp((double)y, -64));  // dnb2_integer_to_float.c:29
                if (s + 1 == cnt) break;
                row++;
                v ^= pre[__ffsll((long long)row) - 1];
            }
        }
    } else {
        const u64 SJ = q.sj;
        const GdnLayout L = gdn_layout(q.d, q.r, SJ);
        const u32 *S = (const u32 *)(tab + L.S) + l * SJ;
        const u64 *perm = tab + L.perm + l * d * kTNmax;
        const u64 *mask2 = tab + L.mask2 + l * kTNmax;
        u32 *dig = sdig + threadIdx.x;  // digit i at dig[i * kBlock]
        // One dimension per wave (units of 64 rows x 1 dimension): the digit
        // counts, LMS rows and permutations are then uniform across the wave,
        // so no lane waits on a smaller base's longer loops.
        const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
        const u32 chunks = (q.rows_per_tile + 63) / 64;
        for (u32 unit = wave; unit < chunks * (u32)d; unit += nwaves) {
            const u32 j = unit % (u32)d;
            const u32 rr = (unit / (u32)d) * 64 + lane;
            if (rr >= q.rows_per_tile) continue;
            const u32 row = (u32)(row0 + rr);
            const u64 *ji = tab + L.jinfo + kJInfo * j;
            const u32 b = (u32)ji[0];
            const int tn = (int)ji[1];
            const int mm = (int)ji[2];
            const u32 *Sj = S + ji[3];
            const u64 M = ji[4];
            const u64 *pj = perm + (u64)j * kTNmax;
            u64 Y = 0;
            if (b == 2) {
                // digit t = parity of (row t of S) & row: one popcount per digit
                for (int tt = 0; tt < tn; tt++) {
                    const u32 xd = (u32)__popcll(mask2[tt] & (u64)row) & 1u;
                    const u64 pc = pj[tt];
                    Y = 2 * Y + (((u32)(pc >> 32) * xd + (u32)pc) & 1u);
                }
            } else {
                u32 k = row;
                for (int i = 0; i < mm; i++) {
                    const u32 qq = fast_div(k, M);
                    dig[i * kBlock] = k - qq * b;
                    k = qq;
                }
                for (int tt = 0; tt < tn; tt++) {
                    u32 acc = 0;  // < 32 * b^2 < 2^27
                    const int kmax = tt < mm ? tt + 1 : mm;
                    const u32 *Srow = Sj + tt * mm;
                    for (int i = 0; i < kmax; i++) acc += Srow[i] * dig[i * kBlock];
                    const u32 xd = acc - fast_div(acc, M) * b;
                    const u64 pc = pj[tt];
                    const u32 v = (u32)(pc >> 32) * xd + (u32)pc;
                    Y = Y * b + (v - fast_div(v, M) * b);
                }
            }
            store_stream(slab + (u64)row * d + j, (double)Y / __longlong_as_double((long long)ji[5]));
        }
    }
}

// ------------------------------------------------------------ the kernels

__global__ __launch_bounds__(kBlock) void batch_persistent_kernel(const Req *__restrict__ reqs, int nreq,
                                                                  u64 total_tiles, unsigned long long *counter,
                                                                  const u64 *__restrict__ ws,
                                                                  double *__restrict__ out) {
    __shared__ u64 s_tile;
    __shared__ int s_req;
    __shared__ u32 sdig[kMaxDigB * kBlock];
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
        run_tile(q, t - q.tile0, ws, out, sdig);
    }
}

__global__ void reset_counter_kernel(unsigned long long *counter) { *counter = 0; }

// one request, one tile per block (the per-call form)
__global__ __launch_bounds__(kBlock) void batch_single_kernel(const Req *__restrict__ req, const u64 *__restrict__ ws,
                                                              double *__restrict__ out) {
    __shared__ u32 sdig[kMaxDigB * kBlock];
    const Req q = *req;
    run_tile(q, blockIdx.x, ws, out, sdig);
}

// ----------------------------------------------------------- host plan

struct Plan {
    std::vector<Req> reqs;
    u64 total_tiles = 0;
    u64 ws_words = 0;
    u64 ring_elems = 0;
    Req *d_reqs = nullptr;
    u64 *d_ws = nullptr;
    unsigned long long *d_counter = nullptr;
    int grid = 0;
    int device = 0;
    hipStream_t cap = nullptr;       // private stream for graph capture
    hipGraphExec_t gexec = nullptr;  // [counter reset, persistent kernel]
    double *gring = nullptr;         // ring the graph was captured for
};

void launch_batch(Plan *P, double *ring, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, P->d_counter);
    hipLaunchKernelGGL(batch_persistent_kernel, dim3(P->grid), dim3(kBlock), 0, s, P->d_reqs, (int)P->reqs.size(),
                       P->total_tiles, P->d_counter, P->d_ws, ring);
}

}  // namespace

extern "C" {

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
            delete P;
            return nullptr;
        }
        const u64 n = 1ull << q.m;
        u64 rows = points_per_tile / (u64)q.d;
        u64 p2 = 1;
        while (p2 * 2 <= rows) p2 *= 2;
        if (p2 > n) p2 = n;
        q.rows_per_tile = (u32)p2;
        q.tiles_per_l = (u32)(n / p2);
        q.tile0 = tile;
        tile += (u64)q.r * q.tiles_per_l;
        q.ddiv = make_div((u32)q.d);
        u64 SJ;
        q.tab = words;
        words += table_words(q, SJ);
        q.sj = SJ;
        q.slabs = words;
        words += q.r;
        P->reqs.push_back(q);
        const u64 slab = n * (u64)q.d;
        if (slab > ring_elems) {
            qmc_fail(QMC_ERR_ARG, "batch: a replication slab of request %d (%llu doubles) exceeds the ring", i,
                     (unsigned long long)slab);
            delete P;
            return nullptr;
        }
    }
    P->total_tiles = tile;
    P->ws_words = words;
    if (hipGetDevice(&P->device) != hipSuccess) { delete P; qmc_fail(QMC_ERR_HIP, "batch: no device"); return nullptr; }
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, P->device);
    P->grid = grid > 0 ? grid : 4 * cus;
    std::vector<u64> ws_host(words, 0);
    // slab bases in the ring: consecutive, wrapping when a slab would run past the end
    u64 pos = 0;
    for (auto &q : P->reqs) {
        const u64 slab = (1ull << q.m) * (u64)q.d;
        for (int l = 0; l < q.r; l++) {
            if (pos + slab > ring_elems) pos = 0;
            ws_host[q.slabs + l] = pos;
            pos += slab;
        }
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_primes), h_primes, sizeof(h_primes)) != hipSuccess ||
        hipMalloc((void **)&P->d_reqs, sizeof(Req) * P->reqs.size()) != hipSuccess ||
        hipMalloc((void **)&P->d_ws, sizeof(u64) * (words ? words : 1)) != hipSuccess ||
        hipMalloc((void **)&P->d_counter, sizeof(unsigned long long)) != hipSuccess) {
        qmc_fail(QMC_ERR_NOMEM, "batch: device allocation failed");
        delete P;
        return nullptr;
    }
    hipMemcpy(P->d_reqs, P->reqs.data(), sizeof(Req) * P->reqs.size(), hipMemcpyHostToDevice);
    hipMemcpy(P->d_ws, ws_host.data(), sizeof(u64) * words, hipMemcpyHostToDevice);
    // tables
    hipLaunchKernelGGL(batch_setup_kernel, dim3(64, (unsigned)P->reqs.size()), dim3(kBlock), 0, 0, P->d_reqs,
                       P->d_ws);
    if (hipDeviceSynchronize() != hipSuccess) {
        qmc_fail(QMC_ERR_HIP, "batch: setup kernel failed");
        delete P;
        return nullptr;
    }
    return P;
}

QMC_EXPORT void qmc_batch_free(void *h) {
    Plan *P = (Plan *)h;
    if (!P) return;
    if (P->gexec) hipGraphExecDestroy(P->gexec);
    if (P->cap) hipStreamDestroy(P->cap);
    hipFree(P->d_reqs);
    hipFree(P->d_ws);
    hipFree(P->d_counter);
    delete P;
}

QMC_EXPORT u64 qmc_batch_total_tiles(void *h) { return ((Plan *)h)->total_tiles; }
QMC_EXPORT u64 qmc_batch_points(void *h) {
    u64 p = 0;
    for (auto &q : ((Plan *)h)->reqs) p += (u64)q.r * (1ull << q.m) * (u64)q.d;
    return p;
}
// element offset in the ring of replication l of request q
QMC_EXPORT u64 qmc_batch_slab_offset(void *h, int q, int l) {
    Plan *P = (Plan *)h;
    u64 v = 0;
    hipMemcpy(&v, P->d_ws + P->reqs[q].slabs + l, sizeof(u64), hipMemcpyDeviceToHost);
    return v;
}

// All requests in one persistent launch (capturable: a fill kernel resets the
// tile counter, no memset node).
QMC_EXPORT int qmc_batch_run_dev(void *h, double *ring, qmc_stream_t s) {
    Plan *P = (Plan *)h;
    if (!P->total_tiles) return QMC_OK;
    launch_batch(P, ring, s);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

// The same launch replayed from a hipGraph (captured once per ring on a
// private stream, then launched on s).  Calls on one plan must not overlap:
// they share the tile counter.
QMC_EXPORT int qmc_batch_graph_run_dev(void *h, double *ring, qmc_stream_t s) {
    Plan *P = (Plan *)h;
    if (!P->total_tiles) return QMC_OK;
    if (!P->gexec || P->gring != ring) {
        if (P->gexec) {
            QMC_HIP_CHECK(hipGraphExecDestroy(P->gexec));
            P->gexec = nullptr;
        }
        if (!P->cap) QMC_HIP_CHECK(hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        QMC_HIP_CHECK(hipStreamBeginCapture(P->cap, hipStreamCaptureModeThreadLocal));
        launch_batch(P, ring, P->cap);
        QMC_HIP_CHECK(hipStreamEndCapture(P->cap, &g));
        QMC_HIP_CHECK(hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0));
        QMC_HIP_CHECK(hipGraphDestroy(g));
        P->gring = ring;
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
        hipLaunchKernelGGL(batch_single_kernel, dim3((unsigned)tiles), dim3(kBlock), 0, s, P->d_reqs + i, P->d_ws,
                           ring);
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
