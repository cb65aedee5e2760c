// Generalized (base-b) digital nets and Halton points on gfx950: generation,
// linear matrix scrambling, digital shift, digit permutation, conversion to
// double, and interlacing.
//
// Digit outputs are laid out [l][i][j][t] with t fastest, so for one
// replication l the rows rho = i*d + j and their digits form one contiguous
// array.  The row kernels give each block R consecutive rows: lane f of the
// block produces digit (rho0 + f / T, f % T), so stores are fully coalesced.
// Generation walks down consecutive points with the same row layout across
// the lanes (see gdn_gen_chain_kernel).
//
// Arithmetic follows the reference's unsigned 64-bit semantics exactly,
// including wrap-around: gdn_gen_natural.c:40-50 reduces after every digit,
// x = (x + dig*C) % b.  Where no step can wrap, the walks use exact modular
// sums; every other element takes the step-by-step path.
#include <mutex>
#include <type_traits>

#include "qmc_common.hpp"

#include <algorithm>

namespace {

constexpr int kRows = 64;        // rows per block in the row kernels

enum { STATUS_BAD_BASE = 1u, STATUS_OOB = 2u };

__device__ __forceinline__ void flag(u32 *status, u32 bit) { atomicOr(status, bit); }

// Row geometry shared by the kernels below: rho -> j via a 32-bit magic
// division when n*d < 2^32.
struct Rows {
    u64 nd, d;
    DivU32 ddiv, tdiv;
    int small;
};

__device__ __forceinline__ u64 j_of(const Rows &g, u64 rho) {
    if (g.small) {
        const u32 q = div_q((u32)rho, g.ddiv);
        return (u32)rho - q * g.ddiv.d;
    }
    return rho % g.d;
}

// ---------------------------------------------------------------- generation

// One output digit as gdn_gen_natural.c:40-50 computes it: over the base-b
// digits dig_m of k, x = (x + dig_m * C[cbase + m*tmax]) % b in wrapping u64
// arithmetic.  Zero digits leave x unchanged and read nothing.
__device__ u64 gdn_digit(u64 k, u64 b, const u64 *__restrict__ C, u64 cbase, u64 tmax, u64 Ctotal, u32 &bad) {
    if (b < 2) {  // the reference loops forever (b = 1) or divides by zero
        bad |= STATUS_BAD_BASE;
        return 0;
    }
    u64 acc = 0;
    for (u64 m = 0; k; m++) {
        u64 q, dg;
        if ((k >> 32) == 0 && (b >> 32) == 0) {
            q = (u32)k / (u32)b;
            dg = (u32)k - (u32)q * (u32)b;
        } else {
            q = k / b;
            dg = k - q * b;
        }
        k = q;
        if (!dg) continue;
        const u64 ci = cbase + m * tmax;
        u64 c = 0;
        if (ci < Ctotal) c = C[ci];
        else bad |= STATUS_OOB;
        acc = (acc + dg * c) % b;
    }
    return acc;
}

// Point walks.  From k to k+1 the c trailing base-b digits equal to b-1 turn
// to 0 and digit c grows by 1, so every output digit moves by
//     Q_c[t] = (C_0[t] + ... + C_c[t]) mod b      (-(b-1) = 1 mod b),
// provided the reference's sums are exact: b < 2^31 and C < 2^32 keep
// x + dig*C below 2^64, and k < b^mmax keeps every digit inside the column's
// mmax rows of C.  gdn_chain_setup_kernel writes Q (u32) and that condition
// per (l, j); the walk kernel takes the row of one point, d*tmax contiguous
// digits, across a block's lanes (every store instruction writes 512
// consecutive bytes per wave) and moves down kChainH points, adding Q_0 (held
// in a register) or, on a carry, Q_c.  Columns failing the condition take
// gdn_digit per point.
constexpr int kChainE = 8;                       // row entries per lane
constexpr u64 kChainF = (u64)kBlock * kChainE;  // row entries per block
constexpr u32 kChainH = 256;                     // points per block

__global__ __launch_bounds__(kBlock) void gdn_chain_setup_kernel(u64 d, u64 mmax, u64 tmax, u64 kmax, u64 r_b,
                                                                 const u64 *__restrict__ bases, u64 b_common,
                                                                 const u64 *__restrict__ C, u32 *__restrict__ Q,
                                                                 u32 *__restrict__ ok) {
    const u64 lj = blockIdx.x;  // l * d + j
    const u64 l = lj / d, j = lj - (lj / d) * d;
    const u64 b = bases ? bases[(l % r_b) * d + j] : b_common;
    const bool b_ok = b >= 2 && (b >> 31) == 0;  // z + Q < 2^32 in the walk
    int big = 0;
    for (u64 t = threadIdx.x; t < tmax; t += blockDim.x) {
        u64 acc = 0;
        for (u64 m = 0; m < mmax; m++) {
            const u64 c = C[(lj * mmax + m) * tmax + t];
            big |= (c >> 32) != 0;
            if (b_ok) acc = (acc + (c & 0xffffffffull)) % b;
            Q[(lj * mmax + m) * tmax + t] = (u32)acc;
        }
    }
    big = __syncthreads_or(big);
    if (threadIdx.x == 0) {
        // every k <= kmax has at most mmax base-b digits
        bool range = false;
        if (b_ok) {
            u64 p = 1;
            for (u64 m = 0; m < mmax && !range; m++) {
                if (p > kmax / b) range = true;  // p * b > kmax
                p *= b;
            }
            range = range || p > kmax;
        }
        ok[lj] = b_ok && !big && range;
    }
}

__global__ __launch_bounds__(kBlock) void gdn_gen_chain_kernel(u64 n, u64 d, u64 tmax, u64 n_start, u64 r_b,
                                                               u64 mmax, const u64 *__restrict__ bases,
                                                               u64 b_common, const u64 *__restrict__ C, u64 Ctotal,
                                                               const u32 *__restrict__ Q,
                                                               const u32 *__restrict__ okf,
                                                               u64 *__restrict__ xdig, u32 *status) {
    const u64 l = blockIdx.z;
    const u64 i0 = (u64)blockIdx.x * kChainH;
    const u32 H = (n - i0) < (u64)kChainH ? (u32)(n - i0) : kChainH;
    const u64 row = d * tmax;
    const u64 k0 = i0 + n_start;
    u64 *out = xdig + (l * n + i0) * row;
    u32 bad = 0;
#pragma unroll 1
    for (int e = 0; e < kChainE; e++) {
        const u64 f = (u64)blockIdx.y * kChainF + (u64)e * kBlock + threadIdx.x;
        if (f >= row) break;
        const u64 j = f / tmax, t = f - (f / tmax) * tmax;
        const u64 lj = l * d + j;
        const u64 b = bases ? bases[(l % r_b) * d + j] : b_common;
        const u64 cbase = lj * mmax * tmax + t;
        u64 *o = out + f;
        if (Q && okf[lj]) {
            const u32 bb = (u32)b;
            const DivU32 bdiv = make_div(bb);
            const u32 *q = Q + lj * mmax * tmax + t;
            const u32 q0 = q[0];
            u32 z = (u32)gdn_digit(k0, b, C, cbase, tmax, Ctotal, bad);
            u32 dlow = (k0 >> 32) ? (u32)(k0 % b) : (u32)k0 - div_q((u32)k0, bdiv) * bb;
            u64 k = k0;
            for (u32 h = 0;; h++) {
                *o = (u64)z;  // cached: rows of 8*d*tmax bytes may share a line with the next point's
                if (h + 1 == H) break;
                u32 inc = q0;
                if (dlow + 1 < bb) {
                    dlow++;
                } else {
                    // k + 1 ends in c >= 1 zero digits: c = 1 + trailing zeros of (k + 1) / b
                    dlow = 0;
                    const u64 k1 = k + 1;
                    u64 x = (k1 >> 32) ? k1 / b : (u64)div_q((u32)k1, bdiv);
                    u32 c = 1;
                    while (x != 0 && c + 1 < (u32)mmax) {  // c < mmax whenever the column qualifies
                        u64 qq;
                        if ((x >> 32) == 0) qq = div_q((u32)x, bdiv);
                        else qq = x / b;
                        if (x - qq * b != 0) break;
                        x = qq;
                        c++;
                    }
                    inc = q[(u64)(c < (u32)mmax ? c : (u32)mmax - 1) * tmax];
                }
                k++;
                z += inc;
                z = z >= bb ? z - bb : z;
                o += row;
            }
        } else {
            for (u32 h = 0; h < H; h++) o[h * row] = gdn_digit(k0 + h, b, C, cbase, tmax, Ctotal, bad);
        }
    }
    if (bad) flag(status, bad);
}

// -------------------------------------------------- digit-wise shift / permute

template <bool PERMUTE>
__global__ __launch_bounds__(kBlock) void gdn_digitwise_kernel(Rows g, u64 r_x, u64 r_b, u64 tmax, u64 tmax_new,
                                                               u64 bmax, const u64 *__restrict__ bases,
                                                               const u64 *__restrict__ tab, u64 tab_total,
                                                               const u64 *__restrict__ xdig,
                                                               u64 *__restrict__ xnew, u32 *status) {
    __shared__ u64 sbase[kRows];
    __shared__ u64 sj[kRows];
    const u64 l = blockIdx.y;
    const u64 rho0 = (u64)blockIdx.x * kRows;
    const u64 nrows = (g.nd - rho0) < (u64)kRows ? (g.nd - rho0) : (u64)kRows;
    if (threadIdx.x < nrows) {
        const u64 j = j_of(g, rho0 + threadIdx.x);
        sj[threadIdx.x] = j;
        if (!PERMUTE) sbase[threadIdx.x] = bases[(l % r_b) * g.d + j];
    }
    __syncthreads();
    const u32 T = (u32)tmax_new;
    const u32 total = (u32)nrows * T;
    const u64 *src = xdig + ((l % r_x) * g.nd + rho0) * tmax;
    u64 *out = xnew + (l * g.nd + rho0) * tmax_new;
    u32 bad = 0;
#pragma unroll 4
    for (u32 f = threadIdx.x; f < total; f += blockDim.x) {
        const u32 row = div_q(f, g.tdiv);
        const u32 t = f - row * T;
        const u64 j = sj[row];
        u64 v;
        if (PERMUTE) {
            // gdn_digital_permutation.c:36-42: perms[l,j,t, digit]; perms[l,j,t,0] past tmax
            const u64 pbase = ((l * g.d + j) * tmax_new + t) * bmax;
            const u64 p = t < tmax ? src[(u64)row * tmax + t] : 0;
            const u64 pi = pbase + p;
            if (pi < tab_total) {
                v = tab[pi];
            } else {
                v = 0;
                bad |= STATUS_OOB;
            }
        } else {
            // gdn_digital_shift.c:37-42: (digit + shift) % b; the raw shift past tmax
            const u64 sh = tab[(l * g.d + j) * tmax_new + t];
            if (t < tmax) {
                const u64 b = sbase[row];
                if (b == 0) {
                    bad |= STATUS_BAD_BASE;
                    v = 0;
                } else {
                    const u64 x = src[(u64)row * tmax + t] + sh;  // wraps as the reference's u64 sum
                    v = x < 2 * b && x >= sh ? (x >= b ? x - b : x) : umod(x, b);
                }
            } else {
                v = sh;
            }
        }
        store_stream(out + f, v);
    }
    if (bad) flag(status, bad);
}

// Digit permutation with one column's tables in LDS: a block takes column j
// of replication l for kPermPts consecutive points, so its gathers
// perms[l][j][t][digit] hit a u32 copy of that column's tmax_new x bmax table
// instead of L2.  Values of 2^32 or more, and digits >= bmax (which the
// reference reads from the following rows), take the global table.  A row of
// 8*tmax_new bytes need not fill whole 128-byte lines, so the stores are
// ordinary (cached) ones: with streaming stores the line a column shares with
// its neighbour left L2 half written (3^12 x 32 x 8, 52 digits: 30.0 ms ->
// 24.9 ms with the column grouping below).
constexpr int kPermPts = 1024;
constexpr u64 kPermLdsEntries = 12288;  // 48 KiB of u32

__global__ __launch_bounds__(kBlock) void gdn_perm_lds_kernel(u64 n, u64 d, u64 r_x, u64 tmax, u64 tmax_new,
                                                              u64 bmax, const u64 *__restrict__ perms, u64 tab_total,
                                                              const u64 *__restrict__ xdig,
                                                              u64 *__restrict__ xnew, u32 *status) {
    extern __shared__ u32 sperm[];
    __shared__ int s_big;
    const u64 l = blockIdx.y;
    // the d column blocks of one point range run on one XCD (block b runs on
    // XCD b % 8), so the lines their 8*tmax_new-byte rows share meet in one L2
    const u64 xcd = blockIdx.x & 7, qd = blockIdx.x >> 3;
    const u64 j = qd % d;
    const u64 i0 = ((qd / d) * 8 + xcd) * (u64)kPermPts;
    if (i0 >= n) return;  // uniform over the block
    const u64 pbase = (l * d + j) * tmax_new * bmax;
    const u32 entries = (u32)(tmax_new * bmax);
    if (threadIdx.x == 0) s_big = 0;
    __syncthreads();
    int big = 0;
    for (u32 e = threadIdx.x; e < entries; e += blockDim.x) {
        const u64 v = perms[pbase + e];
        sperm[e] = (u32)v;
        big |= (v >> 32) != 0;
    }
    if (big) atomicOr(&s_big, 1);
    __syncthreads();
    const bool wide = s_big != 0;
    const u32 T = (u32)tmax_new;
    const u64 npts = (n - i0) < (u64)kPermPts ? (n - i0) : (u64)kPermPts;
    const u32 total = (u32)npts * T;
    const DivU32 tdiv = make_div(T);
    const u64 *src = xdig + ((l % r_x) * n * d + j) * tmax;
    u64 *out = xnew + (l * n * d + j) * tmax_new;
    u32 bad = 0;
#pragma unroll 4
    for (u32 f = threadIdx.x; f < total; f += blockDim.x) {
        const u32 ii = div_q(f, tdiv);
        const u32 t = f - ii * T;
        const u64 i = i0 + ii;
        const u64 p = t < tmax ? src[i * d * tmax + t] : 0;  // gdn_digital_permutation.c:36-42
        u64 v;
        if (p < bmax && !wide) {
            v = sperm[t * bmax + p];
        } else {
            const u64 pi = pbase + (u64)t * bmax + p;
            if (pi < tab_total) {
                v = perms[pi];
            } else {
                v = 0;
                bad |= STATUS_OOB;
            }
        }
        out[i * d * tmax_new + t] = v;  // cached: neighbouring columns complete the shared lines in L2
    }
    if (bad) flag(status, bad);
}

// -------------------------------------------------------------- to double

// gdn_integer_to_float.c:32-40: v += recip*digit; recip /= b, sequential per
// element (compiled without FMA contraction).  The reciprocals depend only on
// the base, so they are computed once per (l % r_b, j) by the same IEEE
// divisions, into rt[(lb*d + j)*tmax + t]; every element then does
// v += rt[t] * digit in the reference's order.
__global__ __launch_bounds__(kBlock) void gdn_recip_kernel(u64 nb, u64 tmax, const u64 *__restrict__ bases,
                                                           double *__restrict__ rt) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb) return;
    const double b = (double)bases[e];
    double recip = 1 / b;
    for (u64 t = 0; t < tmax; t++) {
        rt[e * tmax + t] = recip;
        recip /= b;
    }
}

__global__ __launch_bounds__(kBlock) void gdn_zero_kernel(u64 total, double *__restrict__ x) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) x[e] = 0.;
}

__global__ __launch_bounds__(kBlock) void gdn_to_float_kernel(u64 r, u64 nd, u64 d, u64 r_b, u64 tmax,
                                                              const double *__restrict__ rt,
                                                              const u64 *__restrict__ xdig,
                                                              double *__restrict__ x) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < r * nd; e += stride) {
        const u64 l = e / nd;
        const u64 j = (e - l * nd) % d;
        const double *rec = rt + ((l % r_b) * d + j) * tmax;
        const u64 *src = xdig + e * tmax;
        double v = 0.;
#pragma unroll 8
        for (u64 t = 0; t < tmax; t++) v += rec[t] * (double)src[t];
        store_stream(x + e, v);
    }
}

typedef u64 u64x2 __attribute__((ext_vector_type(2)));

// The same sums with the digit rows staged through LDS.  A one-wave block
// takes kTfPts consecutive elements, whose digits are one contiguous run of
// kTfPts*tmax words: it reads the run with 16-byte loads (each load
// instruction covers 1 KiB of consecutive bytes across the wave, sixteen of
// them in flight per lane), writes it to LDS rows `stride` = tmax | 1 words
// apart (so that the lanes' reads of one digit spread over the banks), and
// each lane then sums its own row.  With RT_LDS the reciprocal table is
// copied to LDS once per block, so the summing phase makes no global loads
// (as gathers from 64 rows they occupy the address unit more than the
// digits' coalesced loads do).  On MI355X, n = 3^12, d = 32, r = 8, tmax =
// 52: 13.7 ms against 20.1 ms for one lane per row.
constexpr int kTfPts = 64;
constexpr u32 kTfMaxT = 64;
constexpr u32 kTfRound = 16;  // 16-byte loads in flight per lane

template <bool RT_LDS>
__global__ __launch_bounds__(kTfPts) void gdn_to_float_staged_kernel(u64 total, u64 nd, u64 d, u64 r_b,
                                                                     DivU32 tdiv, u32 stride, DivU32 nddiv,
                                                                     DivU32 ddiv, const double *__restrict__ rt,
                                                                     const u64 *__restrict__ xdig,
                                                                     double *__restrict__ x) {
    extern __shared__ u64 sd[];
    const u32 lane = threadIdx.x;
    const u32 tmax = tdiv.d;
    double *srt = (double *)(sd + kTfPts * stride);
    if (RT_LDS) {  // visible after the first chunk's barrier
        const u32 nrt = (u32)(r_b * d) * tmax;
        for (u32 e = lane; e < nrt; e += kTfPts) {
            const u32 p = div_q(e, tdiv);
            srt[p * stride + (e - p * tmax)] = rt[e];
        }
    }
    for (u64 e0 = (u64)blockIdx.x * kTfPts; e0 < total; e0 += (u64)gridDim.x * kTfPts) {
        const u32 cnt = (u32)((total - e0) < (u64)kTfPts ? (total - e0) : (u64)kTfPts);
        const u32 words = cnt * tmax;
        const u64 *src = xdig + e0 * tmax;  // e0 * tmax is even: 16-byte aligned
        const u32 pairs = words >> 1;
#pragma unroll 1
        for (u32 base = 0; base < pairs; base += kTfRound * kTfPts) {
            u64x2 v[kTfRound];
#pragma unroll
            for (int k = 0; k < (int)kTfRound; k++) {
                const u32 w = base + k * kTfPts + lane;
                if (w < pairs) v[k] = __builtin_nontemporal_load((const u64x2 *)src + w);
            }
#pragma unroll
            for (int k = 0; k < (int)kTfRound; k++) {
                const u32 w = base + k * kTfPts + lane;
                if (w < pairs) {
                    const u32 p = div_q(2 * w, tdiv);
                    const u32 t = 2 * w - p * tmax;
                    u64 *row = sd + p * stride;
                    row[t] = v[k].x;
                    if (t + 1 < tmax) row[t + 1] = v[k].y;
                    else row[stride] = v[k].y;  // digit 0 of the next row
                }
            }
        }
        if ((words & 1) && lane == 0) sd[(cnt - 1) * stride + tmax - 1] = src[words - 1];
        __syncthreads();
        if (lane < cnt) {
            const u64 e = e0 + lane;
            u64 l, j;
            if (total >> 32) {
                l = e / nd;
                j = (e - l * nd) % d;
            } else {
                l = div_q((u32)e, nddiv);
                const u32 rho = (u32)e - (u32)l * (u32)nd;
                j = rho - div_q(rho, ddiv) * (u32)d;
            }
            const u64 *row = sd + lane * stride;
            double v = 0.;
            if (RT_LDS) {
                const double *rec = srt + ((l % r_b) * d + j) * stride;
#pragma unroll 8
                for (u32 t = 0; t < tmax; t++) v += rec[t] * (double)row[t];
            } else {
                const double *rec = rt + ((l % r_b) * d + j) * tmax;
#pragma unroll 8
                for (u32 t = 0; t < tmax; t++) v += rec[t] * (double)row[t];
            }
            store_stream(x + e, v);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------- matrix operations

// gdn_linear_matrix_scramble.c:32-44
__global__ __launch_bounds__(kBlock) void gdn_lms_kernel(u64 r, u64 d, u64 mmax, u64 r_C, u64 r_b, u64 tmax,
                                                         u64 tmax_new, const u64 *__restrict__ bases,
                                                         const u64 *__restrict__ S, const u64 *__restrict__ C,
                                                         u64 *__restrict__ C_lms, u32 *status) {
    const u64 total = r * d * mmax * tmax_new;
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const u64 t = e % tmax_new;
    const u64 rest = e / tmax_new;
    const u64 k = rest % mmax;
    const u64 lj = rest / mmax;
    const u64 j = lj % d;
    const u64 l = lj / d;
    const u64 b = bases[(l % r_b) * d + j];
    if (b == 0) {
        flag(status, STATUS_BAD_BASE);
        C_lms[e] = 0;
        return;
    }
    const u64 *row = S + ((l * d + j) * tmax_new + t) * tmax;
    const u64 *col = C + (((l % r_C) * d + j) * mmax + k) * tmax;
    u64 acc = 0;
    for (u64 c = 0; c < tmax; c++) acc += umod(row[c] * col[c], b);
    C_lms[e] = acc;
}

// gdn_interlace.c:31-36 (undo == 0) and gdn_undo_interlace.c:31-36 (undo == 1)
__global__ __launch_bounds__(kBlock) void gdn_interlace_kernel(int undo, u64 r, u64 d_out, u64 mmax, u64 t_out,
                                                               u64 d_in, u64 t_in, u64 alpha,
                                                               const u64 *__restrict__ src,
                                                               u64 *__restrict__ dst) {
    const u64 total = r * d_out * mmax * t_out;
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const u64 to = e % t_out;
    const u64 rest = e / t_out;
    const u64 k = rest % mmax;
    const u64 lj = rest / mmax;
    const u64 jo = lj % d_out;
    const u64 l = lj / d_out;
    u64 ji, ti;
    if (!undo) {  // C_alpha[l, ja, k, ta] = C[l, ja*alpha + ta%alpha, k, ta/alpha]
        ji = jo * alpha + to % alpha;
        ti = to / alpha;
    } else {      // C[l, j, k, t] = C_alpha[l, j/alpha, k, t*alpha + j%alpha]
        ji = jo / alpha;
        ti = to * alpha + jo % alpha;
    }
    dst[e] = src[((l * d_in + ji) * mmax + k) * t_in + ti];
}

// ------------------------------------- fused generation, scrambling, float
//
// gdn_gen_natural -> gdn_digital_permutation or gdn_digital_shift (r_x = r)
// -> gdn_integer_to_float in one pass that writes only the doubles.  A block
// takes one column (l, j) and kBlock * kFuseH consecutive points; each lane
// walks kFuseH of them with the column's tmax digits in registers (the walk
// of gdn_gen_chain_kernel, its prefix rows Q in LDS), looks each digit up in
// the column's permutation rows (LDS when they fit as u32) or adds its shift,
// and sums v += recip_t * y_t in the reference's order with the same
// reciprocals.  Columns that fail the walk's exactness condition compute
// every digit from its index (gdn_digit).  The d blocks of one point range
// are consecutive on one XCD (block b runs on XCD b % 8), so the 8-byte
// stores of a point's row meet in that XCD's L2.
constexpr u32 kFuseH = 128;  // consecutive points per lane (64: 1.53 ms against 1.40 on config 2)

struct FuseArgs {
    u64 n, d, r_b, mmax, n_start, bmax, Ctotal, tab_total, ngroups, nchunks;
    const u64 *bases, *C, *tab;
    const u32 *Q, *okf;
    double *x;
    u32 *status;
    u32 tmax, tnew, perm_lds, off_q, off_qs, off_rt, off_tab, off_thr;
};

// Scrambled modes with a value table: V[z][t] = RN(recip_t * (double)y_t(z))
// for z below the table's rows, t < TP (0 past tmax_new), where y_t is the
// permutation perm_t(z) (mode 1) or the shifted digit (z + s_t) mod b, the
// raw s_t past tmax (mode 2): the product the reference rounds before each
// add (gdn_integer_to_float.c:36-38), so a digit costs one LDS load and one
// add.  Mode 1 uses it when bmax rows of TP + 1 doubles fit in
// kFuseTabBytes, mode 2 (bases unknown to the host) with as many rows as fit
// there, a column with a larger base taking the general path; the table sits at LDS address 0 and the walk keeps each
// digit as the byte offset z (TP + 1) 8 of its row, so the load's immediate
// offset is 8 t.  The odd row length in doubles spreads the lanes' random z
// over all LDS bank pairs (rows of TP doubles put every z on one of four).
// Workgroups of kFuseTabBlock threads keep four waves per SIMD with two
// tables per CU.
constexpr u32 kFuseTabBytes = 65536;
constexpr u32 kFuseTabBlock = 512;

// carries of k -> k+1 in the point walk: returns the row of Q to add (the
// number c of trailing digits of k equal to b - 1, capped at mmax - 1) and
// advances the lowest digit dlow
__device__ __forceinline__ u32 walk_carry(u32 &dlow, u64 k, u64 b, u32 bb, const DivU32 &bdiv, u32 mmax) {
    u32 c = 0;
    if (dlow + 1 < bb) {
        dlow++;
    } else {
        dlow = 0;
        // (k + 1) / b: 32-bit reciprocal division while the index fits 32 bits (a wave takes
        // this branch whenever any of its lanes carries)
        const u64 k1 = k + 1;
        u64 xq = (k1 >> 32) ? k1 / b : (u64)div_q((u32)k1, bdiv);
        c = 1;
        while (xq != 0 && c + 1 < mmax) {
            const u64 qq = (xq >> 32) ? xq / b : (u64)div_q((u32)xq, bdiv);
            if (xq - qq * b != 0) break;
            xq = qq;
            c++;
        }
    }
    return c < mmax ? c : mmax - 1;
}

template <int MODE>
__device__ __forceinline__ u64 fuse_scramble(const FuseArgs &a, u32 t, u64 z, u64 b, u64 pbase, const u32 *sP,
                                             const u64 *sS, bool plds, u32 &bad) {
    if (MODE == 0) return z;
    if (MODE == 1) {  // gdn_digital_permutation.c:36-42
        if (plds && z < a.bmax) return sP[(u64)t * a.bmax + z];
        const u64 pi = pbase + (u64)t * a.bmax + z;
        if (pi < a.tab_total) return a.tab[pi];
        bad |= STATUS_OOB;
        return 0;
    }
    // gdn_digital_shift.c:37-42 (the sum wraps as the reference's u64 sum)
    const u64 sh = sS[t];
    if (b == 0) {
        bad |= STATUS_BAD_BASE;
        return 0;
    }
    const u64 xs = z + sh;
    return xs < 2 * b && xs >= sh ? (xs >= b ? xs - b : xs) : umod(xs, b);
}

template <int MODE>
__device__ __forceinline__ u64 fuse_tail(const FuseArgs &a, u32 t, u64 pbase, const u32 *sP, const u64 *sS,
                                         bool plds, u32 &bad) {
    if (MODE == 1) {  // perms[l, j, t, 0] past tmax
        if (plds) return sP[(u64)t * a.bmax];
        const u64 pi = pbase + (u64)t * a.bmax;
        if (pi < a.tab_total) return a.tab[pi];
        bad |= STATUS_OOB;
        return 0;
    }
    return MODE == 2 ? sS[t] : 0;  // the raw shift past tmax
}

template <int TP, int MODE, bool TAB>
__global__ __launch_bounds__(TAB ? kFuseTabBlock : kBlock) void gdn_fused_kernel(FuseArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const u64 bid = blockIdx.x;
    const u64 xcd = bid & 7, qd = bid >> 3;
    const u64 j = qd % a.d, u = qd / a.d;
    const u64 g = u * 8 + xcd;  // (l, point range), one per XCD lane of the grid
    if (g >= a.ngroups) return;  // uniform over the block, before any barrier
    const u64 l = g / a.nchunks, chunk = g - (g / a.nchunks) * a.nchunks;
    const u64 lj = l * a.d + j;
    const u64 b = a.bases[(l % a.r_b) * a.d + j];
    const u32 tmax = a.tmax, tnew = a.tnew;
    const u32 rows = (u32)TP > tnew ? (u32)TP : tnew;  // table rows in LDS, zero past tnew
    const bool ok = a.Q != nullptr && a.okf[lj] != 0;
    // LDS: [value table V (bmax x TP f64) | Q rows scaled by TP * 8] when TAB |
    // Q rows (mmax x TP u32, zero past tmax) | reciprocals (tnew) | the same
    // zeroed past tmax (TP) | perms (rows x bmax u32) or shifts (rows u64)
    u32 *sQ = (u32 *)(smem + a.off_q);
    u32 *sQs = (u32 *)(smem + a.off_qs);
    double *srt = (double *)(smem + a.off_rt);
    double *srtm = srt + tnew;
    u32 *sP = (u32 *)(smem + a.off_tab);
    u64 *sS = (u64 *)(smem + a.off_tab);
    // table path: per-level largest term (as bits), thresholds, walked levels
    unsigned long long *sM = (unsigned long long *)(smem + a.off_thr);
    double *sThr = (double *)(smem + a.off_thr + 8 * 64);
    u32 *sTW = (u32 *)(smem + a.off_thr + 8 * 64 + 8 * 65);
    // base 2: Q rows as bit words (64), the flip word, the flag
    unsigned long long *sB2 = (unsigned long long *)(smem + a.off_thr + 8 * 64 + 8 * 65 + 16);
    u32 bad0 = 0;
    const u64 pbase = lj * tnew * a.bmax;
    if (ok)
        for (u32 e = threadIdx.x; e < (u32)a.mmax * TP; e += blockDim.x) {
            const u32 m = e / TP, t = e - m * TP;
            const u32 qv = t < tmax ? a.Q[(lj * a.mmax + m) * tmax + t] : 0u;
            sQ[e] = qv;
            if (TAB) sQs[e] = qv * (u32)((TP + 1) * 8);
        }
    if (threadIdx.x == 0) {  // gdn_integer_to_float.c:34-38
        const double bd = (double)b;
        double recip = 1 / bd;
        for (u32 t = 0; t < tnew; t++) {
            srt[t] = recip;
            if (t < (u32)TP) srtm[t] = t < tmax ? recip : 0.;
            recip /= bd;
        }
        for (u32 t = tnew; t < (u32)TP; t++) srtm[t] = 0.;
    }
    int flags = 0;  // bit 0: a permutation entry >= 2^32; bit 1: a shift >= b
    if (MODE == 1 && a.perm_lds)
        for (u32 e = threadIdx.x; e < rows * (u32)a.bmax; e += blockDim.x) {
            const u64 v = e < tnew * (u32)a.bmax ? a.tab[pbase + e] : 0;
            sP[e] = (u32)v;
            flags |= (v >> 32) != 0;
        }
    if (MODE == 2)
        for (u32 t = threadIdx.x; t < rows; t += blockDim.x) {
            const u64 v = t < tnew ? a.tab[lj * tnew + t] : 0;
            sS[t] = v;
            flags |= (t < tmax && v >= b) ? 2 : 0;
        }
    if (TAB)
        __syncthreads();  // no flags in this mode; the or-barriers would place LDS before the table
    else
        flags = (__syncthreads_or(flags & 1) ? 1 : 0) | (__syncthreads_or(flags & 2) ? 2 : 0);
    if (TAB) {
        double *V = (double *)smem;
        const u32 vrows = (u32)(b < a.bmax ? b : a.bmax);  // digits z < b
        const DivU32 vdiv = make_div(vrows > 0 ? vrows : 1u);
        // z fastest: a permutation row's entries are contiguous in memory
#pragma unroll 4
        for (u32 e = threadIdx.x; e < vrows * TP; e += blockDim.x) {
            const u32 t = div_q(e, vdiv), zz = e - t * vrows;
            double y = 0.0;
            if (t < tnew) {
                if (MODE == 1) {
                    y = (double)a.tab[pbase + (u64)t * a.bmax + zz];
                } else {  // gdn_digital_shift.c:37-42, the sum wrapping as the reference's u64 sum
                    const u64 sh = sS[t], xs = (u64)zz + sh;
                    y = (double)(t >= tmax ? sh : xs >= sh && xs < 2 * b ? (xs >= b ? xs - b : xs) : umod(xs, b));
                }
            }
            const double term = t < tnew ? srt[t] * y : 0.0;
            V[zz * (TP + 1) + t] = term;
        }
        __syncthreads();
        // the largest term of each level: a scan of the table's column by one
        // thread per level (terms are >= 0, so their bit patterns order as the
        // values; LDS atomics on one word per level serialized the fill);
        // levels past the table have one tail term each
        // four consecutive lanes per level (tnew <= 64, so 4 tnew <= 256 lanes, one pass), each
        // scanning every fourth row, their maxima combined within the lane group
        for (u32 e = threadIdx.x; e < 4 * tnew; e += blockDim.x) {
            const u32 t = e >> 2, part = e & 3;
            unsigned long long mx = 0;
            if (t < (u32)TP) {
                for (u32 zz = part; zz < vrows; zz += 4) {
                    const unsigned long long w = (unsigned long long)__double_as_longlong(V[zz * (TP + 1) + t]);
                    mx = w > mx ? w : mx;
                }
            } else if (part == 0) {
                mx = (unsigned long long)__double_as_longlong(srt[t] *
                                                               (double)fuse_tail<MODE>(a, t, pbase, sP, sS, false, bad0));
            }
            unsigned long long o = __shfl_xor(mx, 1, 4);
            mx = o > mx ? o : mx;
            o = __shfl_xor(mx, 2, 4);
            mx = o > mx ? o : mx;
            if (part == 0) sM[t] = mx;
        }
        __syncthreads();
        // Exact early exit of the sums (as in the batch's gdn class): B_t =
        // the largest term of any level s >= t (terms are >= 0, so their bit
        // patterns order as the values); a partial sum v >= thr_t =
        // 2^(ilogb(B_t) + 54) has ulp(v) / 2 > B_t, so every later add
        // rounds back to v.  Levels from TW on are walked only while some
        // lane of the wave is below the threshold; TW is the first level
        // whose threshold is at most 2^-20.
        if (threadIdx.x == 0) {
            double B = 0.0;
            u32 tw = TP;
            for (int t = (int)tnew - 1; t >= 0; t--) {
                const double mt = __longlong_as_double((long long)sM[t]);
                B = mt > B ? mt : B;
                const double th = B > 0.0 ? ldexp(1.0, ilogb(B) + 54) : 0.0;
                sThr[t] = th;
                if (th <= 0x1p-20 && (u32)t < tw) tw = (u32)t;
            }
            sThr[tnew] = 0.0;
            sTW[0] = tw;
            // Base 2 with every term 0 or recip_t = 2^-(t+1) (a digit map of {0, 1} onto
            // itself; the raw tail values 0 or 1): every partial sum of the reference's
            // loop is a multiple of 2^-tnew below 1, exact for tnew <= 53, so the sum
            // is Y 2^-tnew with y_t the bit tnew - 1 - t of Y, and Y = Z xor F, Z the
            // digits as bits and F the images of digit 0
            unsigned long long okb = 0, F = 0;
            if (b == 2 && tnew <= (u32)TP && tnew <= 53) {
                okb = 1;
                for (u32 t = 0; t < tnew; t++) {
                    const double v0 = V[t], v1 = V[(TP + 1) + t], rt = srt[t];
                    const bool good = t < tmax ? (v0 == 0.0 && v1 == rt) || (v0 == rt && v1 == 0.0)
                                               : (v0 == 0.0 || v0 == rt);
                    okb &= good ? 1ull : 0ull;
                    F |= (unsigned long long)(v0 != 0.0) << (tnew - 1 - t);
                }
            }
            sB2[64] = F;
            sB2[65] = okb;
        }
        if (b == 2 && threadIdx.x >= 64 && threadIdx.x < 64 + (u32)a.mmax) {
            const u32 m = threadIdx.x - 64;
            unsigned long long w = 0;
            for (u32 t = 0; t < tmax && t < tnew && t < (u32)TP; t++)
                w |= (unsigned long long)(sQ[m * TP + t] & 1u) << (tnew - 1 - t);
            sB2[m] = w;
        }
        __syncthreads();
    }
    const bool plds = MODE == 1 && a.perm_lds && !(flags & 1);
    // branch-free digit loop: every digit a walk produces is < b, so it needs
    // a table in LDS and b <= bmax (permutations) or shifts < b (no wrap)
    // the table path addresses V by byte offset from LDS address 0, where
    // the dynamic LDS starts when the kernel has no static LDS
    const bool tab0 = TAB && (u32)(uintptr_t)(__attribute__((address_space(3))) const unsigned char *)smem == 0;
    const bool fast = TAB ? tab0 && b <= a.bmax
                          : MODE == 0 || (MODE == 1 && plds && b <= a.bmax) || (MODE == 2 && !(flags & 2));

    const u64 i_first = chunk * ((u64)blockDim.x * kFuseH) + (u64)threadIdx.x * kFuseH;
    if (i_first >= a.n) return;
    const u32 H = (a.n - i_first) < (u64)kFuseH ? (u32)(a.n - i_first) : kFuseH;
    u64 k = i_first + a.n_start;
    double *xo = a.x + (l * a.n + i_first) * a.d + j;
    u32 bad = 0;
    if (!ok) {
        for (u32 h = 0; h < H; h++, k++) {
            double v = 0.;
            for (u32 t = 0; t < tnew; t++) {
                u64 y;
                if (t < tmax)
                    y = fuse_scramble<MODE>(a, t, gdn_digit(k, b, a.C, lj * a.mmax * tmax + t, tmax, a.Ctotal, bad),
                                            b, pbase, sP, sS, plds, bad);
                else
                    y = fuse_tail<MODE>(a, t, pbase, sP, sS, plds, bad);
                v += srt[t] * (double)y;
            }
            xo[(u64)h * a.d] = v;
        }
        if (bad) flag(a.status, bad);
        return;
    }
    // digits of the first point: z_t = sum_m (dig_m - dig_{m+1}) Q_m[t] mod b
    // (the padded entries t >= tmax stay 0)
    const u32 bb = (u32)b;
    const DivU32 bdiv = make_div(bb);
    u32 z[TP];
#pragma unroll
    for (int t = 0; t < TP; t++) z[t] = 0;
    {
        u64 kk = k;
        u32 dig = (kk >> 32) ? (u32)(kk % b) : (u32)kk - div_q((u32)kk, bdiv) * bb;
        kk = (kk - dig) / b;
        for (u32 m = 0; dig != 0 || kk != 0; m++) {
            const u32 dn = (kk >> 32) ? (u32)(kk % b) : (u32)kk - div_q((u32)kk, bdiv) * bb;
            kk = (kk - dn) / b;
            const u32 e = dig >= dn ? dig - dn : dig + bb - dn;
            const u32 *qm = sQ + m * TP;
            if (bb < (1u << 16)) {
#pragma unroll
                for (int t = 0; t < TP; t++) {
                    const u32 p = e * qm[t];
                    const u32 s2 = z[t] + (p - div_q(p, bdiv) * bb);
                    z[t] = __builtin_elementwise_min(s2, s2 - bb);
                }
            } else {
#pragma unroll
                for (int t = 0; t < TP; t++) {
                    const u32 s2 = z[t] + (u32)(((u64)e * qm[t]) % b);
                    z[t] = __builtin_elementwise_min(s2, s2 - bb);
                }
            }
            dig = dn;
        }
    }
    u32 dlow = (k >> 32) ? (u32)(k % b) : (u32)k - div_q((u32)k, bdiv) * bb;
    if (TAB && fast && b == 2 && sB2[65] != 0) {
        // base 2 (see the prologue): the digits as one bit word, a carry row xor per step
        unsigned long long Z = 0;
#pragma unroll
        for (int t = 0; t < TP; t++)
            if ((u32)t < tnew) Z |= (unsigned long long)(z[t] & 1u) << (tnew - 1 - t);
        const unsigned long long F = sB2[64];
        const double sc = ldexp(1.0, -(int)tnew);
        for (u32 h = 0;; h++) {
            xo[(u64)h * a.d] = (double)(Z ^ F) * sc;
            if (h + 1 == H) break;
            Z ^= sB2[walk_carry(dlow, k, b, bb, bdiv, (u32)a.mmax)];
            k++;
        }
        if (bad | bad0) flag(a.status, bad | bad0);
        return;
    }
    if (TAB && fast) {
        constexpr u32 S = (TP + 1) * 8;
        const u32 tw = __builtin_amdgcn_readfirstlane(sTW[0]);  // levels a point usually needs (uniform)
#pragma unroll
        for (int t = 0; t < TP; t++) z[t] *= S;
        // walk and sum the first TL levels (TL covering tw, below): a wave adds
        // every one of them (terms past a lane's threshold leave its sum
        // unchanged) and checks once, at TL
        auto run = [&](auto tl) {
            constexpr int TL = decltype(tl)::value;
            const double thl = sThr[(u32)TL < tnew ? (u32)TL : tnew];
            for (u32 h = 0;; h++) {
                // the reference adds the first term to 0.0: the same value, as no term is -0.0
                double v = *(__attribute__((address_space(3))) const double *)(uintptr_t)z[0];
#pragma unroll
                for (int t = 1; t < TL; t++)  // + 0 past tmax_new
                    v += *(__attribute__((address_space(3))) const double *)(uintptr_t)(z[t] + 8u * t);
                if ((u32)TL < tnew && __builtin_amdgcn_ballot_w64(v < thl) != 0) {
                    // the levels from TL on, for the lanes still below them:
                    // the digit from the point's index, then the table (or
                    // the tail value past tmax), the reference's product and add
                    for (u32 t = TL; t < tnew; t++) {
                        const bool act = v < sThr[t];
                        if (__builtin_amdgcn_ballot_w64(act) == 0) break;
                        if (!act) continue;
                        u64 y;
                        if (t < tmax)
                            y = fuse_scramble<MODE>(a, t,
                                                    gdn_digit(k, b, a.C, lj * a.mmax * tmax + t, tmax, a.Ctotal, bad),
                                                    b, pbase, sP, sS, plds, bad);
                        else
                            y = fuse_tail<MODE>(a, t, pbase, sP, sS, plds, bad);
                        v += srt[t] * (double)y;
                    }
                }
                xo[(u64)h * a.d] = v;
                if (h + 1 == H) break;
                const u32 *qc = sQs + walk_carry(dlow, k, b, bb, bdiv, (u32)a.mmax) * TP;
#pragma unroll
                for (int t = 0; t < TL; t++) {
                    const u32 s2 = z[t] + qc[t];
                    z[t] = __builtin_elementwise_min(s2, s2 - bb * S);
                }
                k++;
            }
        };
        // TL: the smallest of 12, 14, ..., 28, 32, 36, 40, 44, 48 (below TP) or
        // TP at or above tw (each walked level costs its walk step, load and
        // add on every point: the buckets follow the bases closely; for the
        // first 32 primes at 52 digits they walk 556 levels against 540 needed)
#define QMC_FUSE_TL(L_)                                                  \
    else if (TP > (L_) && tw <= (L_)) run(std::integral_constant<int, (TP > (L_) ? (L_) : TP)>());
        if (false) {
        }
        QMC_FUSE_TL(12) QMC_FUSE_TL(14) QMC_FUSE_TL(16) QMC_FUSE_TL(18) QMC_FUSE_TL(20) QMC_FUSE_TL(22)
        QMC_FUSE_TL(24) QMC_FUSE_TL(26) QMC_FUSE_TL(28) QMC_FUSE_TL(32) QMC_FUSE_TL(36) QMC_FUSE_TL(40)
        QMC_FUSE_TL(44) QMC_FUSE_TL(48) else run(std::integral_constant<int, TP>());
#undef QMC_FUSE_TL
        if (bad | bad0) flag(a.status, bad | bad0);
        return;
    }
    const u32 bm = (u32)a.bmax;
    for (u32 h = 0;; h++) {
        double v = 0.;
        if (fast) {
#pragma unroll
            for (int t = 0; t < TP; t++) {
                u32 y;
                if (MODE == 0) {
                    y = z[t];
                } else if (MODE == 1) {
                    y = sP[t * bm + z[t]];
                } else {
                    const u32 s2 = z[t] + (u32)sS[t];
                    y = __builtin_elementwise_min(s2, s2 - bb);
                }
                v += srtm[t] * (double)y;  // + 0 for the padded digits
            }
        } else {
            for (u32 t = 0; t < tmax; t++) {
                u32 zt = 0;
#pragma unroll
                for (int tt = 0; tt < TP; tt++)
                    if (tt == (int)t) zt = z[tt];
                v += srt[t] * (double)fuse_scramble<MODE>(a, t, zt, b, pbase, sP, sS, plds, bad);
            }
        }
        for (u32 t = tmax; t < tnew; t++) v += srt[t] * (double)fuse_tail<MODE>(a, t, pbase, sP, sS, plds, bad);
        xo[(u64)h * a.d] = v;
        if (h + 1 == H) break;
        const u32 *qc = sQ + walk_carry(dlow, k, b, bb, bdiv, (u32)a.mmax) * TP;
#pragma unroll
        for (int t = 0; t < TP; t++) {
            const u32 s2 = z[t] + qc[t];
            z[t] = __builtin_elementwise_min(s2, s2 - bb);
        }
        k++;
    }
    if (bad) flag(a.status, bad);
}

Rows make_rows(u64 n, u64 d, u64 tdigits) {
    Rows g;
    g.nd = n * d;
    g.d = d;
    g.small = g.nd < (1ull << 32) && d < (1ull << 31);
    g.ddiv = make_div(g.small ? (u32)d : 1u);
    g.tdiv = make_div((u32)tdigits);
    return g;
}

}  // namespace

u32 *qmc_status_ptr();

// Scratch of qmc_gdn_gen_scramble_float_noalloc_dev: the prefix table Q
// (r*d*mmax*tmax words) followed by one walk flag per column (r*d words).
// gdn_chain_setup_kernel writes every word the fused kernel reads in the same
// call, so nothing in it needs clearing.  Each device holds its own copy until
// the process ends; calls that overlap on one device race on it.
constexpr u64 kScratchWords = 1ull << 20;  // 4 MiB
static __device__ __align__(256) u32 scratch[kScratchWords];

namespace {
std::mutex g_scratch_mu;
u32 *g_scratch[64] = {};

// The current device's address of `scratch`, looked up once per device (in
// host code the array's name is only a stand-in for it).
int scratch_for_device(u32 **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return qmc_fail(QMC_ERR_HIP, "no usable HIP device");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        u32 *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return qmc_fail(QMC_ERR_HIP, "cannot find the gdn scratch: %s", hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return QMC_OK;
}
}  // namespace

// Called by qmc_init on the current device.
int gdn_scratch_setup() {
    u32 *p = nullptr;
    return scratch_for_device(&p);
}

extern "C" {

static int gen_common(const char *who, u64 r, u64 n, u64 d, u64 r_b, u64 mmax, u64 tmax, u64 n_start,
                      const u64 *bases, u64 b_common, const u64 *C, u64 *xdig, hipStream_t s) {
    if (!(r && n && d && tmax)) return QMC_OK;
    if (bases && r_b == 0) return qmc_fail(QMC_ERR_ARG, "%s: r_b must be positive", who);
    if (tmax > 4096) return qmc_fail(QMC_ERR_ARG, "%s: tmax = %llu is above the supported 4096", who,
                                     (unsigned long long)tmax);
    if (r > 65535) return qmc_fail(QMC_ERR_ARG, "%s: r above 65535 is not supported", who);
    u32 *status = qmc_status_ptr();
    if (!status) return QMC_ERR_HIP;
    // the prefix table Q (u32 per C entry) and one flag per column; past 1 GiB
    // of table every digit is computed from its index
    const u64 qwords = r * d * mmax * tmax;
    u32 *ws = nullptr;
    const u64 kmax = n_start + (n - 1);
    const bool walk = qwords * 4 <= (1ull << 30) && mmax > 0 && r * d < (1ull << 31) &&
                      kmax >= n_start;  // no index wrap-around
    if (walk) QMC_HIP_CHECK(hipMallocAsync((void **)&ws, sizeof(u32) * (qwords + r * d), s));
    u32 *okf = walk ? ws + qwords : nullptr;
    if (walk)
        hipLaunchKernelGGL(gdn_chain_setup_kernel, dim3((unsigned)(r * d)), dim3(64), 0, s, d, mmax, tmax, kmax,
                           r_b, bases, b_common, C, ws, okf);
    const u64 row = d * tmax;
    dim3 grid((unsigned)((n + kChainH - 1) / kChainH), (unsigned)((row + kChainF - 1) / kChainF), (unsigned)r);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gdn_gen_chain_kernel, grid, dim3(kBlock), 0, s, n, d, tmax, n_start, r_b, mmax, bases,
                           b_common, C, r * d * mmax * tmax, walk ? ws : nullptr, okf, xdig, status);
        e = hipGetLastError();
    }
    if (ws) (void)hipFreeAsync(ws, s);
    if (e != hipSuccess) return qmc_fail_hip(e, who);
    return QMC_OK;
}

int gdn_gen_natural_dev(u64 r, u64 n, u64 d, u64 r_b, u64 mmax, u64 tmax, u64 n_start, const u64 *bases,
                        const u64 *C, u64 *xdig, hipStream_t s) {
    return gen_common("gdn_gen_natural", r, n, d, r_b, mmax, tmax, n_start, bases, 0, C, xdig, s);
}

int gdn_gen_natural_same_base_dev(u64 r, u64 n, u64 d, u64 mmax, u64 tmax, u64 n_start, u64 b, const u64 *C,
                                  u64 *xdig, hipStream_t s) {
    return gen_common("gdn_gen_natural_same_base", r, n, d, 1, mmax, tmax, n_start, nullptr, b, C, xdig, s);
}

int gdn_linear_matrix_scramble_dev(u64 r, u64 d, u64 mmax, u64 r_C, u64 r_b, u64 tmax, u64 tmax_new,
                                   const u64 *bases, const u64 *S, const u64 *C, u64 *C_lms, hipStream_t s) {
    const u64 total = r * d * mmax * tmax_new;
    if (!total) return QMC_OK;
    if (r_C == 0 || r_b == 0) return qmc_fail(QMC_ERR_ARG, "gdn_linear_matrix_scramble: r_C and r_b must be positive");
    u32 *status = qmc_status_ptr();
    if (!status) return QMC_ERR_HIP;
    hipLaunchKernelGGL(gdn_lms_kernel, dim3(grid_for(total, kBlock, 0x7fffffffu)), dim3(kBlock), 0, s, r, d, mmax,
                       r_C, r_b, tmax, tmax_new, bases, S, C, C_lms, status);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

static int digitwise(bool permute, const char *who, u64 r, u64 n, u64 d, u64 r_x, u64 r_b, u64 tmax,
                     u64 tmax_new, u64 bmax, const u64 *bases, const u64 *tab, const u64 *xdig, u64 *xnew,
                     hipStream_t s) {
    if (!(r && n && d && tmax_new)) return QMC_OK;
    if (r_x == 0 || (!permute && r_b == 0)) return qmc_fail(QMC_ERR_ARG, "%s: r_x and r_b must be positive", who);
    if (tmax > tmax_new)
        return qmc_fail(QMC_ERR_ARG, "%s: tmax = %llu > tmax_new = %llu overlaps output rows in the reference",
                        who, (unsigned long long)tmax, (unsigned long long)tmax_new);
    if (tmax_new > 4096) return qmc_fail(QMC_ERR_ARG, "%s: tmax_new above 4096 is not supported", who);
    if (r > 65535) return qmc_fail(QMC_ERR_ARG, "%s: r above 65535 is not supported", who);
    u32 *status = qmc_status_ptr();
    if (!status) return QMC_ERR_HIP;
    Rows g = make_rows(n, d, tmax_new);
    dim3 grid((unsigned)((g.nd + kRows - 1) / kRows), (unsigned)r);
    if (permute && tmax_new * bmax <= kPermLdsEntries && bmax) {
        const u64 chunks = ((n + kPermPts - 1) / kPermPts + 7) & ~7ull;  // whole groups of 8
        if (chunks * d > 0x7fffffffull) return qmc_fail(QMC_ERR_ARG, "%s: n*d too large", who);
        hipLaunchKernelGGL(gdn_perm_lds_kernel, dim3((unsigned)(chunks * d), (unsigned)r), dim3(kBlock),
                           sizeof(u32) * tmax_new * bmax, s, n, d, r_x, tmax, tmax_new, bmax, tab,
                           r * d * tmax_new * bmax, xdig, xnew, status);
    } else if (permute)
        hipLaunchKernelGGL(gdn_digitwise_kernel<true>, grid, dim3(kBlock), 0, s, g, r_x, r_b, tmax, tmax_new, bmax,
                           bases, tab, r * d * tmax_new * bmax, xdig, xnew, status);
    else
        hipLaunchKernelGGL(gdn_digitwise_kernel<false>, grid, dim3(kBlock), 0, s, g, r_x, r_b, tmax, tmax_new, bmax,
                           bases, tab, r * d * tmax_new, xdig, xnew, status);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

// gdn_gen_natural -> (mode 1) gdn_digital_permutation or (mode 2)
// gdn_digital_shift, each with r_x = r -> gdn_integer_to_float, writing only
// x (r, n, d).  Mode 0 converts the generated digits directly (tmax_new must
// equal tmax).  tab: perms (r, d, tmax_new, bmax) or shifts (r, d, tmax_new).
// noalloc: the prefix table lives in `scratch` instead of a stream-ordered
// allocation, and a call needing more than kScratchWords words is rejected.
static int fused_gdn(const char *who, bool noalloc, int mode, u64 r, u64 n, u64 d, u64 r_b, u64 mmax, u64 tmax,
                     u64 tmax_new, u64 n_start, u64 bmax, const u64 *bases, const u64 *C, const u64 *tab, double *x,
                     hipStream_t s) {
    if (!(r && n && d)) return QMC_OK;
    if (mode < 0 || mode > 2) return qmc_fail(QMC_ERR_ARG, "%s: mode must be 0, 1 or 2", who);
    if (r_b == 0) return qmc_fail(QMC_ERR_ARG, "%s: r_b must be positive", who);
    if (mode == 0 && tmax_new != tmax) return qmc_fail(QMC_ERR_ARG, "%s: mode 0 needs tmax_new == tmax", who);
    if (tmax < 1 || tmax > tmax_new || tmax_new > 64)
        return qmc_fail(QMC_ERR_ARG, "%s: needs 1 <= tmax <= tmax_new <= 64", who);
    if (mmax > 64) return qmc_fail(QMC_ERR_ARG, "%s: mmax above 64 is not supported", who);
    if (mode == 1 && bmax == 0) return qmc_fail(QMC_ERR_ARG, "%s: bmax must be positive", who);
    u32 *status = qmc_status_ptr();
    if (!status) return QMC_ERR_HIP;
    const u64 kmax = n_start + (n - 1);
    const u64 qwords = r * d * mmax * tmax;
    const bool walk = mmax > 0 && kmax >= n_start && qwords * 4 <= (1ull << 30) && r * d < (1ull << 31);
    // 52 digits (the double-precision digit count of the bases up to 131
    // with permutations) get rows of their own rather than 56 padded ones
    const int TP = tmax <= 16 ? 16 : tmax <= 32 ? 32 : tmax <= 52 ? 52 : tmax <= 56 ? 56 : 64;
    const u64 rows = (u64)TP > tmax_new ? (u64)TP : tmax_new;
    const bool tabm = walk && ((mode == 1 && bmax * (TP + 1) * 8 <= kFuseTabBytes) || mode == 2);
    const u64 vrows = mode == 1 ? bmax : kFuseTabBytes / ((TP + 1) * 8);  // rows of the value table
    const u64 bs = tabm ? kFuseTabBlock : kBlock;
    FuseArgs a;
    a.n = n;
    a.d = d;
    a.r_b = r_b;
    a.mmax = mmax;
    a.n_start = n_start;
    a.bmax = tabm ? vrows : bmax;
    a.Ctotal = r * d * mmax * tmax;
    a.tab_total = mode == 1 ? r * d * tmax_new * bmax : r * d * tmax_new;
    a.nchunks = (n + bs * kFuseH - 1) / (bs * kFuseH);
    a.ngroups = r * a.nchunks;
    a.bases = bases;
    a.C = C;
    a.tab = tab;
    a.x = x;
    a.status = status;
    a.tmax = (u32)tmax;
    a.tnew = (u32)tmax_new;
    a.perm_lds = mode == 1 && !tabm && rows * bmax * 4 <= 40960;
    const u32 qbytes = (u32)((sizeof(u32) * mmax * TP + 15) & ~(u64)15);
    a.off_q = tabm ? (u32)((vrows * (TP + 1) * 8 + 15) & ~(u64)15) : 0;
    a.off_qs = a.off_q + qbytes;
    a.off_rt = a.off_qs + (tabm ? qbytes : 0);
    a.off_tab = a.off_rt + (u32)(8 * (tmax_new + TP));
    a.off_thr = (u32)((a.off_tab + (mode == 1 ? (a.perm_lds ? 4 * rows * bmax : 0) : mode == 2 ? 8 * rows : 0) + 15) &
                      ~(u64)15);
    const size_t shm = a.off_thr + (tabm ? 8 * 64 + 8 * 65 + 16 + 8 * 66 : 0);
    const u64 nblocks = (a.ngroups + 7) / 8 * 8 * d;
    if (nblocks > 0x7fffffffull) return qmc_fail(QMC_ERR_ARG, "%s: r*n*d too large", who);
    u32 *ws = nullptr;
    if (walk && noalloc) {
        if (qwords + r * d > kScratchWords)
            return qmc_fail(QMC_ERR_ARG,
                            "%s: needs %llu words of scratch (r*d*mmax*tmax + r*d), above the %llu it holds", who,
                            (unsigned long long)(qwords + r * d), (unsigned long long)kScratchWords);
        const int rc = scratch_for_device(&ws);
        if (rc != QMC_OK) return rc;
    } else if (walk) {
        QMC_HIP_CHECK(hipMallocAsync((void **)&ws, sizeof(u32) * (qwords + r * d), s));
    }
    a.Q = ws;
    a.okf = walk ? ws + qwords : nullptr;
    if (walk)
        hipLaunchKernelGGL(gdn_chain_setup_kernel, dim3((unsigned)(r * d)), dim3(64), 0, s, d, mmax, tmax, kmax,
                           r_b, bases, (u64)0, C, ws, ws + qwords);
    if (tabm) {
        // the value table and two Q tables exceed the default 64 KiB of LDS
        static PerDeviceOnce attrs;
        const int most = (int)(kFuseTabBytes + 2 * sizeof(u32) * 64 * 64 + 8 * 128 + 8 * 64 + 2048);
        const int rc = attrs.run([&]() -> int {
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<16, 1, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<32, 1, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<52, 1, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<56, 1, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<64, 1, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<16, 2, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<32, 2, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<52, 2, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<56, 2, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            QMC_HIP_CHECK(hipFuncSetAttribute((const void *)gdn_fused_kernel<64, 2, true>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, most));
            return QMC_OK;
        });
        if (rc != QMC_OK) {
            if (ws && !noalloc) (void)hipFreeAsync(ws, s);
            return rc;
        }
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
#define QMC_FUSE(tp, md, tb)                                                                                   \
    hipLaunchKernelGGL((gdn_fused_kernel<tp, md, tb>), dim3((unsigned)nblocks), dim3((unsigned)bs), shm, s, a)
#define QMC_FUSE_TP(md, tb)         \
    if (TP == 16)                   \
        QMC_FUSE(16, md, tb);       \
    else if (TP == 32)              \
        QMC_FUSE(32, md, tb);       \
    else if (TP == 52)              \
        QMC_FUSE(52, md, tb);       \
    else if (TP == 56)              \
        QMC_FUSE(56, md, tb);       \
    else                            \
        QMC_FUSE(64, md, tb)
        if (mode == 0) {
            QMC_FUSE_TP(0, false);
        } else if (mode == 1 && tabm) {
            QMC_FUSE_TP(1, true);
        } else if (mode == 1) {
            QMC_FUSE_TP(1, false);
        } else if (tabm) {
            QMC_FUSE_TP(2, true);
        } else {
            QMC_FUSE_TP(2, false);
        }
#undef QMC_FUSE_TP
#undef QMC_FUSE
        e = hipGetLastError();
    }
    if (ws && !noalloc) (void)hipFreeAsync(ws, s);
    if (e != hipSuccess) return qmc_fail_hip(e, who);
    return QMC_OK;
}

QMC_EXPORT int qmc_gdn_gen_scramble_float_dev(int mode, u64 r, u64 n, u64 d, u64 r_b, u64 mmax, u64 tmax,
                                              u64 tmax_new, u64 n_start, u64 bmax, const u64 *bases, const u64 *C,
                                              const u64 *tab, double *x, hipStream_t s) {
    return fused_gdn("gdn_gen_scramble_float", false, mode, r, n, d, r_b, mmax, tmax, tmax_new, n_start, bmax, bases,
                     C, tab, x, s);
}

// The same computation with no device allocation (for callers that cannot
// allocate, e.g. under graph capture); the workspace is `scratch`.
QMC_EXPORT int qmc_gdn_gen_scramble_float_noalloc_dev(int mode, u64 r, u64 n, u64 d, u64 r_b, u64 mmax, u64 tmax,
                                                      u64 tmax_new, u64 n_start, u64 bmax, const u64 *bases,
                                                      const u64 *C, const u64 *tab, double *x, hipStream_t s) {
    return fused_gdn("gdn_gen_scramble_float_noalloc", true, mode, r, n, d, r_b, mmax, tmax, tmax_new, n_start, bmax,
                     bases, C, tab, x, s);
}

int gdn_digital_permutation_dev(u64 r, u64 n, u64 d, u64 r_x, u64 r_b, u64 tmax, u64 tmax_new, u64 bmax,
                                const u64 *perms, const u64 *xdig, u64 *xdig_new, hipStream_t s) {
    return digitwise(true, "gdn_digital_permutation", r, n, d, r_x, r_b, tmax, tmax_new, bmax, nullptr, perms,
                     xdig, xdig_new, s);
}

int gdn_digital_shift_dev(u64 r, u64 n, u64 d, u64 r_x, u64 r_b, u64 tmax, u64 tmax_new, const u64 *bases,
                          const u64 *shifts, const u64 *xdig, u64 *xdig_new, hipStream_t s) {
    return digitwise(false, "gdn_digital_shift", r, n, d, r_x, r_b, tmax, tmax_new, 0, bases, shifts, xdig,
                     xdig_new, s);
}

int gdn_integer_to_float_dev(u64 r, u64 n, u64 d, u64 r_b, u64 tmax, const u64 *bases, const u64 *xdig,
                             double *x, hipStream_t s) {
    if (!(r && n && d)) return QMC_OK;
    if (r_b == 0) return qmc_fail(QMC_ERR_ARG, "gdn_integer_to_float: r_b must be positive");
    if (!tmax) {
        hipLaunchKernelGGL(gdn_zero_kernel, dim3(grid_for(r * n * d, kBlock, 2048 * 8)), dim3(kBlock), 0, s,
                           r * n * d, x);
        QMC_HIP_CHECK(hipGetLastError());
        return QMC_OK;
    }
    double *rt = nullptr;
    const u64 nb = r_b * d;
    QMC_HIP_CHECK(hipMallocAsync((void **)&rt, sizeof(double) * nb * tmax, s));
    hipLaunchKernelGGL(gdn_recip_kernel, dim3((unsigned)((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nb, tmax,
                       bases, rt);
    if (((uintptr_t)xdig % 16) == 0 && tmax <= kTfMaxT) {
        const u32 stride = (u32)tmax | 1u;
        const u64 total = r * n * d;
        const u64 chunks = (total + kTfPts - 1) / kTfPts;
        const DivU32 nddiv = make_div((total >> 32) ? 1u : (u32)(n * d));
        const u64 rt_bytes = sizeof(double) * nb * stride;
        const size_t shm = sizeof(u64) * kTfPts * stride;
        if (shm + rt_bytes <= 49152)  // at least three blocks per CU
            hipLaunchKernelGGL(gdn_to_float_staged_kernel<true>, dim3((unsigned)std::min<u64>(chunks, 256 * 8)),
                               dim3(kTfPts), shm + rt_bytes, s, total, n * d, d, r_b, make_div((u32)tmax), stride,
                               nddiv, make_div((u32)d), rt, xdig, x);
        else
            hipLaunchKernelGGL(gdn_to_float_staged_kernel<false>, dim3((unsigned)std::min<u64>(chunks, 2048 * 16)),
                               dim3(kTfPts), shm, s, total, n * d, d, r_b, make_div((u32)tmax), stride, nddiv,
                               make_div((u32)d), rt, xdig, x);
    } else {
        hipLaunchKernelGGL(gdn_to_float_kernel, dim3(grid_for(r * n * d, kBlock, 2048 * 8)), dim3(kBlock), 0, s,
                           r, n * d, d, r_b, tmax, rt, xdig, x);
    }
    const hipError_t e = hipGetLastError();
    (void)hipFreeAsync(rt, s);
    if (e != hipSuccess) return qmc_fail_hip(e, "gdn_integer_to_float");
    return QMC_OK;
}

int gdn_interlace_dev(u64 r, u64 d_alpha, u64 mmax, u64 d, u64 tmax, u64 tmax_alpha, u64 alpha, const u64 *C,
                      u64 *C_alpha, hipStream_t s) {
    const u64 total = r * d_alpha * mmax * tmax_alpha;
    if (!total) return QMC_OK;
    if (alpha == 0) return qmc_fail(QMC_ERR_ARG, "gdn_interlace: alpha must be positive");
    if (tmax_alpha > alpha * tmax || d_alpha * alpha > d)
        return qmc_fail(QMC_ERR_RANGE, "gdn_interlace: tmax_alpha <= alpha*tmax and d_alpha*alpha <= d are required");
    hipLaunchKernelGGL(gdn_interlace_kernel, dim3(grid_for(total, kBlock, 0x7fffffffu)), dim3(kBlock), 0, s, 0, r,
                       d_alpha, mmax, tmax_alpha, d, tmax, alpha, C, C_alpha);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int gdn_undo_interlace_dev(u64 r, u64 d, u64 mmax, u64 d_alpha, u64 tmax, u64 tmax_alpha, u64 alpha,
                           const u64 *C_alpha, u64 *C, hipStream_t s) {
    const u64 total = r * d * mmax * tmax;
    if (!total) return QMC_OK;
    if (alpha == 0) return qmc_fail(QMC_ERR_ARG, "gdn_undo_interlace: alpha must be positive");
    if (tmax * alpha > tmax_alpha || (d + alpha - 1) / alpha > d_alpha)
        return qmc_fail(QMC_ERR_RANGE, "gdn_undo_interlace: tmax*alpha <= tmax_alpha and ceil(d/alpha) <= d_alpha are required");
    hipLaunchKernelGGL(gdn_interlace_kernel, dim3(grid_for(total, kBlock, 0x7fffffffu)), dim3(kBlock), 0, s, 1, r, d,
                       mmax, tmax, d_alpha, tmax_alpha, alpha, C_alpha, C);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

}  // extern "C"
