// Generalized (base-b) digital nets and Halton points on gfx950: generation,
// linear matrix scrambling, digital shift, digit permutation, conversion to
// double, and interlacing.
//
// Digit outputs are laid out [l][i][j][t] with t fastest, so for one
// replication l the rows rho = i*d + j and their digits form one contiguous
// array.  The row kernels give each block R consecutive rows: lane f of the
// block produces digit (rho0 + f / T, f % T), so stores are fully coalesced,
// and anything shared by a row (its base, the base-b digits of its index) is
// computed once into LDS.
//
// Arithmetic follows the reference's unsigned 64-bit semantics exactly,
// including wrap-around: gdn_gen_natural.c:40-50 reduces after every digit,
// x = (x + dig*C) % b.  When b < 2^16 and every C entry used is < 2^32, no
// step can wrap, so the sum is accumulated exactly and reduced once; any
// element that does not meet that bound takes the step-by-step path.
#include "qmc_common.hpp"

#include <algorithm>

namespace {

constexpr int kRows = 64;        // rows per block in the row kernels
constexpr int kMaxDigits = 64;   // base-b digits of a 64-bit index, b >= 2

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

__global__ __launch_bounds__(kBlock) void gdn_gen_kernel(Rows g, u64 n_start, u64 r_b, u64 mmax, u64 tmax,
                                                         const u64 *__restrict__ bases, u64 b_common,
                                                         const u64 *__restrict__ C, u64 Ctotal,
                                                         u64 *__restrict__ xdig, u32 *status) {
    __shared__ u32 sdig[kRows][kMaxDigits];
    __shared__ u64 sbase[kRows];
    __shared__ u64 smagic[kRows];  // ceil(2^64 / b) for b < 2^16
    __shared__ u32 sr32[kRows];    // 2^32 mod b
    __shared__ u32 scount[kRows];
    __shared__ u64 sj[kRows];
    const u64 l = blockIdx.y;
    const u64 rho0 = (u64)blockIdx.x * kRows;
    const u64 nrows = (g.nd - rho0) < (u64)kRows ? (g.nd - rho0) : (u64)kRows;

    // 1. base-b digits of k = i + n_start for each row of the block
    if (threadIdx.x < nrows) {
        const u64 rho = rho0 + threadIdx.x;
        const u64 j = j_of(g, rho);
        const u64 b = bases ? bases[(l % r_b) * g.d + j] : b_common;
        u64 k = rho / g.d + n_start;
        u32 cnt = 0;
        if (b < 2) {
            flag(status, STATUS_BAD_BASE);  // the reference loops forever (b=1) or divides by 0
            k = 0;
        }
        if ((k >> 32) == 0 && (b >> 32) == 0) {
            u32 kk = (u32)k, bb = (u32)b;
            while (kk) {
                const u32 q = kk / bb;
                sdig[threadIdx.x][cnt++] = kk - q * bb;
                kk = q;
            }
        } else {
            while (k) {
                const u64 q = k / b;
                sdig[threadIdx.x][cnt++] = (u32)(k - q * b);
                k = q;
            }
        }
        sbase[threadIdx.x] = b;
        smagic[threadIdx.x] = (b >= 2 && b < (1u << 16)) ? (~0ull) / b + 1 : 0;
        sr32[threadIdx.x] = (b >= 2 && b < (1u << 16)) ? (u32)(0x100000000ull % b) : 0;
        scount[threadIdx.x] = cnt;
        sj[threadIdx.x] = j;
    }
    __syncthreads();

    // 2. one output digit per lane
    const u32 T = (u32)tmax;
    const u32 total = (u32)nrows * T;
    u64 *out = xdig + (l * g.nd + rho0) * tmax;
    const u64 cl = l * g.d * mmax * tmax;
    for (u32 f = threadIdx.x; f < total; f += blockDim.x) {
        const u32 row = div_q(f, g.tdiv);
        const u32 t = f - row * T;
        const u64 b = sbase[row];
        const u32 cnt = scount[row];
        const u64 cbase = cl + sj[row] * mmax * tmax + t;
        u64 acc = 0;
        // fast path: b < 2^16 and every column entry < 2^32, so no step wraps;
        // all loads of the row are independent (checked once, after the sum)
        bool fast = b >= 2 && b < (1u << 16) && (cnt == 0 || cbase + (u64)(cnt - 1) * tmax < Ctotal);
        if (fast) {
            u64 orc = 0;
            const u64 *Cp = C + cbase;
#pragma unroll 4
            for (u32 m = 0; m < cnt; m++) {
                const u64 c = Cp[(u64)m * tmax];
                orc |= c;
                acc += (u64)sdig[row][m] * c;
            }
            fast = (orc >> 32) == 0;
        }
        if (fast) {
            // acc mod b with 32-bit magic divisions: acc = h 2^32 + lo, b < 2^16
            const u64 M = smagic[row];
            const u32 bb = (u32)b;
            const u32 h = (u32)(acc >> 32), lo = (u32)acc;
            const u32 hm = h - (u32)__umul64hi(M, h) * bb;
            const u32 r32 = sr32[row];
            const u32 lm = lo - (u32)__umul64hi(M, lo) * bb;
            const u32 y = hm * r32 + lm;  // < b^2 <= 2^32
            acc = y - (u32)__umul64hi(M, y) * bb;
        } else if (b < 2) {
            acc = 0;
        } else {  // step by step, as gdn_gen_natural.c:46
            acc = 0;
            for (u32 m = 0; m < cnt; m++) {
                const u64 dg = sdig[row][m];
                if (!dg) continue;
                const u64 ci = cbase + (u64)m * tmax;
                u64 c = 0;
                if (ci < Ctotal) c = C[ci];
                else flag(status, STATUS_OOB);
                acc = (acc + dg * c) % b;
            }
        }
        store_stream(out + f, acc);
    }
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
// reference reads from the following rows), take the global table.
constexpr int kPermPts = 1024;
constexpr u64 kPermLdsEntries = 12288;  // 48 KiB of u32

__global__ __launch_bounds__(kBlock) void gdn_perm_lds_kernel(u64 n, u64 d, u64 r_x, u64 tmax, u64 tmax_new,
                                                              u64 bmax, const u64 *__restrict__ perms, u64 tab_total,
                                                              const u64 *__restrict__ xdig,
                                                              u64 *__restrict__ xnew, u32 *status) {
    extern __shared__ u32 sperm[];
    __shared__ int s_big;
    const u64 l = blockIdx.y;
    const u64 j = blockIdx.x % d;
    const u64 i0 = (blockIdx.x / d) * (u64)kPermPts;
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
        store_stream(out + i * d * tmax_new + t, v);
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
    Rows g = make_rows(n, d, tmax);
    dim3 grid((unsigned)((g.nd + kRows - 1) / kRows), (unsigned)r);
    hipLaunchKernelGGL(gdn_gen_kernel, grid, dim3(kBlock), 0, s, g, n_start, r_b, mmax, tmax, bases, b_common, C,
                       r * d * mmax * tmax, xdig, status);
    QMC_HIP_CHECK(hipGetLastError());
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
        const u64 chunks = (n + kPermPts - 1) / kPermPts;
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
