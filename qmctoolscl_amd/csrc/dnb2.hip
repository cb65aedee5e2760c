// Base-2 digital nets on gfx950: generating-matrix conversion, linear matrix
// scrambling (LMS) over GF(2), point generation in natural and Gray order,
// digital shift, conversion to double, interlacing, and the fused
// generate -> shift -> float pipeline.
//
// A point's value is the XOR of the generating-matrix columns selected by the
// bits of its index (natural order: q = row + n_start; Gray order: Gray(k),
// k = row + n_start).  The reference walks all rows serially
// (dnb2_gen_natural.c:53-70); here each thread walks its own run of rows for
// one (replication, dimension) column:
//   Gray order   : next value = value ^ C[ctz(k+1)]            (one XOR per row)
//   natural order: next value = value ^ C[0] ^ ... ^ C[ctz(q+1)]  (the bits that
//                  flip when q is incremented; two columns per row on average)
// so all runs proceed in parallel and lanes write consecutive dimensions.
#include "qmc_common.hpp"

namespace {

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// XOR of the columns selected by the bits of q
__device__ __forceinline__ u64 xor_cols(const u64 *__restrict__ col, u64 q) {
    u64 v = 0;
    while (q) {
        v ^= col[__ffsll((long long)q) - 1];
        q &= q - 1;
    }
    return v;
}

struct RunMap {
    u64 r, n, d, nd, n_start, runs;
    u32 K;      // rows per run
    int check;  // natural order over a range not known to be closed under Gray
};

// thread -> (l, run, j) with j fastest so lanes write consecutive dimensions
__device__ __forceinline__ bool run_of(const RunMap &m, u64 t, u64 &l, u64 &run, u64 &j) {
    if (t >= m.r * m.runs * m.d) return false;
    const u64 rest = t / m.d;
    j = t - rest * m.d;
    l = rest / m.runs;
    run = rest - l * m.runs;
    return true;
}

// Emit xb (and the fused shift/float) for one run.  OP = 0: plain generation.
// OP = 1: fused with digital shift and conversion to double.
struct FuseArgs {
    u64 r_out;              // replications after the shift (>= r of the net)
    const u64 *lshifts;     // per net replication
    const u64 *shiftsb;     // r_out * d
    const u64 *tmaxes;      // per output replication
    u64 *xrb;               // may be null
    double *x;              // may be null
};

template <bool NATURAL, int OP>
__global__ __launch_bounds__(kBlock) void dnb2_gen_kernel(RunMap m, u64 mmax,
                                                          const u64 *__restrict__ C,
                                                          u64 *__restrict__ xb, FuseArgs f) {
    u64 l, run, j;
    if (!run_of(m, (u64)blockIdx.x * blockDim.x + threadIdx.x, l, run, j)) return;
    const u64 *col = C + (l * m.d + j) * mmax;
    const u64 i0 = run * m.K;
    const u64 i1 = (i0 + m.K < m.n) ? i0 + m.K : m.n;
    u64 lsh = 0;
    if (OP == 1) lsh = f.lshifts[l] & 63;  // x86 masks variable shift counts to 6 bits
    u64 k = i0 + m.n_start;
    u64 v = xor_cols(col, NATURAL ? k : gray_of(k));
    for (u64 i = i0;;) {
        bool write = true;
        if (NATURAL && m.check) {
            const u64 kk = gray_inverse(k);
            write = kk >= m.n_start && kk - m.n_start < m.n;
        }
        if (write) {
            const u64 e2 = i * m.d + j;
            if (OP == 0) {
                xb[l * m.nd + e2] = v;
            } else {
                const u64 sv = v << lsh;
                for (u64 lo = l; lo < f.r_out; lo += m.r) {
                    const u64 out = sv ^ f.shiftsb[lo * m.d + j];
                    if (f.xrb) store_stream(f.xrb + lo * m.nd + e2, out);
                    if (f.x) store_stream(f.x + lo * m.nd + e2, ldexp((double)out, -(int)f.tmaxes[lo]));
                }
            }
        }
        if (++i >= i1) break;
        k++;
        const int c = __ffsll((long long)k) - 1;  // ctz(k)
        if (NATURAL) {
            for (int b = 0; b <= c; b++) v ^= col[b];
        } else {
            v ^= col[c];
        }
    }
}

// dnb2_linear_matrix_scramble -> dnb2_gen_<order> -> dnb2_digital_shift (r_x
// = r) -> dnb2_integer_to_float in one launch.  Block (chunk, l) first builds
// the d * mmax scrambled columns of replication l in LDS (each a GF(2)
// product of S's rows, staged in LDS, and one column of C,
// dnb2_linear_matrix_scramble.c:33-48) and their prefix XORs, then computes
// rows [chunk R, (chunk + 1) R), R = 8 floor(256 / d): thread t walks the 8
// rows of run t / d in column t % d (the first value XORs the columns
// selected by the bits of its index -- Gray order: of the index's Gray code --
// and every later one a prefix XOR (natural order) or one column (Gray
// order), as the index grows by one), the shifted integers meet in LDS, and
// the block writes its R * d consecutive elements of both outputs with
// 16-byte stores.  Natural order over a range that Gray codes do not map onto
// itself writes only the rows the reference's walk reaches (the rest stay
// untouched, as there).
constexpr u32 kLmsRun = 8;  // rows per thread

template <bool NATURAL>
__global__ __launch_bounds__(kBlock) void dnb2_lms_gen_kernel(u64 n, u64 d, u64 n_start, u64 mmax, u64 nC,
                                                              u64 tmax_new, u32 R, int check,
                                                              const u64 *__restrict__ S, const u64 *__restrict__ C,
                                                              const u64 *__restrict__ lshifts,
                                                              const u64 *__restrict__ shiftsb,
                                                              const u64 *__restrict__ tmaxes, u64 *__restrict__ xrb,
                                                              double *__restrict__ x) {
    extern __shared__ u64 lds_u64[];
    u64 *cl = lds_u64;               // [d][mmax] scrambled columns
    u64 *pre = cl + d * mmax;        // [d][mmax] prefix XORs
    u64 *srows = pre + d * mmax;     // [d][tmax_new] rows of S
    u64 *stage = srows + d * tmax_new;  // [R * d] shifted integers of the block's rows
    const u64 l = blockIdx.y;
    // Only columns c < nbits are read below: every index of the call is below 2^nbits, so
    // are its Gray codes and the trailing-zero counts of its increments.  When d cb threads
    // suffice (cb = nbits rounded up to a power of two), thread (j, c) = (tid / cb, tid % cb)
    // builds one column: its load goes out with those of S, and the prefix XORs of a
    // dimension's columns are a scan over cb consecutive lanes of one wave.
    const u64 top = n_start + n - 1;
    const int nbits = top ? 64 - __clzll((long long)top) : 1;
    int cb = 1;
    while (cb < nbits) cb <<= 1;
    const bool fast = (u64)d * (u64)cb <= (u64)blockDim.x;
    const u32 fj = threadIdx.x / (u32)cb, fc = threadIdx.x % (u32)cb;
    const bool act = fast && fj < d && fc < (u32)nbits;
    u64 col0 = 0;
    if (act) col0 = C[(l * d * mmax + fj * mmax + fc) % nC];
    for (u64 e = threadIdx.x; e < d * tmax_new; e += blockDim.x) srows[e] = S[l * d * tmax_new + e];
    __syncthreads();
    if (fast) {
        u64 out = 0;
        if (act) {
            const u64 *srow = srows + fj * tmax_new;
            if (tmax_new <= 32 && (col0 >> 32) == 0) {
                // 32-bit columns: the parity of row t is bit tmax_new - 1 - t, shifted in
                const u32 c32 = (u32)col0;
                u32 o = 0;
                for (u64 t = 0; t < tmax_new; t++) o = (o << 1) | (__popc((u32)srow[t] & c32) & 1u);
                out = o;
            } else {
                for (u64 t = 0; t < tmax_new; t++)
                    if (__popcll(srow[t] & col0) & 1) out += 1ull << ((tmax_new - t - 1) & 63);
            }
        }
        u64 acc = out;
        for (u32 off = 1; off < (u32)cb; off <<= 1) {
            const u64 o = __shfl_up(acc, off, 64);
            if (fc >= off) acc ^= o;
        }
        if (act) {
            cl[fj * mmax + fc] = out;
            pre[fj * mmax + fc] = acc;
        }
    } else {
        for (u64 e = threadIdx.x; e < d * mmax; e += blockDim.x) {
            const u64 j = e / mmax;
            const u64 col = C[(l * d * mmax + e) % nC];
            const u64 *srow = srows + j * tmax_new;
            u64 out = 0;
            for (u64 t = 0; t < tmax_new; t++)
                if (__popcll(srow[t] & col) & 1) out += 1ull << ((tmax_new - t - 1) & 63);
            cl[e] = out;
        }
        __syncthreads();
        for (u64 e = threadIdx.x; e < d * mmax; e += blockDim.x) {
            const u64 j = e / mmax;
            u64 acc = 0;
            for (u64 f = j * mmax; f <= e; f++) acc ^= cl[f];
            pre[e] = acc;
        }
    }
    __syncthreads();
    const u64 i0 = (u64)blockIdx.x * R;
    const u32 rows = (u32)(n - i0 < (u64)R ? n - i0 : (u64)R);
    const u64 lsh = lshifts[l] & 63;  // x86 masks variable shift counts to 6 bits
    const u32 run = threadIdx.x / (u32)d, j = threadIdx.x - run * (u32)d;
    if (run * kLmsRun < rows) {
        const u64 *colj = cl + j * mmax, *prej = pre + j * mmax;
        const u64 sh = shiftsb[l * d + j];
        const u32 r0 = run * kLmsRun, r1 = r0 + kLmsRun < rows ? r0 + kLmsRun : rows;
        u64 k = i0 + r0 + n_start;
        u64 v = xor_cols(colj, NATURAL ? k : gray_of(k));
        for (u32 i = r0;;) {
            stage[(u64)i * d + j] = (v << lsh) ^ sh;
            if (++i >= r1) break;
            k++;
            const int c = __ffsll((long long)k) - 1;  // ctz(k)
            v ^= NATURAL ? prej[c] : colj[c];
        }
    }
    __syncthreads();
    const int tm = (int)tmaxes[l];
    const u64 base = l * n * d + i0 * d, cnt = (u64)rows * d;
    for (u64 e = 2 * (u64)threadIdx.x; e < cnt; e += 2 * (u64)blockDim.x) {
        const bool two = e + 1 < cnt;
        bool w0 = true, w1 = two;
        if (NATURAL && check) {
            const u64 k0 = i0 + e / d + n_start, k1 = i0 + (e + 1) / d + n_start;
            const u64 g0 = gray_inverse(k0), g1 = gray_inverse(k1);
            w0 = g0 >= n_start && g0 - n_start < n;
            w1 = two && g1 >= n_start && g1 - n_start < n;
        }
        const u64 a = stage[e], b = two ? stage[e + 1] : 0;
        if (w0 && w1 && ((base + e) & 1) == 0) {
            if (xrb) {
                u64x2 o;
                o.x = a;
                o.y = b;
                store_stream((u64x2 *)(xrb + base + e), o);
            }
            if (x) {
                f64x2 o;
                o.x = ldexp((double)a, -tm);
                o.y = ldexp((double)b, -tm);
                store_stream((f64x2 *)(x + base + e), o);
            }
        } else {
            if (w0) {
                if (xrb) store_stream(xrb + base + e, a);
                if (x) store_stream(x + base + e, ldexp((double)a, -tm));
            }
            if (w1) {
                if (xrb) store_stream(xrb + base + e + 1, b);
                if (x) store_stream(x + base + e + 1, ldexp((double)b, -tm));
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void lsb_to_msb_kernel(u64 r, u64 dm, const u64 *__restrict__ tmaxes,
                                                            const u64 *__restrict__ src,
                                                            u64 *__restrict__ dst) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= r * dm) return;
    const u64 tmax = tmaxes[e / dm];
    u64 v = src[e], out = 0;
    // dnb2_gmat_lsb_to_msb.c:34-40; the shift count is taken mod 64 as x86 does
    for (u64 t = 0; v != 0; t++, v >>= 1)
        if (v & 1) out += 1ull << ((tmax - t - 1) & 63);
    dst[e] = out;
}

// C_lms[l,j,k] = sum_t parity(S[l,j,t] & C[(l,j,k) mod r_C]) << (tmax_new-1-t);
// one block per (l, j): the rows of S are staged in LDS once and every lane
// takes columns k.
__global__ __launch_bounds__(64) void lms2_kernel(u64 mmax, u64 nC, u64 tmax_new,
                                                  const u64 *__restrict__ S,
                                                  const u64 *__restrict__ C,
                                                  u64 *__restrict__ C_lms) {
    __shared__ u64 rows[64];
    const u64 lj = blockIdx.x;  // l * d + j
    for (u64 t = threadIdx.x; t < tmax_new && t < 64; t += blockDim.x) rows[t] = S[lj * tmax_new + t];
    __syncthreads();
    for (u64 k = threadIdx.x; k < mmax; k += blockDim.x) {
        const u64 e = lj * mmax + k;
        const u64 col = C[e % nC];
        u64 out = 0;
        for (u64 t = 0; t < tmax_new; t++) {
            const u64 row = t < 64 ? rows[t] : S[lj * tmax_new + t];
            if (__popcll(row & col) & 1) out += 1ull << ((tmax_new - t - 1) & 63);
        }
        C_lms[e] = out;
    }
}

__global__ __launch_bounds__(kBlock) void shift2_kernel(u64 r, u64 r_x, u64 nd, u64 d,
                                                        const u64 *__restrict__ lshifts,
                                                        const u64 *__restrict__ xb,
                                                        const u64 *__restrict__ shiftsb,
                                                        u64 *__restrict__ xrb) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e2 = (u64)blockIdx.x * blockDim.x + threadIdx.x; e2 < nd; e2 += stride) {
        const u64 j = ((e2 | d) >> 32) ? e2 % d : (u64)((u32)e2 % (u32)d);
        for (u64 lx = 0; lx < r_x && lx < r; lx++) {
            const u64 v = xb[lx * nd + e2] << (lshifts[lx] & 63);
            for (u64 l = lx; l < r; l += r_x) store_stream(xrb + l * nd + e2, v ^ shiftsb[l * d + j]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void to_float2_kernel(u64 r, u64 nd, const u64 *__restrict__ tmaxes,
                                                           const u64 *__restrict__ xb,
                                                           double *__restrict__ x) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 e2 = (u64)blockIdx.x * blockDim.x + threadIdx.x; e2 < nd; e2 += stride)
        for (u64 l = 0; l < r; l++)
            store_stream(x + l * nd + e2, ldexp((double)xb[l * nd + e2], -(int)tmaxes[l]));
}

// dnb2_interlace.c:26-45 (undo == 0) and dnb2_undo_interlace.c:26-45 (undo == 1)
__global__ __launch_bounds__(kBlock) void interlace2_kernel(int undo, u64 r, u64 d_out, u64 mmax, u64 d_in,
                                                            u64 tmax, u64 tmax_alpha, u64 alpha,
                                                            const u64 *__restrict__ src,
                                                            u64 *__restrict__ dst) {
    const u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= r * d_out * mmax) return;
    const u64 k = e % mmax;
    const u64 lj = e / mmax;
    const u64 jo = lj % d_out;
    const u64 l = lj / d_out;
    u64 out = 0;
    if (!undo) {
        for (u64 ta = 0; ta < tmax_alpha; ta++) {
            const u64 t = ta / alpha;
            const u64 j = jo * alpha + ta % alpha;
            if ((src[(l * d_in + j) * mmax + k] >> ((tmax - t - 1) & 63)) & 1)
                out += 1ull << ((tmax_alpha - ta - 1) & 63);
        }
    } else {
        const u64 ja = jo / alpha;
        const u64 tt = jo % alpha;
        for (u64 t = 0; t < tmax; t++) {
            const u64 ta = t * alpha + tt;
            if ((src[(l * d_in + ja) * mmax + k] >> ((tmax_alpha - ta - 1) & 63)) & 1)
                out += 1ull << ((tmax - t - 1) & 63);
        }
    }
    dst[e] = out;
}

bool pow2_or_zero(u64 v) { return (v & (v - 1)) == 0; }

RunMap make_runs(u64 r, u64 n, u64 d, u64 n_start, int check) {
    RunMap m;
    m.r = r;
    m.n = n;
    m.d = d;
    m.nd = n * d;
    m.n_start = n_start;
    m.check = check;
    // aim for >= 64K threads so every CU has several waves, with runs of 16..256 rows
    u64 K = (r * d * n) / (1u << 16);
    if (K < 16) K = 16;
    if (K > 256) K = 256;
    m.K = (u32)K;
    m.runs = (n + K - 1) / K;
    return m;
}

int check_bits(const char *who, u64 n_start, u64 n, u64 mmax) {
    if (mmax < 64 && (n_start > (1ull << mmax) || n > (1ull << mmax) - n_start))
        return qmc_fail(QMC_ERR_RANGE, "%s: indices up to n_start+n-1 = %llu need more than mmax = %llu "
                        "columns; the reference would read outside the generating matrix", who,
                        (unsigned long long)(n_start + n - 1), (unsigned long long)mmax);
    return QMC_OK;
}

template <int OP>
int launch_gen(bool natural, u64 r, u64 n, u64 d, u64 n_start, u64 mmax, const u64 *C, u64 *xb,
               FuseArgs f, hipStream_t s) {
    const int check = natural && !(pow2_or_zero(n_start) && pow2_or_zero(n_start + n));
    RunMap m = make_runs(r, n, d, n_start, check);
    const u64 threads = r * m.runs * d;
    dim3 grid(grid_for(threads, kBlock, 0x7fffffffu));
    if (natural)
        hipLaunchKernelGGL((dnb2_gen_kernel<true, OP>), grid, dim3(kBlock), 0, s, m, mmax, C, xb, f);
    else
        hipLaunchKernelGGL((dnb2_gen_kernel<false, OP>), grid, dim3(kBlock), 0, s, m, mmax, C, xb, f);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

// one element per thread (up to 2^20 workgroups): full grids write faster than capped grid-stride ones
unsigned ew_grid(u64 work) { return grid_for(work, kBlock); }

}  // namespace

extern "C" {

int dnb2_gmat_lsb_to_msb_dev(u64 r, u64 d, u64 mmax, const u64 *tmaxes, const u64 *C_lsb, u64 *C_msb,
                             hipStream_t s) {
    const u64 total = r * d * mmax;
    if (!total) return QMC_OK;
    hipLaunchKernelGGL(lsb_to_msb_kernel, dim3(grid_for(total, kBlock, 0x7fffffffu)), dim3(kBlock), 0, s,
                       r, d * mmax, tmaxes, C_lsb, C_msb);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int dnb2_linear_matrix_scramble_dev(u64 r, u64 d, u64 mmax, u64 r_C, u64 tmax_new, const u64 *S,
                                    const u64 *C, u64 *C_lms, hipStream_t s) {
    if (!(r && d && mmax)) return QMC_OK;
    if (r_C == 0) return qmc_fail(QMC_ERR_ARG, "dnb2_linear_matrix_scramble: r_C must be positive");
    if (r * d > 0x7fffffffull) return qmc_fail(QMC_ERR_ARG, "dnb2_linear_matrix_scramble: r*d too large");
    hipLaunchKernelGGL(lms2_kernel, dim3((unsigned)(r * d)), dim3(64), 0, s, mmax, r_C * d * mmax, tmax_new, S, C, C_lms);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int dnb2_gen_natural_dev(u64 r, u64 n, u64 d, u64 n_start, u64 mmax, const u64 *C, u64 *xb, hipStream_t s) {
    if (!(r && n && d)) return QMC_OK;
    int rc = check_bits("dnb2_gen_natural", n_start, n, mmax);
    if (rc) return rc;
    return launch_gen<0>(true, r, n, d, n_start, mmax, C, xb, FuseArgs{}, s);
}

int dnb2_gen_gray_dev(u64 r, u64 n, u64 d, u64 n_start, u64 mmax, const u64 *C, u64 *xb, hipStream_t s) {
    if (!(r && n && d)) return QMC_OK;
    int rc = check_bits("dnb2_gen_gray", n_start, n, mmax);
    if (rc) return rc;
    return launch_gen<0>(false, r, n, d, n_start, mmax, C, xb, FuseArgs{}, s);
}

int dnb2_digital_shift_dev(u64 r, u64 n, u64 d, u64 r_x, const u64 *lshifts, const u64 *xb,
                           const u64 *shiftsb, u64 *xrb, hipStream_t s) {
    if (!(r && n && d)) return QMC_OK;
    if (r_x == 0) return qmc_fail(QMC_ERR_ARG, "dnb2_digital_shift: r_x must be positive");
    hipLaunchKernelGGL(shift2_kernel, dim3(ew_grid(n * d)), dim3(kBlock), 0, s, r, r_x, n * d, d, lshifts, xb,
                       shiftsb, xrb);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int dnb2_integer_to_float_dev(u64 r, u64 n, u64 d, const u64 *tmaxes, const u64 *xb, double *x,
                              hipStream_t s) {
    if (!(r && n && d)) return QMC_OK;
    hipLaunchKernelGGL(to_float2_kernel, dim3(ew_grid(n * d)), dim3(kBlock), 0, s, r, n * d, tmaxes, xb, x);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int dnb2_interlace_dev(u64 r, u64 d_alpha, u64 mmax, u64 d, u64 tmax, u64 tmax_alpha, u64 alpha,
                       const u64 *C, u64 *C_alpha, hipStream_t s) {
    const u64 total = r * d_alpha * mmax;
    if (!total) return QMC_OK;
    if (alpha == 0) return qmc_fail(QMC_ERR_ARG, "dnb2_interlace: alpha must be positive");
    if (d_alpha * alpha > d)
        return qmc_fail(QMC_ERR_RANGE, "dnb2_interlace: d_alpha*alpha <= d is required");
    hipLaunchKernelGGL(interlace2_kernel, dim3(grid_for(total, kBlock, 0x7fffffffu)), dim3(kBlock), 0, s, 0, r,
                       d_alpha, mmax, d, tmax, tmax_alpha, alpha, C, C_alpha);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int dnb2_undo_interlace_dev(u64 r, u64 d, u64 mmax, u64 d_alpha, u64 tmax, u64 tmax_alpha, u64 alpha,
                            const u64 *C_alpha, u64 *C, hipStream_t s) {
    const u64 total = r * d * mmax;
    if (!total) return QMC_OK;
    if (alpha == 0) return qmc_fail(QMC_ERR_ARG, "dnb2_undo_interlace: alpha must be positive");
    if ((d + alpha - 1) / alpha > d_alpha)
        return qmc_fail(QMC_ERR_RANGE, "dnb2_undo_interlace: ceil(d/alpha) <= d_alpha is required");
    hipLaunchKernelGGL(interlace2_kernel, dim3(grid_for(total, kBlock, 0x7fffffffu)), dim3(kBlock), 0, s, 1, r,
                       d, mmax, d_alpha, tmax, tmax_alpha, alpha, C_alpha, C);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

// Fused dnb2_gen_<order>(r_x nets) -> dnb2_digital_shift(r, r_x) ->
// dnb2_integer_to_float(r): one pass that writes the shifted integers (xrb,
// optional) and/or their doubles (x, optional).  natural: 1 natural, 0 Gray.
int qmc_dnb2_gen_shift_float_dev(int natural, u64 r, u64 n, u64 d, u64 r_x, u64 n_start, u64 mmax,
                                 const u64 *C, const u64 *lshifts, const u64 *shiftsb,
                                 const u64 *tmaxes, u64 *xrb, double *x, hipStream_t s) {
    if (!(r && n && d)) return QMC_OK;
    if (r_x == 0) return qmc_fail(QMC_ERR_ARG, "dnb2_gen_shift_float: r_x must be positive");
    int rc = check_bits("dnb2_gen_shift_float", n_start, n, mmax);
    if (rc) return rc;
    const u64 rx = r_x < r ? r_x : r;
    FuseArgs f{r, lshifts, shiftsb, tmaxes, xrb, x};
    return launch_gen<1>(natural != 0, rx, n, d, n_start, mmax, C, nullptr, f, s);
}

// Fused dnb2_linear_matrix_scramble(r, d, mmax, r_C, tmax_new, S, C) ->
// dnb2_gen_<order>(r) -> dnb2_digital_shift(r, r_x = r) ->
// dnb2_integer_to_float(r) in one launch, writing xrb and/or x.
int qmc_dnb2_lms_gen_shift_float_dev(int natural, u64 r, u64 n, u64 d, u64 r_C, u64 n_start, u64 mmax,
                                     u64 tmax_new, const u64 *S, const u64 *C, const u64 *lshifts,
                                     const u64 *shiftsb, const u64 *tmaxes, u64 *xrb, double *x, hipStream_t s) {
    const char *who = "dnb2_lms_gen_shift_float";
    if (!(r && n && d)) return QMC_OK;
    if (r_C == 0) return qmc_fail(QMC_ERR_ARG, "%s: r_C must be positive", who);
    if (mmax == 0 || mmax > 64 || tmax_new == 0 || tmax_new > 64)
        return qmc_fail(QMC_ERR_ARG, "%s: needs 1 <= mmax <= 64 and 1 <= tmax_new <= 64", who);
    if (d > (u64)kBlock || d * (2 * mmax + tmax_new) > 6144)
        return qmc_fail(QMC_ERR_ARG, "%s: needs d <= %d and d * (2 mmax + tmax_new) <= 6144 (LDS)", who, kBlock);
    if (r > 65535) return qmc_fail(QMC_ERR_ARG, "%s: r above 65535 is not supported", who);
    int rc = check_bits(who, n_start, n, mmax);
    if (rc) return rc;
    const int check = natural && !(pow2_or_zero(n_start) && pow2_or_zero(n_start + n));
    const u32 R = kLmsRun * (u32)(kBlock / d);  // rows per block: one run of kLmsRun rows per thread
    const u64 chunks = (n + R - 1) / R;
    if (chunks > 0x7fffffffull) return qmc_fail(QMC_ERR_ARG, "%s: n too large", who);
    const size_t shm = sizeof(u64) * (d * (2 * mmax + tmax_new) + (u64)R * d);
    const dim3 grid((unsigned)chunks, (unsigned)r);
    if (natural)
        hipLaunchKernelGGL(dnb2_lms_gen_kernel<true>, grid, dim3(kBlock), shm, s, n, d, n_start, mmax, r_C * d * mmax,
                           tmax_new, R, check, S, C, lshifts, shiftsb, tmaxes, xrb, x);
    else
        hipLaunchKernelGGL(dnb2_lms_gen_kernel<false>, grid, dim3(kBlock), shm, s, n, d, n_start, mmax,
                           r_C * d * mmax, tmax_new, R, check, S, C, lshifts, shiftsb, tmaxes, xrb, x);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

}  // extern "C"
