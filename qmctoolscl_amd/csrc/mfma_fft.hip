// FFT of bit-reversed input on the gfx950 matrix cores (reduced precision):
// the matrix-core counterpart of fft_bro_1d_radix2 (orthonormal, input in
// bit-reversed order, output in natural order).
//
// With the input stored bit-reversed, the Cooley-Tukey split N = N1 * 16
// turns into the same kind of transform at every level: a 16-point DFT over
// each group of 16 consecutive stored values (input bit-reversed inside the
// group), a twiddle, then the same problem of length N/16 over the group
// index.  Level s therefore acts on index bits [4s, 4s+4):
//
//     y = W16 x along those bits,   W16[k][j] = w16^(k * brev4(j))
//     y *= w_{Ns}^(brev(c) * k),    Ns = N / 16^s, c = the index bits above
//
// and the output lands in natural order with no explicit permutation.  A
// 16-point complex DFT of 16 columns is two real 16x16x32 MFMA products,
// [Wr | -Wi] [xr; xi] and [Wi | Wr] [xr; xi]; inputs and the FP64-generated
// DFT tile are each split into three bf16 pieces and the six leading cross
// products are accumulated in fp32 (relative error ~1e-7).  Inter-level
// twiddles are applied in fp32; a leftover radix-2/4/8 top level runs in
// fp32 directly.  Tiles of 2^13 complex values stay in LDS; longer vectors
// take one pass per 8 index bits above the first 12, with an fp32
// intermediate.  The 1/sqrt(N) scale is applied once, in float64.
#include "qmc_common.hpp"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

constexpr int kT = 13;          // 2^13 complex fp32 (68 KiB of LDS with padding)
constexpr int kThreads = 512;  // 8 waves; two workgroups (68 KiB of LDS, <= 128 VGPRs each) share a CU

// LDS word of tile index t: one pad word per 16 keeps the 16-point columns of
// the lowest level (addresses 16 apart) on distinct banks
__device__ __forceinline__ int pad(int t) { return t + (t >> 4); }
static inline size_t padded_bytes(int T) { return 2 * sizeof(float) * (((size_t)1 << T) + ((size_t)1 << T) / 16); }

__device__ __forceinline__ void split3f(float x, __bf16 &a, __bf16 &b, __bf16 &c) {
    a = (__bf16)x;
    const float r1 = x - (float)a;
    b = (__bf16)r1;
    c = (__bf16)(r1 - (float)b);
}

__device__ __forceinline__ void split3d(double x, __bf16 &a, __bf16 &b, __bf16 &c) {
    a = (__bf16)(float)x;
    const double r1 = x - (double)(float)a;
    b = (__bf16)(float)r1;
    c = (__bf16)(float)(r1 - (double)(float)b);
}

__device__ __forceinline__ int brev_bits(int v, int bits) { return bits ? (int)(__brev((unsigned)v) >> (32 - bits)) : 0; }

// A operands for one lane: Ar[p] = piece p of [Wr | -Wi], Ai[p] = piece p of [Wi | Wr]
struct Wtile {
    bf16x8 Ar[3], Ai[3];
};

__device__ Wtile make_w(int lane) {
    const int m = lane & 15, g = lane >> 4;
    Wtile w;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = g * 8 + i;
        const int c = k & 15;
        // w16^(m * brev4(c)) = cos(2 pi e/16) - i sin(2 pi e/16), generated in FP64
        const int e = (m * brev_bits(c, 4)) & 15;
        double s, co;
        sincospi(2.0 * e / 16.0, &s, &co);
        const double wr = co, wi = -s;
        const double vr = k < 16 ? wr : -wi;  // [Wr | -Wi]
        const double vi = k < 16 ? wi : wr;   // [Wi | Wr]
        __bf16 a0, a1, a2;
        split3d(vr, a0, a1, a2);
        w.Ar[0][i] = a0;
        w.Ar[1][i] = a1;
        w.Ar[2][i] = a2;
        split3d(vi, a0, a1, a2);
        w.Ai[0][i] = a0;
        w.Ai[1][i] = a1;
        w.Ai[2][i] = a2;
    }
    return w;
}

struct PassGeom {
    int m;       // log2 vector length
    int lr, lc;  // tile = 2^lr rows x 2^lc columns
    int lo, hi;  // tile rows are global index bits [lo, hi) (lo == hi == lc for a contiguous tile)
};

// Global position (within the vector) of tile index t of column block cb:
// columns are bits [0, lc), rows bits [lo, hi); cb supplies bits [lc, lo)
// then [hi, m).
__device__ __forceinline__ u64 gpos(const PassGeom &G, u64 cb, int t) {
    const int nlo = G.lo - G.lc;
    const u64 cb_lo = cb & ((1ull << nlo) - 1);
    const u64 cb_hi = cb >> nlo;
    return (cb_hi << G.hi) | ((u64)(t >> G.lc) << G.lo) | (cb_lo << G.lc) | (u64)(t & ((1 << G.lc) - 1));
}

// Twiddles w_N^E = exp(-2 pi i E / N), N = 2^m, from two FP64 tables:
// E = Eh 2^klo + El, w = twh[Eh] * twl[El] in FP64, rounded once to fp32.
struct Twiddles {
    const f64x2 *lo, *hi;  // (cos, -sin)
    int klo;
};

__global__ __launch_bounds__(kThreads / 4) void twiddle_tables_kernel(int m, int klo, f64x2 *twl, f64x2 *twh) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const u64 nlo = 1ull << klo, nhi = 1ull << (m - klo);
    if (i >= nlo + nhi) return;
    const u64 E = i < nlo ? i : (i - nlo) << klo;
    double s, c;
    sincospi(-2.0 * (double)E * ldexp(1.0, -m), &s, &c);  // exact angle: E / 2^m
    f64x2 w;
    w.x = c;
    w.y = s;
    if (i < nlo)
        twl[i] = w;
    else
        twh[i - nlo] = w;
}

__device__ __forceinline__ void twiddle(const Twiddles &tw, u64 E, float &c, float &s) {
    const f64x2 a = tw.hi[E >> tw.klo], b = tw.lo[E & ((1ull << tw.klo) - 1)];
    c = (float)(a.x * b.x - a.y * b.y);
    s = (float)(a.x * b.y + a.y * b.x);
}

// radix-16 level on tile bits [ta, ta+4) == global bits [ga, ga+4)
__device__ void dft16_level(float *re, float *im, const PassGeom &G, u64 cb, int T, int ta, int ga,
                            const Wtile &W, int wave, int lane, const Twiddles &tw) {
    const int nblk = 1 << (T - 8);
    const int n = lane & 15, g = lane >> 4;
    const int c0 = (g & 1) * 8;
    float *src = g < 2 ? re : im;
    const int hi_bits = G.m - ga - 4;  // index bits above this level
    for (int blk = wave; blk < nblk; blk += (int)(blockDim.x >> 6)) {
        const int col = blk * 16 + n;
        const int lo = col & ((1 << ta) - 1);
        const int base = ((col >> ta) << (ta + 4)) | lo;
        bf16x8 b[3];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            __bf16 s0, s1, s2;
            split3f(src[pad(base + ((c0 + i) << ta))], s0, s1, s2);
            b[0][i] = s0;
            b[1][i] = s1;
            b[2][i] = s2;
        }
        f32x4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
        // sum over pieces p + q <= 2 of W_p * X_q, smallest terms first
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[2], b[0], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[1], b[1], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[0], b[2], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[1], b[0], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[0], b[1], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[0], b[0], yr, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[2], b[0], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[1], b[1], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[0], b[2], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[1], b[0], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[0], b[1], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[0], b[0], yi, 0, 0, 0);
        // lane holds outputs k = 4g + i of column n; twiddle w_{Ns}^(brev(c_hi) * k)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = 4 * g + i;
            const int t = base + (k << ta);
            float vr = yr[i], vi = yi[i];
            if (hi_bits > 0) {
                const u64 p = gpos(G, cb, t);
                const u64 chi = p >> (ga + 4);
                const u64 e = (__brevll(chi) >> (64 - hi_bits)) * (u64)k;  // < Ns = 2^(hi_bits+4)
                float s, c;
                twiddle(tw, e << (G.m - hi_bits - 4), c, s);  // w_Ns^e = w_N^(e N/Ns)
                const float tr = vr * c - vi * s;
                vi = vr * s + vi * c;
                vr = tr;
            }
            re[pad(t)] = vr;
            im[pad(t)] = vi;
        }
    }
}

// leftover top level of R = 2^rb points (rb < 4) along tile bits [ta, ta+rb)
__device__ void dft_small_level(float *re, float *im, int T, int ta, int rb) {
    const int R = 1 << rb;
    const int groups = 1 << (T - rb);
    for (int gi = threadIdx.x; gi < groups; gi += blockDim.x) {
        const int lo = gi & ((1 << ta) - 1);
        const int base = ((gi >> ta) << (ta + rb)) | lo;
        float xr[8], xi[8];
        for (int j = 0; j < R; j++) {
            xr[j] = re[pad(base + (j << ta))];
            xi[j] = im[pad(base + (j << ta))];
        }
        for (int k = 0; k < R; k++) {
            float ar = 0.f, ai = 0.f;
            for (int j = 0; j < R; j++) {
                const int e = (k * brev_bits(j, rb)) & (R - 1);
                float s, c;
                sincospif(-2.0f * (float)e / (float)R, &s, &c);
                ar += xr[j] * c - xi[j] * s;
                ai += xr[j] * s + xi[j] * c;
            }
            re[pad(base + (k << ta))] = ar;
            im[pad(base + (k << ta))] = ai;
        }
    }
}

// levels on global bits [glo, ghi) of tiles; first/last pass move float64
template <int IN64, int OUT64>
__global__ __launch_bounds__(kThreads) void fft_mfma_pass(PassGeom G, int glo, int ghi, const void *__restrict__ sr,
                                                        const void *__restrict__ si, void *__restrict__ dr,
                                                        void *__restrict__ di, double scale, Twiddles tw) {
    extern __shared__ float lds[];
    const int T = G.lr + G.lc;
    float *re = lds;
    float *im = lds + (1 << T) + (1 << T) / 16;
    const u64 cblocks = (1ull << G.m) >> T;
    const u64 v = blockIdx.x / cblocks;
    const u64 cb = blockIdx.x - v * cblocks;
    const u64 vbase = v << G.m;
    // two consecutive values per access (G.lc >= 1: gpos(t + 1) = gpos(t) + 1 for even t)
#pragma unroll 4
    for (int t = 2 * threadIdx.x; t < (1 << T); t += 2 * blockDim.x) {
        const u64 gi = vbase + gpos(G, cb, t);
        if (IN64) {
            const f64x2 a = *(const f64x2 *)((const double *)sr + gi);
            const f64x2 b = *(const f64x2 *)((const double *)si + gi);
            re[pad(t)] = (float)a.x;
            re[pad(t + 1)] = (float)a.y;
            im[pad(t)] = (float)b.x;
            im[pad(t + 1)] = (float)b.y;
        } else {
            const f32x2 a = *(const f32x2 *)((const float *)sr + gi);
            const f32x2 b = *(const f32x2 *)((const float *)si + gi);
            re[pad(t)] = a.x;
            re[pad(t + 1)] = a.y;
            im[pad(t)] = b.x;
            im[pad(t + 1)] = b.y;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const Wtile W = make_w(lane);
    for (int ga = glo; ga < ghi; ga += 4) {
        // tile bit of global bit ga
        const int ta = ga < G.lc ? ga : G.lc + (ga - G.lo);
        if (ghi - ga >= 4) {
            dft16_level(re, im, G, cb, T, ta, ga, W, wave, lane, tw);
        } else {
            dft_small_level(re, im, T, ta, ghi - ga);  // only the top level can be short
        }
        __syncthreads();
    }
#pragma unroll 4
    for (int t = 2 * threadIdx.x; t < (1 << T); t += 2 * blockDim.x) {
        const u64 gi = vbase + gpos(G, cb, t);
        if (OUT64) {
            f64x2 a, b;
            a.x = (double)re[pad(t)] * scale;
            a.y = (double)re[pad(t + 1)] * scale;
            b.x = (double)im[pad(t)] * scale;
            b.y = (double)im[pad(t + 1)] * scale;
            __builtin_nontemporal_store(a, (f64x2 *)((double *)dr + gi));
            __builtin_nontemporal_store(b, (f64x2 *)((double *)di + gi));
        } else {
            f32x2 a, b;
            a.x = re[pad(t)];
            a.y = re[pad(t + 1)];
            b.x = im[pad(t)];
            b.y = im[pad(t + 1)];
            *(f32x2 *)((float *)dr + gi) = a;
            *(f32x2 *)((float *)di + gi) = b;
        }
    }
}

}  // namespace

extern "C" {

// In-place orthonormal FFT of d1*d2 split-complex vectors of length 2*n_half,
// input bit-reversed, output natural (as fft_bro_1d_radix2), on the matrix
// cores.  Lengths 2^8 .. 2^28.  `work` (2 * d1*d2*2*n_half floats) may be
// NULL to let the library allocate it on the stream.
QMC_EXPORT int qmc_fft_bro_mfma_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *xr, double *xi, float *work,
                                    qmc_stream_t s) {
    const u64 V = d1 * d2;
    if (V == 0 || n_half == 0) return QMC_OK;
    const u64 n = 2 * n_half;
    int m = 0;
    while ((1ull << m) < n) m++;
    if ((1ull << m) != n) return qmc_fail(QMC_ERR_ARG, "fft_mfma: length is not a power of two");
    if (m < 8 || m > 28) return qmc_fail(QMC_ERR_ARG, "fft_mfma: length 2^%d outside 2^8 .. 2^28", m);
    if (((uintptr_t)xr | (uintptr_t)xi | (uintptr_t)work) % 16)
        return qmc_fail(QMC_ERR_ARG, "fft_mfma: xr, xi and work must be 16-byte aligned");
    const double scale = (m % 2 == 0) ? ldexp(1.0, -m / 2) : ldexp(1.0, -(m - 1) / 2) / 1.4142135623730951;
    static bool attr_set = false;
    if (!attr_set) {  // padded tiles need more than the default 64 KiB of LDS
        const int bytes = (int)padded_bytes(kT);
        QMC_HIP_CHECK(hipFuncSetAttribute((const void *)fft_mfma_pass<1, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        QMC_HIP_CHECK(hipFuncSetAttribute((const void *)fft_mfma_pass<1, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        QMC_HIP_CHECK(hipFuncSetAttribute((const void *)fft_mfma_pass<0, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        QMC_HIP_CHECK(hipFuncSetAttribute((const void *)fft_mfma_pass<0, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        attr_set = true;
    }
    const size_t shm = padded_bytes(kT);
    // FP64 twiddle tables (one product per twiddle), 2^klo + 2^(m-klo) entries
    Twiddles tw;
    tw.klo = (m + 1) / 2;
    const u64 ntw = (1ull << tw.klo) + (1ull << (m - tw.klo));
    f64x2 *tt = nullptr;
    QMC_HIP_CHECK(hipMallocAsync((void **)&tt, sizeof(f64x2) * ntw, s));
    tw.lo = tt;
    tw.hi = tt + (1ull << tw.klo);
    hipLaunchKernelGGL(twiddle_tables_kernel, dim3((unsigned)((ntw + kThreads / 4 - 1) / (kThreads / 4))),
                       dim3(kThreads / 4), 0, s, m, tw.klo, tt, tt + (1ull << tw.klo));
    float *buf = nullptr;
    if (m <= kT) {
        PassGeom G{m, 0, m, m, m};
        // one vector per workgroup: half the threads keeps more workgroups per CU
        hipLaunchKernelGGL((fft_mfma_pass<1, 1>), dim3((unsigned)V), dim3(kThreads / 2), padded_bytes(m), s, G, 0, m,
                           (const void *)xr, (const void *)xi, (void *)xr, (void *)xi, scale, tw);
    } else {
        buf = work;
        if (!buf) QMC_HIP_CHECK(hipMallocAsync((void **)&buf, 2 * sizeof(float) * V * n, s));
        float *br = buf, *bi = buf + V * n;
        // pass 1: global bits [0, 12) over contiguous tiles of 2^13
        {
            PassGeom G{m, 0, kT, kT, kT};
            hipLaunchKernelGGL((fft_mfma_pass<1, 0>), dim3((unsigned)(V << (m - kT))), dim3(kThreads), shm, s, G, 0,
                               12, (const void *)xr, (const void *)xi, (void *)br, (void *)bi, 1.0, tw);
        }
        // then 8 bits (two levels) per pass: tiles of 2^8 rows x 2^5 columns
        for (int lo = 12; lo < m; lo += 8) {
            const int hi = lo + 8 < m ? lo + 8 : m;
            const int lr = hi - lo, lc = kT - lr;
            PassGeom G{m, lr, lc, lo, hi};
            const bool last = hi == m;
            if (last)
                hipLaunchKernelGGL((fft_mfma_pass<0, 1>), dim3((unsigned)(V << (m - kT))), dim3(kThreads), shm, s, G,
                                   lo, hi, (const void *)br, (const void *)bi, (void *)xr, (void *)xi, scale, tw);
            else
                hipLaunchKernelGGL((fft_mfma_pass<0, 0>), dim3((unsigned)(V << (m - kT))), dim3(kThreads), shm, s, G,
                                   lo, hi, (const void *)br, (const void *)bi, (void *)br, (void *)bi, 1.0, tw);
        }
    }
    hipError_t e = hipGetLastError();
    if (buf && !work) (void)hipFreeAsync(buf, s);
    (void)hipFreeAsync(tt, s);
    if (e != hipSuccess) return qmc_fail_hip(e, "fft_mfma_pass");
    return QMC_OK;
}

}  // extern "C"
