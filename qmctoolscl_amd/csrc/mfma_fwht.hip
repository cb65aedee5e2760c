// Walsh-Hadamard transform on the gfx950 matrix cores (reduced precision).
//
// H_N = H16 (x) H16 (x) ... (x) H_{2^rem}: the transform of a length-2^m vector
// is a 16-point Hadamard product along every 4-bit group of the index (and
// radix-2 butterflies for the leftover bits).  A 16-point product over 16
// columns is one 16x16 output tile of four v_mfma_f32_16x16x4f32:
//
//     D = sum over kb < 4 of H16[:, 4kb..4kb+3] * x[4kb..4kb+3, :]
//
// with the fp32 values as they are (the +-1 entries are exact), so the only
// rounding is the fp32 accumulation.  (A 3 x bf16 split on the bf16 matrix
// cores needs about 100 VALU instructions per tile for the split and ran the
// first pass at 2.9 TB/s, VALU-bound; four fp32 MFMAs take 4 LDS reads and 4
// writes per lane and leave the passes memory-bound.)  Values stay in LDS
// (fp32) between axes; a vector longer than one tile (2^14 values) takes two
// passes over HBM with an fp32 intermediate, against up to five float64
// passes of the radix-2 path.  The orthonormal scale 1/sqrt(n) is applied
// once, in float64, when the result is written.
//
// Accuracy is not bit-exact: every result reports its relative L2 and max
// error against the float64 transform (see tests/test_mfma.py).
#include "qmc_common.hpp"

#include <algorithm>
#include <type_traits>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

constexpr int kTileLog2 = 14;  // 2^14 fp32 (68 KiB of LDS with padding) per workgroup
constexpr int kThreads = 1024;  // 16 waves: two workgroups fill a CU, each with many loads in flight

// LDS word of tile index t: one pad word per 16 keeps the 16-point columns of
// the lowest axis (addresses 16 apart) on distinct banks
__device__ __forceinline__ int pad(int t) { return t + (t >> 4); }
static inline size_t padded_bytes(int T) { return sizeof(float) * (((size_t)1 << T) + ((size_t)1 << T) / 16); }

// 16-point Hadamard along tile index bits [a, a+4) of a 2^T-value fp32 tile
// in LDS.  Each wave owns whole 16-column blocks, so reads and writes of a
// column never cross waves.  v_mfma_f32_16x16x4f32 operands: lane (n, g) =
// (lane % 16, lane / 16) holds A[n][g] and B[g][n] of its k-block and gets
// D[4g + i][n], i < 4.  The sign dither (one +-1 per column, undone on the
// way out) keeps the accumulation's rounding bias from adding up over axes.
// DITHER false: a tile that holds a whole vector (2^14 values at most) showed no
// excess rounding bias without the dither (relative L2 and max error unchanged at
// 2^10 .. 2^14 for normal, smooth and constant inputs), so those kernels skip it.
template <bool DITHER = true>
__device__ __forceinline__ void h16_axis_body(float *tile, int T, int a, int wave, int nwave, int lane, u32 salt) {
    const int nblk = 1 << (T - 8);
    const int n = lane & 15, g = lane >> 4;
    float A[4];  // H16[n][4kb + g]
#pragma unroll
    for (int kb = 0; kb < 4; kb++) A[kb] = (__popc(n & (4 * kb + g)) & 1) ? -1.0f : 1.0f;
    for (int blk = wave; blk < nblk; blk += nwave) {
        const int col = blk * 16 + n;
        const int lo = col & ((1 << a) - 1);
        const int base = ((col >> a) << (a + 4)) | lo;
        const u32 sg = DITHER ? dither_sign((u32)col, salt) : 0u;
        float b[4];
#pragma unroll
        for (int kb = 0; kb < 4; kb++) b[kb] = flip(tile[pad(base + ((4 * kb + g) << a))], sg);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; kb++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[kb], b[kb], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) tile[pad(base + ((g * 4 + i) << a))] = flip(acc[i], sg);
    }
}

// the same with the wave count of the block, for the passes whose geometry is known at run time only
__device__ void h16_axis(float *tile, int T, int a, int wave, int lane, u32 salt) {
    h16_axis_body(tile, T, a, wave, (int)(blockDim.x >> 6), lane, salt);
}

// unnormalized radix-2 butterflies along tile bits [a, a+k), k < 4
__device__ void h2_axes(float *tile, int T, int a, int k) {
    for (int s = a; s < a + k; s++) {
        const int h = 1 << s;
        for (int b = threadIdx.x; b < (1 << (T - 1)); b += blockDim.x) {
            const int p = ((b >> s) << (s + 1)) | (b & (h - 1));
            const float x1 = tile[pad(p)], x2 = tile[pad(p + h)];
            tile[pad(p)] = x1 + x2;
            tile[pad(p + h)] = x1 - x2;
        }
        __syncthreads();
    }
}

// One pass over tiles of 2^lr rows x 2^lc columns of one vector (global
// index = v*2^m + row*row_stride + cb*2^lc + col), applying the Hadamard
// product along tile bits [tb0, tb0 + nb).  IN/OUT: 0 fp32, 1 float64 (the
// output is then scaled by `scale`).  A workgroup walks tiles blockIdx.x,
// +gridDim.x, ..., with all NP value pairs of a thread (2^T = 2 NP
// blockDim.x) loaded at once, the next tile's loads issued as soon as the
// current tile is staged in LDS.
// TT > 0: a whole vector of 2^TT values per tile (lr = 0, lc = tb0 + nb = TT), the
// length at compile time so the axes unroll with their strides fixed.
template <int IN64, int OUT64, int NP, int TT = 0>
__global__ __launch_bounds__(kThreads) void fwht_mfma_pass(int m, int lr, int lc, u64 row_stride, int tb0, int nb,
                                                         const void *__restrict__ src, void *__restrict__ dst,
                                                         double scale, u64 ntiles) {
    typedef typename std::conditional<IN64 != 0, f64x2, f32x2>::type In2;
    extern __shared__ float tile[];
    if (TT) {
        m = TT;
        lr = 0;
        lc = TT;
        tb0 = 0;
        nb = TT;
    }
    const int T = lr + lc;
    const int nwave = TT ? (1 << TT) / (128 * NP) : 0;  // waves of a whole-vector block
    const int tlog = m - T;  // log2 of tiles per vector
    const u64 cmask_tile = (1ull << tlog) - 1;
    const int cmask = (1 << lc) - 1;
    // value pair p of this thread sits at tile index t = 2 tid + 2 p blockDim.x; as 2 blockDim.x and 2^lc
    // are powers of two, its global index is that of pair 0 plus a per-p constant (no carries)
    auto gidx = [&](u64 ti, int tid) -> u64 {
        const int t = 2 * tid;
        return ((ti >> tlog) << m) + (u64)(t >> lc) * row_stride + ((ti & cmask_tile) << lc) + (u64)(t & cmask);
    };
    auto poff = [&](int p) -> u64 {
        const u32 q = 2u * (u32)p * blockDim.x;
        return (u64)(q >> lc) * row_stride + (q & (u32)cmask);
    };
    In2 r[NP];
    auto load = [&](u64 ti) {
        const In2 *s0 = (const In2 *)src + gidx(ti, opaque(threadIdx.x)) / 2;
#pragma unroll
        for (int p = 0; p < NP; p++) {
            // the caller's float64 input streams through; the fp32 intermediate is read with ordinary
            // loads, so the L2 keeps whole lines for the neighbouring tile when a tile row is narrower
            // than a line (2^24: 15% faster)
            if (IN64)
                r[p] = __builtin_nontemporal_load(s0 + poff(p) / 2);
            else
                r[p] = s0[poff(p) / 2];
        }
    };
    if (blockIdx.x < ntiles) load(blockIdx.x);
    for (u64 tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        {
            const int tid = opaque(threadIdx.x);
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const int t = 2 * (tid + p * blockDim.x);
                tile[pad(t)] = (float)r[p].x;
                tile[pad(t + 1)] = (float)r[p].y;
            }
        }
        __syncthreads();
        // the next tile's loads are in flight while this one is transformed
        if (tix + gridDim.x < ntiles) load(tix + gridDim.x);
        int a = tb0;
        if constexpr (TT != 0) {
#pragma unroll
            for (int aa = 0; aa + 4 <= TT; aa += 4) {
                const int ltid = opaque(threadIdx.x);
                h16_axis_body<false>(tile, T, aa, ltid >> 6, nwave, ltid & 63, 0u);
                __syncthreads();
            }
            a = TT & ~3;
        } else {
            for (; a + 4 <= tb0 + nb; a += 4) {
                const int ltid = opaque(threadIdx.x);
                h16_axis(tile, T, a, ltid >> 6, ltid & 63, (u32)tix * 0x632BE5ABu + (u32)a * 0x27D4EB2Fu);
                __syncthreads();
            }
        }
        if (a < tb0 + nb) h2_axes(tile, T, a, tb0 + nb - a);
        const int stid = opaque(threadIdx.x);
        const u64 g0 = gidx(tix, stid);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int t = 2 * (stid + p * blockDim.x);
            const float v0 = tile[pad(t)], v1 = tile[pad(t + 1)];
            const u64 g = g0 + poff(p);
            if (OUT64) {
                f64x2 w;
                w.x = (double)v0 * scale;
                w.y = (double)v1 * scale;
                __builtin_nontemporal_store(w, (f64x2 *)((double *)dst + g));
            } else {
                f32x2 w;
                w.x = v0;
                w.y = v1;
                *(f32x2 *)((float *)dst + g) = w;
            }
        }
        __syncthreads();  // every LDS read of this tile is done before the next one is staged
    }
}

__global__ void mfma_probe_kernel(const float *a, const float *b, float *d) {
    const int l = threadIdx.x;
    bf16x8 av, bv;
    for (int i = 0; i < 8; i++) {
        av[i] = (__bf16)a[l * 8 + i];
        bv[i] = (__bf16)b[l * 8 + i];
    }
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
    for (int i = 0; i < 4; i++) d[l * 4 + i] = c[i];
}

}  // namespace

namespace {
// the whole-vector kernel for length 2^m (m <= kTileLog2): 8 value pairs per thread and the
// length at compile time from 2^10 up, fewer pairs below
const void *whole_fn(int m) {
    switch (m) {
    case 8: return (const void *)fwht_mfma_pass<1, 1, 2>;
    case 9: return (const void *)fwht_mfma_pass<1, 1, 4>;
    case 10: return (const void *)fwht_mfma_pass<1, 1, 8, 10>;
    case 11: return (const void *)fwht_mfma_pass<1, 1, 8, 11>;
    case 12: return (const void *)fwht_mfma_pass<1, 1, 8, 12>;
    case 13: return (const void *)fwht_mfma_pass<1, 1, 8, 13>;
    default: return (const void *)fwht_mfma_pass<1, 1, 8, 14>;
    }
}
}  // namespace

extern "C" {

// Orthonormal FWHT of d1*d2 vectors of length 2*n_half (in place, float64 in
// and out) on the matrix cores.  Needs 2*n_half >= 256.  `work` is an fp32
// scratch of d1*d2*2*n_half values, used when the length exceeds 2^14; pass
// NULL to let the library allocate it on the stream.
QMC_EXPORT int qmc_fwht_mfma_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *x, float *work, qmc_stream_t s) {
    const u64 V = d1 * d2;
    if (V == 0 || n_half == 0) return QMC_OK;
    const u64 n = 2 * n_half;
    int m = 0;
    while ((1ull << m) < n) m++;
    if ((1ull << m) != n) return qmc_fail(QMC_ERR_ARG, "fwht_mfma: length is not a power of two");
    if (m < 8) return qmc_fail(QMC_ERR_ARG, "fwht_mfma: length 2^%d is below the 2^8 minimum", m);
    if (m > 2 * kTileLog2 - 4) return qmc_fail(QMC_ERR_ARG, "fwht_mfma: length 2^%d is above 2^%d", m, 2 * kTileLog2 - 4);
    if (((uintptr_t)x | (uintptr_t)work) % 16)
        return qmc_fail(QMC_ERR_ARG, "fwht_mfma: x and work must be 16-byte aligned");
    // 1/sqrt(2^m), exact for even m
    const double scale = (m % 2 == 0) ? ldexp(1.0, -m / 2) : ldexp(1.0, -(m - 1) / 2) / 1.4142135623730951;
    static PerDeviceOnce attrs;  // tiles of 2^14 padded values need more than the default 64 KiB of LDS
    const int arc = attrs.run([]() -> int {
        const int bytes = (int)padded_bytes(kTileLog2);
        QMC_HIP_CHECK(hipFuncSetAttribute(whole_fn(kTileLog2), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        QMC_HIP_CHECK(hipFuncSetAttribute((const void *)fwht_mfma_pass<1, 0, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        QMC_HIP_CHECK(hipFuncSetAttribute((const void *)fwht_mfma_pass<0, 1, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        return QMC_OK;
    });
    if (arc) return arc;
    if (m <= kTileLog2) {
        // one vector per tile, 16 values per thread (m >= 10), fewer below
        const int np = m >= 10 ? 8 : 1 << (m - 7);
        const unsigned threads = (1u << m) / (2u * np);
        const size_t shm = padded_bytes(m);
        // resident workgroups looping over vectors, each loading its next vector while it transforms
        // the current one (2^12: 1-2% faster than one workgroup per vector; half as many resident
        // workgroups: 10% slower)
        const void *fn = whole_fn(m);
        const unsigned grid =
            (unsigned)std::min<u64>(V, (u64)resident_per_cu(fn, (int)threads, shm) * (u64)device_cus());
        int a_m = m, a_lr = 0, a_lc = m, a_tb0 = 0, a_nb = m;
        u64 a_rs = 0, a_nt = V;
        const void *a_src = x;
        void *a_dst = x;
        double a_scale = scale;
        void *args[] = {&a_m, &a_lr, &a_lc, &a_rs, &a_tb0, &a_nb, &a_src, &a_dst, &a_scale, &a_nt};
        QMC_HIP_CHECK(hipLaunchKernel(fn, dim3(grid), dim3(threads), args, shm, s));
        QMC_HIP_CHECK(hipGetLastError());
        return QMC_OK;
    }
    float *buf = work;
    if (!buf) QMC_HIP_CHECK(hipMallocAsync((void **)&buf, sizeof(float) * V * n, s));
    // pass 1: the low m1 bits over contiguous tiles of 2^m1 values, 8 value pairs per thread; pass 2:
    // the high m2 = m - m1 bits over tiles of 2^m2 rows (stride 2^m1) x 2^(14 - m2) columns, 1024
    // threads.  One workgroup per tile (measured faster here than resident workgroups looping over
    // tiles).  m1 = 13 (four 34 KiB workgroups per CU in pass 1, 4% faster than 14 at 2^20) unless
    // pass 2 would then get rows of fewer than 32 floats.
    int m1 = std::max(13, m - 9);
    if (m1 > kTileLog2) m1 = kTileLog2;
    const int m2 = m - m1;
    const u64 ntiles1 = V << (m - m1), ntiles2 = V << (m - kTileLog2);
    const unsigned grid1 = (unsigned)std::min<u64>(ntiles1, 1u << 30);
    const unsigned grid2 = (unsigned)std::min<u64>(ntiles2, 1u << 30);
    hipLaunchKernelGGL((fwht_mfma_pass<1, 0, 8>), dim3(grid1), dim3((1u << m1) / 16), padded_bytes(m1), s, m, 0, m1,
                       (u64)0, 0, m1, (const void *)x, (void *)buf, 1.0, ntiles1);
    const int lc = kTileLog2 - m2;
    hipLaunchKernelGGL((fwht_mfma_pass<0, 1, 8>), dim3(grid2), dim3(kThreads), padded_bytes(kTileLog2), s, m, m2, lc,
                       (u64)1 << m1, lc, m2, (const void *)buf, (void *)x, scale, ntiles2);
    hipError_t e = hipGetLastError();
    if (!work) (void)hipFreeAsync(buf, s);
    if (e != hipSuccess) return qmc_fail_hip(e, "fwht_mfma_pass");
    return QMC_OK;
}

// D (16x16, lane-major: lane l holds 4 values) = A (16x32) * B (32x16) on
// one wave, with a, b given lane-major (8 values per lane, converted to
// bf16): lets a test check the operand layouts the kernels rely on.
QMC_EXPORT int qmc_mfma_probe_dev(const float *a, const float *b, float *d, qmc_stream_t s) {
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, s, a, b, d);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

}  // extern "C"
