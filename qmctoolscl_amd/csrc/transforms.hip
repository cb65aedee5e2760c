// FP64 radix-2 fast transforms on gfx950: the in-place orthonormal
// Walsh-Hadamard transform, the decimation-in-time FFT on bit-reversed input,
// and the decimation-in-frequency inverse FFT producing bit-reversed output
// (qmctoolscl/c_funcs/fwht_1d_radix2.c, fft_bro_1d_radix2.c,
// ifft_bro_1d_radix2.c).
//
// Every butterfly performs the reference's operations in the reference's
// order -- (a + b)/sqrt(2) with IEEE division, products then sums with no FMA
// contraction -- and every stage is applied in the reference's stage order.
// Butterflies inside one stage are independent, so grouping stages into
// passes changes nothing in the arithmetic: the FWHT is bit-identical to the
// reference, and the FFT/IFFT differ only through the twiddle table (device
// cos/sin of the same double angle, within an ulp of the host libm).
//
// Passes: the stages whose stride is below the tile (4096 values, 32 KiB of
// LDS per component) run in one LDS-resident pass over contiguous tiles, in
// register rounds of four stages; the stages with larger strides run in
// register passes of up to four stages straight from HBM, each lane holding
// 16 elements spaced by the smallest stride of the group, so lanes read
// consecutive addresses.  Division by sqrt(2) is a product and two fma
// corrections that round exactly as the division does.
#include "qmc_common.hpp"

namespace {

constexpr double kSqrt2 = 1.4142135623730951;      // sqrt(2.0), correctly rounded
constexpr double kRcpSqrt2 = 0x1.6a09e667f3bccp-1;  // 1.0 / kSqrt2, correctly rounded
constexpr double kPi = 3.141592653589793;           // acos(-1.0)
constexpr int kTileLog2Real = 12;
constexpr int kTileLog2Cplx = 12;
constexpr int kGroupLog2 = 4;

enum { KIND_FWHT = 0, KIND_FFT = 1, KIND_IFFT = 2 };

// x / sqrt(2) rounded to nearest, as IEEE division rounds it, from a product
// and two Markstein corrections with y = RN(1/s), s = RN(sqrt 2):
//   q0 = RN(x y) is within 1.21 ulp of x/s, so r0 = x - q0 s is exact (fma);
//   q1 = RN(q0 + r0 y) is within 0.5 + 2^-53 ulp, r1 = x - q1 s exact;
//   q2 = RN(q1 + r1 y) = RN(x/s + r1 (y - 1/s)) with |r1 (y - 1/s)| below
//   2^-54.5 ulp, while x/s is never closer than 2^-53.5 ulp to a midpoint m
//   (x - m s is a nonzero multiple of 2^(e-105): s has an odd 53-bit
//   significand, m an odd 54-bit one), so q2 = RN(x/s).
// Zeros, infinities, NaN and |x| outside [2^-960, 2^1000] take the division.
__device__ __forceinline__ double div_sqrt2(double x) {
    const double ax = fabs(x);
    if (__builtin_expect(!(ax >= 0x1p-960 && ax <= 0x1p1000), 0)) return x / kSqrt2;
    const double q0 = x * kRcpSqrt2;
    const double r0 = __fma_rn(-q0, kSqrt2, x);
    const double q1 = __fma_rn(r0, kRcpSqrt2, q0);
    const double r1 = __fma_rn(-q1, kSqrt2, x);
    return __fma_rn(r1, kRcpSqrt2, q1);
}

// fwht_1d_radix2.c:44-47
__device__ __forceinline__ void bfly_fwht(double &a, double &b) {
    const double x1 = a, x2 = b;
    a = div_sqrt2(x1 + x2);
    b = div_sqrt2(x1 - x2);
}

// fft_bro_1d_radix2.c:66-75
__device__ __forceinline__ void bfly_dit(double &ar, double &ai, double &br, double &bi, double c, double s) {
    const double yr = br * c - bi * s;
    const double yi = br * s + bi * c;
    const double xr1 = ar, xi1 = ai;
    ar = div_sqrt2(xr1 + yr);
    ai = div_sqrt2(xi1 + yi);
    br = div_sqrt2(xr1 - yr);
    bi = div_sqrt2(xi1 - yi);
}

// ifft_bro_1d_radix2.c:66-75
__device__ __forceinline__ void bfly_dif(double &ar, double &ai, double &br, double &bi, double c, double s) {
    const double yr = ar - br, yi = ai - bi;
    ar = div_sqrt2(ar + br);
    ai = div_sqrt2(ai + bi);
    br = div_sqrt2(yr * c - yi * s);
    bi = div_sqrt2(yr * s + yi * c);
}

// twiddles as fft_bro_1d_radix2.c:32-42: angle = (sign*pi*i)/n, cos and sin
__global__ __launch_bounds__(kBlock) void twiddle_kernel(u64 n, double sign, double *__restrict__ twr,
                                                         double *__restrict__ twi) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ang = sign * kPi * (double)i / (double)n;
    double s, c;
    sincos(ang, &s, &c);
    twr[i] = c;
    twi[i] = s;
}

// The butterflies of CNT consecutive stages on 2^CNT values held in
// registers: value c sits at index c << lo of a group whose other index bits
// are fixed; stage bit order ascending (FFT) or descending (FWHT, IFFT).
// i0 = the group's index within its vector; stage-major twiddles
// ts[2^s + t] = tw[t << (m-1-s)].
template <int KIND, bool ASC, int CNT>
__device__ __forceinline__ void run_stages(double *re, double *im, u64 i0, int lo, const double *__restrict__ tsr,
                                           const double *__restrict__ tsi) {
#pragma unroll
    for (int st = 0; st < CNT; st++) {
        const int bit = ASC ? st : CNT - 1 - st;
        const int ss = lo + bit;  // stride 2^ss
#pragma unroll
        for (int c1 = 0; c1 < (1 << CNT); c1++) {
            if ((c1 >> bit) & 1) continue;
            const int c2 = c1 | (1 << bit);
            if (KIND == KIND_FWHT) {
                bfly_fwht(re[c1], re[c2]);
            } else {
                const u64 i1 = i0 + ((u64)c1 << lo);
                const u64 tw = (1ull << ss) + (i1 & ((1ull << ss) - 1));
                const double cc = tsr[tw], sn = tsi[tw];
                if (KIND == KIND_FFT)
                    bfly_dit(re[c1], im[c1], re[c2], im[c2], cc, sn);
                else
                    bfly_dif(re[c1], im[c1], re[c2], im[c2], cc, sn);
            }
        }
    }
}

// one register round of the LDS tile: stages on index bits lo .. lo+CNT-1
template <int KIND, bool ASC, int CNT>
__device__ __forceinline__ void tile_round(double *sr, double *si, u64 len, int lo, u64 vbase,
                                           const double *__restrict__ tsr, const double *__restrict__ tsi) {
    auto P = [](u64 t) { return t + (t >> 4); };  // one pad double per 16 values
    const u64 groups = len >> CNT;
    for (u64 g = threadIdx.x; g < groups; g += blockDim.x) {
        const u64 gb = ((g >> lo) << (lo + CNT)) | (g & ((1ull << lo) - 1));
        double re[1 << CNT], im[1 << CNT];
#pragma unroll
        for (int c = 0; c < (1 << CNT); c++) {
            re[c] = sr[P(gb + ((u64)c << lo))];
            if (KIND != KIND_FWHT) im[c] = si[P(gb + ((u64)c << lo))];
        }
        run_stages<KIND, ASC, CNT>(re, im, vbase + gb, lo, tsr, tsi);
#pragma unroll
        for (int c = 0; c < (1 << CNT); c++) {
            sr[P(gb + ((u64)c << lo))] = re[c];
            if (KIND != KIND_FWHT) si[P(gb + ((u64)c << lo))] = im[c];
        }
    }
}

// The nst stages with strides 2^0 .. 2^(nst-1) (all below the tile), over
// contiguous tiles of 2^tlog2 values of the flat array (tiles hold whole
// vectors or aligned parts of one vector), in rounds of up to four stages:
// one LDS round trip and one barrier per round.
// complex tiles (64 KiB of LDS, two per CU) use 512 threads: one 8-value group per lane per round
template <int KIND>
constexpr int tile_threads() { return KIND == KIND_FWHT ? kBlock : 2 * kBlock; }

template <int KIND, bool ASC>
__global__ __launch_bounds__(tile_threads<KIND>()) void tile_kernel(u64 total, int m, int tlog2, int nst, double *__restrict__ xr,
                                                      double *__restrict__ xi, const double *__restrict__ tsr,
                                                      const double *__restrict__ tsi) {
    extern __shared__ double lds[];
    const u64 tile = 1ull << tlog2;
    const u64 base = (u64)blockIdx.x << tlog2;
    const u64 len = (total - base) < tile ? (total - base) : tile;
    double *sr = lds;
    double *si = lds + tile + tile / 16;
    auto P = [](u64 t) { return t + (t >> 4); };
    for (u64 t = threadIdx.x; t < len; t += blockDim.x) {
        sr[P(t)] = xr[base + t];
        if (KIND != KIND_FWHT) si[P(t)] = xi[base + t];
    }
    __syncthreads();
    const u64 vbase = base & ((1ull << m) - 1);  // the tile's offset within its vector
    // complex rounds take three stages (8 values in 32 registers), real ones four
    constexpr int kRound = KIND == KIND_FWHT ? kGroupLog2 : 3;
    for (int done = 0; done < nst;) {
        const int cnt = (nst - done) < kRound ? (nst - done) : kRound;
        const int lo = ASC ? done : nst - done - cnt;
        switch (cnt) {
            case 4: tile_round<KIND, ASC, 4>(sr, si, len, lo, vbase, tsr, tsi); break;
            case 3: tile_round<KIND, ASC, 3>(sr, si, len, lo, vbase, tsr, tsi); break;
            case 2: tile_round<KIND, ASC, 2>(sr, si, len, lo, vbase, tsr, tsi); break;
            default: tile_round<KIND, ASC, 1>(sr, si, len, lo, vbase, tsr, tsi); break;
        }
        __syncthreads();
        done += cnt;
    }
    for (u64 t = threadIdx.x; t < len; t += blockDim.x) {
        xr[base + t] = sr[P(t)];
        if (KIND != KIND_FWHT) xi[base + t] = si[P(t)];
    }
}

// Stages on index bits a .. a+R-1, straight from HBM: one lane owns the 2^R
// values that differ only in those bits (lanes with consecutive indices read
// consecutive addresses).
template <int KIND, bool ASC, int R>
__global__ __launch_bounds__(kBlock) void group_kernel(u64 groups, int m, int a, double *__restrict__ xr,
                                                       double *__restrict__ xi, const double *__restrict__ tsr,
                                                       const double *__restrict__ tsi) {
    const u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) return;
    const u64 L = g & ((1ull << a) - 1);
    const u64 rest = g >> a;
    const int hib = m - a - R;  // bits of the high part within a vector
    const u64 H = rest & ((1ull << hib) - 1);
    const u64 v = rest >> hib;
    const u64 in_vec = (H << (a + R)) | L;  // index of value c = 0 within its vector
    const u64 base = (v << m) + in_vec;
    double re[1 << R], im[1 << R];
#pragma unroll
    for (int c = 0; c < (1 << R); c++) {
        re[c] = xr[base + ((u64)c << a)];
        if (KIND != KIND_FWHT) im[c] = xi[base + ((u64)c << a)];
    }
    run_stages<KIND, ASC, R>(re, im, in_vec, a, tsr, tsi);
#pragma unroll
    for (int c = 0; c < (1 << R); c++) {
        xr[base + ((u64)c << a)] = re[c];
        if (KIND != KIND_FWHT) xi[base + ((u64)c << a)] = im[c];
    }
}

// stage-major copy of the twiddle table: ts[2^s + t] = tw[t << (m-1-s)], t < 2^s
__global__ __launch_bounds__(kBlock) void stage_twiddles_kernel(int m, const double *__restrict__ twr,
                                                                const double *__restrict__ twi,
                                                                double *__restrict__ tsr, double *__restrict__ tsi) {
    const u64 k = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0 || k >= (1ull << m)) return;
    const int s = 63 - __clzll((long long)k);
    const u64 src = (k - (1ull << s)) << (m - 1 - s);
    tsr[k] = twr[src];
    tsi[k] = twi[src];
}

int log2_exact(u64 n) {
    int m = 0;
    while ((1ull << m) < n) m++;
    return ((1ull << m) == n) ? m : -1;
}

template <int KIND, bool ASC>
int launch_tile(u64 total, int m, int tlog2, double *xr, double *xi, const double *tsr, const double *tsi,
                hipStream_t s) {
    const u64 tiles = (total + (1ull << tlog2) - 1) >> tlog2;
    const size_t words = ((size_t)1 << tlog2) + ((size_t)1 << tlog2) / 16;  // padded
    const size_t shm = (size_t)(KIND == KIND_FWHT ? 1 : 2) * sizeof(double) * words;
    static PerDeviceOnce attrs;  // padded 64 KiB tiles need a little more than the default LDS limit
    const int arc = attrs.run([]() -> int {
        const size_t most = (size_t)(KIND == KIND_FWHT ? 1 : 2) * sizeof(double) *
                            (((size_t)1 << (KIND == KIND_FWHT ? kTileLog2Real : kTileLog2Cplx)) * 17 / 16);
        QMC_HIP_CHECK(hipFuncSetAttribute((const void *)tile_kernel<KIND, ASC>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)most));
        return QMC_OK;
    });
    if (arc) return arc;
    hipLaunchKernelGGL((tile_kernel<KIND, ASC>), dim3((unsigned)tiles), dim3(tile_threads<KIND>()), shm, s, total, m,
                       tlog2, tlog2,
                       xr, xi, tsr, tsi);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

template <int KIND, bool ASC>
int launch_group(u64 total, int m, int a, int R, double *xr, double *xi, const double *tsr, const double *tsi,
                 hipStream_t s) {
    const u64 groups = total >> R;
    const dim3 grid((unsigned)((groups + kBlock - 1) / kBlock));
    switch (R) {
        case 4: hipLaunchKernelGGL((group_kernel<KIND, ASC, 4>), grid, dim3(kBlock), 0, s, groups, m, a, xr, xi, tsr, tsi); break;
        case 3: hipLaunchKernelGGL((group_kernel<KIND, ASC, 3>), grid, dim3(kBlock), 0, s, groups, m, a, xr, xi, tsr, tsi); break;
        case 2: hipLaunchKernelGGL((group_kernel<KIND, ASC, 2>), grid, dim3(kBlock), 0, s, groups, m, a, xr, xi, tsr, tsi); break;
        default: hipLaunchKernelGGL((group_kernel<KIND, ASC, 1>), grid, dim3(kBlock), 0, s, groups, m, a, xr, xi, tsr, tsi); break;
    }
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

template <int KIND>
int run_transform(const char *who, u64 d1, u64 d2, u64 n_half, double *twr, double *twi, double *xr, double *xi,
                  hipStream_t s) {
    const u64 V = d1 * d2;
    if (V == 0 || n_half == 0) return QMC_OK;
    const u64 n = 2 * n_half;
    const int m = log2_exact(n);
    if (m < 0) return qmc_fail(QMC_ERR_ARG, "%s: 2*n_half = %llu is not a power of two", who, (unsigned long long)n);
    if (m > 31) return qmc_fail(QMC_ERR_ARG, "%s: length 2^%d exceeds the reference's int stride (1<<s)", who, m);
    if (V > (1ull << 40) / n) return qmc_fail(QMC_ERR_ARG, "%s: d1*d2*n too large", who);
    const u64 total = V * n;
    double *tsr = nullptr, *tsi = nullptr;
    if (KIND != KIND_FWHT) {
        hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, n,
                           KIND == KIND_FFT ? -2.0 : 2.0, twr, twi);
        QMC_HIP_CHECK(hipGetLastError());
        QMC_HIP_CHECK(hipMallocAsync((void **)&tsr, 2 * sizeof(double) * n, s));
        tsi = tsr + n;
        hipLaunchKernelGGL(stage_twiddles_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, m,
                           twr, twi, tsr, tsi);
    }
    const int tcap = KIND == KIND_FWHT ? kTileLog2Real : kTileLog2Cplx;
    const int tl = m < tcap ? m : tcap;  // LDS stages: strides 2^0 .. 2^(tl-1)
    int rc = QMC_OK;
    if (KIND == KIND_FFT) {
        // strides ascend: LDS stages first, then register groups
        rc = launch_tile<KIND, true>(total, m, tl, xr, xi, tsr, tsi, s);
        for (int a = tl; a < m && !rc; a += kGroupLog2) {
            const int R = (m - a) < kGroupLog2 ? (m - a) : kGroupLog2;
            rc = launch_group<KIND, true>(total, m, a, R, xr, xi, tsr, tsi, s);
        }
    } else {
        // strides descend: register groups from the top, then LDS stages
        int top = m;
        while (top > tl && !rc) {
            const int R = (top - tl) < kGroupLog2 ? (top - tl) : kGroupLog2;
            rc = launch_group<KIND, false>(total, m, top - R, R, xr, xi, tsr, tsi, s);
            top -= R;
        }
        if (!rc) rc = launch_tile<KIND, false>(total, m, tl, xr, xi, tsr, tsi, s);
    }
    if (tsr) (void)hipFreeAsync(tsr, s);
    return rc;
}

}  // namespace

extern "C" {

int fwht_1d_radix2_dev(u64 d1, u64 d2, u64 n_half, double *x, hipStream_t s) {
    return run_transform<KIND_FWHT>("fwht_1d_radix2", d1, d2, n_half, nullptr, nullptr, x, nullptr, s);
}

int fft_bro_1d_radix2_dev(u64 d1, u64 d2, u64 n_half, double *twiddler, double *twiddlei, double *xr, double *xi,
                          hipStream_t s) {
    return run_transform<KIND_FFT>("fft_bro_1d_radix2", d1, d2, n_half, twiddler, twiddlei, xr, xi, s);
}

int ifft_bro_1d_radix2_dev(u64 d1, u64 d2, u64 n_half, double *twiddler, double *twiddlei, double *xr, double *xi,
                           hipStream_t s) {
    return run_transform<KIND_IFFT>("ifft_bro_1d_radix2", d1, d2, n_half, twiddler, twiddlei, xr, xi, s);
}

}  // extern "C"
