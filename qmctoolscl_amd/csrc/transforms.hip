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
// Passes: the stages whose stride is below the tile (8192 reals / 4096
// complex values, 64 KiB of LDS) run in one LDS-resident pass over contiguous
// tiles; the stages with larger strides run in register passes of up to four
// stages, each lane holding 16 elements spaced by the smallest stride of the
// group, so lanes read consecutive addresses.
#include "qmc_common.hpp"

namespace {

constexpr double kSqrt2 = 1.4142135623730951;  // sqrt(2.0), correctly rounded
constexpr double kPi = 3.141592653589793;      // acos(-1.0)
constexpr int kTileLog2Real = 13;
constexpr int kTileLog2Cplx = 12;
constexpr int kGroupLog2 = 4;

enum { KIND_FWHT = 0, KIND_FFT = 1, KIND_IFFT = 2 };

__device__ __forceinline__ void bfly_fwht(double &a, double &b) {
    const double x1 = a, x2 = b;
    a = (x1 + x2) / kSqrt2;
    b = (x1 - x2) / kSqrt2;
}

// fft_bro_1d_radix2.c:66-75
__device__ __forceinline__ void bfly_dit(double &ar, double &ai, double &br, double &bi, double c, double s) {
    const double yr = br * c - bi * s;
    const double yi = br * s + bi * c;
    const double xr1 = ar, xi1 = ai;
    ar = (xr1 + yr) / kSqrt2;
    ai = (xi1 + yi) / kSqrt2;
    br = (xr1 - yr) / kSqrt2;
    bi = (xi1 - yi) / kSqrt2;
}

// ifft_bro_1d_radix2.c:66-75
__device__ __forceinline__ void bfly_dif(double &ar, double &ai, double &br, double &bi, double c, double s) {
    const double yr = ar - br, yi = ai - bi;
    ar = (ar + br) / kSqrt2;
    ai = (ai + bi) / kSqrt2;
    br = (yr * c - yi * s) / kSqrt2;
    bi = (yr * s + yi * c) / kSqrt2;
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

// Stages with strides 2^s_first .. 2^s_last (stepping by dir = +1 or -1),
// all below the tile, over contiguous tiles of 2^tlog2 values of the flat
// array (tiles hold whole vectors or aligned parts of one vector).
template <int KIND>
__global__ __launch_bounds__(kBlock) void tile_kernel(u64 total, int m, int tlog2, int s_first, int s_last,
                                                      int dir, double *__restrict__ xr, double *__restrict__ xi,
                                                      const double *__restrict__ twr,
                                                      const double *__restrict__ twi) {
    extern __shared__ double lds[];
    const u64 tile = 1ull << tlog2;
    const u64 base = (u64)blockIdx.x << tlog2;
    const u64 len = (total - base) < tile ? (total - base) : tile;
    double *sr = lds;
    double *si = lds + tile + tile / 16;
    const u64 nmask = (1ull << m) - 1;
    // one pad double per 16 keeps 16-value groups of the lowest bits off a single bank
    auto P = [](u64 t) { return t + (t >> 4); };
    for (u64 t = threadIdx.x; t < len; t += blockDim.x) {
        sr[P(t)] = xr[base + t];
        if (KIND != KIND_FWHT) si[P(t)] = xi[base + t];
    }
    __syncthreads();
    // Rounds of up to four consecutive stages: a lane loads the 16 values that
    // differ in the round's index bits into registers, runs the stages there
    // (same butterflies, same order) and stores them back: one LDS round trip
    // and one barrier per four stages.
    int s = s_first;
    const int nstages = (s_last - s_first) * dir + 1;
    for (int done = 0; done < nstages;) {
        const int cnt = (nstages - done) < kGroupLog2 ? (nstages - done) : kGroupLog2;
        const int lo = dir > 0 ? s : s - cnt + 1;  // lowest index bit of the round
        const u64 groups = len >> cnt;
        for (u64 g = threadIdx.x; g < groups; g += blockDim.x) {
            const u64 gb = ((g >> lo) << (lo + cnt)) | (g & ((1ull << lo) - 1));
            double re[1 << kGroupLog2], im[1 << kGroupLog2];
#pragma unroll
            for (int c = 0; c < (1 << kGroupLog2); c++) {
                if (c < (1 << cnt)) {
                    re[c] = sr[P(gb + ((u64)c << lo))];
                    if (KIND != KIND_FWHT) im[c] = si[P(gb + ((u64)c << lo))];
                }
            }
#pragma unroll
            for (int st = 0; st < kGroupLog2; st++) {
                if (st >= cnt) break;
                const int bit = dir > 0 ? st : cnt - 1 - st;
                const int ss = lo + bit;  // this stage's stride is 2^ss
#pragma unroll
                for (int c1 = 0; c1 < (1 << kGroupLog2); c1++) {
                    if (c1 >= (1 << cnt) || ((c1 >> bit) & 1)) continue;
                    const int c2 = c1 | (1 << bit);
                    if (KIND == KIND_FWHT) {
                        bfly_fwht(re[c1], re[c2]);
                    } else {
                        const u64 p = gb + ((u64)c1 << lo);
                        const u64 tw = ((base + p) & nmask & ((1ull << ss) - 1)) << (m - 1 - ss);
                        const double cc = twr[tw], sn = twi[tw];
                        if (KIND == KIND_FFT)
                            bfly_dit(re[c1], im[c1], re[c2], im[c2], cc, sn);
                        else
                            bfly_dif(re[c1], im[c1], re[c2], im[c2], cc, sn);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < (1 << kGroupLog2); c++) {
                if (c < (1 << cnt)) {
                    sr[P(gb + ((u64)c << lo))] = re[c];
                    if (KIND != KIND_FWHT) si[P(gb + ((u64)c << lo))] = im[c];
                }
            }
        }
        __syncthreads();
        done += cnt;
        s += cnt * dir;
    }
    for (u64 t = threadIdx.x; t < len; t += blockDim.x) {
        xr[base + t] = sr[P(t)];
        if (KIND != KIND_FWHT) xi[base + t] = si[P(t)];
    }
}

// Stages on index bits a .. a+R-1 (R <= 4), applied in ascending bit order
// when ascending != 0 (FFT) and descending otherwise (FWHT, IFFT).  One lane
// owns the 2^R values that differ only in those bits.
template <int KIND>
__global__ __launch_bounds__(kBlock) void group_kernel(u64 groups, int m, int a, int R, int ascending,
                                                       double *__restrict__ xr, double *__restrict__ xi,
                                                       const double *__restrict__ twr,
                                                       const double *__restrict__ twi) {
    const u64 g = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) return;
    const u64 L = g & ((1ull << a) - 1);
    const u64 rest = g >> a;
    const int hib = m - a - R;  // bits of the high part within a vector
    const u64 H = rest & ((1ull << hib) - 1);
    const u64 v = rest >> hib;
    const u64 in_vec = (H << (a + R)) | L;  // index of value c = 0 within its vector
    const u64 base = (v << m) + in_vec;
    const int cnt = 1 << R;
    double re[1 << kGroupLog2], im[1 << kGroupLog2];
#pragma unroll
    for (int c = 0; c < (1 << kGroupLog2); c++) {
        if (c < cnt) {
            re[c] = xr[base + ((u64)c << a)];
            if (KIND != KIND_FWHT) im[c] = xi[base + ((u64)c << a)];
        }
    }
#pragma unroll
    for (int st = 0; st < kGroupLog2; st++) {
        if (st >= R) break;
        const int bit = ascending ? st : R - 1 - st;  // bit of c handled by this stage
        const int s = a + bit;                         // stride 2^s
#pragma unroll
        for (int c1 = 0; c1 < (1 << kGroupLog2); c1++) {
            if (c1 >= cnt || (c1 >> bit) & 1) continue;
            const int c2 = c1 | (1 << bit);
            if (KIND == KIND_FWHT) {
                bfly_fwht(re[c1], re[c2]);
            } else {
                // stage-major table: entry 2^s + t holds twiddle t << (m-1-s), so
                // lanes with consecutive indices read consecutive entries
                const u64 i1 = in_vec + ((u64)c1 << a);
                const u64 tw = (1ull << s) + (i1 & ((1ull << s) - 1));
                const double cc = twr[tw], sn = twi[tw];
                if (KIND == KIND_FFT)
                    bfly_dit(re[c1], im[c1], re[c2], im[c2], cc, sn);
                else
                    bfly_dif(re[c1], im[c1], re[c2], im[c2], cc, sn);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < (1 << kGroupLog2); c++) {
        if (c < cnt) {
            xr[base + ((u64)c << a)] = re[c];
            if (KIND != KIND_FWHT) xi[base + ((u64)c << a)] = im[c];
        }
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

template <int KIND>
int launch_tile(u64 total, int m, int tlog2, int s_first, int s_last, int dir, double *xr, double *xi,
                const double *twr, const double *twi, hipStream_t s) {
    const u64 tiles = (total + (1ull << tlog2) - 1) >> tlog2;
    const size_t words = ((size_t)1 << tlog2) + ((size_t)1 << tlog2) / 16;  // padded
    const size_t shm = (size_t)(KIND == KIND_FWHT ? 1 : 2) * sizeof(double) * words;
    static bool attr_set = false;
    if (!attr_set) {  // padded 64 KiB tiles need a little more than the default LDS limit
        const size_t most = (size_t)(KIND == KIND_FWHT ? 1 : 2) * sizeof(double) *
                            (((size_t)1 << (KIND == KIND_FWHT ? kTileLog2Real : kTileLog2Cplx)) * 17 / 16);
        QMC_HIP_CHECK(hipFuncSetAttribute((const void *)tile_kernel<KIND>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)most));
        attr_set = true;
    }
    hipLaunchKernelGGL(tile_kernel<KIND>, dim3((unsigned)tiles), dim3(kBlock), shm, s, total, m, tlog2, s_first,
                       s_last, dir, xr, xi, twr, twi);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

template <int KIND>
int launch_group(u64 total, int m, int a, int R, int ascending, double *xr, double *xi, const double *twr,
                 const double *twi, hipStream_t s) {
    const u64 groups = total >> R;
    hipLaunchKernelGGL(group_kernel<KIND>, dim3((unsigned)((groups + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       groups, m, a, R, ascending, xr, xi, twr, twi);
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
    if (KIND != KIND_FWHT) {
        hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, n,
                           KIND == KIND_FFT ? -2.0 : 2.0, twr, twi);
        QMC_HIP_CHECK(hipGetLastError());
    }
    const int tcap = KIND == KIND_FWHT ? kTileLog2Real : kTileLog2Cplx;
    const int tl = m < tcap ? m : tcap;  // LDS stages: strides 2^0 .. 2^(tl-1)
    double *tsr = nullptr, *tsi = nullptr;
    if (KIND != KIND_FWHT && m > tl) {
        QMC_HIP_CHECK(hipMallocAsync((void **)&tsr, 2 * sizeof(double) * n, s));
        tsi = tsr + n;
        hipLaunchKernelGGL(stage_twiddles_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, m,
                           twr, twi, tsr, tsi);
    }
    int rc = QMC_OK;
    if (KIND == KIND_FFT) {
        // strides ascend: LDS stages first, then register groups
        rc = launch_tile<KIND>(total, m, tl, 0, tl - 1, +1, xr, xi, twr, twi, s);
        for (int a = tl; a < m && !rc; a += kGroupLog2) {
            const int R = (m - a) < kGroupLog2 ? (m - a) : kGroupLog2;
            rc = launch_group<KIND>(total, m, a, R, 1, xr, xi, tsr, tsi, s);
        }
    } else {
        // strides descend: register groups from the top, then LDS stages
        int top = m;
        while (top > tl && !rc) {
            const int R = (top - tl) < kGroupLog2 ? (top - tl) : kGroupLog2;
            rc = launch_group<KIND>(total, m, top - R, R, 0, xr, xi, tsr, tsi, s);
            top -= R;
        }
        if (!rc) rc = launch_tile<KIND>(total, m, tl, tl - 1, 0, -1, xr, xi, twr, twi, s);
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
