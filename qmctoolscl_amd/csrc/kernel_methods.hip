// Fast QMC kernel methods: the Gram matrices of shift-invariant kernels on
// rank-1 lattice points and of digitally shift-invariant kernels on base-2
// digital net points, both in natural order, are diagonalized by the
// bit-reversed FFT and by the Walsh-Hadamard transform respectively.  These
// entry points evaluate the kernel's first column in registers straight from
// the generating vector / matrices -- the d-dimensional points are never
// written -- and run the float64 transforms of this library on it, as the
// reference's documentation chains them (docs/index.md:1825-1883 for
// fft_bro/ifft_bro, :1929-1977 for fwht): kt = fft(k), then
// K y = ifft(fft(y) kt) sqrt(n) and K^-1 y = ifft(fft(y) / kt) / sqrt(n).
//
// Lattice (natural order, n_start = 0, as lat_gen_natural):
//   x_ij = frac(g_j phi2(i)),  k_i = prod_j (1 + c_j B(x_ij)),
//   B(x) = (x - 1) x + 1/6 (alpha = 1, Bernoulli B_2) or ((x - 1) x)^2 - 1/30
//   (alpha = 2, B_4); c_j = gamma_j (2 pi)^(2 alpha) (-1)^(alpha+1) / (2 alpha)!
//   gives the kernel with Fourier weights gamma_j |h|^(-2 alpha).
// Base-2 net (natural order, n_start = 0, as dnb2_gen_natural):
//   X_ij = C_j i (tmax-bit integer), k_i = prod_j (1 + c_j w(X_ij)),
//   w(0) = 1, w(X) = 1 - 3 2^(bitlen(X) - 1 - tmax)  (the order-1 digitally
//   shift-invariant kernel, 1 - 6 2^(floor(log2 x) - 1) at x = X 2^-tmax).
// Every column value is computed with exactly the IEEE steps written above,
// left to right, no contraction, so it is reproducible bit for bit on the
// host.  x_0 = 0, so k is the kernel matrix's first column; shifts and
// digital shifts leave the matrix unchanged, linear matrix scrambling enters
// through C.
#include "qmc_common.hpp"

namespace {

constexpr double kSixth = 1.0 / 6.0;
constexpr double kThirtieth = 1.0 / 30.0;
constexpr int kRun = 16;  // consecutive indices per thread in the net column

__device__ __forceinline__ double phi2(u64 q) { return (double)__brevll(q) * 0x1p-64; }

// k[l n + i]; blockIdx.y = l, so g and coef rows are uniform per block
template <int ALPHA>
__global__ __launch_bounds__(kBlock) void lat_column_kernel(u64 n, u64 d, const u64 *__restrict__ g,
                                                            const double *__restrict__ coef, double *__restrict__ k) {
    const u64 l = blockIdx.y;
    const u64 *gl = g + l * d;
    const double *cl = coef + l * d;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        const double f = phi2(i);
        double p = 1.0;
        for (u64 j = 0; j < d; j++) {
            const double x = frac_mod1((double)gl[j] * f);
            double b = (x - 1.0) * x;
            if (ALPHA == 1)
                b = b + kSixth;
            else
                b = b * b - kThirtieth;
            p = p * (1.0 + cl[j] * b);
        }
        k[l * n + i] = p;
    }
}

// k[l n + i] for a run of kRun consecutive i per thread: the first X_ij from
// the set bits of i, the next ones by one XOR with a prefix XOR of the low
// columns (i -> i+1 flips bits 0 .. ctz(i+1)).
__global__ __launch_bounds__(kBlock) void dnb2_column_kernel(u64 n, u64 d, u64 mmax, int tmax,
                                                             const u64 *__restrict__ C,
                                                             const double *__restrict__ coef, double *__restrict__ k) {
    const u64 l = blockIdx.y;
    const u64 runs = (n + kRun - 1) / kRun;
    for (u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x; t < runs; t += (u64)gridDim.x * blockDim.x) {
        const u64 i0 = t * kRun;
        double p[kRun];
#pragma unroll
        for (int s = 0; s < kRun; s++) p[s] = 1.0;
        for (u64 j = 0; j < d; j++) {
            const u64 *Cj = C + (l * d + j) * mmax;
            const double c = coef[l * d + j];
            u64 x = 0;
            for (u64 b = 0, q = i0; q; b++, q >>= 1)
                if (q & 1) x ^= Cj[b];
            const u64 c0 = Cj[0];
            const u64 P1 = c0 ^ (mmax > 1 ? Cj[1] : 0);
            const u64 P2 = P1 ^ (mmax > 2 ? Cj[2] : 0);
            const u64 P3 = P2 ^ (mmax > 3 ? Cj[3] : 0);
#pragma unroll
            for (int s = 0; s < kRun; s++) {
                const double w = x ? 1.0 - ldexp(3.0, (64 - __clzll((long long)x)) - 1 - tmax) : 1.0;
                p[s] = p[s] * (1.0 + c * w);
                // i0 + s -> i0 + s + 1 flips bits 0 .. ctz(s + 1) (i0 is a multiple of 16)
                const int z = __builtin_ctz(s + 1);
                x ^= z == 0 ? c0 : z == 1 ? P1 : z == 2 ? P2 : P3;
            }
        }
#pragma unroll
        for (int s = 0; s < kRun; s++)
            if (i0 + s < n) k[l * n + i0 + s] = p[s];
    }
}

// a fill kernel rather than a memset node, so that callers may capture these entry points in a graph
__global__ __launch_bounds__(kBlock) void zero_kernel(u64 total, double *__restrict__ x) {
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (u64)gridDim.x * blockDim.x) x[e] = 0.0;
}

// y = z * kt * sc (mode 0) or z / kt / sc (mode 1), complex or real (ki == nullptr)
__global__ __launch_bounds__(kBlock) void pointwise_kernel(u64 total, int mode, double sc, const double *__restrict__ kr,
                                                           const double *__restrict__ ki, double *__restrict__ yr,
                                                           double *__restrict__ yi) {
    for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (u64)gridDim.x * blockDim.x) {
        const double a = yr[e], b = yi ? yi[e] : 0.0;
        const double c = kr[e], s = ki ? ki[e] : 0.0;
        double ur, ui;
        if (mode == 0) {
            ur = (a * c - b * s) * sc;
            ui = (a * s + b * c) * sc;
        } else {
            const double den = (c * c + s * s) * sc;
            ur = (a * c + b * s) / den;
            ui = (b * c - a * s) / den;
        }
        yr[e] = ur;
        if (yi) yi[e] = ui;
    }
}

int log2_pow2(u64 n) {
    int m = 0;
    while ((1ull << m) < n) m++;
    return (1ull << m) == n ? m : -1;
}

int check_sizes(const char *who, u64 r, u64 n) {
    if (log2_pow2(n) < 1) return qmc_fail(QMC_ERR_ARG, "%s: n = %llu is not a power of two >= 2", who, (unsigned long long)n);
    if (r > 65535) return qmc_fail(QMC_ERR_ARG, "%s: r = %llu above 65535", who, (unsigned long long)r);
    return QMC_OK;
}

}  // namespace

extern "C" {

QMC_EXPORT int qmc_lat_kernel_column_dev(int alpha, qmc_u64 r, qmc_u64 n, qmc_u64 d, const qmc_u64 *g,
                                         const double *coef, double *k, qmc_stream_t s) {
    if (r == 0 || n == 0) return QMC_OK;
    if (int rc = check_sizes("lat_kernel_column", r, n)) return rc;
    if (alpha != 1 && alpha != 2) return qmc_fail(QMC_ERR_ARG, "lat_kernel_column: alpha must be 1 or 2");
    if (n > (1ull << 32)) return qmc_fail(QMC_ERR_RANGE, "lat_kernel_column: n above 2^32");
    const dim3 grid(grid_for(n, kBlock, 8192), (unsigned)r);
    if (alpha == 1)
        hipLaunchKernelGGL(lat_column_kernel<1>, grid, dim3(kBlock), 0, s, n, d, g, coef, k);
    else
        hipLaunchKernelGGL(lat_column_kernel<2>, grid, dim3(kBlock), 0, s, n, d, g, coef, k);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

QMC_EXPORT int qmc_dnb2_kernel_column_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 mmax, int tmax, const qmc_u64 *C,
                                          const double *coef, double *k, qmc_stream_t s) {
    if (r == 0 || n == 0) return QMC_OK;
    if (int rc = check_sizes("dnb2_kernel_column", r, n)) return rc;
    if (tmax < 1 || tmax > 64) return qmc_fail(QMC_ERR_ARG, "dnb2_kernel_column: tmax must be in 1..64");
    if (mmax < 64 && n > (1ull << mmax))
        return qmc_fail(QMC_ERR_RANGE, "dnb2_kernel_column: n = %llu needs more than mmax = %llu columns",
                        (unsigned long long)n, (unsigned long long)mmax);
    const u64 runs = (n + kRun - 1) / kRun;
    const dim3 grid(grid_for(runs, kBlock, 8192), (unsigned)r);
    hipLaunchKernelGGL(dnb2_column_kernel, grid, dim3(kBlock), 0, s, n, d, mmax, tmax, C, coef, k);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

QMC_EXPORT int qmc_lat_kernel_fft_dev(int alpha, qmc_u64 r, qmc_u64 n, qmc_u64 d, const qmc_u64 *g,
                                      const double *coef, double *twiddler, double *twiddlei, double *kr, double *ki,
                                      qmc_stream_t s) {
    if (r == 0 || n == 0) return QMC_OK;
    if (int rc = qmc_lat_kernel_column_dev(alpha, r, n, d, g, coef, kr, s)) return rc;
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(r * n, kBlock, 8192)), dim3(kBlock), 0, s, r * n, ki);
    QMC_HIP_CHECK(hipGetLastError());
    return fft_bro_1d_radix2_dev(1, r, n / 2, twiddler, twiddlei, kr, ki, s);
}

QMC_EXPORT int qmc_dnb2_kernel_fwht_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 mmax, int tmax, const qmc_u64 *C,
                                        const double *coef, double *kt, qmc_stream_t s) {
    if (r == 0 || n == 0) return QMC_OK;
    if (int rc = qmc_dnb2_kernel_column_dev(r, n, d, mmax, tmax, C, coef, kt, s)) return rc;
    return fwht_1d_radix2_dev(1, r, n / 2, kt, s);
}

QMC_EXPORT int qmc_lat_kernel_apply_dev(int mode, qmc_u64 r, qmc_u64 n, const double *kr, const double *ki,
                                        double *twiddler, double *twiddlei, double *yr, double *yi, qmc_stream_t s) {
    if (r == 0 || n == 0) return QMC_OK;
    if (int rc = check_sizes("lat_kernel_apply", r, n)) return rc;
    if (mode != 0 && mode != 1) return qmc_fail(QMC_ERR_ARG, "lat_kernel_apply: mode must be 0 (K y) or 1 (K^-1 y)");
    if (int rc = fft_bro_1d_radix2_dev(1, r, n / 2, twiddler, twiddlei, yr, yi, s)) return rc;
    hipLaunchKernelGGL(pointwise_kernel, dim3(grid_for(r * n, kBlock, 8192)), dim3(kBlock), 0, s, r * n, mode,
                       sqrt((double)n), kr, ki, yr, yi);
    QMC_HIP_CHECK(hipGetLastError());
    return ifft_bro_1d_radix2_dev(1, r, n / 2, twiddler, twiddlei, yr, yi, s);
}

QMC_EXPORT int qmc_dnb2_kernel_apply_dev(int mode, qmc_u64 r, qmc_u64 n, const double *kt, double *y,
                                         qmc_stream_t s) {
    if (r == 0 || n == 0) return QMC_OK;
    if (int rc = check_sizes("dnb2_kernel_apply", r, n)) return rc;
    if (mode != 0 && mode != 1) return qmc_fail(QMC_ERR_ARG, "dnb2_kernel_apply: mode must be 0 (K y) or 1 (K^-1 y)");
    if (int rc = fwht_1d_radix2_dev(1, r, n / 2, y, s)) return rc;
    hipLaunchKernelGGL(pointwise_kernel, dim3(grid_for(r * n, kBlock, 8192)), dim3(kBlock), 0, s, r * n, mode,
                       sqrt((double)n), kt, (const double *)nullptr, y, (double *)nullptr);
    QMC_HIP_CHECK(hipGetLastError());
    return fwht_1d_radix2_dev(1, r, n / 2, y, s);
}

}  // extern "C"
