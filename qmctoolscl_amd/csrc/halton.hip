// Generalized Halton points (halton_qrng; qmctoolscl/c_funcs/halton_qrng.c).
//
// Point i, dimension j (base b = p_{dvec[j]}, factor f = 1 or the generalized
// factor, shift digits s_k = (int)(b * randu[32 j + k])):
//
//     u = 0;  for k = 31 .. 0:  u = (u + ((f c_k + s_k) mod b)) / b
//
// with c_k the base-b digits of i (halton_qrng.c:122-150), in IEEE double in
// exactly that order, so results match the reference bit for bit.  Above the
// K digits of i every c_k is 0, so the first 32 - K steps depend only on the
// dimension: a set-up kernel runs that chain once per dimension and stores u
// after each step (lead[K]); a point then starts from lead[K] and does its K
// digit steps (<= 31 for base 2, 13 for base 3 at i < 2^20, ...).  Point 0 of
// a call with n0 == 0 uses the unreduced s_k (halton_qrng.c:81-95).
// Unlike the reference, no state is kept between calls (its perm array is a
// static), so concurrent calls are safe.
#include <array>

#include "halton_tables.h"
#include "qmc_common.hpp"

namespace {

constexpr int kDigits = 32;

// the first QMC_HALTON_DIMS primes, at compile time
constexpr std::array<unsigned, QMC_HALTON_DIMS> first_primes() {
    std::array<unsigned, QMC_HALTON_DIMS> p{};
    int found = 0;
    for (unsigned v = 2; found < QMC_HALTON_DIMS; v++) {
        bool prime = true;
        for (int i = 0; i < found && p[i] * p[i] <= v; i++)
            if (v % p[i] == 0) {
                prime = false;
                break;
            }
        if (prime) p[found++] = v;
    }
    return p;
}
constexpr std::array<unsigned, QMC_HALTON_DIMS> kPrimes = first_primes();
static_assert(kPrimes[0] == 2 && kPrimes[359] == 2423, "first 360 primes");

__constant__ unsigned c_prime[QMC_HALTON_DIMS];
__constant__ unsigned short c_factor[QMC_HALTON_DIMS];

struct Dim {
    u32 base, factor;
    DivU32 div;
    double rcp;                // RN(1 / base)
    u32 shift[kDigits];        // s_k as the reference's unsigned int
    double lead[kDigits + 1];  // u after the steps k = 31 .. K with digit 0 (lead[32] = 0)
    double first;              // point 0 of a call with n0 == 0
};

// x / b rounded as IEEE division rounds it, for an integer 2 <= b < 2^50 and
// x >= 0 (no overflow or subnormals): q0 = RN(x y) with y = RN(1/b), then one
// correction q1 = RN(q0 + r0 y), r0 = fma(-q0, b, x).  q0 + r0 y lies within
// about 1.5 2^-53 ulp of x/b, while for odd b the quotient is never closer
// than ulp/(2b) to a rounding midpoint (x = m b for a midpoint m would need
// the odd 54-bit significand of m times b to fit in 53 bits), so q1 = RN(x/b);
// for b = 2 every step is exact.  tests/test_div_small.py checks the sequence
// against division for every odd divisor below 4096.
__device__ __forceinline__ double div_small(double x, double b, double y) {
    const double q0 = x * y;
    const double r0 = __fma_rn(-q0, b, x);
    return __fma_rn(r0, y, q0);
}

// one thread per dimension
__global__ void halton_setup_kernel(int d, int generalized, const double *__restrict__ randu,
                                    const int *__restrict__ dvec, Dim *__restrict__ dims, u32 *status) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d) return;
    int idx = dvec[j];
    if (idx < 0 || idx >= QMC_HALTON_DIMS) {  // the reference reads past its tables: flag it, fill base 2
        atomicOr(status, 2u);
        idx = 0;
    }
    Dim &D = dims[j];
    const u32 b = c_prime[idx];
    D.base = b;
    D.factor = generalized ? c_factor[idx] : 1u;
    D.div = make_div(b);
    D.rcp = 1.0 / (double)b;
    for (int k = 0; k < kDigits; k++) D.shift[k] = (u32)(int)((double)(int)b * randu[j * kDigits + k]);
    double u = 0.0, v = 0.0;
    D.lead[kDigits] = 0.0;
    for (int k = kDigits - 1; k >= 0; k--) {
        u += (double)(D.shift[k] % b);
        u /= (double)b;
        D.lead[k] = u;
        v += (double)D.shift[k];
        v /= (double)b;
    }
    D.first = v;
}

// res[j n + i], i < n, is the point n0 + i; blockIdx.y = j
__global__ __launch_bounds__(kBlock) void halton_points_kernel(int n, int n0, const Dim *__restrict__ dims,
                                                              double *__restrict__ res) {
    const int j = blockIdx.y;
    const Dim &D = dims[j];
    const u32 b = D.base, f = D.factor;
    const DivU32 dv = D.div;
    const double bd = (double)b, y = D.rcp;
    // digits of the largest index of the call: no index has more (uniform over the grid's rows)
    int kall = 0;
    for (u32 t = (u32)n0 + (u32)(n - 1); t != 0; t = div_q(t, dv)) kall++;
    double *out = res + (u64)j * (u64)n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const u32 idx = (u32)n0 + (u32)i;
        double u;
        if (idx == 0) {
            u = D.first;  // only reached with n0 == 0
        } else {
            u32 c[kDigits];
            u32 t = idx;
            int K = 0;
#pragma unroll
            for (int k = 0; k < kDigits; k++) {
                if (k < kall) {
                    const u32 q = div_q(t, dv);
                    c[k] = t - q * b;
                    if (t != 0) K = k + 1;
                    t = q;
                }
            }
            u = D.lead[K];
#pragma unroll
            for (int k = kDigits - 1; k >= 0; k--) {
                if (k < kall && k < K) {
                    // halton_qrng.c: u = (u + (f c + s) % b) / b in unsigned int arithmetic;
                    // the multiply-high division is exact for every 32-bit numerator
                    const u32 num = f * c[k] + D.shift[k];
                    u = div_small(u + (double)(num - div_q(num, dv) * b), bd, y);
                }
            }
        }
        out[i] = u;
    }
}

}  // namespace

u32 *qmc_status_ptr();

extern "C" {

// Device form of halton_qrng: res (d x n float64, res[j n + i] = point n0 + i,
// dimension j), randu (d x 32) and dvec (d ints, each < 360) in device memory.
QMC_EXPORT int qmc_halton_qrng_dev(int n, int d, int n0, int generalized, double *res, const double *randu_d_32,
                                   const int *dvec, qmc_stream_t s) {
    if (n < 0 || d < 0 || n0 < 0) return qmc_fail(QMC_ERR_ARG, "halton_qrng: negative size");
    if ((long long)n + (long long)n0 > 2147483647LL)
        return qmc_fail(QMC_ERR_RANGE, "halton_qrng: n0 + n exceeds the reference's int index range");
    if (n == 0 || d == 0) return QMC_OK;
    static PerDeviceOnce tables;  // the prime and factor tables live in each device's constant memory
    const int trc = tables.run([]() -> int {
        unsigned p[QMC_HALTON_DIMS];
        for (int i = 0; i < QMC_HALTON_DIMS; i++) p[i] = kPrimes[i];
        QMC_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_prime), p, sizeof p));
        QMC_HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_factor), qmc_halton_factor, sizeof qmc_halton_factor));
        return QMC_OK;
    });
    if (trc) return trc;
    u32 *status = qmc_status_ptr();
    if (!status) return QMC_ERR_HIP;
    Dim *dims = nullptr;
    QMC_HIP_CHECK(hipMallocAsync((void **)&dims, sizeof(Dim) * (size_t)d, s));
    hipLaunchKernelGGL(halton_setup_kernel, dim3((unsigned)((d + 63) / 64)), dim3(64), 0, s, d, generalized,
                       randu_d_32, dvec, dims, status);
    const unsigned gx = grid_for((u64)n, kBlock, 4096);
    hipLaunchKernelGGL(halton_points_kernel, dim3(gx, (unsigned)d), dim3(kBlock), 0, s, n, n0, dims, res);
    const hipError_t e = hipGetLastError();
    (void)hipFreeAsync(dims, s);
    if (e != hipSuccess) return qmc_fail_hip(e, "halton_qrng");
    return QMC_OK;
}

}  // extern "C"
