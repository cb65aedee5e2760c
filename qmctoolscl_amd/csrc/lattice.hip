// Rank-1 lattices on gfx950: generation in linear, natural and Gray order,
// shift mod 1, and the fused generate+shift used for randomized lattices.
//
// Every value is produced with the reference's own floating-point steps
// (qmctoolscl/c_funcs/lat_gen_*.c, lat_shift_mod_1.c): the generator entry is
// converted to double, multiplied by the index fraction, and reduced with an
// exact fmod(., 1).  The index fraction is exact in all three orders:
//   linear : i / (double)n                      (IEEE division, lat_gen_linear.c:25)
//   natural: phi2(q),        q = row + n_start  (lat_gen_natural.c:37-45)
//   gray   : phi2(Gray(k)),  k = row + n_start  (lat_gen_gray.c:37-43)
// where phi2 is the base-2 radical inverse.  The reference walks k serially
// and carries phi2(Gray(k)) as v / 2^p; here every element computes it in
// closed form with one bit reversal, so rows are independent and the kernels
// are write-bound.
#include "qmc_common.hpp"

namespace {

enum { ORDER_LINEAR = 0, ORDER_NATURAL = 1, ORDER_GRAY = 2 };

// radical inverse of q in base 2: reverse the 64 bits and scale by 2^-64;
// exact while q < 2^53 (the reversed pattern spans at most 53 bits)
__device__ __forceinline__ double phi2(u64 q) { return (double)__brevll(q) * 0x1p-64; }

struct RowMap {
    u64 n, d, nd, n_start;
    DivU32 ddiv;   // division by d when nd < 2^32
    int small;     // nd < 2^32
    int check;     // natural order over a range not known to be Gray-closed
};

// row index and index fraction of element e2 of one replication slab; returns
// false when the reference never writes this row (natural order only)
template <int ORDER>
__device__ __forceinline__ bool row_value(const RowMap &m, u64 e2, u64 &j, double &frac) {
    u64 i;
    if (m.small) {
        u32 q = div_q((u32)e2, m.ddiv);
        i = q;
        j = (u32)e2 - q * m.ddiv.d;
    } else {
        split2(e2, m.d, i, j);
    }
    if (ORDER == ORDER_LINEAR) {
        frac = (double)i / (double)m.n;
    } else if (ORDER == ORDER_GRAY) {
        frac = phi2(gray_of(i + m.n_start));
    } else {
        const u64 q = i + m.n_start;
        if (m.check) {
            const u64 k = gray_inverse(q);
            if (k < m.n_start || k - m.n_start >= m.n) return false;
        }
        frac = phi2(q);
    }
    return true;
}

typedef double double2_t __attribute__((ext_vector_type(2)));

// Each thread owns the element pair (2p, 2p+1) of a replication slab and
// writes both with one 16-byte store when the slab length is even (so every
// pair is 16-byte aligned) and both elements are written; otherwise with two
// 8-byte stores.  Pairs halve the store instructions, which lets the kernels
// reach the write bandwidth of a plain fill.
struct Pair {
    u64 j0, j1;
    double f0, f1;
    bool ok0, ok1;  // the reference writes this element (natural order may skip rows)
};

template <int ORDER>
__device__ __forceinline__ Pair pair_at(const RowMap &m, u64 e2) {
    Pair P;
    P.ok0 = row_value<ORDER>(m, e2, P.j0, P.f0);
    P.ok1 = false;
    P.j1 = 0;
    P.f1 = 0.0;
    if (e2 + 1 < m.nd) {
        if (P.j0 + 1 < m.d) {  // same row
            P.j1 = P.j0 + 1;
            P.f1 = P.f0;
            P.ok1 = P.ok0;
        } else {
            P.ok1 = row_value<ORDER>(m, e2 + 1, P.j1, P.f1);
        }
    }
    return P;
}

template <bool NT = false>
__device__ __forceinline__ void store_pair(double *p, bool vec, bool ok0, bool ok1, double v0, double v1) {
    if (vec && ok0 && ok1) {
        double2_t v = {v0, v1};
        if (NT)
            __builtin_nontemporal_store(v, (double2_t *)p);
        else
            *(double2_t *)p = v;
    } else {
        if (ok0) p[0] = v0;
        if (ok1) p[1] = v1;
    }
}

// Replication groups.  A thread computes its element pair's index fraction
// once and writes the pair into kRepGroup consecutive replication slabs;
// blockIdx.y selects the group and x-fastest dispatch order makes the grid
// group-major, so the waves in flight write kRepGroup slabs, each over one
// contiguous window.  Each replication reads its g and shift entries again
// (L1 hits) just before its store.  Measured for n = 2^20, d = 64, r = 16:
//   all 16 slabs per thread (one group)                     742 Gpts/s
//   groups of 8                                             795
//   groups of 4                                             791
//   one slab per thread                                     650
//   groups of 8 with g read once per thread                 758
//   all 16 shifts loaded before the first store             565
//   g and shifts staged in LDS per workgroup                655-713
// against 868 for a plain 16-byte fill of the same 8 GiB.  Fewer windows
// open at once and fewer stores in flight per wave both write faster.  The
// fused kernel's 16-byte stores are streaming ones (786 against 768-774
// with cached stores in the same run; workgroups of 128 or 512 threads
// change nothing).  The generator alone keeps cached stores: streaming ones
// made it slower (737 against 762 Gpts/s, r = 16, same run).
constexpr u32 kRepGroup = 8;

struct RepGrid {
    dim3 grid;
    u64 groups;
};

// x: one pair per thread (up to 2^20 workgroups, then the kernels'
// grid-stride loops take the rest).  A full grid writes faster than a capped
// one looping: the waves in flight then cover one contiguous window of each
// slab, where a grid-stride loop sends each wave's successive stores
// gridDim*blockDim pairs apart (710 Gpts/s with 65536 workgroups, 745 with
// the full 131072, before the groups).  y: the replication groups, looped
// beyond 65535.
RepGrid rep_grid(u64 pairs, u64 reps) {
    RepGrid G;
    G.groups = (reps + kRepGroup - 1) / kRepGroup;
    G.grid = dim3(grid_for(pairs, kBlock), (unsigned)(G.groups < 65535 ? G.groups : 65535));
    return G;
}

// x[l, row, j] = fmod((double)g[l, j] * frac(row), 1) for all r generating vectors
template <int ORDER>
__global__ __launch_bounds__(kBlock) void lat_gen_kernel(RowMap m, u64 r, u64 groups, bool vec,
                                                         const u64 *__restrict__ g,
                                                         double *__restrict__ x) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 pairs = (m.nd + 1) / 2;
    for (u64 grp = blockIdx.y; grp < groups; grp += gridDim.y) {
        const u64 l0 = grp * kRepGroup, l1 = l0 + kRepGroup < r ? l0 + kRepGroup : r;
        for (u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += stride) {
            const u64 e2 = 2 * p;
            const Pair P = pair_at<ORDER>(m, e2);
            for (u64 l = l0; l < l1; l++)
                store_pair(x + l * m.nd + e2, vec, P.ok0, P.ok1, frac_mod1((double)g[l * m.d + P.j0] * P.f0),
                           P.ok1 ? frac_mod1((double)g[l * m.d + P.j1] * P.f1) : 0.0);
        }
    }
}

// xr[l, row, j] = fmod(fmod((double)g[l % r_g, j] * frac(row), 1) + shifts[l, j], 1):
// lat_gen_* followed by lat_shift_mod_1 with r_x = r_g, one pass over HBM.
template <int ORDER>
__global__ __launch_bounds__(kBlock) void lat_gen_shift_kernel(RowMap m, u64 r, u64 r_g, u64 groups, bool vec,
                                                               const u64 *__restrict__ g,
                                                               const double *__restrict__ shifts,
                                                               double *__restrict__ xr) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 pairs = (m.nd + 1) / 2;
    for (u64 grp = blockIdx.y; grp < groups; grp += gridDim.y) {
        const u64 l0 = grp * kRepGroup, l1 = l0 + kRepGroup < r ? l0 + kRepGroup : r;
        const u64 lg0 = l0 % r_g;
        for (u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += stride) {
            const u64 e2 = 2 * p;
            const Pair P = pair_at<ORDER>(m, e2);
            u64 lg = lg0;
            for (u64 l = l0; l < l1; l++) {
                const double x0 = frac_mod1((double)g[lg * m.d + P.j0] * P.f0);
                const double x1 = P.ok1 ? frac_mod1((double)g[lg * m.d + P.j1] * P.f1) : 0.0;
                store_pair<true>(xr + l * m.nd + e2, vec, P.ok0, P.ok1, frac_mod1(x0 + shifts[l * m.d + P.j0]),
                                 frac_mod1(x1 + shifts[l * m.d + P.j1]));
                if (++lg == r_g) lg = 0;
            }
        }
    }
}

// x[i] = v, one 16-byte store per thread over a full grid: the store pattern
// of the generators at its simplest, a reference for their write rate
__global__ __launch_bounds__(kBlock) void fill_pairs_kernel(u64 pairs, double v, double *__restrict__ x) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += stride) {
        double2_t w = {v, v};
        *(double2_t *)(x + 2 * p) = w;
    }
}

// xr[l, i, j] = fmod(x[l % r_x, i, j] + shifts[l, j], 1): each x element is read
// once and feeds the r / r_x replications that share it.
__global__ __launch_bounds__(kBlock) void lat_shift_kernel(RowMap m, u64 r, u64 r_x, bool vec,
                                                           const double *__restrict__ x,
                                                           const double *__restrict__ shifts,
                                                           double *__restrict__ xr) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    const u64 pairs = (m.nd + 1) / 2;
    for (u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x; t < r_x * pairs; t += stride) {
        const u64 lx = t / pairs;
        const u64 e2 = 2 * (t - lx * pairs);
        u64 j0, i0;
        if (m.small) {
            const u32 q = div_q((u32)e2, m.ddiv);
            i0 = q;
            j0 = (u32)e2 - q * m.ddiv.d;
        } else {
            split2(e2, m.d, i0, j0);
        }
        (void)i0;
        const bool two = e2 + 1 < m.nd;
        const u64 j1 = (j0 + 1 < m.d) ? j0 + 1 : 0;
        const double *src = x + lx * m.nd + e2;
        double x0, x1 = 0.0;
        if (vec && two) {
            const double2_t v = *(const double2_t *)src;
            x0 = v.x;
            x1 = v.y;
        } else {
            x0 = src[0];
            if (two) x1 = src[1];
        }
        for (u64 l = lx; l < r; l += r_x)
            store_pair(xr + l * m.nd + e2, vec, true, two, frac_mod1(x0 + shifts[l * m.d + j0]),
                       frac_mod1(x1 + shifts[l * m.d + j1]));
    }
}

RowMap make_map(u64 n, u64 d, u64 n_start, int check) {
    RowMap m;
    m.n = n;
    m.d = d;
    m.nd = n * d;
    m.n_start = n_start;
    m.small = (m.nd < (1ull << 32)) && d < (1ull << 31);
    m.ddiv = make_div(m.small ? (u32)d : 1u);
    m.check = check;
    return m;
}

bool is_pow2_or_zero(u64 v) { return (v & (v - 1)) == 0; }

// The natural-order output rows are exactly [n_start, n_start+n) when both ends
// are 0 or powers of two (the reference's own precondition, util.py:84-86);
// otherwise each row checks whether the reference would write it.
int natural_check(u64 n_start, u64 n) { return !(is_pow2_or_zero(n_start) && is_pow2_or_zero(n_start + n)); }

// the reference builds phi2(Gray(k)) with a 32-bit int shift, valid for k < 2^31
int check_k_range(u64 n_start, u64 n, const char *who) {
    if (n_start > (1ull << 31) || n > (1ull << 31) - n_start)
        return qmc_fail(QMC_ERR_RANGE, "%s: n_start+n = %llu exceeds 2^31, where the reference's "
                        "int shift (lat_gen_natural.c:30,57) is undefined", who,
                        (unsigned long long)(n_start + n));
    return QMC_OK;
}

bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// 16-byte stores need even slabs (pairs never straddle replications) and aligned bases
bool use_vec(u64 nd, const void *a, const void *b = nullptr) {
    return (nd % 2 == 0) && aligned16(a) && (b == nullptr || aligned16(b));
}

template <int ORDER>
int launch_gen(u64 r, u64 n, u64 d, u64 n_start, const u64 *g, double *x, hipStream_t s, int check) {
    RowMap m = make_map(n, d, n_start, check);
    const RepGrid G = rep_grid((m.nd + 1) / 2, r);
    hipLaunchKernelGGL(lat_gen_kernel<ORDER>, G.grid, dim3(kBlock), 0, s, m, r, G.groups, use_vec(m.nd, x), g, x);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

template <int ORDER>
int launch_gen_shift(u64 r, u64 n, u64 d, u64 rg, u64 n_start, const u64 *g, const double *shifts, double *xr,
                     hipStream_t s, int check) {
    RowMap m = make_map(n, d, n_start, check);
    const RepGrid G = rep_grid((m.nd + 1) / 2, r);
    hipLaunchKernelGGL(lat_gen_shift_kernel<ORDER>, G.grid, dim3(kBlock), 0, s, m, r, rg, G.groups,
                       use_vec(m.nd, xr), g, shifts, xr);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

}  // namespace

extern "C" {

// x[0 .. n) = v (x 16-byte aligned, n even): the plain-store write rate the
// generators are measured against
QMC_EXPORT int qmc_fill_dev(double *x, qmc_u64 n, double v, qmc_stream_t s) {
    if (n == 0) return QMC_OK;
    if ((n % 2) || !aligned16(x)) return qmc_fail(QMC_ERR_ARG, "fill: x must be 16-byte aligned and n even");
    hipLaunchKernelGGL(fill_pairs_kernel, dim3(grid_for(n / 2, kBlock, 0x7fffffffu)), dim3(kBlock), 0, s, n / 2, v, x);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

int lat_gen_linear_dev(u64 r, u64 n, u64 d, const u64 *g, double *x, hipStream_t s) {
    if (r == 0 || n == 0 || d == 0) return QMC_OK;
    return launch_gen<ORDER_LINEAR>(r, n, d, 0, g, x, s, 0);
}

int lat_gen_natural_dev(u64 r, u64 n, u64 d, u64 n_start, const u64 *g, double *x, hipStream_t s) {
    if (r == 0 || n == 0 || d == 0) return QMC_OK;
    int rc = check_k_range(n_start, n, "lat_gen_natural");
    if (rc) return rc;
    return launch_gen<ORDER_NATURAL>(r, n, d, n_start, g, x, s, natural_check(n_start, n));
}

int lat_gen_gray_dev(u64 r, u64 n, u64 d, u64 n_start, const u64 *g, double *x, hipStream_t s) {
    if (r == 0 || n == 0 || d == 0) return QMC_OK;
    int rc = check_k_range(n_start, n, "lat_gen_gray");
    if (rc) return rc;
    return launch_gen<ORDER_GRAY>(r, n, d, n_start, g, x, s, 0);
}

int lat_shift_mod_1_dev(u64 r, u64 n, u64 d, u64 r_x, const double *x, const double *shifts,
                        double *xr, hipStream_t s) {
    if (r == 0 || n == 0 || d == 0) return QMC_OK;
    if (r_x == 0) return qmc_fail(QMC_ERR_ARG, "lat_shift_mod_1: r_x must be positive");
    RowMap m = make_map(n, d, 0, 0);
    const u64 rx = r_x < r ? r_x : r;  // replications of x beyond r are never read
    hipLaunchKernelGGL(lat_shift_kernel, dim3(grid_for(rx * ((m.nd + 1) / 2), kBlock)), dim3(kBlock), 0, s, m,
                       r, rx, use_vec(m.nd, x, xr), x, shifts, xr);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

// order: 0 linear, 1 natural, 2 gray.  Same result as lat_gen_<order>(r_g, ...)
// into a temporary followed by lat_shift_mod_1(r, ..., r_x = r_g, ...).
int qmc_lat_gen_shift_dev(int order, u64 r, u64 n, u64 d, u64 r_g, u64 n_start, const u64 *g,
                          const double *shifts, double *xr, hipStream_t s) {
    if (r == 0 || n == 0 || d == 0) return QMC_OK;
    if (r_g == 0) return qmc_fail(QMC_ERR_ARG, "lat_gen_shift: r_g must be positive");
    const u64 rg = r_g < r ? r_g : r;
    if (order == ORDER_LINEAR) return launch_gen_shift<ORDER_LINEAR>(r, n, d, rg, 0, g, shifts, xr, s, 0);
    if (order != ORDER_NATURAL && order != ORDER_GRAY)
        return qmc_fail(QMC_ERR_ARG, "lat_gen_shift: order must be 0 (linear), 1 (natural) or 2 (gray)");
    int rc = check_k_range(n_start, n, "lat_gen_shift");
    if (rc) return rc;
    if (order == ORDER_NATURAL)
        return launch_gen_shift<ORDER_NATURAL>(r, n, d, rg, n_start, g, shifts, xr, s, natural_check(n_start, n));
    return launch_gen_shift<ORDER_GRAY>(r, n, d, rg, n_start, g, shifts, xr, s, 0);
}

}  // extern "C"
