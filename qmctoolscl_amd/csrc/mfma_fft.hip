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
//
// Twiddles.  The exponent brev(c) k of a level splits over the bits of c: the
// bits inside the tile give a factor w_{2^q}^j (q <= 13) that depends only on
// the position in the tile, read from two small LDS tables (w_{2^13}^J =
// t1[J >> 7] t2[J & 127]); the bits fixed for the whole tile give one factor
// per (level, k), computed once per tile from the FP64 tables of w_N.  All
// table entries are FP64 values rounded once to fp32.
#include "qmc_common.hpp"

#include <algorithm>
#include <type_traits>

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
struct TwLds;
static inline size_t shm_bytes(int T);

__device__ __forceinline__ void split3f(float x, __bf16 &a, __bf16 &b, __bf16 &c) {
    a = (__bf16)x;
    const float r1 = x - (float)a;
    b = (__bf16)r1;
    c = (__bf16)(r1 - (float)b);
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// two floats rounded to bf16 (one v_cvt_pk_bf16_f32), as one 32-bit word
__device__ __forceinline__ u32 pack_bf16(float x0, float x1) {
    f32x2 v;
    v.x = x0;
    v.y = x1;
    return __builtin_bit_cast(u32, __builtin_convertvector(v, bf16x2));
}

// split3f of a pair of floats, the pieces packed as the MFMA operands hold
// them: the bf16 of the low half is word << 16, of the high half word & ~0xffff
__device__ __forceinline__ void split3_pair(float x0, float x1, u32 &a, u32 &b, u32 &c) {
    a = pack_bf16(x0, x1);
    const float r0 = x0 - __uint_as_float(a << 16), r1 = x1 - __uint_as_float(a & 0xffff0000u);
    b = pack_bf16(r0, r1);
    c = pack_bf16(r0 - __uint_as_float(b << 16), r1 - __uint_as_float(b & 0xffff0000u));
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

// forward: W[m][c] = w16^(m brev4(c)); inverse: the conjugate transpose,
// W'[m][c] = w16^(-c brev4(m))
__device__ Wtile make_w(int lane, bool inverse) {
    const int m = lane & 15, g = lane >> 4;
    Wtile w;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = g * 8 + i;
        const int c = k & 15;
        // w16^e = cos(2 pi e/16) - i sin(2 pi e/16), generated in FP64
        const int e = inverse ? (c * brev_bits(m, 4)) & 15 : (m * brev_bits(c, 4)) & 15;
        double s, co;
        sincospi(2.0 * e / 16.0, &s, &co);
        const double wr = co, wi = inverse ? s : -s;
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

// LDS twiddle tables after the two value tiles (the DFT-16 operands stay in
// each lane's registers: 6 KiB less LDS per workgroup, four workgroups per CU
// at 2^12 instead of three)
struct TwLds {
    f32x2 t1[64];    // w_64^a = w_{2^13}^(128 a)
    f32x2 t2[128];   // w_{2^13}^b
    f32x2 t3[256];   // w_256^c: one load for the levels with at most 4 tile bits above them
    f32x2 f[4][16];  // per level of the pass: the tile's fixed factor for k = 0..15
};

static inline size_t shm_bytes(int T) { return padded_bytes(T) + sizeof(TwLds); }

// complex product, one rounding fewer per part than mul-mul-add (the file is
// built without contraction, so the fused forms are written out)
__device__ __forceinline__ f32x2 cmul(f32x2 a, f32x2 b) {
    f32x2 r;
    r.x = __builtin_fmaf(a.x, b.x, -(a.y * b.y));
    r.y = __builtin_fmaf(a.x, b.y, a.y * b.x);
    return r;
}

// Twiddle of level `lev` (tile bits [ta, ta+4), b = T - ta - 4 tile bits above
// it) for tile position t: row k = the level's digit of t, column bits c = t's
// bits above the level; w = w_{2^(b+4)}^(brev_b(c) k) * f[lev][k], the factor f
// skipped when the tile fixes no bits above the level (it is exactly 1 then).
__device__ __forceinline__ f32x2 level_twiddle(const TwLds &tl, int T, int ta, int lev, bool fone, int t) {
    const int k = (t >> ta) & 15;
    const int b = T - ta - 4;
    // b = 0: J = 0 and t1[0] t2[0] = 1 exactly; without branches the lookups of
    // several outputs can all be in flight together
    const int rb = b > 0 ? brev_bits(t >> (ta + 4), b) : 0;
    f32x2 w;
    if (b <= 4) {
        w = tl.t3[((rb * k) & ((16 << b) - 1)) << (4 - b)];  // w_{2^(b+4)}^j = w_256^(j 2^(4-b))
    } else {
        const int J = ((rb * k) & ((16 << b) - 1)) << (9 - b);  // w_{2^(b+4)}^j = w_{2^13}^(j 2^(9-b))
        w = cmul(tl.t1[J >> 7], tl.t2[J & 127]);
    }
    return fone ? w : cmul(w, tl.f[lev][k]);
}

// w_{2^(b+4)}^x (x < 2^(b+4)) from the tables, times f[lev][k] unless fone
__device__ __forceinline__ f32x2 twiddle_of(const TwLds &tl, int b, int x, int lev, int k, bool fone) {
    f32x2 w;
    if (b <= 4) {
        w = tl.t3[x << (4 - b)];
    } else {
        const int J = x << (9 - b);
        w = cmul(tl.t1[J >> 7], tl.t2[J & 127]);
    }
    return fone ? w : cmul(w, tl.f[lev][k]);
}

// level_twiddle for the four outputs k = 4 g + i (i < 4) of the column whose
// tile position with k = 0 is `base` (no bits in [ta, ta+4)), with the index
// products strength-reduced (checked against level_twiddle on the CPU for
// every tile size and level): forward, the level's own twiddles (one column
// c, so brev_b(c) k = rb 4g + rb i); inverse, level lev - 1's, on tile bits
// [ta-4, ta) (one k' per column; the outputs' bits [ta, ta+4) are the low
// bits of c', so brev(c') = R0 + brev4(4g+i) 2^(b'-4) with brev4(4g+i) =
// 4 brev2(i) + brev2(g))
template <bool INV>
__device__ __forceinline__ void out_twiddles(const TwLds &tl, int T, int ta, int lev, bool fone, int base, int g,
                                             f32x2 w[4]) {
    if (!INV) {
        const int b = T - ta - 4;
        const int rb = b > 0 ? brev_bits(base >> (ta + 4), b) : 0;
        const int M = (16 << b) - 1, x0 = rb * 4 * g;
#pragma unroll
        for (int i = 0; i < 4; i++) w[i] = twiddle_of(tl, b, (x0 + i * rb) & M, lev, 4 * g + i, fone);
    } else {
        const int tap = ta - 4, b = T - ta;  // level lev - 1: tile bits [ta - 4, ta), b >= 4 bits above
        const int kp = (base >> tap) & 15;
        const int x0 = brev_bits(base >> ta, b) * kp + (((g >> 1 | (g & 1) << 1) * kp) << (b - 4));
        const int U = kp << (b - 2), M = (16 << b) - 1;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int bi = (i >> 1) | ((i & 1) << 1);  // brev2(i)
            w[i] = twiddle_of(tl, b, (x0 + bi * U) & M, lev - 1, kp, fone);
        }
    }
}

// radix-16 level on tile bits [ta, ta+4), level `lev` of the pass.  Forward,
// with tw: output k of the column with in-tile bits c above the level (b of
// them) is multiplied by the level's twiddle (level_twiddle).  Inverse (the
// transposed network): the conjugate twiddles of a level act on its inputs;
// with tw_in they are applied here on the way in, and with tw_next this
// level's outputs take the conjugate twiddles of level lev - 1 (the next one
// run), which then reads its inputs as they are -- so each element is
// twiddled once per level either way, by the lane that holds both its parts.
template <bool INV, bool WHOLE>
__device__ __forceinline__ void dft16_level(float *re, float *im, int T, int ta, bool tw, bool tw_in, bool tw_next, bool fone_,
                            bool fone_next_, int wave, int nwave, int lane, const Wtile &W, const TwLds &tl, int lev, u32 salt) {
    // a tile that is a whole vector fixes no index bits: every factor f is 1
    const bool fone = WHOLE || fone_, fone_next = WHOLE || fone_next_;
    const int nblk = 1 << (T - 8);
    const int n = lane & 15, g = lane >> 4;
    const int c0 = (g & 1) * 8;
    float *src = g < 2 ? re : im;
    // LDS word of row c of a column: pad(base + (c << ta)) = pad(base) + c * st for ta = 0 or
    // ta >= 4 (no carry out of the low four bits), which every radix-16 level has
    const int st = (1 << ta) + ((1 << ta) >> 4);
    for (int blk = wave; blk < nblk; blk += nwave) {
        const int col = blk * 16 + n;
        const int lo = col & ((1 << ta) - 1);
        const int base = ((col >> ta) << (ta + 4)) | lo;
        const int pin = pad(base + (c0 << ta));
        // sign dither against the accumulation's rounding bias, which adds up
        // over the long sums of multi-pass transforms; a tile that is a whole
        // vector (2^13 values at most) showed no such excess without it
        // (relative L2 and max error unchanged at 2^10, 2^12 and 2^13 for
        // normal, smooth and constant inputs)
        const u32 sg = WHOLE ? 0u : dither_sign((u32)col, salt);
        // input twiddles (inverse, tw_in): rows k = c0 + i of one column, so
        // brev_b(c) k = rb c0 + rb i
        const int bin = T - ta - 4;
        const int rbin = INV && tw_in && bin > 0 ? brev_bits(base >> (ta + 4), bin) : 0;
        float v[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            if (INV && tw_in) {
                float vr = re[pin + i * st], vi = im[pin + i * st];
                const f32x2 w = twiddle_of(tl, bin, (rbin * (c0 + i)) & ((16 << bin) - 1), lev, c0 + i, fone);
                const float tr = __builtin_fmaf(vr, w.x, vi * w.y);
                vi = __builtin_fmaf(vi, w.x, -(vr * w.y));
                vr = tr;
                v[i] = flip(g < 2 ? vr : vi, sg);
            } else {
                v[i] = flip(src[pin + i * st], sg);
            }
        }
        u32x4 pa, pb, pc;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            u32 wa, wb, wc;
            split3_pair(v[2 * i], v[2 * i + 1], wa, wb, wc);
            pa[i] = wa;
            pb[i] = wb;
            pc[i] = wc;
        }
        bf16x8 x[3];
        x[0] = __builtin_bit_cast(bf16x8, pa);
        x[1] = __builtin_bit_cast(bf16x8, pb);
        x[2] = __builtin_bit_cast(bf16x8, pc);
        f32x4 yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
        // sum over pieces p + q <= 2 of W_p * X_q, smallest terms first
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[2], x[0], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[1], x[1], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[0], x[2], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[1], x[0], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[0], x[1], yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ar[0], x[0], yr, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[2], x[0], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[1], x[1], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[0], x[2], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[1], x[0], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[0], x[1], yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.Ai[0], x[0], yi, 0, 0, 0);
        // lane holds outputs k = 4g + i of column n; the twiddles of all four first (their
        // lookups in flight together), then the products
        const int pout = pad(base + ((4 * g) << ta));
        const bool twout = INV ? tw_next : tw;
        f32x2 w[4];
        if (twout) out_twiddles<INV>(tl, T, ta, lev, INV ? fone_next : fone, base, g, w);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float vr = flip(yr[i], sg), vi = flip(yi[i], sg);
            if (twout) {
                // forward: times w; inverse: times conj(w)
                const float wy = INV ? -w[i].y : w[i].y;
                const float tr = __builtin_fmaf(vr, w[i].x, -(vi * wy));
                vi = __builtin_fmaf(vr, wy, vi * w[i].x);
                vr = tr;
            }
            re[pout + i * st] = vr;
            im[pout + i * st] = vi;
        }
    }
}

// leftover top level of R = 2^rb points (rb < 4) along tile bits [ta, ta+rb)
template <bool INV>
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
                // forward W[k][j] = w_R^(k brev(j)); inverse W'[k][j] = w_R^(-j brev(k))
                const int e = INV ? (j * brev_bits(k, rb)) & (R - 1) : (k * brev_bits(j, rb)) & (R - 1);
                float s, c;
                sincospif((INV ? 2.0f : -2.0f) * (float)e / (float)R, &s, &c);
                ar += xr[j] * c - xi[j] * s;
                ai += xr[j] * s + xi[j] * c;
            }
            re[pad(base + (k << ta))] = ar;
            im[pad(base + (k << ta))] = ai;
        }
    }
}

// Levels on global bits [glo, ghi) of tiles; the first/last pass move
// float64.  Workgroups are persistent: each walks tiles blockIdx.x,
// +gridDim.x, ... with all NP value pairs (of re and of im) of a thread
// (2^T = 2 NP blockDim.x) loaded at once.
// TT > 0 (whole vectors of 2^TT only): the tile size at compile time, so the
// levels unroll with their strides, twiddle cases and loop counts fixed.
template <int IN64, int OUT64, int NP, bool INV, bool WHOLE = false, int TT = 0>
__global__ __launch_bounds__(kThreads) void fft_mfma_pass(PassGeom G, int glo, int ghi, const void *__restrict__ sr,
                                                        const void *__restrict__ si, void *__restrict__ dr,
                                                        void *__restrict__ di, double scale, Twiddles tw, u64 ntiles) {
    typedef typename std::conditional<IN64 != 0, f64x2, f32x2>::type In2;
    extern __shared__ float lds[];
    static_assert(TT == 0 || WHOLE, "a compile-time tile size is for whole vectors");
    const int T = TT ? TT : G.lr + G.lc;
    const int P = (1 << T) + (1 << T) / 16;
    float *re = lds;
    float *im = lds + P;
    TwLds &tl = *(TwLds *)(lds + 2 * P);
    const int tlog = G.m - T;
    const u64 cmask_tile = (1ull << tlog) - 1;
    if (blockIdx.x >= ntiles) return;
    const Wtile W = make_w(threadIdx.x & 63, INV);
    for (int i = threadIdx.x; i < 448; i += blockDim.x) {
        double sn, cs;
        if (i < 64)
            sincospi(-2.0 * (double)i / 64.0, &sn, &cs);
        else if (i < 192)
            sincospi(-2.0 * (double)(i - 64) / 8192.0, &sn, &cs);
        else
            sincospi(-2.0 * (double)(i - 192) / 256.0, &sn, &cs);
        f32x2 w;
        w.x = (float)cs;
        w.y = (float)sn;
        if (i < 64)
            tl.t1[i] = w;
        else if (i < 192)
            tl.t2[i - 64] = w;
        else
            tl.t3[i - 192] = w;
    }
    // value pair p of this thread: tile index 2 tid + 2 p blockDim.x (two consecutive values share a
    // row as G.lc >= 1); as 2 blockDim.x and 2^lc are powers of two, its global position is that of
    // pair 0 plus a per-p constant
    auto g0 = [&](u64 ti, int tid) -> u64 { return ((ti >> tlog) << G.m) + gpos(G, ti & cmask_tile, 2 * tid); };
    auto poff = [&](int p) -> u64 {
        const u32 q = 2u * (u32)p * blockDim.x;
        return ((u64)(q >> G.lc) << G.lo) + (q & ((1u << G.lc) - 1));
    };
    const int nlev = TT ? (TT + 3) / 4 : (ghi - glo + 3) / 4;
    for (u64 tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const u64 cb = tix & cmask_tile;
        const int tid = opaque(threadIdx.x);
        {
            const u64 a0 = g0(tix, tid);
            In2 qa[NP], qc[NP];
#pragma unroll
            for (int p = 0; p < NP; p++) {
                qa[p] = __builtin_nontemporal_load((const In2 *)sr + (a0 + poff(p)) / 2);
                qc[p] = __builtin_nontemporal_load((const In2 *)si + (a0 + poff(p)) / 2);
            }
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const int t = 2 * (tid + p * blockDim.x);
                re[pad(t)] = (float)qa[p].x;
                re[pad(t + 1)] = (float)qa[p].y;
                im[pad(t)] = (float)qc[p].x;
                im[pad(t + 1)] = (float)qc[p].y;
            }
        }
        // the tile's fixed twiddle factors: level lev, output k (none for a whole vector)
        for (int idx = tid; idx < (WHOLE ? 0 : 16 * nlev); idx += blockDim.x) {
            const int lev = idx >> 4, k = idx & 15;
            const int ga = glo + 4 * lev;
            const int hi_bits = G.m - ga - 4;
            if (ghi - ga < 4 || hi_bits <= 0) continue;  // no twiddle after this level
            const int ta = ga < G.lc ? ga : G.lc + (ga - G.lo);
            const int b = T - ta - 4;
            const int nf = hi_bits - b;  // index bits above the level that the tile fixes
            const u64 chi_f = gpos(G, cb, 0) >> (ga + 4 + b);
            const u64 bf = nf > 0 ? __brevll(chi_f) >> (64 - nf) : 0;
            const u64 e = (bf * (u64)k) & ((2ull << (hi_bits + 3)) - 1);  // mod 2^(hi_bits + 4)
            float c, sn;
            twiddle(tw, e << ga, c, sn);  // w_{2^(hi_bits+4)}^e = w_N^(e 2^ga)
            f32x2 w;
            w.x = c;
            w.y = sn;
            tl.f[lev][k] = w;
        }
        __syncthreads();
        // forward: levels upward; inverse (the transposed network): downward, each level's
        // conjugate twiddles applied to its inputs -- by the level run before it when that is a
        // radix-16 level of this pass, else on the way in.  The levels of a pass sit on
        // consecutive tile bits (ta(lev - 1) = ta(lev) - 4).
        // a whole vector of 2^TT: global bits [0, TT) are tile bits [0, TT)
        const int gm = TT ? TT : G.m, gl = TT ? 0 : glo, gh = TT ? TT : ghi;
        const int glc = TT ? TT : G.lc, glo_ = TT ? TT : G.lo;
        const int nwave = TT ? (1 << TT) / (128 * NP) : (int)(blockDim.x >> 6);
        bool pre = false;  // this step's input twiddles were applied by the previous step
        auto level_step = [&](int step) {
            const int lev = INV ? nlev - 1 - step : step;
            const int ga = gl + 4 * lev;
            const int ltid = opaque(threadIdx.x);
            const int ta = ga < glc ? ga : glc + (ga - glo_);  // tile bit of global bit ga
            const bool tw = gm - ga - 4 > 0;                   // index bits above the level
            const bool fone = (gm - ga - 4) - (T - ta - 4) <= 0;  // the tile fixes none of them
            if (gh - ga >= 4) {
                const bool next = INV && lev > 0;  // level lev - 1 runs next (it always has bits above)
                const bool fone_next = next && (gm - ga) - (T - ta) <= 0;
                dft16_level<INV, WHOLE>(re, im, T, ta, tw, INV && tw && !pre, next, fone, fone_next, ltid >> 6, nwave,
                                        ltid & 63, W, tl, lev, (u32)tix * 0x632BE5ABu + (u32)ga * 0x27D4EB2Fu);
                pre = next;
            } else {
                dft_small_level<INV>(re, im, T, ta, gh - ga);  // only the top level can be short
                pre = false;
            }
            __syncthreads();
        };
        if constexpr (TT != 0) {
#pragma unroll
            for (int step = 0; step < (TT + 3) / 4; step++) level_step(step);
        } else {
            for (int step = 0; step < nlev; step++) level_step(step);
        }
        const int stid = opaque(threadIdx.x);
        const u64 a0 = g0(tix, stid);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int t = 2 * (stid + p * blockDim.x);
            const u64 gi = a0 + poff(p);
            if (OUT64) {
                f64x2 a, c;
                a.x = (double)re[pad(t)] * scale;
                a.y = (double)re[pad(t + 1)] * scale;
                c.x = (double)im[pad(t)] * scale;
                c.y = (double)im[pad(t + 1)] * scale;
                __builtin_nontemporal_store(a, (f64x2 *)((double *)dr + gi));
                __builtin_nontemporal_store(c, (f64x2 *)((double *)di + gi));
            } else {
                f32x2 a, c;
                a.x = re[pad(t)];
                a.y = re[pad(t + 1)];
                c.x = im[pad(t)];
                c.y = im[pad(t + 1)];
                *(f32x2 *)((float *)dr + gi) = a;
                *(f32x2 *)((float *)di + gi) = c;
            }
        }
        __syncthreads();  // every LDS read of this tile is done before the next one is staged
    }
}

// launch one pass of direction INV with float64 input (in) / output (out) chosen at run time
template <bool INV>
int launch_pass(bool in, bool out, int threads, size_t shm, unsigned grid, hipStream_t s, const PassGeom &G, int glo,
                int ghi, const void *sr, const void *si, void *dr, void *di, double scale, const Twiddles &tw,
                u64 ntiles) {
    if (in && out)
        hipLaunchKernelGGL((fft_mfma_pass<1, 1, 8, INV>), dim3(grid), dim3(threads), shm, s, G, glo, ghi, sr, si, dr,
                           di, scale, tw, ntiles);
    else if (in)
        hipLaunchKernelGGL((fft_mfma_pass<1, 0, 8, INV>), dim3(grid), dim3(threads), shm, s, G, glo, ghi, sr, si, dr,
                           di, scale, tw, ntiles);
    else if (out)
        hipLaunchKernelGGL((fft_mfma_pass<0, 1, 8, INV>), dim3(grid), dim3(threads), shm, s, G, glo, ghi, sr, si, dr,
                           di, scale, tw, ntiles);
    else
        hipLaunchKernelGGL((fft_mfma_pass<0, 0, 8, INV>), dim3(grid), dim3(threads), shm, s, G, glo, ghi, sr, si, dr,
                           di, scale, tw, ntiles);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

template <bool INV>
const void *pass_fn(bool in, bool out) {
    return in && out ? (const void *)fft_mfma_pass<1, 1, 8, INV>
           : in      ? (const void *)fft_mfma_pass<1, 0, 8, INV>
           : out     ? (const void *)fft_mfma_pass<0, 1, 8, INV>
                     : (const void *)fft_mfma_pass<0, 0, 8, INV>;
}

// the whole-vector kernel for length 2^m (m <= kT): 8 value pairs per thread
// and the length at compile time from 2^10 up, fewer pairs below
template <bool INV>
const void *whole_fn(int m) {
    switch (m) {
    case 8: return (const void *)fft_mfma_pass<1, 1, 2, INV, true>;
    case 9: return (const void *)fft_mfma_pass<1, 1, 4, INV, true>;
    case 10: return (const void *)fft_mfma_pass<1, 1, 8, INV, true, 10>;
    case 11: return (const void *)fft_mfma_pass<1, 1, 8, INV, true, 11>;
    case 12: return (const void *)fft_mfma_pass<1, 1, 8, INV, true, 12>;
    default: return (const void *)fft_mfma_pass<1, 1, 8, INV, true, 13>;
    }
}

// The forward transform (INV = false, fft_bro_1d_radix2's) runs its passes
// from the lowest index bits up; the inverse (ifft_bro_1d_radix2's) runs the
// conjugate transpose of the same network: the same passes in reverse order,
// each level's conjugated twiddles applied to its inputs and the conjugate
// transposed DFT-16 tiles.
template <bool INV>
int run_mfma_fft(const char *who, u64 d1, u64 d2, u64 n_half, double *xr, double *xi, float *work, hipStream_t s) {
    const u64 V = d1 * d2;
    if (V == 0 || n_half == 0) return QMC_OK;
    const u64 n = 2 * n_half;
    int m = 0;
    while ((1ull << m) < n) m++;
    if ((1ull << m) != n) return qmc_fail(QMC_ERR_ARG, "%s: length is not a power of two", who);
    if (m < 8 || m > 28) return qmc_fail(QMC_ERR_ARG, "%s: length 2^%d outside 2^8 .. 2^28", who, m);
    if (((uintptr_t)xr | (uintptr_t)xi | (uintptr_t)work) % 16)
        return qmc_fail(QMC_ERR_ARG, "%s: xr, xi and work must be 16-byte aligned", who);
    const double scale = (m % 2 == 0) ? ldexp(1.0, -m / 2) : ldexp(1.0, -(m - 1) / 2) / 1.4142135623730951;
    static PerDeviceOnce attrs;  // padded tiles need more than the default 64 KiB of LDS
    const int arc = attrs.run([]() -> int {
        const int bytes = (int)shm_bytes(kT);
        for (int k = 0; k < 4; k++)
            QMC_HIP_CHECK(hipFuncSetAttribute(pass_fn<INV>(k & 1, k >> 1), hipFuncAttributeMaxDynamicSharedMemorySize,
                                              bytes));
        QMC_HIP_CHECK(hipFuncSetAttribute(whole_fn<INV>(kT), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        return QMC_OK;
    });
    if (arc) return arc;
    const u64 cus = (u64)device_cus();
    // FP64 twiddle tables (one product per twiddle), 2^klo + 2^(m-klo) entries, for the per-tile
    // factors of the multi-pass lengths (a whole-vector tile has none)
    Twiddles tw{};
    f64x2 *tt = nullptr;
    if (m > kT) {
        tw.klo = (m + 1) / 2;
        const u64 ntw = (1ull << tw.klo) + (1ull << (m - tw.klo));
        QMC_HIP_CHECK(hipMallocAsync((void **)&tt, sizeof(f64x2) * ntw, s));
        tw.lo = tt;
        tw.hi = tt + (1ull << tw.klo);
        hipLaunchKernelGGL(twiddle_tables_kernel, dim3((unsigned)((ntw + kThreads / 4 - 1) / (kThreads / 4))),
                           dim3(kThreads / 4), 0, s, m, tw.klo, tt, tt + (1ull << tw.klo));
    }
    // persistent grid: as many workgroups as stay resident, at most one per tile
    auto grid_of = [&](const void *fn, int threads, size_t shm, u64 ntiles) -> unsigned {
        return (unsigned)std::min<u64>(ntiles, (u64)resident_per_cu(fn, threads, shm) * cus);
    };
    float *buf = nullptr;
    int rc = QMC_OK;
    if (m <= kT) {
        PassGeom G{m, 0, m, m, m};
        // one vector per tile, 8 value pairs per thread (m >= 10), fewer below
        const int np = m >= 10 ? 8 : 1 << (m - 7);
        const int threads = (1 << m) / (2 * np);
        const size_t shm = shm_bytes(m);
        const void *fn = whole_fn<INV>(m);
        const unsigned grid = grid_of(fn, threads, shm, V);
        int glo = 0, ghi = m;
        const void *sr = xr, *si = xi;
        void *dr = xr, *di = xi;
        double sc = scale;
        u64 nt = V;
        void *args[] = {&G, &glo, &ghi, &sr, &si, &dr, &di, &sc, &tw, &nt};
        if (hipLaunchKernel(fn, dim3(grid), dim3(threads), args, shm, s) != hipSuccess)
            rc = qmc_fail(QMC_ERR_HIP, "%s: launch failed", who);
    } else {
        buf = work;
        if (!buf) QMC_HIP_CHECK(hipMallocAsync((void **)&buf, 2 * sizeof(float) * V * n, s));
        float *br = buf, *bi = buf + V * n;
        const size_t shm = shm_bytes(kT);
        const u64 ntiles = V << (m - kT);
        // passes of the forward order: global bits [0, 12) over contiguous tiles of 2^13, then 8 bits (two
        // levels) per pass over tiles of 2^8 rows x 2^5 columns
        struct P {
            PassGeom G;
            int glo, ghi;
        };
        P passes[4];
        int np = 0;
        passes[np++] = P{PassGeom{m, 0, kT, kT, kT}, 0, 12};
        for (int lo = 12; lo < m; lo += 8) {
            const int hi = lo + 8 < m ? lo + 8 : m;
            const int lr = hi - lo, lc = kT - lr;
            passes[np++] = P{PassGeom{m, lr, lc, lo, hi}, lo, hi};
        }
        for (int step = 0; step < np && !rc; step++) {
            const P &q = passes[INV ? np - 1 - step : step];
            const bool in = step == 0, out = step == np - 1;
            const unsigned grid = grid_of(pass_fn<INV>(in, out), kThreads, shm, ntiles);
            rc = launch_pass<INV>(in, out, kThreads, shm, grid, s, q.G, q.glo, q.ghi,
                                  in ? (const void *)xr : (const void *)br, in ? (const void *)xi : (const void *)bi,
                                  out ? (void *)xr : (void *)br, out ? (void *)xi : (void *)bi, out ? scale : 1.0,
                                  tw, ntiles);
        }
    }
    if (buf && !work) (void)hipFreeAsync(buf, s);
    if (tt) (void)hipFreeAsync(tt, s);
    return rc;
}

}  // namespace

extern "C" {

// In-place orthonormal FFT of d1*d2 split-complex vectors of length 2*n_half,
// input bit-reversed, output natural (as fft_bro_1d_radix2), on the matrix
// cores.  Lengths 2^8 .. 2^28.  `work` (2 * d1*d2*2*n_half floats) may be
// NULL to let the library allocate it on the stream.
QMC_EXPORT int qmc_fft_bro_mfma_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *xr, double *xi, float *work,
                                    qmc_stream_t s) {
    return run_mfma_fft<false>("fft_mfma", d1, d2, n_half, xr, xi, work, s);
}

// In-place orthonormal inverse FFT with natural-order input and bit-reversed
// output (as ifft_bro_1d_radix2), on the matrix cores; same arguments.
QMC_EXPORT int qmc_ifft_bro_mfma_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *xr, double *xi, float *work,
                                     qmc_stream_t s) {
    return run_mfma_fft<true>("ifft_mfma", d1, d2, n_half, xr, xi, work, s);
}

}  // extern "C"
