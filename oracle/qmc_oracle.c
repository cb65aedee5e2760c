/*
 * CPU oracle for the QMC point-generation and transform path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker, never the thing measured or shipped.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so.  The product (qmctoolscl_amd) never links or calls it.
 *
 * Each function restates one reference C kernel loop for loop, with the same
 * operation order, so integer outputs and IEEE-754 double outputs match the
 * reference bit for bit.  Build with -O2 -ffp-contract=off (no FMA contraction).
 * Citations are to qmctoolscl/c_funcs/<file>.c in the reference tree.
 *
 * The reference's bs_* batch-size arguments are dropped: the C backend is
 * always called with batch size == size (qmctoolscl/util.py:107), i.e. one
 * "work item" that owns the whole index space, so l0 = i0 = j0 = 0 throughout.
 *
 * How each function is pinned (see tests/test_oracle.py):
 *   - every function: the doctest golden vectors of docs/index.md, transcribed
 *     into tests/golden/doctest_vectors.json;
 *   - every function: bit-for-bit agreement with the reference C compiled
 *     from its own sources into oracle/_ref/ (when that build is present);
 *   - closed forms and invariants: lattice = frac(g*phi2(Gray(k))),
 *     natural[Gray(i)] == gray[i], XOR linearity, S = I leaves C unchanged,
 *     undo(interlace(C)) == C, Halton identity C gives radical inverses,
 *     fwht(fwht(x)) == x and fwht == sympy-style Hadamard / sqrt(n),
 *     fft_bro == numpy.fft.fft(x[bitrev], norm="ortho"), ifft(fft(x)) == x.
 */
#include <math.h>
#include <stdint.h>
#include "qmc_oracle.h"
#include "../qmctoolscl_amd/csrc/halton_tables.h"  /* data only: the generalized Halton factors */

/* index of the lowest set bit; k > 0 (the reference's `while(!((k>>b)&1)) b++`) */
static u64 low_bit_index(u64 k) {
    u64 b = 0;
    while (!((k >> b) & 1)) b++;
    return b;
}

/* ------------------------------------------------------------------ lattices */

/* lat_gen_linear.c:17-33: x[l,i,j] = fmod((double)g[l,j] * (i / (double)n), 1) */
void oracle_lat_gen_linear(u64 r, u64 n, u64 d, const u64 *g, double *x) {
    const double nd = (double)n;
    for (u64 i = 0; i < n; i++) {
        const double frac_i = i / nd;
        for (u64 j = 0; j < d; j++)
            for (u64 l = 0; l < r; l++)
                x[l * n * d + i * d + j] = fmod((double)g[l * d + j] * frac_i, 1.0);
    }
}

/* Shared walk of lat_gen_natural.c:23-58 and lat_gen_gray.c:23-56.  The pair
 * (v, p) carries v / 2^p = radical inverse of Gray(k) in base 2; the reference
 * builds the bit masks with an `int` shift (1 << (p-b-1)), reproduced here,
 * which is why k must stay below 2^31. */
static void lat_walk(u64 r, u64 n, u64 d, u64 n_start, const u64 *g, double *x,
                     int natural) {
    u64 p = (u64)ceil(log2((double)n_start + 1));
    u64 v = 0;
    {
        u64 gray0 = n_start ^ (n_start >> 1);
        for (u64 b = 0; gray0 > 0; b++, gray0 >>= 1)
            if (gray0 & 1) v += (u64)(1 << (p - b - 1));
    }
    for (u64 i = 0; i < n; i++) {
        const double phi = ldexp((double)v, -(int)p);
        const u64 k = i + n_start;
        const u64 row = natural ? ((k ^ (k >> 1)) - n_start) : i;
        for (u64 j = 0; j < d; j++)
            for (u64 l = 0; l < r; l++)
                x[l * n * d + row * d + j] = fmod((double)g[l * d + j] * phi, 1.0);
        const u64 kn = k + 1;
        if ((kn & (kn - 1)) == 0) { /* kn is a power of two: one more bit */
            p += 1;
            v <<= 1;
        }
        v ^= (u64)(1 << (p - low_bit_index(kn) - 1));
    }
}

/* lat_gen_natural.c:15-58 */
void oracle_lat_gen_natural(u64 r, u64 n, u64 d, u64 n_start, const u64 *g, double *x) {
    lat_walk(r, n, d, n_start, g, x, 1);
}

/* lat_gen_gray.c:15-56 */
void oracle_lat_gen_gray(u64 r, u64 n, u64 d, u64 n_start, const u64 *g, double *x) {
    lat_walk(r, n, d, n_start, g, x, 0);
}

/* lat_shift_mod_1.c:16-34: xr = fmod(x[idx % (r_x*n*d)] + shifts[l,j], 1) */
void oracle_lat_shift_mod_1(u64 r, u64 n, u64 d, u64 r_x, const double *x,
                            const double *shifts, double *xr) {
    const u64 nx = r_x * n * d;
    for (u64 l = 0; l < r; l++)
        for (u64 i = 0; i < n; i++)
            for (u64 j = 0; j < d; j++) {
                const u64 e = l * n * d + i * d + j;
                xr[e] = fmod(x[e % nx] + shifts[l * d + j], 1.0);
            }
}

/* ------------------------------------------------------- base-2 digital nets */

/* dnb2_gmat_lsb_to_msb.c:15-44: reverse each column integer into tmaxes[l] bits */
void oracle_dnb2_gmat_lsb_to_msb(u64 r, u64 d, u64 mmax, const u64 *tmaxes,
                                 const u64 *C_lsb, u64 *C_msb) {
    for (u64 l = 0; l < r; l++) {
        const u64 tmax = tmaxes[l];
        for (u64 j = 0; j < d; j++)
            for (u64 k = 0; k < mmax; k++) {
                const u64 e = l * d * mmax + j * mmax + k;
                u64 src = C_lsb[e], out = 0;
                for (u64 t = 0; src != 0; t++, src >>= 1)
                    if (src & 1) out += (u64)1 << (tmax - t - 1);
                C_msb[e] = out;
            }
    }
}

/* dnb2_linear_matrix_scramble.c:17-51: C_lms = S * C over GF(2); row t of S
 * dotted with column C[k] gives bit (tmax_new-1-t) of the new column. */
void oracle_dnb2_linear_matrix_scramble(u64 r, u64 d, u64 mmax, u64 r_C, u64 tmax_new,
                                        const u64 *S, const u64 *C, u64 *C_lms) {
    const u64 nC = r_C * d * mmax;
    for (u64 l = 0; l < r; l++)
        for (u64 j = 0; j < d; j++)
            for (u64 k = 0; k < mmax; k++) {
                const u64 e = l * d * mmax + j * mmax + k;
                const u64 col = C[e % nC];
                u64 out = 0;
                for (u64 t = 0; t < tmax_new; t++) {
                    u64 w = S[l * d * tmax_new + j * tmax_new + t] & col;
                    u64 ones = 0;
                    while (w) { ones++; w &= w - 1; }
                    if (ones & 1) out += (u64)1 << (tmax_new - t - 1);
                }
                C_lms[e] = out;
            }
}

/* Shared walk of dnb2_gen_natural.c:23-70 and dnb2_gen_gray.c:23-69: the first
 * point XORs the columns selected by Gray(n_start); every later point is the
 * previous one XOR column ctz(k). */
static void dnb2_walk(u64 r, u64 n, u64 d, u64 n_start, u64 mmax, const u64 *C,
                      u64 *xb, int natural) {
    if (n == 0) return;
    const u64 gray0 = n_start ^ (n_start >> 1);
    u64 prev = natural ? (gray0 - n_start) * d : 0;
    for (u64 j = 0; j < d; j++)
        for (u64 l = 0; l < r; l++) xb[l * n * d + prev + j] = 0;
    {
        u64 bits = gray0;
        for (u64 b = 0; bits > 0; b++, bits >>= 1)
            if (bits & 1)
                for (u64 j = 0; j < d; j++)
                    for (u64 l = 0; l < r; l++)
                        xb[l * n * d + prev + j] ^= C[l * d * mmax + j * mmax + b];
    }
    for (u64 i = 1; i < n; i++) {
        const u64 k = i + n_start;
        const u64 cur = natural ? ((k ^ (k >> 1)) - n_start) * d : i * d;
        const u64 b = low_bit_index(k);
        for (u64 j = 0; j < d; j++)
            for (u64 l = 0; l < r; l++)
                xb[l * n * d + cur + j] =
                    xb[l * n * d + prev + j] ^ C[l * d * mmax + j * mmax + b];
        prev = cur;
    }
}

/* dnb2_gen_natural.c:16-70 */
void oracle_dnb2_gen_natural(u64 r, u64 n, u64 d, u64 n_start, u64 mmax,
                             const u64 *C, u64 *xb) {
    dnb2_walk(r, n, d, n_start, mmax, C, xb, 1);
}

/* dnb2_gen_gray.c:16-69 */
void oracle_dnb2_gen_gray(u64 r, u64 n, u64 d, u64 n_start, u64 mmax,
                          const u64 *C, u64 *xb) {
    dnb2_walk(r, n, d, n_start, mmax, C, xb, 0);
}

/* dnb2_digital_shift.c:17-35: xrb = (xb[idx % (r_x*n*d)] << lshifts[l % r_x]) ^ shiftsb[l,j] */
void oracle_dnb2_digital_shift(u64 r, u64 n, u64 d, u64 r_x, const u64 *lshifts,
                               const u64 *xb, const u64 *shiftsb, u64 *xrb) {
    const u64 nx = r_x * n * d;
    for (u64 l = 0; l < r; l++)
        for (u64 i = 0; i < n; i++)
            for (u64 j = 0; j < d; j++) {
                const u64 e = l * n * d + i * d + j;
                xrb[e] = (xb[e % nx] << lshifts[l % r_x]) ^ shiftsb[l * d + j];
            }
}

/* dnb2_integer_to_float.c:15-32: x = ldexp((double)xb, -tmaxes[l]) */
void oracle_dnb2_integer_to_float(u64 r, u64 n, u64 d, const u64 *tmaxes,
                                  const u64 *xb, double *x) {
    for (u64 l = 0; l < r; l++)
        for (u64 i = 0; i < n; i++)
            for (u64 j = 0; j < d; j++) {
                const u64 e = l * n * d + i * d + j;
                x[e] = ldexp((double)xb[e], -(int)tmaxes[l]);
            }
}

/* dnb2_interlace.c:18-45: bit t_a of C_alpha[l,ja,k] = bit t_a/alpha of
 * C[l, ja*alpha + t_a%alpha, k] (bits counted from the most significant). */
void oracle_dnb2_interlace(u64 r, u64 d_alpha, u64 mmax, u64 d, u64 tmax,
                           u64 tmax_alpha, u64 alpha, const u64 *C, u64 *C_alpha) {
    for (u64 l = 0; l < r; l++)
        for (u64 ja = 0; ja < d_alpha; ja++)
            for (u64 k = 0; k < mmax; k++) {
                u64 out = 0;
                for (u64 ta = 0; ta < tmax_alpha; ta++) {
                    const u64 t = ta / alpha;
                    const u64 j = ja * alpha + ta % alpha;
                    if ((C[l * d * mmax + j * mmax + k] >> (tmax - t - 1)) & 1)
                        out += (u64)1 << (tmax_alpha - ta - 1);
                }
                C_alpha[l * d_alpha * mmax + ja * mmax + k] = out;
            }
}

/* dnb2_undo_interlace.c:18-45: inverse of the above */
void oracle_dnb2_undo_interlace(u64 r, u64 d, u64 mmax, u64 d_alpha, u64 tmax,
                                u64 tmax_alpha, u64 alpha, const u64 *C_alpha, u64 *C) {
    for (u64 l = 0; l < r; l++)
        for (u64 j = 0; j < d; j++)
            for (u64 k = 0; k < mmax; k++) {
                u64 out = 0;
                for (u64 t = 0; t < tmax; t++) {
                    const u64 ja = j / alpha;
                    const u64 ta = t * alpha + j % alpha;
                    if ((C_alpha[l * d_alpha * mmax + ja * mmax + k] >> (tmax_alpha - ta - 1)) & 1)
                        out += (u64)1 << (tmax - t - 1);
                }
                C[l * d * mmax + j * mmax + k] = out;
            }
}

/* ------------------------------------------------- generalized digital nets */

/* gdn_gen_natural.c:19-53: digits of k = i + n_start in base b select columns
 * of C[l,j]; x[t] = (x[t] + dig * C[l,j,m,t]) % b after every digit m. */
void oracle_gdn_gen_natural(u64 r, u64 n, u64 d, u64 r_b, u64 mmax, u64 tmax,
                            u64 n_start, const u64 *bases, const u64 *C, u64 *xdig) {
    for (u64 l = 0; l < r; l++)
        for (u64 i = 0; i < n; i++)
            for (u64 j = 0; j < d; j++) {
                u64 *out = xdig + l * n * d * tmax + i * d * tmax + j * tmax;
                for (u64 t = 0; t < tmax; t++) out[t] = 0;
                const u64 b = bases[(l % r_b) * d + j];
                u64 rest = i + n_start;
                for (u64 m = 0; rest > 0; m++) {
                    const u64 dig = rest % b;
                    rest = (rest - dig) / b;
                    if (dig > 0) {
                        const u64 *col = C + l * d * mmax * tmax + j * mmax * tmax + m * tmax;
                        for (u64 t = 0; t < tmax; t++) out[t] = (out[t] + dig * col[t]) % b;
                    }
                }
            }
}

/* gdn_gen_natural_same_base.c:18-63: as above with one base for every dimension */
void oracle_gdn_gen_natural_same_base(u64 r, u64 n, u64 d, u64 mmax, u64 tmax,
                                      u64 n_start, u64 b, const u64 *C, u64 *xdig) {
    for (u64 e = 0; e < r * n * d * tmax; e++) xdig[e] = 0;
    for (u64 i = 0; i < n; i++) {
        u64 rest = i + n_start;
        for (u64 m = 0; rest > 0; m++) {
            const u64 dig = rest % b;
            rest = (rest - dig) / b;
            if (dig == 0) continue;
            for (u64 l = 0; l < r; l++)
                for (u64 j = 0; j < d; j++) {
                    u64 *out = xdig + l * n * d * tmax + i * d * tmax + j * tmax;
                    const u64 *col = C + l * d * mmax * tmax + j * mmax * tmax + m * tmax;
                    for (u64 t = 0; t < tmax; t++) out[t] = (out[t] + dig * col[t]) % b;
                }
        }
    }
}

/* gdn_linear_matrix_scramble.c:20-46: C_lms[l,j,k,t] = sum_c (S[l,j,t,c]*C[l%r_C,j,k,c]) % b,
 * with each product reduced but the sum left unreduced. */
void oracle_gdn_linear_matrix_scramble(u64 r, u64 d, u64 mmax, u64 r_C, u64 r_b,
                                       u64 tmax, u64 tmax_new, const u64 *bases,
                                       const u64 *S, const u64 *C, u64 *C_lms) {
    for (u64 l = 0; l < r; l++)
        for (u64 j = 0; j < d; j++) {
            const u64 b = bases[(l % r_b) * d + j];
            for (u64 k = 0; k < mmax; k++) {
                const u64 *col = C + (l % r_C) * d * mmax * tmax + j * mmax * tmax + k * tmax;
                u64 *out = C_lms + l * d * mmax * tmax_new + j * mmax * tmax_new + k * tmax_new;
                for (u64 t = 0; t < tmax_new; t++) {
                    const u64 *row = S + l * d * tmax_new * tmax + j * tmax_new * tmax + t * tmax;
                    u64 acc = 0;
                    for (u64 c = 0; c < tmax; c++) acc += (row[c] * col[c]) % b;
                    out[t] = acc;
                }
            }
        }
}

/* gdn_digital_permutation.c:20-45: digit t goes through perms[l,j,t,:];
 * digits t >= tmax take perms[l,j,t,0]. */
void oracle_gdn_digital_permutation(u64 r, u64 n, u64 d, u64 r_x, u64 r_b, u64 tmax,
                                    u64 tmax_new, u64 bmax, const u64 *perms,
                                    const u64 *xdig, u64 *xdig_new) {
    (void)r_b;
    for (u64 l = 0; l < r; l++)
        for (u64 i = 0; i < n; i++)
            for (u64 j = 0; j < d; j++) {
                const u64 *src = xdig + (l % r_x) * n * d * tmax + i * d * tmax + j * tmax;
                u64 *out = xdig_new + l * n * d * tmax_new + i * d * tmax_new + j * tmax_new;
                const u64 *pm = perms + l * d * tmax_new * bmax + j * tmax_new * bmax;
                for (u64 t = 0; t < tmax; t++) out[t] = pm[t * bmax + src[t]];
                for (u64 t = tmax; t < tmax_new; t++) out[t] = pm[t * bmax];
            }
}

/* gdn_digital_shift.c:20-45: (digit + shift) % b for t < tmax, raw shift after */
void oracle_gdn_digital_shift(u64 r, u64 n, u64 d, u64 r_x, u64 r_b, u64 tmax,
                              u64 tmax_new, const u64 *bases, const u64 *shifts,
                              const u64 *xdig, u64 *xdig_new) {
    for (u64 l = 0; l < r; l++)
        for (u64 i = 0; i < n; i++)
            for (u64 j = 0; j < d; j++) {
                const u64 *src = xdig + (l % r_x) * n * d * tmax + i * d * tmax + j * tmax;
                u64 *out = xdig_new + l * n * d * tmax_new + i * d * tmax_new + j * tmax_new;
                const u64 *sh = shifts + l * d * tmax_new + j * tmax_new;
                const u64 b = bases[(l % r_b) * d + j];
                for (u64 t = 0; t < tmax; t++) out[t] = (src[t] + sh[t]) % b;
                for (u64 t = tmax; t < tmax_new; t++) out[t] = sh[t];
            }
}

/* gdn_integer_to_float.c:17-43: v += recip*digit; recip /= b, in digit order */
void oracle_gdn_integer_to_float(u64 r, u64 n, u64 d, u64 r_b, u64 tmax,
                                 const u64 *bases, const u64 *xdig, double *x) {
    for (u64 l = 0; l < r; l++)
        for (u64 i = 0; i < n; i++)
            for (u64 j = 0; j < d; j++) {
                const u64 *src = xdig + l * n * d * tmax + i * d * tmax + j * tmax;
                const double b = (double)bases[(l % r_b) * d + j];
                double recip = 1 / b, v = 0.;
                for (u64 t = 0; t < tmax; t++) {
                    v += recip * (double)src[t];
                    recip /= b;
                }
                x[l * n * d + i * d + j] = v;
            }
}

/* gdn_interlace.c:18-39 */
void oracle_gdn_interlace(u64 r, u64 d_alpha, u64 mmax, u64 d, u64 tmax,
                          u64 tmax_alpha, u64 alpha, const u64 *C, u64 *C_alpha) {
    for (u64 l = 0; l < r; l++)
        for (u64 ja = 0; ja < d_alpha; ja++)
            for (u64 k = 0; k < mmax; k++)
                for (u64 ta = 0; ta < tmax_alpha; ta++) {
                    const u64 t = ta / alpha;
                    const u64 j = ja * alpha + ta % alpha;
                    C_alpha[l * d_alpha * mmax * tmax_alpha + ja * mmax * tmax_alpha + k * tmax_alpha + ta] =
                        C[l * d * mmax * tmax + j * mmax * tmax + k * tmax + t];
                }
}

/* gdn_undo_interlace.c:18-39 */
void oracle_gdn_undo_interlace(u64 r, u64 d, u64 mmax, u64 d_alpha, u64 tmax,
                               u64 tmax_alpha, u64 alpha, const u64 *C_alpha, u64 *C) {
    for (u64 l = 0; l < r; l++)
        for (u64 j = 0; j < d; j++)
            for (u64 k = 0; k < mmax; k++)
                for (u64 t = 0; t < tmax; t++) {
                    const u64 ja = j / alpha;
                    const u64 ta = t * alpha + j % alpha;
                    C[l * d * mmax * tmax + j * mmax * tmax + k * tmax + t] =
                        C_alpha[l * d_alpha * mmax * tmax_alpha + ja * mmax * tmax_alpha + k * tmax_alpha + ta];
                }
}

/* ----------------------------------------------------------- fast transforms */

/* butterfly partner pair for butterfly i of a stage with stride 2^s
 * (fwht_1d_radix2.c:31-38, fft_bro_1d_radix2.c:50-57) */
static void pair_for(u64 i, u64 s, u64 n_half, u64 *a, u64 *b) {
    const u64 f = (u64)(1 << s);
    if ((i >> s) & 1) {
        *b = i + n_half;
        *a = *b ^ f;
    } else {
        *a = i;
        *b = i ^ f;
    }
}

/* fwht_1d_radix2.c:15-52: m stages, largest stride first, (a +- b)/sqrt(2) */
void oracle_fwht_1d_radix2(u64 d1, u64 d2, u64 n_half, double *x) {
    const u64 n = 2 * n_half;
    const u64 m = (u64)log2((double)n);
    const double rt2 = sqrt(2.0);
    for (u64 stage = 0; stage < m; stage++) {
        const u64 s = m - stage - 1;
        for (u64 i = 0; i < n_half; i++) {
            u64 a, b;
            pair_for(i, s, n_half, &a, &b);
            for (u64 v = 0; v < d1 * d2; v++) {
                double *row = x + v * n;
                const double xa = row[a], xb = row[b];
                row[a] = (xa + xb) / rt2;
                row[b] = (xa - xb) / rt2;
            }
        }
    }
}

/* twiddle table of fft_bro_1d_radix2.c:32-42 / ifft_bro_1d_radix2.c:32-42;
 * sign = -2 for the forward transform, +2 for the inverse */
static void twiddles(u64 n_half, double sign, double *twr, double *twi) {
    const double pi = acos(-1.);
    const u64 n = 2 * n_half;
    for (u64 i = 0; i < n; i++) {
        const double ang = sign * pi * i / n;
        twr[i] = cos(ang);
        twi[i] = sin(ang);
    }
}

/* fft_bro_1d_radix2.c:18-79: decimation in time, bit-reversed input, stride
 * 2^k at stage k, twiddle index (a mod 2^k) << (m-1-k). */
void oracle_fft_bro_1d_radix2(u64 d1, u64 d2, u64 n_half, double *twr, double *twi,
                              double *xr, double *xi) {
    const u64 n = 2 * n_half;
    const u64 m = (u64)log2((double)n);
    const double rt2 = sqrt(2.0);
    twiddles(n_half, -2, twr, twi);
    for (u64 k = 0; k < m; k++) {
        const u64 s = m - k - 1;
        for (u64 i = 0; i < n_half; i++) {
            u64 a, b;
            pair_for(i, k, n_half, &a, &b);
            const u64 tw = (a % (u64)(1 << k)) * ((u64)1 << s);
            const double c = twr[tw], sn = twi[tw];
            for (u64 v = 0; v < d1 * d2; v++) {
                double *re = xr + v * n, *im = xi + v * n;
                const double ra = re[a], rb = re[b], ia = im[a], ib = im[b];
                const double yr = rb * c - ib * sn;
                const double yi = rb * sn + ib * c;
                re[a] = (ra + yr) / rt2;
                im[a] = (ia + yi) / rt2;
                re[b] = (ra - yr) / rt2;
                im[b] = (ia - yi) / rt2;
            }
        }
    }
}

/* ifft_bro_1d_radix2.c:18-80: decimation in frequency, natural input,
 * bit-reversed output, stride 2^(m-1-k) at stage k, twiddle (a mod 2^s) << k. */
void oracle_ifft_bro_1d_radix2(u64 d1, u64 d2, u64 n_half, double *twr, double *twi,
                               double *xr, double *xi) {
    const u64 n = 2 * n_half;
    const u64 m = (u64)log2((double)n);
    const double rt2 = sqrt(2.0);
    twiddles(n_half, 2, twr, twi);
    for (u64 k = 0; k < m; k++) {
        const u64 s = m - k - 1;
        for (u64 i = 0; i < n_half; i++) {
            u64 a, b;
            pair_for(i, s, n_half, &a, &b);
            const u64 tw = (a % (u64)(1 << s)) * ((u64)1 << k);
            const double c = twr[tw], sn = twi[tw];
            for (u64 v = 0; v < d1 * d2; v++) {
                double *re = xr + v * n, *im = xi + v * n;
                const double ra = re[a], rb = re[b], ia = im[a], ib = im[b];
                const double yr = ra - rb, yi = ia - ib;
                re[a] = (ra + rb) / rt2;
                im[a] = (ia + ib) / rt2;
                re[b] = (yr * c - yi * sn) / rt2;
                im[b] = (yr * sn + yi * c) / rt2;
            }
        }
    }
}

/* -------------------------------------------------------- generalized Halton */

/* halton_qrng.c:3-162.  res[j*n + i - n0] for the points i in [n0, n0 + n):
 * u = 0; for k = 31..0: u += (f*c_k + s_k) % b; u /= b, with c_k the base-b
 * digits of i (at most 32, halton_qrng.c:126-138), s_k = (int)(b * randu[32 j
 * + k]) kept as unsigned int (halton_qrng.c:86-92), f = 1 or the generalized
 * factor (halton_qrng.c:100-112).  With n0 == 0, point 0 comes from the set-up
 * loop with the unreduced s_k (halton_qrng.c:88-95) and the main loop starts at
 * 1 (halton_qrng.c:113-121).  Unlike the reference, nothing persists between
 * calls (its perm[] is static) and n == 0 writes nothing. */
void oracle_halton_qrng(int n, int d, int n0, int generalized, double *res,
                        const double *randu_d_32, const int *dvec) {
    unsigned primes[QMC_HALTON_DIMS];
    int found = 0;
    for (unsigned v = 2; found < QMC_HALTON_DIMS; v++) {
        int prime = 1;
        for (int t = 0; t < found && primes[t] * primes[t] <= v; t++)
            if (v % primes[t] == 0) { prime = 0; break; }
        if (prime) primes[found++] = v;
    }
    if (n <= 0) return;
    for (int j = 0; j < d; j++) {
        const int base = (int)primes[dvec[j]];
        const unsigned f = generalized ? qmc_halton_factor[dvec[j]] : 1u;
        unsigned sh[32];
        double u = 0;
        for (int k = 31; k >= 0; k--) {
            sh[k] = (unsigned)(int)(base * randu_d_32[j * 32 + k]);
            u += sh[k];
            u /= base;
        }
        if (n0 == 0) res[(long)j * n] = u;
        for (int i = (n0 == 0 ? 1 : n0); i < n + n0; i++) {
            unsigned coeff[32] = {0};
            unsigned tmp = (unsigned)i;
            for (int k = 0; tmp > 0 && k < 32; k++) {
                coeff[k] = tmp % (unsigned)base;
                tmp /= (unsigned)base;
            }
            double v = 0.0;
            for (int k = 31; k >= 0; k--) {
                v += (f * coeff[k] + sh[k]) % (unsigned)base;
                v /= base;
            }
            res[(long)j * n + (i - n0)] = v;
        }
    }
}
