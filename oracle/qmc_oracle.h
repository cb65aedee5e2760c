/*
 * CPU oracle for the QMC point-generation and transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library or the
 * Python package) may link, load or call this code; only the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark use it, as the checker.
 *
 * Every function is a loop-for-loop restatement of one reference C kernel
 * (qmctoolscl/c_funcs/<name>.c), with the same operation order so that
 * integer results and IEEE double results are bit-identical.  The reference's
 * batch-size arguments are dropped: the reference C backend always receives
 * batch size == problem size (qmctoolscl/util.py:107), so one call covers the
 * whole problem, which is what these functions do.
 */
#ifndef QMC_ORACLE_H
#define QMC_ORACLE_H
#include <stdint.h>

typedef uint64_t u64;

/* lattices */
void oracle_lat_gen_linear(u64 r, u64 n, u64 d, const u64 *g, double *x);
void oracle_lat_gen_natural(u64 r, u64 n, u64 d, u64 n_start, const u64 *g, double *x);
void oracle_lat_gen_gray(u64 r, u64 n, u64 d, u64 n_start, const u64 *g, double *x);
void oracle_lat_shift_mod_1(u64 r, u64 n, u64 d, u64 r_x, const double *x,
                            const double *shifts, double *xr);

/* base-2 digital nets */
void oracle_dnb2_gmat_lsb_to_msb(u64 r, u64 d, u64 mmax, const u64 *tmaxes,
                                 const u64 *C_lsb, u64 *C_msb);
void oracle_dnb2_linear_matrix_scramble(u64 r, u64 d, u64 mmax, u64 r_C, u64 tmax_new,
                                        const u64 *S, const u64 *C, u64 *C_lms);
void oracle_dnb2_gen_natural(u64 r, u64 n, u64 d, u64 n_start, u64 mmax,
                             const u64 *C, u64 *xb);
void oracle_dnb2_gen_gray(u64 r, u64 n, u64 d, u64 n_start, u64 mmax,
                          const u64 *C, u64 *xb);
void oracle_dnb2_digital_shift(u64 r, u64 n, u64 d, u64 r_x, const u64 *lshifts,
                               const u64 *xb, const u64 *shiftsb, u64 *xrb);
void oracle_dnb2_integer_to_float(u64 r, u64 n, u64 d, const u64 *tmaxes,
                                  const u64 *xb, double *x);
void oracle_dnb2_interlace(u64 r, u64 d_alpha, u64 mmax, u64 d, u64 tmax,
                           u64 tmax_alpha, u64 alpha, const u64 *C, u64 *C_alpha);
void oracle_dnb2_undo_interlace(u64 r, u64 d, u64 mmax, u64 d_alpha, u64 tmax,
                                u64 tmax_alpha, u64 alpha, const u64 *C_alpha, u64 *C);

/* generalized digital nets */
void oracle_gdn_gen_natural(u64 r, u64 n, u64 d, u64 r_b, u64 mmax, u64 tmax,
                            u64 n_start, const u64 *bases, const u64 *C, u64 *xdig);
void oracle_gdn_gen_natural_same_base(u64 r, u64 n, u64 d, u64 mmax, u64 tmax,
                                      u64 n_start, u64 b, const u64 *C, u64 *xdig);
void oracle_gdn_linear_matrix_scramble(u64 r, u64 d, u64 mmax, u64 r_C, u64 r_b,
                                       u64 tmax, u64 tmax_new, const u64 *bases,
                                       const u64 *S, const u64 *C, u64 *C_lms);
void oracle_gdn_digital_permutation(u64 r, u64 n, u64 d, u64 r_x, u64 r_b, u64 tmax,
                                    u64 tmax_new, u64 bmax, const u64 *perms,
                                    const u64 *xdig, u64 *xdig_new);
void oracle_gdn_digital_shift(u64 r, u64 n, u64 d, u64 r_x, u64 r_b, u64 tmax,
                              u64 tmax_new, const u64 *bases, const u64 *shifts,
                              const u64 *xdig, u64 *xdig_new);
void oracle_gdn_integer_to_float(u64 r, u64 n, u64 d, u64 r_b, u64 tmax,
                                 const u64 *bases, const u64 *xdig, double *x);
void oracle_gdn_interlace(u64 r, u64 d_alpha, u64 mmax, u64 d, u64 tmax,
                          u64 tmax_alpha, u64 alpha, const u64 *C, u64 *C_alpha);
void oracle_gdn_undo_interlace(u64 r, u64 d, u64 mmax, u64 d_alpha, u64 tmax,
                               u64 tmax_alpha, u64 alpha, const u64 *C_alpha, u64 *C);

/* fast transforms */
void oracle_fwht_1d_radix2(u64 d1, u64 d2, u64 n_half, double *x);
void oracle_fft_bro_1d_radix2(u64 d1, u64 d2, u64 n_half, double *twr, double *twi,
                              double *xr, double *xi);
void oracle_ifft_bro_1d_radix2(u64 d1, u64 d2, u64 n_half, double *twr, double *twi,
                               double *xr, double *xi);

/* generalized Halton (halton_qrng.c, int arguments as in the reference) */
void oracle_halton_qrng(int n, int d, int n0, int generalized, double *res,
                        const double *randu_d_32, const int *dvec);

#endif
