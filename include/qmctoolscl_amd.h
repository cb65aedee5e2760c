/*
 * libqmctoolscl_amd.so -- C ABI of the MI355X (gfx950) QMC point-generation
 * and transform library.
 *
 * Two families of entry points:
 *
 *  1. Drop-ins.  The 23 kernels of QMCToolsCL's C backend, with the exact
 *     prototypes of qmctoolscl/c_funcs/<name>.c (the same symbol names the
 *     reference's generated ctypes layer binds, pdm_build.py:79-103):
 *     host pointers, three sizes, three batch sizes (accepted and ignored:
 *     the GPU picks its own tiling), then the remaining arguments.  Arrays
 *     are staged to the current HIP device and back; the call returns when
 *     the outputs are in the caller's arrays.  They return void like the
 *     reference; a failure is reported through qmc_last_error_code() and
 *     qmc_last_error() (thread-local; every drop-in call clears it first).
 *
 *  2. Device entry points (*_dev and qmc_*_dev).  Same arguments without the
 *     batch sizes, device pointers, and a HIP stream; they enqueue and return
 *     a status code (QMC_OK on success) without synchronizing.
 *
 * Numerics: integer outputs (lattice index arithmetic, base-2 and base-b
 * digits) and the double outputs of the lattice, integer_to_float and FWHT
 * kernels are bit-identical to the reference C backend; FFT/IFFT outputs agree
 * to ~1e-15 relative (the twiddle table comes from device cos/sin).
 *
 * Inputs the reference leaves undefined (reads past an input array, bases
 * below 2, natural-order indices needing more than mmax columns, lattice
 * indices at or above 2^31) are rejected with QMC_ERR_RANGE instead.
 */
#ifndef QMCTOOLSCL_AMD_H
#define QMCTOOLSCL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QMC_EXPORT __attribute__((visibility("default")))

typedef struct ihipStream_t *qmc_stream_t; /* == hipStream_t */
typedef uint64_t qmc_u64;

enum {
    QMC_OK = 0,
    QMC_ERR_ARG = 1,
    QMC_ERR_RANGE = 2,
    QMC_ERR_HIP = 3,
    QMC_ERR_NOMEM = 4
};

/* ---------------------------------------------------------------- runtime */
QMC_EXPORT int qmc_init(void);                 /* set up per-device state on the current device */
QMC_EXPORT const char *qmc_version(void);
QMC_EXPORT const char *qmc_last_error(void);
QMC_EXPORT int qmc_last_error_code(void);
QMC_EXPORT void qmc_clear_error(void);
/* read (and optionally clear) the current device's status bits that kernels
 * set on inputs the reference leaves undefined; synchronizes the device */
QMC_EXPORT int qmc_device_status(unsigned *bits, int clear);
/* clear the status bits on stream s, ordered with the caller's launches (the
 * Python tensor path clears them before a call and reads them after it) */
QMC_EXPORT int qmc_status_reset_dev(qmc_stream_t s);

/* ------------------------------------------------- drop-ins: lattices
 * lat_gen_linear.c:3-13, lat_gen_natural.c:3-14, lat_gen_gray.c:3-14,
 * lat_shift_mod_1.c:3-15 */
QMC_EXPORT void lat_gen_linear(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                               const qmc_u64 *g, double *x);
QMC_EXPORT void lat_gen_natural(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                qmc_u64 n_start, const qmc_u64 *g, double *x);
QMC_EXPORT void lat_gen_gray(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                             qmc_u64 n_start, const qmc_u64 *g, double *x);
QMC_EXPORT void lat_shift_mod_1(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                qmc_u64 r_x, const double *x, const double *shifts, double *xr);

/* ------------------------------------------- drop-ins: base-2 digital nets
 * dnb2_gmat_lsb_to_msb.c:3-14, dnb2_linear_matrix_scramble.c:3-16,
 * dnb2_gen_natural.c:3-15, dnb2_gen_gray.c:3-15, dnb2_digital_shift.c:3-16,
 * dnb2_integer_to_float.c:3-14, dnb2_interlace.c:3-17, dnb2_undo_interlace.c:3-17 */
QMC_EXPORT void dnb2_gmat_lsb_to_msb(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 bs_r, qmc_u64 bs_d,
                                     qmc_u64 bs_mmax, const qmc_u64 *tmaxes, const qmc_u64 *C_lsb, qmc_u64 *C_msb);
QMC_EXPORT void dnb2_linear_matrix_scramble(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 bs_r, qmc_u64 bs_d,
                                            qmc_u64 bs_mmax, qmc_u64 r_C, qmc_u64 tmax_new, const qmc_u64 *S,
                                            const qmc_u64 *C, qmc_u64 *C_lms);
QMC_EXPORT void dnb2_gen_natural(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                 qmc_u64 n_start, qmc_u64 mmax, const qmc_u64 *C, qmc_u64 *xb);
QMC_EXPORT void dnb2_gen_gray(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                              qmc_u64 n_start, qmc_u64 mmax, const qmc_u64 *C, qmc_u64 *xb);
QMC_EXPORT void dnb2_digital_shift(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                   qmc_u64 r_x, const qmc_u64 *lshifts, const qmc_u64 *xb, const qmc_u64 *shiftsb,
                                   qmc_u64 *xrb);
QMC_EXPORT void dnb2_integer_to_float(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                      const qmc_u64 *tmaxes, const qmc_u64 *xb, double *x);
QMC_EXPORT void dnb2_interlace(qmc_u64 r, qmc_u64 d_alpha, qmc_u64 mmax, qmc_u64 bs_r, qmc_u64 bs_d_alpha,
                               qmc_u64 bs_mmax, qmc_u64 d, qmc_u64 tmax, qmc_u64 tmax_alpha, qmc_u64 alpha,
                               const qmc_u64 *C, qmc_u64 *C_alpha);
QMC_EXPORT void dnb2_undo_interlace(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 bs_r, qmc_u64 bs_d, qmc_u64 bs_mmax,
                                    qmc_u64 d_alpha, qmc_u64 tmax, qmc_u64 tmax_alpha, qmc_u64 alpha,
                                    const qmc_u64 *C_alpha, qmc_u64 *C);

/* ----------------------------------- drop-ins: generalized digital nets
 * gdn_gen_natural.c:3-18, gdn_gen_natural_same_base.c:3-17,
 * gdn_linear_matrix_scramble.c:3-19, gdn_digital_permutation.c:3-19,
 * gdn_digital_shift.c:3-19, gdn_integer_to_float.c:3-16, gdn_interlace.c:3-17,
 * gdn_undo_interlace.c:3-17 */
QMC_EXPORT void gdn_gen_natural(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                qmc_u64 r_b, qmc_u64 mmax, qmc_u64 tmax, qmc_u64 n_start, const qmc_u64 *bases,
                                const qmc_u64 *C, qmc_u64 *xdig);
QMC_EXPORT void gdn_gen_natural_same_base(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n,
                                          qmc_u64 bs_d, qmc_u64 mmax, qmc_u64 tmax, qmc_u64 n_start, qmc_u64 b,
                                          const qmc_u64 *C, qmc_u64 *xdig);
QMC_EXPORT void gdn_linear_matrix_scramble(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 bs_r, qmc_u64 bs_d,
                                           qmc_u64 bs_mmax, qmc_u64 r_C, qmc_u64 r_b, qmc_u64 tmax,
                                           qmc_u64 tmax_new, const qmc_u64 *bases, const qmc_u64 *S,
                                           const qmc_u64 *C, qmc_u64 *C_lms);
QMC_EXPORT void gdn_digital_permutation(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                        qmc_u64 r_x, qmc_u64 r_b, qmc_u64 tmax, qmc_u64 tmax_new, qmc_u64 bmax,
                                        const qmc_u64 *perms, const qmc_u64 *xdig, qmc_u64 *xdig_new);
QMC_EXPORT void gdn_digital_shift(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                  qmc_u64 r_x, qmc_u64 r_b, qmc_u64 tmax, qmc_u64 tmax_new, const qmc_u64 *bases,
                                  const qmc_u64 *shifts, const qmc_u64 *xdig, qmc_u64 *xdig_new);
QMC_EXPORT void gdn_integer_to_float(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 bs_r, qmc_u64 bs_n, qmc_u64 bs_d,
                                     qmc_u64 r_b, qmc_u64 tmax, const qmc_u64 *bases, const qmc_u64 *xdig, double *x);
QMC_EXPORT void gdn_interlace(qmc_u64 r, qmc_u64 d_alpha, qmc_u64 mmax, qmc_u64 bs_r, qmc_u64 bs_d_alpha,
                              qmc_u64 bs_mmax, qmc_u64 d, qmc_u64 tmax, qmc_u64 tmax_alpha, qmc_u64 alpha,
                              const qmc_u64 *C, qmc_u64 *C_alpha);
QMC_EXPORT void gdn_undo_interlace(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 bs_r, qmc_u64 bs_d, qmc_u64 bs_mmax,
                                   qmc_u64 d_alpha, qmc_u64 tmax, qmc_u64 tmax_alpha, qmc_u64 alpha,
                                   const qmc_u64 *C_alpha, qmc_u64 *C);

/* --------------------------------------------- drop-ins: fast transforms
 * fwht_1d_radix2.c:3-14, fft_bro_1d_radix2.c:3-17, ifft_bro_1d_radix2.c:3-17 */
QMC_EXPORT void fwht_1d_radix2(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, qmc_u64 bs_d1, qmc_u64 bs_d2,
                               qmc_u64 bs_n_half, double *x);
QMC_EXPORT void fft_bro_1d_radix2(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, qmc_u64 bs_d1, qmc_u64 bs_d2,
                                  qmc_u64 bs_n_half, double *twiddler, double *twiddlei, double *xr, double *xi);
QMC_EXPORT void ifft_bro_1d_radix2(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, qmc_u64 bs_d1, qmc_u64 bs_d2,
                                   qmc_u64 bs_n_half, double *twiddler, double *twiddlei, double *xr, double *xi);

/* ------------------------------------------- drop-in: generalized Halton
 * halton_qrng.c:3-24 (bound by the reference's pdm_build.py:85-94 as
 * halton_qrng_c, argtypes c_int x 4, double *, double *, int *).  res is
 * d x n (res[j n + i] = point n0 + i in dimension j), randu_d_32 is d x 32,
 * dvec holds d prime indices below 360.  No state is kept between calls. */
QMC_EXPORT void halton_qrng(int n, int d, int n0, int generalized, double *res, double *randu_d_32, int *dvec);

/* ----------------------------------------------- device entry points
 * Each replaces the drop-in of the same name (same citation) for callers
 * whose buffers already live on the GPU. */
QMC_EXPORT int lat_gen_linear_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, const qmc_u64 *g, double *x, qmc_stream_t s);
QMC_EXPORT int lat_gen_natural_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 n_start, const qmc_u64 *g, double *x,
                                   qmc_stream_t s);
QMC_EXPORT int lat_gen_gray_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 n_start, const qmc_u64 *g, double *x,
                                qmc_stream_t s);
QMC_EXPORT int lat_shift_mod_1_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_x, const double *x,
                                   const double *shifts, double *xr, qmc_stream_t s);
QMC_EXPORT int dnb2_gmat_lsb_to_msb_dev(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, const qmc_u64 *tmaxes,
                                        const qmc_u64 *C_lsb, qmc_u64 *C_msb, qmc_stream_t s);
QMC_EXPORT int dnb2_linear_matrix_scramble_dev(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 r_C, qmc_u64 tmax_new,
                                               const qmc_u64 *S, const qmc_u64 *C, qmc_u64 *C_lms, qmc_stream_t s);
QMC_EXPORT int dnb2_gen_natural_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 n_start, qmc_u64 mmax,
                                    const qmc_u64 *C, qmc_u64 *xb, qmc_stream_t s);
QMC_EXPORT int dnb2_gen_gray_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 n_start, qmc_u64 mmax, const qmc_u64 *C,
                                 qmc_u64 *xb, qmc_stream_t s);
QMC_EXPORT int dnb2_digital_shift_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_x, const qmc_u64 *lshifts,
                                      const qmc_u64 *xb, const qmc_u64 *shiftsb, qmc_u64 *xrb, qmc_stream_t s);
QMC_EXPORT int dnb2_integer_to_float_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, const qmc_u64 *tmaxes,
                                         const qmc_u64 *xb, double *x, qmc_stream_t s);
QMC_EXPORT int dnb2_interlace_dev(qmc_u64 r, qmc_u64 d_alpha, qmc_u64 mmax, qmc_u64 d, qmc_u64 tmax,
                                  qmc_u64 tmax_alpha, qmc_u64 alpha, const qmc_u64 *C, qmc_u64 *C_alpha,
                                  qmc_stream_t s);
QMC_EXPORT int dnb2_undo_interlace_dev(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 d_alpha, qmc_u64 tmax,
                                       qmc_u64 tmax_alpha, qmc_u64 alpha, const qmc_u64 *C_alpha, qmc_u64 *C,
                                       qmc_stream_t s);
QMC_EXPORT int gdn_gen_natural_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_b, qmc_u64 mmax, qmc_u64 tmax,
                                   qmc_u64 n_start, const qmc_u64 *bases, const qmc_u64 *C, qmc_u64 *xdig,
                                   qmc_stream_t s);
QMC_EXPORT int gdn_gen_natural_same_base_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 mmax, qmc_u64 tmax,
                                             qmc_u64 n_start, qmc_u64 b, const qmc_u64 *C, qmc_u64 *xdig,
                                             qmc_stream_t s);
QMC_EXPORT int gdn_linear_matrix_scramble_dev(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 r_C, qmc_u64 r_b,
                                              qmc_u64 tmax, qmc_u64 tmax_new, const qmc_u64 *bases,
                                              const qmc_u64 *S, const qmc_u64 *C, qmc_u64 *C_lms, qmc_stream_t s);
QMC_EXPORT int gdn_digital_permutation_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_x, qmc_u64 r_b, qmc_u64 tmax,
                                           qmc_u64 tmax_new, qmc_u64 bmax, const qmc_u64 *perms,
                                           const qmc_u64 *xdig, qmc_u64 *xdig_new, qmc_stream_t s);
QMC_EXPORT int gdn_digital_shift_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_x, qmc_u64 r_b, qmc_u64 tmax,
                                     qmc_u64 tmax_new, const qmc_u64 *bases, const qmc_u64 *shifts,
                                     const qmc_u64 *xdig, qmc_u64 *xdig_new, qmc_stream_t s);
QMC_EXPORT int gdn_integer_to_float_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_b, qmc_u64 tmax,
                                        const qmc_u64 *bases, const qmc_u64 *xdig, double *x, qmc_stream_t s);
QMC_EXPORT int gdn_interlace_dev(qmc_u64 r, qmc_u64 d_alpha, qmc_u64 mmax, qmc_u64 d, qmc_u64 tmax,
                                 qmc_u64 tmax_alpha, qmc_u64 alpha, const qmc_u64 *C, qmc_u64 *C_alpha,
                                 qmc_stream_t s);
QMC_EXPORT int gdn_undo_interlace_dev(qmc_u64 r, qmc_u64 d, qmc_u64 mmax, qmc_u64 d_alpha, qmc_u64 tmax,
                                      qmc_u64 tmax_alpha, qmc_u64 alpha, const qmc_u64 *C_alpha, qmc_u64 *C,
                                      qmc_stream_t s);
QMC_EXPORT int fwht_1d_radix2_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *x, qmc_stream_t s);
QMC_EXPORT int fft_bro_1d_radix2_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *twiddler, double *twiddlei,
                                     double *xr, double *xi, qmc_stream_t s);
QMC_EXPORT int ifft_bro_1d_radix2_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *twiddler, double *twiddlei,
                                      double *xr, double *xi, qmc_stream_t s);
QMC_EXPORT int qmc_halton_qrng_dev(int n, int d, int n0, int generalized, double *res, const double *randu_d_32,
                                   const int *dvec, qmc_stream_t s);

/* ------------------------------------------------- fast QMC kernel methods
 * No single reference symbol: the reference's documentation chains its
 * transforms on a kernel matrix's first column this way (docs/index.md:
 * 1825-1883 with fft_bro/ifft_bro, 1929-1977 with fwht).  The column k of
 * the Gram matrix of a product kernel on natural-order points (lattice:
 * prod_j (1 + coef_j B_2alpha(x_ij)), alpha = 1, 2; base-2 net: prod_j
 * (1 + coef_j w(X_ij)), w(X) = 1 - 3 2^(bitlen X - 1 - tmax), w(0) = 1) is
 * evaluated from g or C without writing the points.  *_fft / *_fwht write the
 * orthonormal transform kt of k (r vectors of n); *_apply overwrite y with
 * K y (mode 0) or K^-1 y (mode 1) as ifft(fft(y) kt) sqrt(n), ifft(fft(y)/kt)
 * / sqrt(n) (fwht for nets).  n is a power of two, r <= 65535. */
QMC_EXPORT int qmc_lat_kernel_column_dev(int alpha, qmc_u64 r, qmc_u64 n, qmc_u64 d, const qmc_u64 *g,
                                         const double *coef, double *k, qmc_stream_t s);
QMC_EXPORT int qmc_dnb2_kernel_column_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 mmax, int tmax, const qmc_u64 *C,
                                          const double *coef, double *k, qmc_stream_t s);
QMC_EXPORT int qmc_lat_kernel_fft_dev(int alpha, qmc_u64 r, qmc_u64 n, qmc_u64 d, const qmc_u64 *g,
                                      const double *coef, double *twiddler, double *twiddlei, double *kr, double *ki,
                                      qmc_stream_t s);
QMC_EXPORT int qmc_dnb2_kernel_fwht_dev(qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 mmax, int tmax, const qmc_u64 *C,
                                        const double *coef, double *kt, qmc_stream_t s);
QMC_EXPORT int qmc_lat_kernel_apply_dev(int mode, qmc_u64 r, qmc_u64 n, const double *kr, const double *ki,
                                        double *twiddler, double *twiddlei, double *yr, double *yi, qmc_stream_t s);
QMC_EXPORT int qmc_dnb2_kernel_apply_dev(int mode, qmc_u64 r, qmc_u64 n, const double *kt, double *y,
                                         qmc_stream_t s);

/* ------------------------------------------------------- fused pipelines
 * No single reference symbol: each equals a chain of reference calls as the
 * doctests chain them (docs/index.md:64-208, 419-677), in one HBM pass. */

/* x[0 .. n) = v with one 16-byte store per thread over a full grid (x 16-byte
 * aligned, n even): no reference counterpart; the plain-store write rate the
 * generators are measured against. */
QMC_EXPORT int qmc_fill_dev(double *x, qmc_u64 n, double v, qmc_stream_t s);

/* lat_gen_<order>(r_g, ...) then lat_shift_mod_1(r, ..., r_x = r_g, ...);
 * order 0 linear, 1 natural, 2 gray */
QMC_EXPORT int qmc_lat_gen_shift_dev(int order, qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_g, qmc_u64 n_start,
                                     const qmc_u64 *g, const double *shifts, double *xr, qmc_stream_t s);
/* dnb2_gen_<natural|gray>(r_x, ...) -> dnb2_digital_shift(r, ..., r_x, ...) ->
 * dnb2_integer_to_float(r, ...); xrb and/or x may be NULL */
QMC_EXPORT int qmc_dnb2_gen_shift_float_dev(int natural, qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_x,
                                            qmc_u64 n_start, qmc_u64 mmax, const qmc_u64 *C,
                                            const qmc_u64 *lshifts, const qmc_u64 *shiftsb,
                                            const qmc_u64 *tmaxes, qmc_u64 *xrb, double *x, qmc_stream_t s);
/* dnb2_linear_matrix_scramble(r, d, mmax, r_C, tmax_new, S, C) ->
 * dnb2_gen_<natural|gray>(r, ...) -> dnb2_digital_shift(r, ..., r_x = r, ...) ->
 * dnb2_integer_to_float(r, ...) in one launch (the scrambled columns never
 * leave LDS); S (r, d, tmax_new), C (r_C, d, mmax); needs mmax, tmax_new <= 64,
 * d <= 256 and d * (2 mmax + tmax_new) <= 6144; xrb and/or x may be NULL.  Replaces the chain of
 * dnb2_linear_matrix_scramble.c:3-16, dnb2_gen_natural.c:3-15 (or
 * dnb2_gen_gray.c:3-15), dnb2_digital_shift.c:3-16, dnb2_integer_to_float.c:3-14. */
QMC_EXPORT int qmc_dnb2_lms_gen_shift_float_dev(int natural, qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_C,
                                                qmc_u64 n_start, qmc_u64 mmax, qmc_u64 tmax_new,
                                                const qmc_u64 *S, const qmc_u64 *C, const qmc_u64 *lshifts,
                                                const qmc_u64 *shiftsb, const qmc_u64 *tmaxes, qmc_u64 *xrb,
                                                double *x, qmc_stream_t s);

/* gdn_gen_natural(r, ...) -> mode 1: gdn_digital_permutation(r, ..., r_x = r, ...)
 * or mode 2: gdn_digital_shift(r, ..., r_x = r, ...) -> gdn_integer_to_float,
 * writing only x (r, n, d); mode 0 converts the generated digits directly
 * (tmax_new == tmax).  tab: perms (r, d, tmax_new, bmax) or shifts
 * (r, d, tmax_new).  Needs 1 <= tmax <= tmax_new <= 64 and mmax <= 64. */
QMC_EXPORT int qmc_gdn_gen_scramble_float_dev(int mode, qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_b,
                                              qmc_u64 mmax, qmc_u64 tmax, qmc_u64 tmax_new, qmc_u64 n_start,
                                              qmc_u64 bmax, const qmc_u64 *bases, const qmc_u64 *C,
                                              const qmc_u64 *tab, double *x, qmc_stream_t s);
/* No counterpart in the reference.  The same call as
 * qmc_gdn_gen_scramble_float_dev with identical results, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops): it calls no hipMalloc/hipFree and no allocating library routine.
 * Its workspace is a 4 MiB array per device that the library holds until the
 * process ends; a call needing more than 2^20 words of it
 * (r*d*mmax*tmax + r*d) fails with QMC_ERR_ARG before any launch.  Calls that
 * overlap on one device share that array and race. */
QMC_EXPORT int qmc_gdn_gen_scramble_float_noalloc_dev(int mode, qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_b,
                                                      qmc_u64 mmax, qmc_u64 tmax, qmc_u64 tmax_new,
                                                      qmc_u64 n_start, qmc_u64 bmax, const qmc_u64 *bases,
                                                      const qmc_u64 *C, const qmc_u64 *tab, double *x,
                                                      qmc_stream_t s);

/* ------------------------------------------- nested uniform scrambling
 * Replaces the pure-Python dnb2_nested_uniform_scramble (rand_funcs.py:162-266)
 * and gdn_nested_uniform_scramble (rand_funcs.py:274-356).  A tree per
 * (replication, dimension) is grown on the host in point order; *_insert
 * returns the sizes of the random draws the new leaves need, in the order the
 * reference draws them (base 2: bit widths, for rng.integers(0, 2**w); base b:
 * row counts, for that many rng.permutation(b) rows); the caller appends the
 * draw values (the root's first) and *_apply_dev scrambles on the GPU.
 * Base-b rows are 16-bit (b <= 2^16).  Trees persist across calls until
 * freed; each *_apply_dev keeps the packed device copy of the last set of
 * trees it was given and uploads again only when a tree grew.
 *
 * Base b with numpy's SFC64 generators: qmc_nusb_grow inserts a call's points
 * into all r*d trees on host threads and replays each tree's generator
 * (Generator.permutation's draws, rand_funcs.py:22-32) from its state, so no
 * row is drawn in Python; a tree materializes only its internal nodes' rows
 * and keeps one generator state per leaf, whose rows the device regenerates.
 * With other bit generators the caller draws the rows (*_insert counts,
 * *_add_rows) and all of them are materialized. */
QMC_EXPORT void *qmc_nus2_tree_new(int tmax_new);
QMC_EXPORT void qmc_nus2_tree_free(void *tree);
QMC_EXPORT qmc_u64 qmc_nus2_tree_nodes(void *tree);
QMC_EXPORT qmc_u64 qmc_nus2_tree_draws(void *tree);
QMC_EXPORT int64_t qmc_nus2_insert(void *tree, const qmc_u64 *v, qmc_u64 n, qmc_u64 stride, int tdelta,
                                   uint32_t *widths);
QMC_EXPORT int qmc_nus2_add_draws(void *tree, const qmc_u64 *vals, qmc_u64 count);
QMC_EXPORT int qmc_nus2_apply_dev(void **trees, qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_x, int tmax,
                                  int tmax_new, const qmc_u64 *xb, qmc_u64 *xrb, qmc_stream_t s);
QMC_EXPORT void *qmc_nusb_tree_new(int tmax, int tmax_new, qmc_u64 b);
QMC_EXPORT void qmc_nusb_tree_free(void *tree);
QMC_EXPORT qmc_u64 qmc_nusb_tree_rows(void *tree);
QMC_EXPORT qmc_u64 qmc_nusb_tree_rows_needed(void *tree);
QMC_EXPORT int64_t qmc_nusb_insert(void *tree, const qmc_u64 *v, qmc_u64 n, qmc_u64 stride_pt,
                                   qmc_u64 stride_dig, uint32_t *counts);
QMC_EXPORT int qmc_nusb_add_rows(void *tree, const uint16_t *rows, qmc_u64 nrows);
QMC_EXPORT qmc_u64 qmc_nusb_tree_states(void *tree);
QMC_EXPORT qmc_u64 qmc_nusb_tree_nodes(void *tree);
QMC_EXPORT qmc_u64 qmc_nusb_tree_bytes(void *tree);
QMC_EXPORT int qmc_nusb_root_rows(void *tree, uint16_t *out);
/* trees[l*d + j] takes xdig[(l % r_x), i, j, 0..tmax) for i < n; where
 * replay[k] != 0 the new rows come from the SFC64 state states[5k..5k+5)
 * (numpy's s[0..3], has_uint32 << 32 | uinteger), advanced in place. */
QMC_EXPORT int qmc_nusb_grow(void **trees, qmc_u64 nt, const qmc_u64 *xdig, qmc_u64 r, qmc_u64 n, qmc_u64 d,
                             qmc_u64 r_x, qmc_u64 *states, const uint8_t *replay, int threads);
/* count Generator.permutation(b) draws from SFC64 state w (advanced in place)
 * into out (count x b): the replay qmc_nusb_grow uses, exported for checks */
QMC_EXPORT void qmc_nus_sfc64_permutations(qmc_u64 *w, qmc_u64 b, qmc_u64 count, uint16_t *out);
QMC_EXPORT int qmc_nusb_apply_dev(void **trees, qmc_u64 r, qmc_u64 n, qmc_u64 d, qmc_u64 r_x, qmc_u64 r_b,
                                  int tmax, int tmax_new, const qmc_u64 *bases, const qmc_u64 *xdig,
                                  qmc_u64 *xrdig, qmc_stream_t s);

/* ------------------------------------------------ matrix-core transforms
 * Reduced-precision counterpart of fwht_1d_radix2 (fwht_1d_radix2.c:3-14):
 * 16-point Hadamard tiles on bf16 MFMA from a 3 x bf16 split, fp32
 * accumulation; relative L2 error ~1e-7 vs float64.  Length 2^8 .. 2^24;
 * `work` (fp32, d1*d2*2*n_half) may be NULL. */
QMC_EXPORT int qmc_fwht_mfma_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *x, float *work, qmc_stream_t s);
/* Reduced-precision counterpart of fft_bro_1d_radix2 (fft_bro_1d_radix2.c:3-17):
 * bit-reversed input, natural output, orthonormal; radix-16 DFT tiles
 * (FP64-generated, 3 x bf16 split) on MFMA, fp32 accumulation and fp32
 * inter-level twiddles.  Length 2^8 .. 2^28; `work` (fp32, 2*d1*d2*2*n_half)
 * may be NULL.  Writes no twiddle table. */
QMC_EXPORT int qmc_fft_bro_mfma_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *xr, double *xi, float *work,
                                    qmc_stream_t s);
QMC_EXPORT int qmc_ifft_bro_mfma_dev(qmc_u64 d1, qmc_u64 d2, qmc_u64 n_half, double *xr, double *xi, float *work,
                                     qmc_stream_t s); /* ifft_bro_1d_radix2.c:3-17, reduced precision */
/* one 16x16x32 bf16 MFMA on lane-major operands (layout check) */
QMC_EXPORT int qmc_mfma_probe_dev(const float *a, const float *b, float *d, qmc_stream_t s);

/* ------------------------------------------------------- batched requests
 * No reference counterpart: the reference runs one call per kernel per
 * request (util.py:99-144).  A plan holds nreq requests (method 0 lattice +
 * shifts, 1 dnb2 LMS + digital shift, 2 dnb2 + hash-based nested scramble,
 * 3 Halton-type base-b LMS to 52 digits + uniform digit permutations; r x 2^m
 * x d float64 points, natural order, counter-based random inputs from
 * seeds[q]).  Outputs go to a ring of ring_elems doubles, each replication
 * slab contiguous at qmc_batch_slab_offset(plan, q, l).  points_per_tile
 * and gdn_points_per_tile (0 = the same) size the tiles.  gdn_level_cap > 0
 * limits the permutation levels gdn tiles read from their tables (the output
 * does not change; a test hook for the other path).  Calls on one plan must
 * not overlap. */
QMC_EXPORT void *qmc_batch_plan(int nreq, const int32_t *methods, const int32_t *m, const int32_t *d,
                                const int32_t *r, const qmc_u64 *seeds, qmc_u64 points_per_tile, int grid,
                                qmc_u64 ring_elems, qmc_u64 gdn_points_per_tile, int gdn_level_cap);
/* the same with a tile size per method: tile_points[method], method 0..3 */
QMC_EXPORT void *qmc_batch_plan_tiles(int nreq, const int32_t *methods, const int32_t *m, const int32_t *d,
                                      const int32_t *r, const qmc_u64 *seeds, const qmc_u64 *tile_points, int grid,
                                      qmc_u64 ring_elems, int gdn_level_cap);
QMC_EXPORT void qmc_batch_free(void *plan);
QMC_EXPORT qmc_u64 qmc_batch_total_tiles(void *plan);
QMC_EXPORT qmc_u64 qmc_batch_points(void *plan);
QMC_EXPORT qmc_u64 qmc_batch_slab_offset(void *plan, int q, int l);
/* CUs given to the gdn class while the store-bound class runs beside it */
QMC_EXPORT int qmc_batch_set_gdn_cus(void *plan, int n);
QMC_EXPORT int qmc_batch_run_dev(void *plan, double *ring, qmc_stream_t s);          /* persistent launch */
QMC_EXPORT int qmc_batch_graph_run_dev(void *plan, double *ring, qmc_stream_t s);    /* same, hipGraph replay */
QMC_EXPORT int qmc_batch_run_per_call_dev(void *plan, double *ring, qmc_stream_t s); /* one launch per request */
/* Digests: qmc_batch_slabs() = sum of r; with qmc_batch_set_digest(plan, D) (D:
 * that many device u64, zeroed by the caller; NULL turns it off) later runs add
 * into D[slab] the wrap-around sum of the stored float64 bit patterns of each
 * replication slab, so batched and per-call runs can be compared exactly at
 * any size, whatever the ring placement. */
QMC_EXPORT qmc_u64 qmc_batch_slabs(void *plan);
QMC_EXPORT int qmc_batch_set_digest(void *plan, unsigned long long *digest);
QMC_EXPORT void qmc_batch_lattice_inputs(qmc_u64 seed, int m, int d, int r, qmc_u64 *g, double *shifts);
QMC_EXPORT void qmc_batch_dnb2_inputs(qmc_u64 seed, int method, int d, int r, qmc_u64 *C, qmc_u64 *S,
                                      qmc_u64 *shifts);
/* gdn request inputs for columns [j0, j0 + dj) in the reference's layouts:
 * bases[dj], S[r][dj][52][m] (gdn_get_linear_scramble_matrix, tmax = m),
 * perms[r][dj][52][bmax] (gdn_get_digital_permutations); identity C. */
QMC_EXPORT int qmc_batch_gdn_inputs(qmc_u64 seed, int m, int j0, int dj, int r, qmc_u64 *bases, qmc_u64 *S,
                                    qmc_u64 *perms);
/* per-base constants of the gdn sums: rec[52], early-exit thresholds thr[52],
 * the levels kept in LDS and the first level that can end a sum */
QMC_EXPORT int qmc_batch_gdn_base(int j, double *rec, double *thr, int *levels, int *tmin);
QMC_EXPORT qmc_u64 qmc_owen_hash64(qmc_u64 x, qmc_u64 seed);

#ifdef __cplusplus
}
#endif
#endif /* QMCTOOLSCL_AMD_H */
