// C-ABI of libqmctoolscl_amd.so: error reporting, the per-device status word,
// and the 23 host-pointer drop-in symbols with the reference C backend's exact
// prototypes (qmctoolscl/c_funcs/<name>.c).  A drop-in call copies its arrays
// to the current HIP device, runs the *_dev entry point on a library stream,
// waits, and copies the outputs back into the caller's arrays, so host code
// written against the reference's c_lib keeps working unchanged.  Errors are
// reported through qmc_last_error_code() / qmc_last_error(), since the
// reference prototypes return void.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "qmc_common.hpp"

namespace {

thread_local char g_err[1024] = "";
thread_local int g_err_code = QMC_OK;

std::mutex g_mu;
constexpr int kMaxDevices = 64;
u32 *g_status[kMaxDevices] = {};
hipStream_t g_stream[kMaxDevices] = {};

}  // namespace

int gdn_scratch_setup();

int qmc_fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    g_err_code = code;
    return code;
}

int qmc_fail_hip(hipError_t e, const char *what) {
    return qmc_fail(QMC_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
}

// Device word that kernels OR status bits into when they meet input the
// reference leaves undefined (a base below 2, a read past the end of an input).
// It is a device symbol, zero when the code object loads, so no entry point
// allocates memory for it; each device holds its own copy, found once per
// device with hipGetSymbolAddress.
static __device__ __align__(256) u32 status_word[64];

u32 *qmc_status_ptr() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        qmc_fail(QMC_ERR_HIP, "no usable HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_status[dev]) {
        u32 *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(status_word));
        if (e != hipSuccess || !p) {
            qmc_fail(QMC_ERR_HIP, "cannot find the device status word: %s", hipGetErrorString(e));
            return nullptr;
        }
        g_status[dev] = p;
    }
    return g_status[dev];
}

__global__ void status_reset_kernel(u32 *p) { p[0] = 0; }

static hipStream_t lib_stream() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_stream[dev]) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        g_stream[dev] = s;
    }
    return g_stream[dev];
}

extern "C" {

QMC_EXPORT const char *qmc_last_error(void) { return g_err; }
QMC_EXPORT int qmc_last_error_code(void) { return g_err_code; }
QMC_EXPORT void qmc_clear_error(void) {
    g_err[0] = 0;
    g_err_code = QMC_OK;
}
QMC_EXPORT const char *qmc_version(void) { return "qmctoolscl_amd 0.1.0 (gfx950)"; }

QMC_EXPORT int qmc_init(void) {
    if (!qmc_status_ptr()) return g_err_code ? g_err_code : QMC_ERR_HIP;
    if (!lib_stream()) return qmc_fail(QMC_ERR_HIP, "cannot create the library stream");
    return gdn_scratch_setup();
}

// Clears the status word on stream s (a one-thread kernel, so the clear is
// ordered with the caller's launches and can sit in a captured graph).
QMC_EXPORT int qmc_status_reset_dev(qmc_stream_t s) {
    u32 *p = qmc_status_ptr();
    if (!p) return g_err_code;
    hipLaunchKernelGGL(status_reset_kernel, dim3(1), dim3(1), 0, (hipStream_t)s, p);
    QMC_HIP_CHECK(hipGetLastError());
    return QMC_OK;
}

QMC_EXPORT int qmc_device_status(unsigned *bits, int clear) {
    u32 *p = qmc_status_ptr();
    if (!p) return g_err_code;
    u32 v = 0;
    QMC_HIP_CHECK(hipDeviceSynchronize());
    QMC_HIP_CHECK(hipMemcpy(&v, p, sizeof v, hipMemcpyDeviceToHost));
    if (clear) QMC_HIP_CHECK(hipMemset(p, 0, sizeof v));
    if (bits) *bits = v;
    return QMC_OK;
}

}  // extern "C"

// ------------------------------------------------------------ host drop-ins

namespace {

// Stages host arrays through stream-ordered device allocations.
class Staging {
  public:
    Staging() : s_(lib_stream()) {
        if (!s_) {
            fail_ = qmc_fail(QMC_ERR_HIP, "no HIP stream (is a GPU visible?)");
            return;
        }
        // status bits left by earlier launches are not this call's
        u32 *st = qmc_status_ptr();
        if (st) hipLaunchKernelGGL(status_reset_kernel, dim3(1), dim3(1), 0, s_, st);
    }
    ~Staging() {
        for (auto &b : bufs_) (void)hipFreeAsync(b.dev, s_);
        if (s_) (void)hipStreamSynchronize(s_);
    }
    hipStream_t stream() const { return s_; }
    bool ok() const { return fail_ == QMC_OK; }

    template <typename T>
    T *in(const T *host, u64 count) { return stage(const_cast<T *>(host), count, true, false); }
    template <typename T>
    T *out(T *host, u64 count) { return stage(host, count, false, true); }
    template <typename T>
    T *out_keep(T *host, u64 count) { return stage(host, count, true, true); }  // unwritten rows keep their values
    template <typename T>
    T *inout(T *host, u64 count) { return stage(host, count, true, true); }

    // wait, report undefined-input status bits, copy outputs back
    void finish(int rc) {
        if (!ok()) return;
        if (rc != QMC_OK) return;
        hipError_t e = hipStreamSynchronize(s_);
        if (e != hipSuccess) { qmc_fail_hip(e, "kernel"); return; }
        u32 *st = qmc_status_ptr();
        u32 bits = 0;
        if (st && hipMemcpyAsync(&bits, st, sizeof bits, hipMemcpyDeviceToHost, s_) == hipSuccess &&
            hipStreamSynchronize(s_) == hipSuccess && bits) {
            (void)hipMemsetAsync(st, 0, sizeof bits, s_);
            qmc_fail(QMC_ERR_RANGE, "input outside the reference's defined domain (status 0x%x): "
                     "a base below 2 or an index that reads past an input array", bits);
        }
        for (auto &b : bufs_) {
            if (!b.copy_back) continue;
            e = hipMemcpyAsync(b.host, b.dev, b.bytes, hipMemcpyDeviceToHost, s_);
            if (e != hipSuccess) { qmc_fail_hip(e, "copy back"); return; }
        }
        e = hipStreamSynchronize(s_);
        if (e != hipSuccess) qmc_fail_hip(e, "copy back");
    }

  private:
    struct Buf {
        void *host;
        void *dev;
        size_t bytes;
        bool copy_back;
    };
    template <typename T>
    T *stage(T *host, u64 count, bool copy_in, bool copy_back) {
        if (!ok() || count == 0 || host == nullptr) return nullptr;
        const size_t bytes = (size_t)count * sizeof(T);
        void *dev = nullptr;
        hipError_t e = hipMallocAsync(&dev, bytes, s_);
        if (e != hipSuccess) { fail_ = qmc_fail(QMC_ERR_NOMEM, "device allocation of %zu bytes failed", bytes); return nullptr; }
        bufs_.push_back(Buf{(void *)host, dev, bytes, copy_back});
        if (copy_in) {
            e = hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s_);
            if (e != hipSuccess) { fail_ = qmc_fail_hip(e, "copy in"); return nullptr; }
        }
        return (T *)dev;
    }
    hipStream_t s_;
    int fail_ = QMC_OK;
    std::vector<Buf> bufs_;
};

#define DROPIN_BEGIN  \
    qmc_clear_error(); \
    Staging st;        \
    if (!st.ok()) return;
#define DROPIN_RUN(call) \
    if (!st.ok()) return; \
    st.finish(call)

}  // namespace

extern "C" {

// lattices ---------------------------------------------------------------
QMC_EXPORT void lat_gen_linear(u64 r, u64 n, u64 d, u64, u64, u64, const u64 *g, double *x) {
    DROPIN_BEGIN
    auto dg = st.in(g, r * d);
    auto dx = st.out(x, r * n * d);
    DROPIN_RUN(lat_gen_linear_dev(r, n, d, dg, dx, st.stream()));
}
QMC_EXPORT void lat_gen_natural(u64 r, u64 n, u64 d, u64, u64, u64, u64 n_start, const u64 *g, double *x) {
    DROPIN_BEGIN
    auto dg = st.in(g, r * d);
    auto dx = st.out_keep(x, r * n * d);
    DROPIN_RUN(lat_gen_natural_dev(r, n, d, n_start, dg, dx, st.stream()));
}
QMC_EXPORT void lat_gen_gray(u64 r, u64 n, u64 d, u64, u64, u64, u64 n_start, const u64 *g, double *x) {
    DROPIN_BEGIN
    auto dg = st.in(g, r * d);
    auto dx = st.out(x, r * n * d);
    DROPIN_RUN(lat_gen_gray_dev(r, n, d, n_start, dg, dx, st.stream()));
}
QMC_EXPORT void lat_shift_mod_1(u64 r, u64 n, u64 d, u64, u64, u64, u64 r_x, const double *x, const double *shifts,
                                double *xr) {
    DROPIN_BEGIN
    auto dx = st.in(x, r_x * n * d);
    auto ds = st.in(shifts, r * d);
    auto dxr = st.out(xr, r * n * d);
    DROPIN_RUN(lat_shift_mod_1_dev(r, n, d, r_x, dx, ds, dxr, st.stream()));
}

// base-2 digital nets -----------------------------------------------------
QMC_EXPORT void dnb2_gmat_lsb_to_msb(u64 r, u64 d, u64 mmax, u64, u64, u64, const u64 *tmaxes, const u64 *C_lsb,
                                     u64 *C_msb) {
    DROPIN_BEGIN
    auto dt = st.in(tmaxes, r);
    auto dc = st.in(C_lsb, r * d * mmax);
    auto dm = st.out(C_msb, r * d * mmax);
    DROPIN_RUN(dnb2_gmat_lsb_to_msb_dev(r, d, mmax, dt, dc, dm, st.stream()));
}
QMC_EXPORT void dnb2_linear_matrix_scramble(u64 r, u64 d, u64 mmax, u64, u64, u64, u64 r_C, u64 tmax_new,
                                            const u64 *S, const u64 *C, u64 *C_lms) {
    DROPIN_BEGIN
    auto dS = st.in(S, r * d * tmax_new);
    auto dC = st.in(C, r_C * d * mmax);
    auto dO = st.out(C_lms, r * d * mmax);
    DROPIN_RUN(dnb2_linear_matrix_scramble_dev(r, d, mmax, r_C, tmax_new, dS, dC, dO, st.stream()));
}
QMC_EXPORT void dnb2_gen_natural(u64 r, u64 n, u64 d, u64, u64, u64, u64 n_start, u64 mmax, const u64 *C, u64 *xb) {
    DROPIN_BEGIN
    auto dC = st.in(C, r * d * mmax);
    auto dx = st.out_keep(xb, r * n * d);
    DROPIN_RUN(dnb2_gen_natural_dev(r, n, d, n_start, mmax, dC, dx, st.stream()));
}
QMC_EXPORT void dnb2_gen_gray(u64 r, u64 n, u64 d, u64, u64, u64, u64 n_start, u64 mmax, const u64 *C, u64 *xb) {
    DROPIN_BEGIN
    auto dC = st.in(C, r * d * mmax);
    auto dx = st.out(xb, r * n * d);
    DROPIN_RUN(dnb2_gen_gray_dev(r, n, d, n_start, mmax, dC, dx, st.stream()));
}
QMC_EXPORT void dnb2_digital_shift(u64 r, u64 n, u64 d, u64, u64, u64, u64 r_x, const u64 *lshifts, const u64 *xb,
                                   const u64 *shiftsb, u64 *xrb) {
    DROPIN_BEGIN
    auto dl = st.in(lshifts, r_x);
    auto dx = st.in(xb, r_x * n * d);
    auto ds = st.in(shiftsb, r * d);
    auto dO = st.out(xrb, r * n * d);
    DROPIN_RUN(dnb2_digital_shift_dev(r, n, d, r_x, dl, dx, ds, dO, st.stream()));
}
QMC_EXPORT void dnb2_integer_to_float(u64 r, u64 n, u64 d, u64, u64, u64, const u64 *tmaxes, const u64 *xb,
                                      double *x) {
    DROPIN_BEGIN
    auto dt = st.in(tmaxes, r);
    auto dx = st.in(xb, r * n * d);
    auto dO = st.out(x, r * n * d);
    DROPIN_RUN(dnb2_integer_to_float_dev(r, n, d, dt, dx, dO, st.stream()));
}
QMC_EXPORT void dnb2_interlace(u64 r, u64 d_alpha, u64 mmax, u64, u64, u64, u64 d, u64 tmax, u64 tmax_alpha,
                               u64 alpha, const u64 *C, u64 *C_alpha) {
    DROPIN_BEGIN
    auto dC = st.in(C, r * d * mmax);
    auto dO = st.out(C_alpha, r * d_alpha * mmax);
    DROPIN_RUN(dnb2_interlace_dev(r, d_alpha, mmax, d, tmax, tmax_alpha, alpha, dC, dO, st.stream()));
}
QMC_EXPORT void dnb2_undo_interlace(u64 r, u64 d, u64 mmax, u64, u64, u64, u64 d_alpha, u64 tmax, u64 tmax_alpha,
                                    u64 alpha, const u64 *C_alpha, u64 *C) {
    DROPIN_BEGIN
    auto dC = st.in(C_alpha, r * d_alpha * mmax);
    auto dO = st.out(C, r * d * mmax);
    DROPIN_RUN(dnb2_undo_interlace_dev(r, d, mmax, d_alpha, tmax, tmax_alpha, alpha, dC, dO, st.stream()));
}

// generalized digital nets --------------------------------------------------
QMC_EXPORT void gdn_gen_natural(u64 r, u64 n, u64 d, u64, u64, u64, u64 r_b, u64 mmax, u64 tmax, u64 n_start,
                                const u64 *bases, const u64 *C, u64 *xdig) {
    DROPIN_BEGIN
    auto db = st.in(bases, r_b * d);
    auto dC = st.in(C, r * d * mmax * tmax);
    auto dO = st.out(xdig, r * n * d * tmax);
    DROPIN_RUN(gdn_gen_natural_dev(r, n, d, r_b, mmax, tmax, n_start, db, dC, dO, st.stream()));
}
QMC_EXPORT void gdn_gen_natural_same_base(u64 r, u64 n, u64 d, u64, u64, u64, u64 mmax, u64 tmax, u64 n_start,
                                          u64 b, const u64 *C, u64 *xdig) {
    DROPIN_BEGIN
    auto dC = st.in(C, r * d * mmax * tmax);
    auto dO = st.out(xdig, r * n * d * tmax);
    DROPIN_RUN(gdn_gen_natural_same_base_dev(r, n, d, mmax, tmax, n_start, b, dC, dO, st.stream()));
}
QMC_EXPORT void gdn_linear_matrix_scramble(u64 r, u64 d, u64 mmax, u64, u64, u64, u64 r_C, u64 r_b, u64 tmax,
                                           u64 tmax_new, const u64 *bases, const u64 *S, const u64 *C, u64 *C_lms) {
    DROPIN_BEGIN
    auto db = st.in(bases, r_b * d);
    auto dS = st.in(S, r * d * tmax_new * tmax);
    auto dC = st.in(C, r_C * d * mmax * tmax);
    auto dO = st.out(C_lms, r * d * mmax * tmax_new);
    DROPIN_RUN(gdn_linear_matrix_scramble_dev(r, d, mmax, r_C, r_b, tmax, tmax_new, db, dS, dC, dO, st.stream()));
}
QMC_EXPORT void gdn_digital_permutation(u64 r, u64 n, u64 d, u64, u64, u64, u64 r_x, u64 r_b, u64 tmax,
                                        u64 tmax_new, u64 bmax, const u64 *perms, const u64 *xdig, u64 *xdig_new) {
    DROPIN_BEGIN
    auto dp = st.in(perms, r * d * tmax_new * bmax);
    auto dx = st.in(xdig, r_x * n * d * tmax);
    auto dO = st.out(xdig_new, r * n * d * tmax_new);
    DROPIN_RUN(gdn_digital_permutation_dev(r, n, d, r_x, r_b, tmax, tmax_new, bmax, dp, dx, dO, st.stream()));
}
QMC_EXPORT void gdn_digital_shift(u64 r, u64 n, u64 d, u64, u64, u64, u64 r_x, u64 r_b, u64 tmax, u64 tmax_new,
                                  const u64 *bases, const u64 *shifts, const u64 *xdig, u64 *xdig_new) {
    DROPIN_BEGIN
    auto db = st.in(bases, r_b * d);
    auto ds = st.in(shifts, r * d * tmax_new);
    auto dx = st.in(xdig, r_x * n * d * tmax);
    auto dO = st.out(xdig_new, r * n * d * tmax_new);
    DROPIN_RUN(gdn_digital_shift_dev(r, n, d, r_x, r_b, tmax, tmax_new, db, ds, dx, dO, st.stream()));
}
QMC_EXPORT void gdn_integer_to_float(u64 r, u64 n, u64 d, u64, u64, u64, u64 r_b, u64 tmax, const u64 *bases,
                                     const u64 *xdig, double *x) {
    DROPIN_BEGIN
    auto db = st.in(bases, r_b * d);
    auto dx = st.in(xdig, r * n * d * tmax);
    auto dO = st.out(x, r * n * d);
    DROPIN_RUN(gdn_integer_to_float_dev(r, n, d, r_b, tmax, db, dx, dO, st.stream()));
}
QMC_EXPORT void gdn_interlace(u64 r, u64 d_alpha, u64 mmax, u64, u64, u64, u64 d, u64 tmax, u64 tmax_alpha,
                              u64 alpha, const u64 *C, u64 *C_alpha) {
    DROPIN_BEGIN
    auto dC = st.in(C, r * d * mmax * tmax);
    auto dO = st.out(C_alpha, r * d_alpha * mmax * tmax_alpha);
    DROPIN_RUN(gdn_interlace_dev(r, d_alpha, mmax, d, tmax, tmax_alpha, alpha, dC, dO, st.stream()));
}
QMC_EXPORT void gdn_undo_interlace(u64 r, u64 d, u64 mmax, u64, u64, u64, u64 d_alpha, u64 tmax, u64 tmax_alpha,
                                   u64 alpha, const u64 *C_alpha, u64 *C) {
    DROPIN_BEGIN
    auto dC = st.in(C_alpha, r * d_alpha * mmax * tmax_alpha);
    auto dO = st.out(C, r * d * mmax * tmax);
    DROPIN_RUN(gdn_undo_interlace_dev(r, d, mmax, d_alpha, tmax, tmax_alpha, alpha, dC, dO, st.stream()));
}

// transforms ---------------------------------------------------------------
QMC_EXPORT void fwht_1d_radix2(u64 d1, u64 d2, u64 n_half, u64, u64, u64, double *x) {
    DROPIN_BEGIN
    auto dx = st.inout(x, d1 * d2 * 2 * n_half);
    DROPIN_RUN(fwht_1d_radix2_dev(d1, d2, n_half, dx, st.stream()));
}
QMC_EXPORT void fft_bro_1d_radix2(u64 d1, u64 d2, u64 n_half, u64, u64, u64, double *twiddler, double *twiddlei,
                                  double *xr, double *xi) {
    DROPIN_BEGIN
    auto tr = st.out(twiddler, 2 * n_half);
    auto ti = st.out(twiddlei, 2 * n_half);
    auto dr = st.inout(xr, d1 * d2 * 2 * n_half);
    auto di = st.inout(xi, d1 * d2 * 2 * n_half);
    DROPIN_RUN(fft_bro_1d_radix2_dev(d1, d2, n_half, tr, ti, dr, di, st.stream()));
}
QMC_EXPORT void ifft_bro_1d_radix2(u64 d1, u64 d2, u64 n_half, u64, u64, u64, double *twiddler, double *twiddlei,
                                   double *xr, double *xi) {
    DROPIN_BEGIN
    auto tr = st.out(twiddler, 2 * n_half);
    auto ti = st.out(twiddlei, 2 * n_half);
    auto dr = st.inout(xr, d1 * d2 * 2 * n_half);
    auto di = st.inout(xi, d1 * d2 * 2 * n_half);
    DROPIN_RUN(ifft_bro_1d_radix2_dev(d1, d2, n_half, tr, ti, dr, di, st.stream()));
}

// generalized Halton -------------------------------------------------------
QMC_EXPORT void halton_qrng(int n, int d, int n0, int generalized, double *res, double *randu_d_32, int *dvec) {
    DROPIN_BEGIN
    if (n <= 0 || d <= 0) {
        if (n < 0 || d < 0 || n0 < 0) qmc_fail(QMC_ERR_ARG, "halton_qrng: negative size");
        return;
    }
    auto du = st.in(randu_d_32, (u64)d * 32);
    auto dv = st.in(dvec, (u64)d);
    auto dr = st.out(res, (u64)d * (u64)n);
    DROPIN_RUN(qmc_halton_qrng_dev(n, d, n0, generalized, dr, du, dv, st.stream()));
}

}  // extern "C"
