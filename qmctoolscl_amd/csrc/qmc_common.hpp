// Shared device/host helpers for the gfx950 QMC kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/qmctoolscl_amd.h"  // status codes and the exported prototypes

typedef uint64_t u64;
typedef uint32_t u32;

#define QMC_HIP_CHECK(expr)                                      \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return qmc_fail_hip(_e, #expr);    \
    } while (0)

int qmc_fail_hip(hipError_t e, const char *what);
int qmc_fail(int code, const char *fmt, ...);

// ---------------------------------------------------------------- launch shape
constexpr int kBlock = 256;

static inline unsigned grid_for(u64 work, u64 per_block, unsigned cap = 1u << 20) {
    u64 g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Workgroups of `fn` that stay resident on one CU at this launch shape
// (looked up once per kernel, shape and device; at least 1).
static inline int resident_per_cu(const void *fn, int threads, size_t shm) {
    struct Entry {
        const void *fn;
        int threads, dev;
        size_t shm;
        int per_cu;
    };
    static Entry cache[128];
    static int used = 0;
    static std::mutex mu;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    for (int i = 0; i < used; i++)
        if (cache[i].fn == fn && cache[i].threads == threads && cache[i].dev == dev && cache[i].shm == shm)
            return cache[i].per_cu;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, shm) != hipSuccess || per_cu < 1) per_cu = 1;
    if (used < 128) cache[used++] = Entry{fn, threads, dev, shm, per_cu};
    return per_cu;
}

// Set-up that must happen once on every device a process uses (constant
// tables, kernel attributes): run(f) calls f with the device current the first
// time per device, under a mutex, and again after a failure.
struct PerDeviceOnce {
    std::mutex mu;
    bool done[64] = {};
    template <class F>
    int run(F f) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return f();
        std::lock_guard<std::mutex> lock(mu);
        if (done[dev]) return QMC_OK;
        const int rc = f();
        if (rc == QMC_OK) done[dev] = true;
        return rc;
    }
};

// Compute units of the current device (looked up once per device).
static inline int device_cus() {
    static int cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int c = __atomic_load_n(&cache[dev], __ATOMIC_RELAXED);
    if (c <= 0) {
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        __atomic_store_n(&cache[dev], c, __ATOMIC_RELAXED);
    }
    return c;
}

// v, hidden from the optimizer.  Persistent loops read the thread index
// through this once per iteration, so per-thread addresses derived from it are
// recomputed each iteration instead of being hoisted out of the loop and kept
// in registers for its whole length.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Sign bit (0 or 0x80000000) of a pseudo-random +-1 for a (column, salt)
// pair.  The matrix cores' fp32 accumulation rounds with a small bias; the
// transforms flip the sign of each 16-point column's inputs and outputs by
// this (an exact linear change), which turns that bias into noise instead of
// letting it add up coherently in the lowest frequencies.
__device__ __forceinline__ u32 dither_sign(u32 col, u32 salt) {
    u32 h = (col ^ salt) * 0x9E3779B1u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    return h & 0x80000000u;
}
__device__ __forceinline__ float flip(float v, u32 sbit) { return __uint_as_float(__float_as_uint(v) ^ sbit); }

// --------------------------------------------------------- exact small helpers

// fmod(y, 1.0) for finite y: the fractional part is exact in binary floating
// point (y and trunc(y) share the grid of ulp(y)), and fmod keeps the sign of y.
__device__ __forceinline__ double frac_mod1(double y) {
    return copysign(y - trunc(y), y);
}

// Unsigned division by a run-time 32-bit divisor via a 64-bit fractional
// reciprocal M = floor(2^64 / d) + 1 (Lemire, Kaser, Kurz 2019): exact for all
// 32-bit numerators, d >= 1.
struct DivU32 {
    u64 M;
    u32 d;
};

__host__ __device__ static inline DivU32 make_div(u32 d) {
    DivU32 r;
    r.d = d;
    r.M = (d == 1) ? 0 : (~(u64)0) / d + 1;
    return r;
}

__device__ __forceinline__ u32 div_q(u32 n, DivU32 v) {
    return v.d == 1 ? n : (u32)__umul64hi(v.M, (u64)n);
}

// index decomposition helpers for 64-bit flat indices (used when a slab does
// not fit 32 bits)
__device__ __forceinline__ void split2(u64 e, u64 d, u64 &i, u64 &j) {
    i = e / d;
    j = e - i * d;
}

// v mod b for a 64-bit v and a run-time b: a 32-bit fast path when both fit.
__device__ __forceinline__ u64 umod(u64 v, u64 b) {
    if ((v >> 32) == 0 && (b >> 32) == 0) return (u64)((u32)v % (u32)b);
    return v % b;
}

__device__ __forceinline__ u64 gray_of(u64 k) { return k ^ (k >> 1); }

__device__ __forceinline__ u64 gray_inverse(u64 g) {
    g ^= g >> 1;
    g ^= g >> 2;
    g ^= g >> 4;
    g ^= g >> 8;
    g ^= g >> 16;
    g ^= g >> 32;
    return g;
}

// streaming stores for large outputs: they bypass the caches on the way to HBM
template <typename T>
__device__ __forceinline__ void store_stream(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}
