// common.h -- shared host helpers of libsssi: status codes, thread-local error
// text, HIP error checks, device buffers, the staggered coefficients and the
// CPML damping profiles (formed in FP64 on the host, rounded once to FP32).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sssi.h"

namespace sssi {

// Exception carrying an ABI status code; converted at the C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const char *msg);
const char *last_error_text();

[[noreturn]] inline void fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

#define SSSI_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            ::sssi::fail(e_ == hipErrorOutOfMemory ? SSSI_ENOMEM : SSSI_EDEVICE,        \
                         "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),         \
                         __FILE__, __LINE__);                                           \
    } while (0)

// Stream of the stream-ordered allocator for the calling thread's device
// allocations (hipMallocAsync / hipFreeAsync from the device's memory pool),
// or null for hipMalloc / hipFree.  The serving loop's helper threads set it:
// hipFree waits for the whole device, which would stall every resident job
// each time one finishes.
hipStream_t &alloc_stream();

// Owning device buffer.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipStream_t as = nullptr;  // allocated on this stream's pool (null: hipMalloc)
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n), as(o.as) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; as = o.as;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        as = alloc_stream();
        if (!count) return;
        if (as) SSSI_HIP(hipMallocAsync((void **)&p, count * sizeof(T), as));
        else SSSI_HIP(hipMalloc((void **)&p, count * sizeof(T)));
    }
    void release() {
        if (p) {
            // freed on the releasing thread's allocation stream if it has one
            // (the memory's users have finished), else on the allocating one
            if (as) (void)hipFreeAsync(p, alloc_stream() ? alloc_stream() : as);
            else (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Non-blocking streams from a per-device pool.  Creating a stream takes about
// 3 ms on this runtime, so a handle or a one-shot call takes one from the pool
// and gives it back (idle) when it is done; the pool keeps them until the
// process ends.
hipStream_t stream_acquire();
void stream_release(hipStream_t s);

// Pool stream held by its holder: given back with it, also when the holder's
// constructor throws after taking it.  Declared before the device buffers
// whose work it orders, it is given back after them.
struct OwnedStream {
    hipStream_t s = nullptr;
    OwnedStream() = default;
    OwnedStream(const OwnedStream &) = delete;
    OwnedStream &operator=(const OwnedStream &) = delete;
    ~OwnedStream() { stream_release(s); }
    void create() { s = stream_acquire(); }
    operator hipStream_t() const { return s; }
};

// v (column-major FP64, z fastest: v[(k*nx + j)*nz + i]) -> a padded FP32
// device field: dst[(i + hz)*plane + (k + hy)*pitch + j + hx] = (v dt)^2 when
// square, else v dt, each rounded once from FP64 (the arithmetic of a host
// conversion).  The FP64 model goes up in chunks through a device staging
// buffer and is converted there; the rest of dst is zero-filled first.
void upload_model(const double *v, int64_t nz, int64_t nx, int64_t ny, double dt, bool square,
                  float *dst, size_t dst_n, int64_t plane, int64_t pitch, int hz, int hy, int hx,
                  hipStream_t s);

// Owning pinned host buffer (hipHostMalloc): copies between it and the device
// run at link speed rather than through the runtime's pageable staging.
template <class T>
struct HostBuf {
    T *p = nullptr;
    size_t n = 0;
    HostBuf() = default;
    explicit HostBuf(size_t count) { alloc(count); }
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) SSSI_HIP(hipHostMalloc((void **)&p, count * sizeof(T), hipHostMallocDefault));
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// f(begin, end) over a split of [0, n) on up to 16 host threads; one thread
// when n / min_chunk < 2.  f must not throw.
template <class F>
void parallel_ranges(int64_t n, F f, int64_t min_chunk = 4096) {
    const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>({16, hw, n / std::max<int64_t>(1, min_chunk)}));
    if (nt <= 1) {
        if (n > 0) f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int64_t k = 1; k < nt; ++k) th.emplace_back(f, n * k / nt, n * (k + 1) / nt);
    f((int64_t)0, n / nt);
    for (auto &t : th) t.join();
}

// Staggered coefficients c_j (dCoef(N,'s'), finiteDifference.c:24-82): exact
// rational c_j = (1/(2j-1)) prod_{m!=j} (2m-1)^2/((2m-1)^2-(2j-1)^2), reduced
// and rounded once (numerator and denominator stay below 2^53 for N <= 8).
void staggered_coefficients(int N, double *c);

// 2-D CPML b-profiles as in fwdTimeCpmlFor2dAw_mex.c:100-145 (FP64), returned
// as b-1 for the strips only (compact layouts):
//   zbm1[(i-(nz-nB))*nx + j]   rows nz-nB..nz-1, all columns
//   xbm1[i*2nB + xs(j)]        columns 0..nB-1 and nx-nB..nx-1, all rows
void cpml2d_bm1(const double *v, int64_t nz, int64_t nx, int nB, double dz, double dx, double dt,
                std::vector<float> &zbm1, std::vector<float> &xbm1);
// Element q of those tables (q < nz*2nB for x, q < nB*nx for z).
float cpml2d_bm1_x(const double *v, int64_t nz, int64_t nx, int nB, double dx, double dt,
                   int64_t q);
float cpml2d_bm1_z(const double *v, int64_t nz, int64_t nx, int nB, double dz, double dt,
                   int64_t q);

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// Level 5 MAT-files of real double arrays (matfile.cpp).  mat_write writes
// nvar variables (dims concatenated, ndims[v] each) atomically (temporary file
// renamed over the target); mat_read returns false if the file holds no such
// variable, and fills ndims/dims (up to 8) and, when out is given, n values.
void mat_write(const char *path, int nvar, const char *const *names, const double *const *data,
               const int *ndims, const int64_t *dims);
bool mat_read(const char *path, const char *name, int *ndims, int64_t *dims, double *out,
              int64_t n);
bool file_exists(const char *path);

// Argument checks shared by every entry point (before any device work).
// dim 3 enforces 2*nB <= ny also for ny = 1 (a 3-D call with a 2-D model).
void validate_common(int dim, int64_t nz, int64_t nx, int64_t ny, int64_t nt, int N, int nB,
                     double dz, double dx, double dy, double dt, bool need_N);
// Every velocity finite and > 0 (the reference accepts anything and returns
// NaN/Inf traces); scanned on up to 16 host threads.
void validate_velocity(const double *v, int64_t n, const char *name);

}  // namespace sssi
