// common.cpp -- host helpers shared by the propagators (see common.h).
#include "common.h"

#include <atomic>
#include <cstring>
#include <map>
#include <mutex>

namespace sssi {

static thread_local std::string g_last_error;
static thread_local hipStream_t g_alloc_stream = nullptr;

hipStream_t &alloc_stream() { return g_alloc_stream; }

void set_last_error(const char *msg) { g_last_error = msg ? msg : ""; }

const char *last_error_text() { return g_last_error.c_str(); }

static int64_t gcd64(int64_t a, int64_t b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) {
        const int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

void staggered_coefficients(int N, double *c) {
    if (N < 1 || N > 8) fail(SSSI_EINVAL, "nDiffOrder must be in 1..8 (got %d)", N);
    for (int j = 1; j <= N; ++j) {
        int64_t num = 1, den = 2 * j - 1;
        bool neg = false;
        for (int m = 1; m <= N; ++m) {
            if (m == j) continue;
            const int64_t a = (int64_t)(2 * m - 1) * (2 * m - 1);
            int64_t b = a - (int64_t)(2 * j - 1) * (2 * j - 1);
            if (b < 0) { neg = !neg; b = -b; }
            const int64_t g1 = gcd64(a, den), g2 = gcd64(num, b);
            num = (num / g2) * (a / g1);
            den = (den / g1) * (b / g2);
            const int64_t g = gcd64(num, den);
            num /= g;
            den /= g;
        }
        c[j - 1] = (neg ? -1.0 : 1.0) * ((double)num / (double)den);
    }
}

// fwdTimeCpmlFor2dAw_mex.c:100-145 with dampPml (finiteDifference.c:253-280):
// d0 = -(3 v)/(2 L) log(1e-6), d = d0 (u/L)(u/L), b = exp(-d dt).
static inline float bm1_of(double v, double u, double L, double dt) {
    const double d0 = -(3.0 * v) / (2 * L) * std::log(1e-6);
    const double d = d0 * (u / L) * (u / L);
    return (float)(std::exp(-d * dt) - 1.0);
}

float cpml2d_bm1_x(const double *v, int64_t nz, int64_t nx, int nB, double dx, double dt,
                   int64_t q) {
    const int64_t i = q / (2 * nB), xs = q % (2 * nB);
    const double L = nB * dx;
    if (xs < nB) return bm1_of(v[xs * nz + i], (nB - xs) * dx, L, dt);         // left
    const int64_t j = xs - nB;
    return bm1_of(v[(nx - nB + j) * nz + i], (j + 1) * dx, L, dt);              // right
}

float cpml2d_bm1_z(const double *v, int64_t nz, int64_t nx, int nB, double dz, double dt,
                   int64_t q) {
    const int64_t i = q / nx, j = q % nx;
    return bm1_of(v[j * nz + (nz - nB + i)], (i + 1) * dz, nB * dz, dt);      // bottom
}

void cpml2d_bm1(const double *v, int64_t nz, int64_t nx, int nB, double dz, double dx, double dt,
                std::vector<float> &zbm1, std::vector<float> &xbm1) {
    zbm1.assign((size_t)nB * nx, 0.f);
    xbm1.assign((size_t)nz * 2 * nB, 0.f);
    if (nB == 0) return;
    for (int64_t q = 0; q < (int64_t)xbm1.size(); ++q)
        xbm1[q] = cpml2d_bm1_x(v, nz, nx, nB, dx, dt, q);
    for (int64_t q = 0; q < (int64_t)zbm1.size(); ++q)
        zbm1[q] = cpml2d_bm1_z(v, nz, nx, nB, dz, dt, q);
}

void validate_common(int dim, int64_t nz, int64_t nx, int64_t ny, int64_t nt, int N, int nB,
                     double dz, double dx, double dy, double dt, bool need_N) {
    if (nz < 1 || nx < 1 || ny < 1)
        fail(SSSI_EINVAL, "grid must be non-empty (nz=%lld nx=%lld ny=%lld)", (long long)nz,
             (long long)nx, (long long)ny);
    if (nz > (1 << 24) || nx > (1 << 24) || ny > (1 << 24) || nz * nx > (int64_t)1 << 30)
        fail(SSSI_EINVAL, "grid too large for the 2-D index space");
    if (nt < 1) fail(SSSI_EINVAL, "nt must be >= 1 (got %lld)", (long long)nt);
    if (need_N && (N < 1 || N > 8)) fail(SSSI_EINVAL, "nDiffOrder must be in 1..8 (got %d)", N);
    if (nB < 0) fail(SSSI_EINVAL, "nBoundary must be >= 0 (got %d)", nB);
    if ((int64_t)nB > nz) fail(SSSI_EINVAL, "nBoundary=%d exceeds nz=%lld", nB, (long long)nz);
    if (2 * (int64_t)nB > nx)
        fail(SSSI_EINVAL, "2*nBoundary=%d exceeds nx=%lld", 2 * nB, (long long)nx);
    if ((dim == 3 || ny > 1) && 2 * (int64_t)nB > ny)
        fail(SSSI_EINVAL, "2*nBoundary=%d exceeds ny=%lld", 2 * nB, (long long)ny);
    if (!(dz > 0) || !(dx > 0) || !(dy > 0) || !(dt > 0) || !std::isfinite(dz) ||
        !std::isfinite(dx) || !std::isfinite(dy) || !std::isfinite(dt))
        fail(SSSI_EINVAL, "dz, dx, dy and dt must be positive and finite");
}

void validate_velocity(const double *v, int64_t n, const char *name) {
    if (!v) fail(SSSI_EINVAL, "%s must not be NULL", name);
    std::atomic<int64_t> bad{-1};
    parallel_ranges(n, [&](int64_t a, int64_t b) {
        for (int64_t q = a; q < b; ++q)
            if (!(v[q] > 0.0) || !std::isfinite(v[q])) {
                int64_t want = -1;
                bad.compare_exchange_strong(want, q);
                return;
            }
    }, 1 << 16);
    const int64_t q = bad.load();
    if (q >= 0)
        fail(SSSI_EINVAL, "%s must be finite and positive (element %lld is %g)", name,
             (long long)q, v[q]);
}

namespace {
std::mutex g_stream_mu;
std::map<int, std::vector<hipStream_t>> g_streams;  // idle pool streams per device
}  // namespace

hipStream_t stream_acquire() {
    int dev = 0;
    SSSI_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        auto &v = g_streams[dev];
        if (!v.empty()) {
            hipStream_t s = v.back();
            v.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    SSSI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void stream_release(hipStream_t s) {
    if (!s) return;
    int dev = 0;
    // an idle stream goes back to its own device's list; one that cannot be
    // drained or placed is destroyed
    if (hipStreamSynchronize(s) != hipSuccess || hipStreamGetDevice(s, &dev) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return;
    }
    std::lock_guard<std::mutex> lk(g_stream_mu);
    g_streams[dev].push_back(s);
}

}  // namespace sssi
