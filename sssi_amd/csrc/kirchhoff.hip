// kirchhoff.hip -- traveltimes by the reference's fast-sweeping eikonal solver
// (src/fd_cmex/eikonal2d_mex.cpp:27-267) and Kirchhoff migration
// (src/migrate.m:1-26, src/shot2RecTime.m:1-6) on gfx950.
//
// Eikonal.  Each of the reference's four sweeps per iteration reads only the
// previous level and writes the next, so a sweep is one Jacobi pass over the
// grid whatever its loop order: one launch per pass updates every cell of
// every source's table at once (grid-stride over cells x sources, ping-pong
// levels).  The arithmetic is FP64 in the reference's operation order, with
// floating-point contraction off, so each pass reproduces the reference's
// values bit for bit: the min-plus update branches on |a-b| < s*dx, and a
// rounding difference could take the other branch.
//
// Migration.  The traveltime of each table is turned once into its 1-based
// sample index round(T/dt)+1 (MATLAB's round: half away from zero).  A thread
// owns one image pixel and, shot by shot, sums the record samples at index
// K_shot + K_receiver over the receivers in order -- the reference's
// accumulation order, in FP64, so the stack is bit-identical to the
// reference's `Stacked = Stacked + M`.  An index outside the record is an
// error, as MATLAB's indexing is.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "common.h"

namespace sssi {

constexpr double kT0 = 65535;  // initial traveltime (eikonal2d_mex.cpp:31)

static unsigned grid_blocks(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256));
}

// One Jacobi pass for nsrc tables [k][j][i] (column-major per table).
__global__ void __launch_bounds__(256) eik_pass(const double *__restrict__ slow,
                                                const double *__restrict__ told,
                                                double *__restrict__ tnew, int64_t nz, int64_t nx,
                                                int nsrc, const int64_t *__restrict__ src_cell,
                                                double dx) {
#pragma clang fp contract(off)
    const int64_t nn = nz * nx;
    const int64_t total = nn * nsrc;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(g / nn);
        const int64_t q = g - (int64_t)k * nn;
        const int64_t j = q / nz, i = q - j * nz;
        const double *T = told + (int64_t)k * nn;
        const double *c = T + j * nz;
        double a, b, TT;
        if (i == 0) a = c[1];
        else if (i == nz - 1) a = c[nz - 2];
        else a = fmin(c[i - 1], c[i + 1]);
        if (j == 0) b = T[i + nz];
        else if (j == nx - 1) b = T[i + (nx - 2) * nz];
        else b = fmin(T[i + (j - 1) * nz], T[i + (j + 1) * nz]);
        const double s = slow[q];
        if (fabs(a - b) < s * dx)
            TT = 0.5 * (a + b + sqrt((2 * s * s * dx * dx) - (a - b) * (a - b)));
        else
            TT = fmin(a, b) + (s * dx);
        const double t = fmin(c[i], TT);
        tnew[g] = q == src_cell[k] ? 0.0 : t;
    }
}

// TOld of the first pass (src_cell given: 0 at the source) or, with no
// passes, the reference's output level TNew, 65535 everywhere (src_cell null)
__global__ void eik_init(double *t, int64_t nn, int nsrc, const int64_t *__restrict__ src_cell) {
    const int64_t total = nn * nsrc;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(g / nn);
        t[g] = src_cell && g - (int64_t)k * nn == src_cell[k] ? 0.0 : kT0;
    }
}

// 1-based sample index of each traveltime: round(T/dt) + 1
__global__ void eik_index(const double *__restrict__ t, int64_t n, double dt,
                          int *__restrict__ kidx) {
#pragma clang fp contract(off)
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x) {
        const double r = round(t[g] / dt) + 1;
        kidx[g] = r < 2147483647.0 ? (int)r : 2147483647;
    }
}

// Stack (and per-shot images) over nshot shots: pixel q, shot s (table ts[s]),
// receivers r in order (table tr[r], record column xr[r]).
__global__ void __launch_bounds__(256) kirch_stack(const int *__restrict__ kidx, int64_t nn,
                                                   const int *__restrict__ ts, int nshot,
                                                   const int *__restrict__ tr,
                                                   const int *__restrict__ xr, int nrec,
                                                   const double *__restrict__ shots, int64_t nt,
                                                   int64_t ncol, double *__restrict__ stack,
                                                   double *__restrict__ images,
                                                   int *__restrict__ bad) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nn;
         q += (int64_t)gridDim.x * blockDim.x) {
        double st = 0.0;
        bool oob = false;
        for (int s = 0; s < nshot; ++s) {
            const int ks = kidx[(int64_t)ts[s] * nn + q];
            const double *rec = shots + (int64_t)s * nt * ncol;
            double m = 0.0;
            for (int r = 0; r < nrec; ++r) {
                const int64_t it = (int64_t)ks + kidx[(int64_t)tr[r] * nn + q];  // 1-based
                if (it < 1 || it > nt) {
                    oob = true;
                    continue;
                }
                m = m + rec[(int64_t)xr[r] * nt + (it - 1)];
            }
            if (images) images[(int64_t)s * nn + q] = m;
            st = st + m;
        }
        stack[q] = st;
        if (oob) atomicOr(bad, 1);
    }
}

// ---------------------------------------------------------------------------
// Traveltime tables of sources (1-based (sz, sx)) on one model, resident.
class Eikonal2D {
   public:
    int64_t nz = 0, nx = 0;
    int nsrc = 0;
    hipStream_t stream = nullptr;
    DevBuf<double> slow, t[2];
    DevBuf<int64_t> src;
    int cur = 0;

    Eikonal2D(const double *v, int64_t nz_, int64_t nx_, double dx, int nsrc_, const int *sz,
              const int *sx, int iMax)
        : nz(nz_), nx(nx_), nsrc(nsrc_) {
        if (nz < 2 || nx < 2) fail(SSSI_EINVAL, "eikonal grid must be at least 2 x 2");
        if (nsrc < 1 || !sz || !sx) fail(SSSI_EINVAL, "eikonal needs at least one source");
        if (iMax < 0) fail(SSSI_EINVAL, "iMax must be >= 0");
        const int64_t nn = nz * nx;
        if ((double)nn * nsrc * 8 * 2 > 200e9) fail(SSSI_ENOMEM, "traveltime tables too large");
        std::vector<int64_t> cells(nsrc);
        for (int k = 0; k < nsrc; ++k) {
            if (sz[k] < 1 || sz[k] > nz || sx[k] < 1 || sx[k] > nx)
                fail(SSSI_EINVAL, "source %d at (%d, %d) outside the %lldx%lld grid (1-based)", k,
                     sz[k], sx[k], (long long)nz, (long long)nx);
            cells[k] = (int64_t)(sz[k] - 1) + (int64_t)(sx[k] - 1) * nz;
        }
        SSSI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        std::vector<double> hs(nn);
        for (int64_t q = 0; q < nn; ++q) hs[q] = 1 / v[q];  // the reference's slowness
        slow.alloc(nn);
        src.alloc(nsrc);
        t[0].alloc((size_t)nn * nsrc);
        t[1].alloc((size_t)nn * nsrc);
        SSSI_HIP(hipMemcpyAsync(slow.p, hs.data(), slow.bytes(), hipMemcpyHostToDevice, stream));
        SSSI_HIP(hipMemcpyAsync(src.p, cells.data(), src.bytes(), hipMemcpyHostToDevice, stream));
        const int64_t total = nn * nsrc;
        hipLaunchKernelGGL(eik_init, dim3(grid_blocks(total)), dim3(256), 0, stream, t[0].p, nn,
                           nsrc, iMax > 0 ? src.p : nullptr);
        SSSI_HIP(hipGetLastError());
        cur = 0;
        const double dxi = (double)(int)dx;  // eikonal2d_mex.cpp:35 reads dx as an int
        for (int p = 0; p < 4 * iMax; ++p) {
            hipLaunchKernelGGL(eik_pass, dim3(grid_blocks(total)), dim3(256), 0, stream, slow.p,
                               t[cur].p, t[cur ^ 1].p, nz, nx, nsrc, src.p, dxi);
            SSSI_HIP(hipGetLastError());
            cur ^= 1;
        }
        SSSI_HIP(hipStreamSynchronize(stream));
    }
    ~Eikonal2D() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    const double *table() const { return t[cur].p; }
    void download(double *T) {
        SSSI_HIP(hipMemcpyAsync(T, t[cur].p, (size_t)nz * nx * nsrc * sizeof(double),
                                hipMemcpyDeviceToHost, stream));
        SSSI_HIP(hipStreamSynchronize(stream));
    }
};

// Receiver tables: position of each xRecGrid entry in xShotAndRecGrid (0-based)
static std::vector<int> receiver_tables(const int *xrec, int nrec, const int *xsr, int nsr,
                                        int64_t ntab, int64_t nx) {
    std::vector<int> tr(nrec);
    for (int r = 0; r < nrec; ++r) {
        int k = 0;
        while (k < nsr && xsr[k] != xrec[r]) ++k;
        if (k == nsr)
            fail(SSSI_EINVAL, "receiver %d (x = %d) is not in xShotAndRecGrid", r, xrec[r]);
        if (k >= ntab) fail(SSSI_EINVAL, "receiver %d has no traveltime table", r);
        if (xrec[r] < 1 || xrec[r] > nx)
            fail(SSSI_EINVAL, "receiver %d at x = %d outside the record's %lld columns", r,
                 xrec[r], (long long)nx);
        tr[r] = k;
    }
    return tr;
}

// Migration of nshot records against resident index tables.
static void kirchhoff_run(const int *kidx_dev, int64_t nz, int64_t nx, int64_t ntab,
                          const int *xrec, int nrec, const int *xsr, int nsr,
                          const double *shots, int64_t nt, int nshot, const int *ixs,
                          double *stack, double *images, hipStream_t s) {
    if (nt < 1 || nshot < 1 || !shots || !ixs || !stack || nrec < 0 || (nrec > 0 && !xrec))
        fail(SSSI_EINVAL, "kirchhoff: bad arguments");
    std::vector<int> tr = receiver_tables(xrec, nrec, xsr, nsr, ntab, nx);
    std::vector<int> ts(nshot), xr0(std::max(nrec, 1));
    for (int q = 0; q < nshot; ++q) {
        if (ixs[q] < 1 || ixs[q] > ntab)
            fail(SSSI_EINVAL, "shot %d: ixs = %d has no traveltime table (1..%lld)", q, ixs[q],
                 (long long)ntab);
        ts[q] = ixs[q] - 1;
    }
    for (int r = 0; r < nrec; ++r) xr0[r] = xrec[r] - 1;
    const int64_t nn = nz * nx;
    DevBuf<int> dts(nshot), dtr(std::max(nrec, 1)), dxr(std::max(nrec, 1)), bad(1);
    DevBuf<double> dshots((size_t)nt * nx * nshot), dstack(nn), dimg(images ? (size_t)nn * nshot : 0);
    SSSI_HIP(hipMemcpyAsync(dts.p, ts.data(), dts.bytes(), hipMemcpyHostToDevice, s));
    if (nrec > 0) {
        SSSI_HIP(hipMemcpyAsync(dtr.p, tr.data(), (size_t)nrec * 4, hipMemcpyHostToDevice, s));
        SSSI_HIP(hipMemcpyAsync(dxr.p, xr0.data(), (size_t)nrec * 4, hipMemcpyHostToDevice, s));
    }
    SSSI_HIP(hipMemcpyAsync(dshots.p, shots, dshots.bytes(), hipMemcpyHostToDevice, s));
    SSSI_HIP(hipMemsetAsync(bad.p, 0, 4, s));
    hipLaunchKernelGGL(kirch_stack, dim3(grid_blocks(nn)), dim3(256), 0, s, kidx_dev, nn, dts.p,
                       nshot, dtr.p, dxr.p, nrec, dshots.p, nt, nx, dstack.p,
                       images ? dimg.p : nullptr, bad.p);
    SSSI_HIP(hipGetLastError());
    int hbad = 0;
    SSSI_HIP(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, s));
    SSSI_HIP(hipMemcpyAsync(stack, dstack.p, dstack.bytes(), hipMemcpyDeviceToHost, s));
    if (images)
        SSSI_HIP(hipMemcpyAsync(images, dimg.p, dimg.bytes(), hipMemcpyDeviceToHost, s));
    SSSI_HIP(hipStreamSynchronize(s));
    if (hbad)
        fail(SSSI_EINVAL, "a traveltime sample index falls outside the %lld-sample record "
             "(the reference's migrate indexes past the shot there)", (long long)nt);
}

// Traveltimes of surface positions and their sample indices, resident.
class Kirchhoff2D {
   public:
    int64_t nz = 0, nx = 0;
    std::vector<int> xsr;
    double dt = 0;
    std::unique_ptr<Eikonal2D> eik;
    DevBuf<int> kidx;

    Kirchhoff2D(const double *v, int64_t nz_, int64_t nx_, double dx, const int *xsr_, int nsr,
                int iMax, double dt_)
        : nz(nz_), nx(nx_), xsr(xsr_, xsr_ + nsr), dt(dt_) {
        if (!(dt > 0) || !std::isfinite(dt)) fail(SSSI_EINVAL, "dt must be positive");
        std::vector<int> sz(nsr, 1);  // sources at the surface row ([1 xs], the app's call)
        eik.reset(new Eikonal2D(v, nz, nx, dx, nsr, sz.data(), xsr_, iMax));
        const int64_t n = nz * nx * (int64_t)nsr;
        kidx.alloc(n);
        hipLaunchKernelGGL(eik_index, dim3(grid_blocks(n)), dim3(256), 0, eik->stream,
                           eik->table(), n, dt, kidx.p);
        SSSI_HIP(hipGetLastError());
        SSSI_HIP(hipStreamSynchronize(eik->stream));
    }
};

}  // namespace sssi

using namespace sssi;

struct sssi_kirchhoff {
    std::unique_ptr<Kirchhoff2D> impl;
};

namespace {
template <class F>
int kguard(F &&f) {
    try {
        f();
        set_last_error("");
        return SSSI_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return SSSI_ENOMEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return SSSI_EDEVICE;
    }
}
void check_model(const double *v, int64_t nz, int64_t nx) {
    if (!v) fail(SSSI_EINVAL, "V must not be NULL");
    if (nz < 2 || nx < 2 || nz > (1 << 24) || nx > (1 << 24) || nz * nx > ((int64_t)1 << 30))
        fail(SSSI_EINVAL, "eikonal grid %lldx%lld out of range (at least 2 x 2)", (long long)nz,
             (long long)nx);
    validate_velocity(v, nz * nx, "V");
}
}  // namespace

extern "C" {

int sssi_eikonal2d(const double *V, int64_t nz, int64_t nx, double dx, int sz, int sx, int iMax,
                   double *T) {
    return kguard([&] {
        check_model(V, nz, nx);
        if (!T) fail(SSSI_EINVAL, "T must not be NULL");
        Eikonal2D e(V, nz, nx, dx, 1, &sz, &sx, iMax);
        e.download(T);
    });
}

int sssi_eikonal2d_batch(const double *V, int64_t nz, int64_t nx, double dx, int nsrc,
                         const int *sz, const int *sx, int iMax, double *T) {
    return kguard([&] {
        check_model(V, nz, nx);
        if (!T) fail(SSSI_EINVAL, "T must not be NULL");
        Eikonal2D e(V, nz, nx, dx, nsrc, sz, sx, iMax);
        e.download(T);
    });
}

int sssi_kirchhoff_migrate(const double *travelTime, int64_t nz, int64_t nx, int64_t ntab,
                           const int *xRecGrid, int nrec, const int *xShotAndRecGrid, int nsr,
                           const double *shot, int64_t nt, double dt, int ixs, double *m) {
    return kguard([&] {
        if (!travelTime || !xShotAndRecGrid || !m) fail(SSSI_EINVAL, "NULL argument");
        if (nz < 1 || nx < 1 || ntab < 1 || nsr < 1) fail(SSSI_EINVAL, "empty tables");
        if (!(dt > 0) || !std::isfinite(dt)) fail(SSSI_EINVAL, "dt must be positive");
        hipStream_t s;
        SSSI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            const int64_t n = nz * nx * ntab;
            DevBuf<double> tt(n);
            DevBuf<int> kidx(n);
            SSSI_HIP(hipMemcpyAsync(tt.p, travelTime, tt.bytes(), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(eik_index, dim3(grid_blocks(n)), dim3(256), 0, s, tt.p, n, dt,
                               kidx.p);
            SSSI_HIP(hipGetLastError());
            kirchhoff_run(kidx.p, nz, nx, ntab, xRecGrid, nrec, xShotAndRecGrid, nsr, shot, nt, 1,
                          &ixs, m, nullptr, s);
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
    });
}

int sssi_kirchhoff_create(const double *V, int64_t nz, int64_t nx, double dx,
                          const int *xShotAndRecGrid, int nsr, int iMax, double dt,
                          sssi_kirchhoff **out) {
    return kguard([&] {
        check_model(V, nz, nx);
        if (!out || !xShotAndRecGrid) fail(SSSI_EINVAL, "NULL argument");
        auto h = std::make_unique<sssi_kirchhoff>();
        h->impl.reset(new Kirchhoff2D(V, nz, nx, dx, xShotAndRecGrid, nsr, iMax, dt));
        *out = h.release();
    });
}

int sssi_kirchhoff_travel_time(sssi_kirchhoff *h, double *travelTime) {
    return kguard([&] {
        if (!h || !h->impl || !travelTime) fail(SSSI_EINVAL, "NULL argument");
        h->impl->eik->download(travelTime);
    });
}

int sssi_kirchhoff_migrate_shots(sssi_kirchhoff *h, const int *xRecGrid, int nrec,
                                 const double *shots, int64_t nt, int nshot, const int *ixs,
                                 double *stack, double *images) {
    return kguard([&] {
        if (!h || !h->impl) fail(SSSI_EINVAL, "NULL handle");
        Kirchhoff2D &k = *h->impl;
        kirchhoff_run(k.kidx.p, k.nz, k.nx, (int64_t)k.xsr.size(), xRecGrid, nrec, k.xsr.data(),
                      (int)k.xsr.size(), shots, nt, nshot, ixs, stack, images, k.eik->stream);
    });
}

void sssi_kirchhoff_destroy(sssi_kirchhoff *h) { delete h; }

}  // extern "C"
