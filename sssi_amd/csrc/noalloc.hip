// noalloc.hip -- sssi_fwd_cpml_2d_points_noalloc: the 2-D CPML forward drop-in
// (sssi_fwd_cpml_2d_points) for callers that cannot allocate device memory.
//
// Every device buffer the propagation needs beyond the caller's host inputs
// and outputs -- the two wavefield levels, v^2 dt^2, the damping-strip memory
// variables and b-1 tables, the source CSR and amplitudes, the receiver
// record, a snapshot slice, the launch order, the watchdog word and an FP64
// staging area -- lives in one file-scope device array, `scratch`, of
// kScratchBytes.  The library holds it on every GPU it runs on until the
// process ends.  The call makes no hipMalloc / hipFree / hipHostMalloc and no
// host heap allocation: host <-> device transfers go through the staging area
// in chunks, with the format conversions (FP64 <-> FP32, column-major <->
// padded rows) done by small kernels.
//
// The time step is the same kernel as the allocating entry point's, on the
// tile path in natural block order (no interior march, which is bit-identical
// to it), so the results are identical bit for bit.
//
// Calls on one device share its array and are serialised by a per-device
// lock; each call re-initialises everything it reads.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <utility>

#include "cpml2d.h"
#include "devbuf.h"
#include "health.h"

namespace sssi {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per device
constexpr size_t kStageBytes = (size_t)2 << 20;     // FP64 staging area within it
constexpr int kNaMaxSrc = 64;                        // point sources per call
constexpr int kNaMaxDev = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// --- conversion kernels (grid-stride loops) ----------------------------------
__global__ void na_fill_f32(float *p, size_t n, float v) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        p[q] = v;
}
__global__ void na_fill_i32(int *p, size_t n, int v) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        p[q] = v;
}
// columns j0 .. j0+ncols-1 of v (column-major FP64 in the staging area) ->
// v^2 dt^2 in the padded FP32 layout, rounded once as the allocating path does
__global__ void na_vdt2(const double *vc, int64_t nz, int64_t j0, int64_t ncols, double dt,
                        float *vdt2, int64_t pitch) {
    const size_t n = (size_t)nz * ncols;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        const int64_t c = (int64_t)(q / nz), i = (int64_t)(q % nz);
        const double t = vc[q] * dt;
        vdt2[(size_t)(i + kHZ) * pitch + (j0 + c + kHX)] = (float)(t * t);
    }
}
struct NaPerm {
    int k[kNaMaxSrc];
};
// wavelet steps t0 .. t0+nts-1 (FP64 [t][k] in staging) -> FP32 amplitudes
// [t][sorted k]
__global__ void na_amp(const double *w, int64_t t0, int64_t nts, int nsrc, NaPerm perm,
                       float *amp) {
    const size_t n = (size_t)nts * nsrc;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        const int64_t t = (int64_t)(q / nsrc);
        const int k = (int)(q % nsrc);
        amp[(size_t)(t0 + t) * nsrc + k] = (float)w[(size_t)t * nsrc + perm.k[k]];
    }
}
// CSR pointer over wave tiles: ptr[t] = number of (tile-sorted) sources whose
// tile is below t, t = 0 .. ntiles
__global__ void na_src_ptr(int *ptr, int ntiles, NaPerm tiles, int nsrc) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles; t += gridDim.x * blockDim.x) {
        int c = 0;
        for (int k = 0; k < nsrc; ++k) c += tiles.k[k] < t;
        ptr[t] = c;
    }
}
__global__ void na_f32_to_f64(const float *a, size_t n, double *out) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        out[q] = a[q];
}
// columns j0 .. j0+ncols-1 of a compact x-fastest slice -> column-major FP64
__global__ void na_slice_out(const float *slice, int64_t nz, int64_t nx, int64_t j0,
                             int64_t ncols, double *out) {
    const size_t n = (size_t)nz * ncols;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        const int64_t c = (int64_t)(q / nz), i = (int64_t)(q % nz);
        out[q] = slice[(size_t)i * nx + j0 + c];
    }
}
__global__ void na_order(int2 *order, int n) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x)
        order[b] = make_int2(0, b);
}

static unsigned na_blocks(size_t n) {
    return (unsigned)std::max<size_t>(1, std::min<size_t>(1024, (n + 255) / 256));
}

// Byte offsets of every buffer in the scratch array (256-byte aligned).
struct NaLayout {
    size_t p0, p1, vdt2, zphi0, zphi1, zpsi, zbm1, xphi0, xphi1, xpsi, xbm1;
    size_t src_ptr, src_cell, src_amp, rec, snap, order, health, stage, total;
};

static NaLayout na_layout(int64_t nz, int64_t nx, int64_t nt, int nsrc, int nB, bool snap,
                          int64_t pitch, int64_t rows, int ntiles, int nblk) {
    NaLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const size_t lev = (size_t)pitch * rows * 4;
    L.p0 = take(lev);
    L.p1 = take(lev);
    L.vdt2 = take(lev);
    const size_t zs = (size_t)nB * nx * 4, xs = (size_t)nz * 2 * nB * 4;
    L.zphi0 = take(zs);
    L.zphi1 = take(zs);
    L.zpsi = take(zs);
    L.zbm1 = take(zs);
    L.xphi0 = take(xs);
    L.xphi1 = take(xs);
    L.xpsi = take(xs);
    L.xbm1 = take(xs);
    L.src_ptr = take((size_t)(ntiles + 1) * 4);
    L.src_cell = take((size_t)std::max(nsrc, 1) * 4);
    L.src_amp = take((size_t)nt * std::max(nsrc, 1) * 4);
    L.rec = take((size_t)nt * nx * 4);
    L.snap = take(snap ? (size_t)nz * nx * 4 : 0);
    L.order = take((size_t)nblk * sizeof(int2));
    L.health = take(4);
    L.stage = take(kStageBytes);
    L.total = off;
    return L;
}

// Device address of `scratch` on each device, found once (hipGetSymbolAddress
// with that device current); the array's name is only a host-side stand-in.
static unsigned char *g_scratch[kNaMaxDev];
static std::mutex g_na_lock[kNaMaxDev];

static unsigned char *scratch_on(int dev) {
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        SSSI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) fail(SSSI_EDEVICE, "no device address for the scratch array");
        g_scratch[dev] = (unsigned char *)ptr;
    }
    return g_scratch[dev];
}

static void na_run(const double *v, int64_t nz, int64_t nx, int64_t nt, int nsrc,
                   const int *src_iz, const int *src_ix, const double *wavelet, int N, int nB,
                   double dz, double dx, double dt, double *data, double *snapshot,
                   int64_t snap_stride) {
    // ---- arguments (the checks of the allocating entry, without its host threads)
    validate_common(2, nz, nx, 1, nt, N, nB, dz, dx, 1.0, dt, true);
    if (!v) fail(SSSI_EINVAL, "v must not be NULL");
    for (int64_t q = 0; q < nz * nx; ++q)
        if (!(v[q] > 0.0) || !std::isfinite(v[q]))
            fail(SSSI_EINVAL, "v must be finite and positive (element %lld is %g)", (long long)q,
                 v[q]);
    if (snapshot && snap_stride < 1) fail(SSSI_EINVAL, "snap_stride must be >= 1");
    if (nsrc < 0 || nsrc > kNaMaxSrc)
        fail(SSSI_EINVAL, "at most %d point sources without allocation (got %d)", kNaMaxSrc,
             nsrc);
    if (nsrc > 0 && (!src_iz || !src_ix || !wavelet))
        fail(SSSI_EINVAL, "source positions and wavelet must not be NULL");
    for (int k = 0; k < nsrc; ++k)
        if (src_iz[k] < 0 || src_iz[k] >= nz || src_ix[k] < 0 || src_ix[k] >= nx)
            fail(SSSI_EINVAL, "source %d at (%d,%d) outside the %lldx%lld grid", k, src_iz[k],
                 src_ix[k], (long long)nz, (long long)nx);
    if ((int64_t)nB * nx * 4 >= ((int64_t)1 << 31) || nz * 2 * nB * 4 >= ((int64_t)1 << 31))
        fail(SSSI_EINVAL, "2-D damping-strip state exceeds 2 GiB per array");

    // ---- geometry of the tile path (4-row tiles, no interior march)
    constexpr int rz = 4;
    const int64_t pitch = round_up(kHX + round_up(nx, kTX) + 32, 64);
    const int64_t rows = kHZ + round_up(nz, kTZmax) + kHZbot;
    const int ntx = (int)((nx + kTX - 1) / kTX), nwz = (int)((nz + rz - 1) / rz);
    const int ntz = (int)((nz + 4 * rz - 1) / (4 * rz));
    const int nblk = ntx * ntz, ntiles = ntx * nwz;
    const NaLayout L = na_layout(nz, nx, nt, nsrc, nB, snapshot != nullptr, pitch, rows, ntiles,
                                 nblk);
    if (L.total > kScratchBytes)
        fail(SSSI_EINVAL,
             "problem needs %.1f MiB of device scratch, above the %zu MiB this entry point holds "
             "(%lldx%lld grid, nt %lld): use sssi_fwd_cpml_2d_points",
             L.total / 1048576.0, kScratchBytes >> 20, (long long)nz, (long long)nx,
             (long long)nt);
    const int64_t stage_n = (int64_t)(kStageBytes / 8);
    if (nz > stage_n || nx > stage_n || nsrc > stage_n)
        fail(SSSI_EINVAL, "grid dimension above %lld for the staging area", (long long)stage_n);

    int dev = 0;
    SSSI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDev) fail(SSSI_EDEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_na_lock[dev]);
    unsigned char *S = scratch_on(dev);
    hipStream_t st = nullptr;  // the null stream: nothing to create
    auto F = [&](size_t o) { return (float *)(S + o); };
    auto I = [&](size_t o) { return (int *)(S + o); };
    double *stage = (double *)(S + L.stage);

    // ---- state: zero levels and memory variables, arm the watchdog
    auto zero = [&](size_t o, size_t bytes) {
        hipLaunchKernelGGL(na_fill_f32, dim3(na_blocks(bytes / 4)), dim3(256), 0, st, F(o),
                           bytes / 4, 0.f);
        SSSI_HIP(hipGetLastError());
    };
    const size_t lev = (size_t)pitch * rows * 4;
    zero(L.p0, lev);
    zero(L.p1, lev);
    zero(L.vdt2, lev);
    const size_t zs = (size_t)nB * nx * 4, xs = (size_t)nz * 2 * nB * 4;
    for (size_t o : {L.zphi0, L.zphi1, L.zpsi}) zero(o, zs);
    for (size_t o : {L.xphi0, L.xphi1, L.xpsi}) zero(o, xs);
    zero(L.rec, (size_t)nt * nx * 4);
    hipLaunchKernelGGL(na_fill_i32, dim3(1), dim3(64), 0, st, I(L.health), (size_t)1, kHealthy);
    hipLaunchKernelGGL(na_order, dim3(na_blocks(nblk)), dim3(256), 0, st, (int2 *)(S + L.order),
                       nblk);
    SSSI_HIP(hipGetLastError());

    // ---- v^2 dt^2: column chunks of v through the staging area
    {
        const int64_t cols = std::max<int64_t>(1, stage_n / nz);
        for (int64_t j0 = 0; j0 < nx; j0 += cols) {
            const int64_t nc = std::min(cols, nx - j0);
            SSSI_HIP(hipMemcpy(stage, v + j0 * nz, (size_t)nc * nz * 8, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(na_vdt2, dim3(na_blocks((size_t)nc * nz)), dim3(256), 0, st,
                               stage, nz, j0, nc, dt, F(L.vdt2), pitch);
            SSSI_HIP(hipGetLastError());
        }
    }
    // ---- b-1 tables: formed on the host (the allocating path's arithmetic) in
    // chunks of a stack buffer
    {
        float buf[4096];
        auto put = [&](size_t o, int64_t n, bool xtab) {
            for (int64_t q0 = 0; q0 < n; q0 += 4096) {
                const int64_t m = std::min<int64_t>(4096, n - q0);
                for (int64_t q = 0; q < m; ++q)
                    buf[q] = xtab ? cpml2d_bm1_x(v, nz, nx, nB, dx, dt, q0 + q)
                                  : cpml2d_bm1_z(v, nz, nx, nB, dz, dt, q0 + q);
                SSSI_HIP(hipMemcpy(F(o) + q0, buf, (size_t)m * 4, hipMemcpyHostToDevice));
            }
        };
        if (nB > 0) {
            put(L.xbm1, nz * 2 * nB, true);
            put(L.zbm1, (int64_t)nB * nx, false);
        }
    }
    // ---- point sources: stable sort by wave tile (coincident points add in
    // source order), CSR pointer on the device, amplitudes in chunks
    if (nsrc > 0) {
        NaPerm perm{}, tiles{};
        int cell[kNaMaxSrc];
        for (int k = 0; k < nsrc; ++k) perm.k[k] = k;
        auto tile_of = [&](int k) { return (src_iz[k] / rz) * ntx + src_ix[k] / kTX; };
        for (int a = 1; a < nsrc; ++a)  // insertion sort: stable
            for (int b = a; b > 0 && tile_of(perm.k[b - 1]) > tile_of(perm.k[b]); --b)
                std::swap(perm.k[b - 1], perm.k[b]);
        for (int k = 0; k < nsrc; ++k) {
            tiles.k[k] = tile_of(perm.k[k]);
            cell[k] = src_iz[perm.k[k]] * (int)nx + src_ix[perm.k[k]];
        }
        SSSI_HIP(hipMemcpy(I(L.src_cell), cell, (size_t)nsrc * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(na_src_ptr, dim3(na_blocks((size_t)ntiles + 1)), dim3(256), 0, st,
                           I(L.src_ptr), ntiles, tiles, nsrc);
        SSSI_HIP(hipGetLastError());
        const int64_t steps = std::max<int64_t>(1, stage_n / nsrc);
        for (int64_t t0 = 0; t0 < nt; t0 += steps) {
            const int64_t m = std::min(steps, nt - t0);
            SSSI_HIP(hipMemcpy(stage, wavelet + t0 * nsrc, (size_t)m * nsrc * 8,
                               hipMemcpyHostToDevice));
            hipLaunchKernelGGL(na_amp, dim3(na_blocks((size_t)m * nsrc)), dim3(256), 0, st, stage,
                               t0, m, nsrc, perm, F(L.src_amp));
            SSSI_HIP(hipGetLastError());
        }
    }

    // ---- descriptor
    Job2D J;
    std::memset(&J, 0, sizeof(J));
    stencil_2d(N, dz, dx, J);
    J.nz = (int)nz; J.nx = (int)nx; J.pitch = (int)pitch; J.rows = (int)rows;
    J.ntx = ntx; J.nwz = nwz; J.nB = nB;
    J.start_step = 0; J.nt = (int)nt; J.reverse = 0;
    J.p[0] = F(L.p0); J.p[1] = F(L.p1);
    J.vdt2 = F(L.vdt2);
    if (nB > 0) {
        J.zphi[0] = F(L.zphi0); J.zphi[1] = F(L.zphi1); J.zpsi = F(L.zpsi); J.zbm1 = F(L.zbm1);
        J.xphi[0] = F(L.xphi0); J.xphi[1] = F(L.xphi1); J.xpsi = F(L.xpsi); J.xbm1 = F(L.xbm1);
    }
    if (nsrc > 0) {
        J.src_ptr = I(L.src_ptr); J.src_cell = I(L.src_cell); J.src_amp = F(L.src_amp);
        J.nsrc = nsrc; J.src_nt = (int)nt;
    }
    J.rec = F(L.rec);
    J.snap_stride = 1; J.ck_stride = 1; J.im_stride = 1;
    J.health = I(L.health);
    const int2 *order = (const int2 *)(S + L.order);

    // ---- time loop; with snapshots, segment k runs steps k*stride .. and
    // slice k is written at its first step, then converted and copied out
    const int64_t stride = snapshot ? snap_stride : nt;
    for (int64_t t0 = 0, k = 0; t0 < nt; t0 += stride, ++k) {
        const int64_t t1 = std::min(nt, t0 + stride);
        Job2D Jk = J;
        if (snapshot) {
            Jk.snap = F(L.snap);
            Jk.snap_stride = (int)stride;
            Jk.snap_slot0 = (int)k;
        }
        for (int64_t t = t0; t < t1; ++t)
            launch_cpml2d_single(N, rz, 0, Jk, order, nblk, (int)t, st);
        if (snapshot) {
            const int64_t cols = std::max<int64_t>(1, stage_n / nz);
            double *dst = snapshot + (size_t)k * nz * nx;
            for (int64_t j0 = 0; j0 < nx; j0 += cols) {
                const int64_t nc = std::min(cols, nx - j0);
                hipLaunchKernelGGL(na_slice_out, dim3(na_blocks((size_t)nc * nz)), dim3(256), 0,
                                   st, F(L.snap), nz, nx, j0, nc, stage);
                SSSI_HIP(hipGetLastError());
                SSSI_HIP(hipMemcpy(dst + j0 * nz, stage, (size_t)nc * nz * 8,
                                   hipMemcpyDeviceToHost));
            }
        }
    }
    // ---- receivers out ([t][j] is data(nx, nt)), then the watchdog
    if (data) {
        const int64_t steps = std::max<int64_t>(1, stage_n / nx);
        for (int64_t t0 = 0; t0 < nt; t0 += steps) {
            const int64_t m = std::min(steps, nt - t0);
            hipLaunchKernelGGL(na_f32_to_f64, dim3(na_blocks((size_t)m * nx)), dim3(256), 0, st,
                               F(L.rec) + t0 * nx, (size_t)m * nx, stage);
            SSSI_HIP(hipGetLastError());
            SSSI_HIP(hipMemcpy(data + t0 * nx, stage, (size_t)m * nx * 8, hipMemcpyDeviceToHost));
        }
    }
    launch_nonfinite_scan(F(nt % 2 == 0 ? L.p0 : L.p1), (size_t)pitch * rows, I(L.health),
                          (int)nt, st);
    const int h = read_health(I(L.health), st);
    if (h != kHealthy)
        fail(SSSI_EUNSTABLE,
             "non-finite wavefield (NaN/Inf) at step %d of %lld: the time step is likely beyond "
             "the scheme's stability limit", h + 1, (long long)nt);
}

}  // namespace sssi

using namespace sssi;

extern "C" int sssi_fwd_cpml_2d_points_noalloc(const double *v, int64_t nz, int64_t nx,
                                               int64_t nt, int nsrc, const int *src_iz,
                                               const int *src_ix, const double *wavelet,
                                               int nDiffOrder, int nBoundary, double dz,
                                               double dx, double dt, double *data,
                                               double *snapshot, int64_t snap_stride) {
    try {
        na_run(v, nz, nx, nt, nsrc, src_iz, src_ix, wavelet, nDiffOrder, nBoundary, dz, dx, dt,
               data, snapshot, snap_stride);
        set_last_error("");
        return SSSI_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return SSSI_EDEVICE;
    }
}
