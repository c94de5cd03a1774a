// abi.cpp -- the C ABI of libsssi (include/sssi.h): argument validation,
// host<->device conversion (column-major FP64 <-> padded x-fastest FP32) and
// the drop-in one-shot entry points built on the device-resident handles.
#include <algorithm>
#include <cstring>
#include <memory>

#include "prop.h"
#include "tune.h"

using namespace sssi;

struct sssi_prop {
    std::unique_ptr<Prop> impl;
};

namespace {

template <class F>
int guarded(F &&f) {
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

void need(const void *p, const char *name) {
    if (!p) fail(SSSI_EINVAL, "%s must not be NULL", name);
}

// Dense source (nz,nx[,ny],nt) -> point list over its support (cells that are
// nonzero at any step) with one wavelet column per point.  Adding 0 at every
// other cell is what the dense form does, so the arithmetic is unchanged.
struct PointList {
    std::vector<int> iz, ix, iy;
    std::vector<double> w;  // [t][k]
};

PointList dense_to_points(const double *src, int64_t nz, int64_t nx, int64_t ny, int64_t nt) {
    const int64_t nn = nz * nx * ny;
    std::vector<uint8_t> mark(nn, 0);
    // the support: each host thread scans its own cells over every step (the
    // dense array is nn * nt doubles, 230 MB at 120 x 240 x 1000)
    const int64_t chunk = std::max<int64_t>(256, ((int64_t)1 << 22) / std::max<int64_t>(1, nt));
    parallel_ranges(nn, [&](int64_t a, int64_t b) {
        for (int64_t t = 0; t < nt; ++t) {
            const double *s = src + t * nn;
            for (int64_t q = a; q < b; ++q)
                if (s[q] != 0.0) mark[q] = 1;
        }
    }, chunk);
    PointList pl;
    std::vector<int64_t> idx;
    for (int64_t q = 0; q < nn; ++q)
        if (mark[q]) idx.push_back(q);
    const int64_t n = (int64_t)idx.size();
    for (int64_t q : idx) {
        const int64_t i = q % nz, j = (q / nz) % nx, k = q / (nz * nx);
        pl.iz.push_back((int)i);
        pl.ix.push_back((int)j);
        pl.iy.push_back((int)k);
    }
    pl.w.resize((size_t)nt * n);
    for (int64_t t = 0; t < nt; ++t)
        for (int64_t k = 0; k < n; ++k) pl.w[t * n + k] = src[t * nn + idx[k]];
    return pl;
}

// Run a whole propagation and bring data / snapshots back in the reference
// layout.  Snapshots go through a device staging buffer of at most ~256 MB.
void run_oneshot(Prop &P, int64_t nt, bool reverse, double *data, double *snap, int64_t stride) {
    const int64_t cells = P.cells();
    const int64_t nsnap = (nt + stride - 1) / stride;
    if (!snap) {
        P.run(reverse ? nt - 1 : 0, nt, reverse, data != nullptr, nullptr, nullptr);
    } else {
        const int64_t per = std::max<int64_t>(1, (int64_t)(256ll << 20) / (cells * 4));
        const int64_t K = std::min(per, nsnap);
        DevBuf<float> stage((size_t)K * cells);
        HostBuf<float> host((size_t)K * cells);
        int64_t done = 0;  // steps executed
        while (done < nt) {
            // slots covered by this chunk
            int64_t t0 = reverse ? nt - 1 - done : done;
            int64_t s_lo, s_hi, count;
            if (!reverse) {
                s_lo = (t0 + stride - 1) / stride;
                s_hi = std::min(nsnap - 1, s_lo + K - 1);
                const int64_t t_end = std::min(nt, (s_hi + 1) * stride);  // exclusive
                count = t_end - t0;
            } else {
                s_hi = t0 / stride;
                s_lo = std::max<int64_t>(0, s_hi - K + 1);
                const int64_t t_stop = s_lo * stride;  // inclusive lowest t
                count = t0 - t_stop + 1;
            }
            SnapSink sink;
            sink.dev = stage.p;
            sink.stride = stride;
            sink.slot0 = s_lo;
            P.run(t0, count, reverse, data != nullptr, &sink, nullptr);
            const int64_t ns = s_hi - s_lo + 1;
            SSSI_HIP(hipMemcpyAsync(host.p, stage.p, (size_t)ns * cells * sizeof(float),
                                    hipMemcpyDeviceToHost, P.stream));
            SSSI_HIP(hipStreamSynchronize(P.stream));
            // device slice is x fastest, [z][y][x]: (i*ny + k)*nx + j;
            // reference layout is z fastest: (k*nx + j)*nz + i.  Split over
            // host threads by (slice, y, 64-column block); each item writes
            // whole z columns.
            const int64_t nz = P.nz, nx = P.nx, ny = P.ny, nxb = (nx + 63) / 64;
            parallel_ranges(ns * ny * nxb, [&](int64_t a, int64_t b) {
                for (int64_t it = a; it < b; ++it) {
                    const int64_t s = it / (ny * nxb), k = (it / nxb) % ny, jb = it % nxb;
                    const float *src = host.p + s * cells;
                    double *dst = snap + (s_lo + s) * cells;
                    for (int64_t j = jb * 64; j < std::min(nx, jb * 64 + 64); ++j)
                        for (int64_t i = 0; i < nz; ++i)
                            dst[(k * nx + j) * nz + i] = src[(i * ny + k) * nx + j];
                }
            }, 1);
            done += count;
        }
    }
    P.scan_field();
    P.sync();
    if (data) P.get_data(data, nt);
}

}  // namespace

extern "C" {

const char *sssi_last_error(void) { return last_error_text(); }

int sssi_version(void) { return 100; }

int sssi_device_count(int *count) {
    return guarded([&] {
        need(count, "count");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        *count = (e == hipSuccess) ? n : 0;
    });
}

int sssi_set_device(int device) {
    return guarded([&] { SSSI_HIP(hipSetDevice(device)); });
}

int sssi_set_autotune(int on) {
    return guarded([&] { set_autotune(on != 0); });
}

int sssi_fwd_cpml_2d_points(const double *v, int64_t nz, int64_t nx, int64_t nt, int nsrc,
                            const int *src_iz, const int *src_ix, const double *wavelet,
                            int nDiffOrder, int nBoundary, double dz, double dx, double dt,
                            double *data, double *snapshot, int64_t snap_stride) {
    return guarded([&] {
        need(v, "v");
        validate_common(2, nz, nx, 1, nt, nDiffOrder, nBoundary, dz, dx, 1.0, dt, true);
        validate_velocity(v, nz * nx, "v");
        if (snapshot && snap_stride < 1) fail(SSSI_EINVAL, "snap_stride must be >= 1");
        std::unique_ptr<Prop> P(make_prop2d(v, nz, nx, SSSI_ABC_CPML, nDiffOrder, nBoundary, dz,
                                            dx, dt));
        P->set_points(nsrc, src_iz, src_ix, nullptr, wavelet, nt);
        run_oneshot(*P, nt, false, data, snapshot, snapshot ? snap_stride : 1);
        P->check_health();
    });
}

int sssi_fwdTimeCpmlFor2dAw(const double *v, int64_t nz, int64_t nx, const double *source,
                            int64_t nt, int nDiffOrder, int nBoundary, double dz, double dx,
                            double dt, double *data, double *snapshot) {
    return guarded([&] {
        need(v, "v");
        need(source, "source");
        validate_common(2, nz, nx, 1, nt, nDiffOrder, nBoundary, dz, dx, 1.0, dt, true);
        validate_velocity(v, nz * nx, "v");
        PointList pl = dense_to_points(source, nz, nx, 1, nt);
        std::unique_ptr<Prop> P(make_prop2d(v, nz, nx, SSSI_ABC_CPML, nDiffOrder, nBoundary, dz,
                                            dx, dt));
        P->set_points((int)pl.iz.size(), pl.iz.data(), pl.ix.data(), nullptr, pl.w.data(), nt);
        run_oneshot(*P, nt, false, data, snapshot, 1);
        P->check_health();
    });
}

int sssi_rvs_cpml_2d_ex(const double *v, int64_t nz, int64_t nx, const double *data, int64_t nt,
                        int nDiffOrder, int nBoundary, double dz, double dx, double dt,
                        double *model, double *snapshot, int64_t snap_stride) {
    return guarded([&] {
        need(v, "v");
        need(data, "data");
        validate_common(2, nz, nx, 1, nt, nDiffOrder, nBoundary, dz, dx, 1.0, dt, true);
        validate_velocity(v, nz * nx, "v");
        if (snapshot && snap_stride < 1) fail(SSSI_EINVAL, "snap_stride must be >= 1");
        std::unique_ptr<Prop> P(make_prop2d(v, nz, nx, SSSI_ABC_CPML, nDiffOrder, nBoundary, dz,
                                            dx, dt));
        P->set_surface(data, nt);
        run_oneshot(*P, nt, true, nullptr, snapshot, snapshot ? snap_stride : 1);
        if (model) P->get_model(model);
        P->check_health();
    });
}

int sssi_rvsTimeCpmlFor2dAw(const double *v, int64_t nz, int64_t nx, const double *data,
                            int64_t nt, int nDiffOrder, int nBoundary, double dz, double dx,
                            double dt, double *model, double *snapshot) {
    return sssi_rvs_cpml_2d_ex(v, nz, nx, data, nt, nDiffOrder, nBoundary, dz, dx, dt, model,
                               snapshot, 1);
}

int sssi_fwd_sponge_2d_points(const double *v, int64_t nz, int64_t nx, int64_t nt, int nsrc,
                              const int *src_iz, const int *src_ix, const double *wavelet,
                              int nBoundary, double dz, double dx, double dt, double *data,
                              double *snapshot, int64_t snap_stride) {
    return guarded([&] {
        need(v, "v");
        validate_common(2, nz, nx, 1, nt, 1, nBoundary, dz, dx, 1.0, dt, false);
        validate_sponge_grid(nz, nx);
        validate_velocity(v, nz * nx, "v");
        if (snapshot && snap_stride < 1) fail(SSSI_EINVAL, "snap_stride must be >= 1");
        std::unique_ptr<Prop> P(make_prop2d(v, nz, nx, SSSI_ABC_SPONGE, 1, nBoundary, dz, dx, dt));
        P->set_points(nsrc, src_iz, src_ix, nullptr, wavelet, nt);
        run_oneshot(*P, nt, false, data, snapshot, snapshot ? snap_stride : 1);
        P->check_health();
    });
}

int sssi_fwdTimeSpongeFor2dAw(const double *v, int64_t nz, int64_t nx, const double *source,
                              int64_t nt, int nBoundary, double dz, double dx, double dt,
                              double *data, double *snapshot) {
    return guarded([&] {
        need(v, "v");
        need(source, "source");
        validate_common(2, nz, nx, 1, nt, 1, nBoundary, dz, dx, 1.0, dt, false);
        validate_sponge_grid(nz, nx);
        validate_velocity(v, nz * nx, "v");
        PointList pl = dense_to_points(source, nz, nx, 1, nt);
        std::unique_ptr<Prop> P(make_prop2d(v, nz, nx, SSSI_ABC_SPONGE, 1, nBoundary, dz, dx, dt));
        P->set_points((int)pl.iz.size(), pl.iz.data(), pl.ix.data(), nullptr, pl.w.data(), nt);
        run_oneshot(*P, nt, false, data, snapshot, 1);
        P->check_health();
    });
}

int sssi_rvs_sponge_2d_ex(const double *v, int64_t nz, int64_t nx, const double *data,
                          int64_t nt, int nBoundary, double dz, double dx, double dt,
                          double *model, double *snapshot, int64_t snap_stride) {
    return guarded([&] {
        need(v, "v");
        need(data, "data");
        validate_common(2, nz, nx, 1, nt, 1, nBoundary, dz, dx, 1.0, dt, false);
        validate_sponge_grid(nz, nx);
        validate_velocity(v, nz * nx, "v");
        if (snapshot && snap_stride < 1) fail(SSSI_EINVAL, "snap_stride must be >= 1");
        std::unique_ptr<Prop> P(make_prop2d(v, nz, nx, SSSI_ABC_SPONGE, 1, nBoundary, dz, dx, dt));
        P->set_surface(data, nt);
        run_oneshot(*P, nt, true, nullptr, snapshot, snapshot ? snap_stride : 1);
        if (model) P->get_model(model);
        P->check_health();
    });
}

int sssi_rvsTimeSpongeFor2dAw(const double *v, int64_t nz, int64_t nx, const double *data,
                              int64_t nt, int nBoundary, double dz, double dx, double dt,
                              double *model, double *snapshot) {
    return sssi_rvs_sponge_2d_ex(v, nz, nx, data, nt, nBoundary, dz, dx, dt, model, snapshot, 1);
}

int sssi_fwd_cpml_3d_points(const double *v, int64_t nz, int64_t nx, int64_t ny, int64_t nt,
                            int nsrc, const int *src_iz, const int *src_ix, const int *src_iy,
                            const double *wavelet, int nDiffOrder, int nBoundary, double dz,
                            double dx, double dy, double dt, double *data, double *snapshot,
                            int64_t snap_stride) {
    return guarded([&] {
        need(v, "v");
        validate_common(3, nz, nx, ny, nt, nDiffOrder, nBoundary, dz, dx, dy, dt, true);
        validate_velocity(v, nz * nx * ny, "v");
        if (snapshot && snap_stride < 1) fail(SSSI_EINVAL, "snap_stride must be >= 1");
        std::unique_ptr<Prop> P(make_prop3d(v, nz, nx, ny, nDiffOrder, nBoundary, dz, dx, dy, dt));
        P->set_points(nsrc, src_iz, src_ix, src_iy, wavelet, nt);
        run_oneshot(*P, nt, false, data, snapshot, snapshot ? snap_stride : 1);
        P->check_health();
    });
}

int sssi_fwdTimeCpmlFor3dAw(const double *v, int64_t nz, int64_t nx, int64_t ny,
                            const double *source, int64_t nt, int nDiffOrder, int nBoundary,
                            double dz, double dx, double dy, double dt, double *data,
                            double *snapshot) {
    return guarded([&] {
        need(v, "v");
        need(source, "source");
        validate_common(3, nz, nx, ny, nt, nDiffOrder, nBoundary, dz, dx, dy, dt, true);
        validate_velocity(v, nz * nx * ny, "v");
        PointList pl = dense_to_points(source, nz, nx, ny, nt);
        std::unique_ptr<Prop> P(make_prop3d(v, nz, nx, ny, nDiffOrder, nBoundary, dz, dx, dy, dt));
        P->set_points((int)pl.iz.size(), pl.iz.data(), pl.ix.data(), pl.iy.data(), pl.w.data(),
                      nt);
        run_oneshot(*P, nt, false, data, snapshot, 1);
        P->check_health();
    });
}

// ---------------------------------------------------------------- handles
int sssi_prop2d_create(const double *v, int64_t nz, int64_t nx, int abc, int nDiffOrder,
                       int nBoundary, double dz, double dx, double dt, sssi_prop **out) {
    return guarded([&] {
        need(v, "v");
        need(out, "out");
        if (abc != SSSI_ABC_CPML && abc != SSSI_ABC_SPONGE) fail(SSSI_EINVAL, "unknown abc %d", abc);
        validate_common(2, nz, nx, 1, 1, nDiffOrder, nBoundary, dz, dx, 1.0, dt,
                        abc == SSSI_ABC_CPML);
        if (abc == SSSI_ABC_SPONGE) validate_sponge_grid(nz, nx);
        validate_velocity(v, nz * nx, "v");
        auto h = std::make_unique<sssi_prop>();
        h->impl.reset(make_prop2d(v, nz, nx, abc, abc == SSSI_ABC_CPML ? nDiffOrder : 1,
                                  nBoundary, dz, dx, dt));
        *out = h.release();
    });
}

int sssi_prop3d_create(const double *v, int64_t nz, int64_t nx, int64_t ny, int nDiffOrder,
                       int nBoundary, double dz, double dx, double dy, double dt,
                       sssi_prop **out) {
    return guarded([&] {
        need(v, "v");
        need(out, "out");
        validate_common(3, nz, nx, ny, 1, nDiffOrder, nBoundary, dz, dx, dy, dt, true);
        validate_velocity(v, nz * nx * ny, "v");
        auto h = std::make_unique<sssi_prop>();
        h->impl.reset(make_prop3d(v, nz, nx, ny, nDiffOrder, nBoundary, dz, dx, dy, dt));
        *out = h.release();
    });
}

void sssi_prop_destroy(sssi_prop *p) { delete p; }

static Prop &impl(sssi_prop *p) {
    if (!p || !p->impl) fail(SSSI_EINVAL, "NULL handle");
    return *p->impl;
}

int sssi_prop_set_points(sssi_prop *p, int nsrc, const int *src_iz, const int *src_ix,
                         const int *src_iy, const double *wavelet, int64_t nt) {
    return guarded([&] { impl(p).set_points(nsrc, src_iz, src_ix, src_iy, wavelet, nt); });
}

int sssi_prop_set_surface(sssi_prop *p, const double *data, int64_t nt) {
    return guarded([&] { impl(p).set_surface(data, nt); });
}

int sssi_prop_reset(sssi_prop *p) {
    return guarded([&] { impl(p).reset(); });
}

int sssi_prop_run(sssi_prop *p, int64_t nt, int reverse, int record) {
    return guarded([&] {
        if (nt < 1) fail(SSSI_EINVAL, "nt must be >= 1");
        impl(p).run(reverse ? nt - 1 : 0, nt, reverse != 0, record != 0, nullptr, nullptr);
    });
}

int sssi_prop_run_steps(sssi_prop *p, int64_t t_first, int64_t count, int reverse, int record) {
    return guarded([&] {
        if (count < 1) fail(SSSI_EINVAL, "count must be >= 1");
        if (t_first < 0 || (reverse && t_first - count + 1 < 0))
            fail(SSSI_EINVAL, "steps outside t >= 0 (t_first=%lld count=%lld)",
                 (long long)t_first, (long long)count);
        impl(p).run(t_first, count, reverse != 0, record != 0, nullptr, nullptr);
    });
}

int sssi_prop_run_timed(sssi_prop *p, int64_t nt, int reverse, int record, double *ms,
                        int64_t *launches) {
    return guarded([&] {
        need(ms, "ms");
        if (nt < 1) fail(SSSI_EINVAL, "nt must be >= 1");
        Prop &P = impl(p);
        hipEvent_t e0, e1;
        SSSI_HIP(hipEventCreate(&e0));
        SSSI_HIP(hipEventCreate(&e1));
        LaunchStats st;
        SSSI_HIP(hipEventRecord(e0, P.stream));
        P.run(reverse ? nt - 1 : 0, nt, reverse != 0, record != 0, nullptr, &st);
        SSSI_HIP(hipEventRecord(e1, P.stream));
        SSSI_HIP(hipEventSynchronize(e1));
        float f = 0.f;
        SSSI_HIP(hipEventElapsedTime(&f, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = f;
        if (launches) *launches = st.launches;
        P.check_health();
    });
}

int sssi_prop_sync(sssi_prop *p) {
    return guarded([&] {
        impl(p).sync();
        impl(p).check_health();
    });
}

int sssi_prop_get_data(sssi_prop *p, double *data, int64_t nt) {
    return guarded([&] {
        need(data, "data");
        impl(p).get_data(data, nt);
        impl(p).check_health();
    });
}

int sssi_prop_get_field(sssi_prop *p, double *field) {
    return guarded([&] {
        need(field, "field");
        impl(p).scan_field();
        impl(p).get_field(field);
        impl(p).check_health();
    });
}

int sssi_prop_get_model(sssi_prop *p, double *model) {
    return guarded([&] {
        need(model, "model");
        impl(p).get_model(model);
    });
}

int sssi_prop_alg_bytes_per_step(sssi_prop *p, double *bytes, double *cells) {
    return guarded([&] {
        need(bytes, "bytes");
        need(cells, "cells");
        impl(p).alg_bytes(bytes, cells);
    });
}

int sssi_prop_layout(sssi_prop *p, int *a, int *b) {
    return guarded([&] {
        need(a, "a");
        need(b, "b");
        impl(p).layout(a, b);
    });
}

}  // extern "C"
