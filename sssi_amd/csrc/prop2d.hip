// prop2d.hip -- device-resident 2-D propagation handles: CPML (the fused step
// of cpml2d.hip; fwdTimeCpmlFor2dAw_mex.c / rvsTimeCpmlFor2dAw_mex.c) and the
// Cerjan sponge (src/fwdTimeSpongeFor2dAw.m:69-106, src/rvsTimeSpongeFor2dAw.m).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>

#include "cpml2d.h"
#include "devbuf.h"
#include "tune.h"

namespace sssi {

// Sponge step: p_new = b p - g p_old + ax (pR+pL) + az (pD+pU) - v^2dt^2 s with
// b = 2 - 2ax - 2az, then multiplied by g.  ax, az, v^2dt^2 are formed from
// vdt = v*dt in-kernel.
// The reference damps both time levels after the rotation
// (fwdTimeSpongeFor2dAw.m:83-92); the stored level carries one factor and the
// second is applied when it is read back as p_old.
struct Sponge2dArgs {
    const float *pcur;
    float *pold;
    const float *vdt;
    const float *g;  // g[m], m < nB
    int nz, nx, pitch, rows, nB;
    float idx2, idz2;  // 1/dx^2, 1/dz^2
    const int *src_cell;
    const float *src_amp;
    int nsrc;
    const float *surf;
    float *rec;
    float *snap;
    int *health;  // non-finite watchdog word (top row), or null
    int step;     // launch step reported to it
};

// Rows per wavefront of the sponge step; a block is 4 wavefronts stacked in z.
// At 1040 x 3080 with the XCD-banded tile order below, us per step at 1, 2, 4,
// 8, 16 rows: 13.9, 9.2, 7.4, 7.9, 9.4.
constexpr int kSpRZ = 4;

// A wavefront owns 64 columns x kSpRZ rows (lane = column).  Every load is a
// raw buffer load with the wave-uniform row offset in the scalar offset: the
// column window p(i0-1 .. i0+kSpRZ) once, the x neighbours as loads shifted by
// one column, then p_old and v*dt.  Rows i0-1 .. i0+kSpRZ and columns j-1 ..
// j+1 of every wavefront that has a row inside the model lie in the padded
// array (16 rows above, at least 32 below the interior rounded up to 64 rows,
// 32 columns each side).  Lanes outside the model store nothing, so the zero
// padding read as the boundary stays zero.
// The launch is 1-D over ntx * ntz block tiles (x fastest).  Block b runs on
// XCD b % 8, so block b takes tile t of XCD b % 8's contiguous band of the
// tile list: row neighbours then share one L2 for the rows between them.
template <int RZ>
__global__ void __launch_bounds__(256) sponge2d_step(Sponge2dArgs a, int ntx) {
    const int lane = threadIdx.x & 63, w = wave_id();
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb / 8, rem = nb % 8, x = b % 8;
    const int t = x * q + min(x, rem) + b / 8;
    const int ty = t / ntx, tx = t - ty * ntx;
    const int j = tx * 64 + lane;
    const int i0 = (ty * 4 + w) * RZ;
    if (i0 >= a.nz) return;
    const bool jin = j < a.nx;
    const uint32_t rowb = (uint32_t)a.pitch * 4u;
    const uint32_t nbytes = rowb * (uint32_t)a.rows;
    const __amdgpu_buffer_rsrc_t rP = buf_rsrc(a.pcur, nbytes);
    const __amdgpu_buffer_rsrc_t rO = buf_rsrc(a.pold, nbytes);
    const __amdgpu_buffer_rsrc_t rV = buf_rsrc(a.vdt, nbytes);
    const uint32_t vcol = (uint32_t)(j + kHX) * 4u;
    const uint32_t srow = (uint32_t)(i0 + kHZ) * rowb;

    float col[RZ + 2], cl[RZ], cr[RZ], po[RZ], vd[RZ];
#pragma unroll
    for (int r = 0; r < RZ + 2; ++r) col[r] = buf_ld(rP, vcol, srow + (r - 1) * rowb);
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        cl[r] = buf_ld(rP, vcol - 4u, srow + r * rowb);
        cr[r] = buf_ld(rP, vcol + 4u, srow + r * rowb);
        po[r] = buf_ld(rO, vcol, srow + r * rowb);
        vd[r] = buf_ld(rV, vcol, srow + r * rowb);
    }

    // damping factor: the bottom strip's by row, else the side strips' by column
    const int nB = a.nB;
    float gx = 1.f;
    if (jin && j < nB) gx = a.g[j];
    else if (jin && j >= a.nx - nB) gx = a.g[a.nx - 1 - j];
    // source terms, summed per cell in source order, then surface injection
    float s[RZ];
#pragma unroll
    for (int r = 0; r < RZ; ++r) s[r] = 0.f;
    for (int k = 0; k < a.nsrc; ++k) {
        const int cell = ld_uniform(a.src_cell + k);
        const int ci = cell / a.nx, cj = cell - ci * a.nx;
        if (ci >= i0 && ci < i0 + RZ && cj == j) {
            const float amp = ld_uniform(a.src_amp + k);
#pragma unroll
            for (int r = 0; r < RZ; ++r)
                if (ci == i0 + r) s[r] += amp;
        }
    }
    if (a.surf && i0 == 0 && jin) s[0] += a.surf[j];

    float pn0 = 0.f;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        const int i = i0 + r;
        if (i >= a.nz) break;
        const float gf = i >= a.nz - nB ? ld_uniform(a.g + (a.nz - 1 - i)) : gx;
        const float v2 = vd[r] * vd[r];
        const float ax = v2 * a.idx2, az = v2 * a.idz2;
        // b p - p_old + ax (pR+pL) + az (pD+pU) regrouped as
        // p + (p - p_old) + ax((pR-p)+(pL-p)) + az((pD-p)+(pU-p)): the FP32
        // rounding then scales with the increments instead of with 2p.
        const float pc = col[r + 1];
        const float lap = ax * ((cr[r] - pc) + (cl[r] - pc)) + az * ((col[r + 2] - pc) + (col[r] - pc));
        float pn = pc + ((pc - gf * po[r]) + (lap - v2 * s[r]));
        pn *= gf;
        buf_st(pn, rO, jin ? vcol : kOOB, srow + r * rowb);
        if (r == 0) pn0 = pn;
        if (jin) {
            if (a.rec && i == 0) a.rec[j] = pn;
            if (a.snap) a.snap[(size_t)i * a.nx + j] = pn;
        }
    }
    if (a.health && i0 == 0) report_nonfinite(a.health, jin && nonfinite(pn0), a.step);
}

// ---------------------------------------------------------------------------
class Prop2D final : public Prop {
   public:
    // CPML
    std::unique_ptr<Cpml2dGrid> g;
    DevSources srcs;
    DevBuf<int2> order;
    int nblk = 0;
    StepGraph graph;   // captured time loop (re-captured when the descriptor changes)
    int64_t done = 0;  // steps since reset (parity of the current level)
    // sponge
    int64_t pitch = 0, rows = 0;
    DevBuf<float> sp[2], vdt, gtab;
    DevBuf<int> sp_health;
    std::vector<int> h_cells;
    int cur = 0;

    Prop2D(const double *v, int64_t nz_, int64_t nx_, int abc_, int N_, int nB_, double dz_,
           double dx_, double dt_) {
        dim = 2;
        abc = abc_;
        nz = nz_; nx = nx_; ny = 1; N = N_; nB = nB_; dz = dz_; dx = dx_; dt = dt_;
        SSSI_HIP(hipGetDevice(&device));
        stream = stream_acquire();
        {
            double vmax = 0.0;
            for (int64_t q = 0; q < nz * nx; ++q) vmax = std::max(vmax, v[q]);
            courant = vmax * dt * std::sqrt(1.0 / (dz * dz) + 1.0 / (dx * dx));
        }
        if (abc == SSSI_ABC_CPML) {
            g.reset(new Cpml2dGrid(v, nz, nx, N, nB, dz, dx, dt));
            if (autotune_enabled()) tune_layout();
            build_order();
        } else {
            validate_sponge_grid(nz, nx);
            pitch = round_up(kHX + round_up(nx, kTX) + 32, 64);
            rows = kHZ + round_up(nz, kTZmax) + kHZbot;
            const size_t np = (size_t)pitch * rows;
            sp[0].alloc(np);
            sp[1].alloc(np);
            vdt.alloc(np);
            upload_model(v, nz, nx, 1, dt, false, vdt.p, vdt.n, pitch, 0, kHZ, 0, kHX, stream);
            std::vector<float> gg(std::max(nB, 1), 1.f);
            const double alpha = 0.015;
            for (int m = 0; m < nB; ++m) {
                const double e = std::exp(-((alpha * (nB - m)) * (alpha * (nB - m))));
                gg[m] = (float)std::pow(e, 10);
            }
            gtab.alloc(gg.size());
            SSSI_HIP(hipMemcpy(gtab.p, gg.data(), gtab.bytes(), hipMemcpyHostToDevice));
            sp_health.alloc(1);
        }
        reset();
    }

    ~Prop2D() override {
        graph.clear();  // before ~Prop destroys the stream its replays ran on
    }

    void build_order() {
        std::vector<int2> ord;
        build_block_order({g.get()}, ord);
        nblk = (int)ord.size();
        order.alloc(ord.size());
        SSSI_HIP(hipMemcpy(order.p, ord.data(), order.bytes(), hipMemcpyHostToDevice));
    }

    // Autotuner: time every instantiated layout (tile height, march groups) on
    // this grid -- kTuneSteps launches after a short warm-up, HIP events on the
    // handle's stream -- keep the fastest, and cache it for the shape.
    void tune_layout() {
        const TuneKey key{2, nz, nx, 1, N, nB};
        static const int cand[][2] = {{4, 0}, {4, 2}, {4, 3}, {4, 4}, {4, 8},
                                      {8, 0}, {8, 2}, {8, 4}};
        int a = 0, b = 0;
        if (tune_lookup(key, a, b)) {
            // a cache entry is used only if it names an instantiated layout
            for (const auto &c : cand)
                if (c[0] == a && c[1] == b) {
                    g->set_layout(a, b);
                    return;
                }
        }
        // Layouts within a few percent of each other: every candidate is timed
        // over kTuneSteps launches in each of two passes (best pass kept), after
        // a first warm-up long enough for the clocks to settle.
        constexpr int kWarm = 20, kFirstWarm = 400, kTuneSteps = 200, kPasses = 2;
        hipEvent_t e0, e1;
        SSSI_HIP(hipEventCreate(&e0));
        SSSI_HIP(hipEventCreate(&e1));
        float best_ms = 1e30f;
        int best_a = g->rz, best_b = g->ng;
        try {
            for (int pass = 0; pass < kPasses; ++pass)
                for (const auto &c : cand) {
                    g->set_layout(c[0], c[1]);
                    build_order();
                    const Job2D J = g->job();  // no sources, receivers or snapshots
                    const int warm = pass == 0 && &c == &cand[0] ? kFirstWarm : kWarm;
                    for (int q = 0; q < warm; ++q)
                        launch_cpml2d_single(*g, J, order.p, nblk, q, stream);
                    SSSI_HIP(hipEventRecord(e0, stream));
                    for (int q = 0; q < kTuneSteps; ++q)
                        launch_cpml2d_single(*g, J, order.p, nblk, warm + q, stream);
                    SSSI_HIP(hipEventRecord(e1, stream));
                    SSSI_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    SSSI_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best_a = c[0];
                        best_b = c[1];
                    }
                }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        g->set_layout(best_a, best_b);
        tune_store(key, best_a, best_b);
    }

    void layout(int *a, int *b) const override {
        *a = g ? g->rz : 0;
        *b = g ? g->ng : 0;
    }

    void use_stream(hipStream_t s) override {
        graph.clear();  // its replays ran on the stream being replaced
        Prop::use_stream(s);
    }

    int64_t cells() const override { return nz * nx; }

    void reset() override {
        if (g) {
            g->zero_state(stream);
        } else {
            SSSI_HIP(hipMemsetAsync(sp[0].p, 0, sp[0].bytes(), stream));
            SSSI_HIP(hipMemsetAsync(sp[1].p, 0, sp[1].bytes(), stream));
            arm_health(sp_health.p, stream);
        }
        done = 0;
        steps_done = 0;
        cur = 0;
        SSSI_HIP(hipStreamSynchronize(stream));
    }

    void set_points(int n, const int *iz, const int *ix, const int *, const double *wavelet,
                    int64_t nt) override {
        if (g) {
            srcs.set(*g, n, iz, ix, wavelet, nt);
            nsrc = srcs.n;
            src_nt = nt;
            return;
        }
        if (n < 0 || nt < 1 || (n > 0 && (!iz || !ix || !wavelet)))
            fail(SSSI_EINVAL, "set_points: bad arguments");
        for (int k = 0; k < n; ++k)
            if (iz[k] < 0 || iz[k] >= nz || ix[k] < 0 || ix[k] >= nx)
                fail(SSSI_EINVAL, "set_points: source %d at (%d,%d) outside the %lldx%lld grid",
                     k, iz[k], ix[k], (long long)nz, (long long)nx);
        nsrc = n;
        src_nt = nt;
        std::vector<int> cell(std::max(n, 1), 0);
        for (int k = 0; k < n; ++k) cell[k] = iz[k] * (int)nx + ix[k];
        std::vector<float> amp((size_t)nt * std::max(n, 1), 0.f);
        for (int64_t t = 0; t < nt; ++t)
            for (int k = 0; k < n; ++k) amp[(size_t)t * n + k] = (float)wavelet[t * n + k];
        src_cell.alloc(cell.size());
        src_amp.alloc(amp.size());
        SSSI_HIP(hipMemcpy(src_cell.p, cell.data(), src_cell.bytes(), hipMemcpyHostToDevice));
        SSSI_HIP(hipMemcpy(src_amp.p, amp.data(), src_amp.bytes(), hipMemcpyHostToDevice));
    }

    void set_surface(const double *data, int64_t nt) override {
        if (!data || nt < 1) fail(SSSI_EINVAL, "set_surface: bad arguments");
        std::vector<float> h((size_t)nt * nx);
        for (size_t q = 0; q < h.size(); ++q) h[q] = (float)data[q];
        surf.alloc(h.size());
        surf_nt = nt;
        SSSI_HIP(hipMemcpy(surf.p, h.data(), surf.bytes(), hipMemcpyHostToDevice));
    }


    // Descriptor for `count` steps from time index t_first (see the kernel's
    // t / parity mapping): the level current now sits at parity rel_0 & 1.
    Job2D make_job(int64_t t_first, bool reverse, bool record, const SnapSink *sink,
                   int *step0) {
        Job2D J = g->job();
        const int curlev = (int)(done & 1);
        int rel0;
        if (!reverse) {
            J.start_step = (int)(done - t_first);
            J.nt = (int)(t_first + 1);
            rel0 = (int)t_first;
        } else {
            J.start_step = (int)done;
            J.nt = (int)(t_first + 1);
            rel0 = 0;
        }
        J.reverse = reverse ? 1 : 0;
        const int pa = rel0 & 1;
        J.p[pa] = g->p[curlev].p;
        J.p[pa ^ 1] = g->p[curlev ^ 1].p;
        J.zphi[pa] = g->zphi[curlev].p;
        J.zphi[pa ^ 1] = g->zphi[curlev ^ 1].p;
        J.xphi[pa] = g->xphi[curlev].p;
        J.xphi[pa ^ 1] = g->xphi[curlev ^ 1].p;
        if (srcs.n > 0) {
            J.src_ptr = srcs.ptr.p;
            J.src_cell = srcs.cell.p;
            J.src_amp = srcs.amp.p;
            J.nsrc = srcs.n;
            J.src_nt = srcs.nt;
        }
        if (surf.p) {
            J.surf = surf.p;
            J.surf_nt = (int)surf_nt;
        }
        if (record) J.rec = rec.p;
        if (sink && sink->dev) {
            J.snap = sink->dev;
            J.snap_stride = (int)sink->stride;
            J.snap_slot0 = (int)sink->slot0;
        }
        *step0 = (int)done;
        return J;
    }

    int *health_word() const override { return g ? g->health.p : sp_health.p; }

    void scan_field() override {
        if (g) g->scan_nonfinite((int)(done & 1), (int)done, stream);
        else launch_nonfinite_scan(sp[cur].p, sp[cur].n, sp_health.p, (int)done, stream);
    }

    void step_sponge(int64_t t, bool reverse, bool record, float *snap, int step) {
        Sponge2dArgs a{};
        a.health = sp_health.p;
        a.step = step;
        a.pcur = sp[cur].p;
        a.pold = sp[cur ^ 1].p;
        a.vdt = vdt.p;
        a.g = gtab.p;
        a.nz = (int)nz; a.nx = (int)nx; a.pitch = (int)pitch; a.rows = (int)rows; a.nB = nB;
        a.idx2 = (float)(1.0 / (dx * dx));
        a.idz2 = (float)(1.0 / (dz * dz));
        if (!reverse && nsrc > 0 && t < src_nt) {
            a.src_cell = src_cell.p;
            a.src_amp = src_amp.p + (size_t)t * nsrc;
            a.nsrc = nsrc;
        }
        if (reverse && surf.p && t < surf_nt) a.surf = surf.p + (size_t)t * nx;
        if (record) a.rec = rec.p + (size_t)t * nx;
        a.snap = snap;
        const int ntx = (int)((nx + 63) / 64);
        const int64_t ntz = (nz + 4 * kSpRZ - 1) / (4 * kSpRZ);
        hipLaunchKernelGGL(sponge2d_step<kSpRZ>, dim3((unsigned)(ntx * ntz)), dim3(256), 0, stream,
                           a, ntx);
        SSSI_HIP(hipGetLastError());
        cur ^= 1;
    }

    void run(int64_t t_first, int64_t count, bool reverse, bool record, const SnapSink *sink,
             LaunchStats *st) override {
        if (count < 0) fail(SSSI_EINVAL, "run: negative step count");
        if (count == 0) return;
        if (record) ensure_rec(reverse ? t_first + 1 : t_first + count, nx);
        if (g) {
            int step0;
            const Job2D J = make_job(t_first, reverse, record, sink, &step0);
            graph.run(&J, sizeof(J), step0, count, stream,
                      [&](int step, const int *base, hipStream_t s) {
                          launch_cpml2d_single(*g, J, order.p, nblk, step, s, base);
                      });
            if (st) st->launches += count;
            done += count;
            steps_done = done;
            return;
        }
        for (int64_t q = 0; q < count; ++q) {
            const int64_t t = reverse ? t_first - q : t_first + q;
            float *snap = nullptr;
            if (sink && sink->dev && t % sink->stride == 0)
                snap = sink->dev + (size_t)(t / sink->stride - sink->slot0) * nz * nx;
            step_sponge(t, reverse, record, snap, (int)(done + q));
            if (st) st->launches++;
        }
        done += count;
        steps_done = done;
    }

    void get_data(double *data, int64_t nt) override {
        if (nt > rec_nt || !rec.p)
            fail(SSSI_ESTATE, "get_data: no recorded traces for %lld steps", (long long)nt);
        HostBuf<float> h((size_t)nt * nx);
        SSSI_HIP(hipMemcpyAsync(h.p, rec.p, h.bytes(), hipMemcpyDeviceToHost, stream));
        SSSI_HIP(hipStreamSynchronize(stream));
        parallel_ranges((int64_t)h.n, [&](int64_t a, int64_t b) {
            for (int64_t q = a; q < b; ++q) data[q] = h.p[q];  // [t][j] == data(nx,nt)
        }, 1 << 16);
    }

    // sponge levels: interior of sp[which], optionally times the second damping factor
    void download_sponge(int which, double *out, bool padded_model) {
        std::vector<float> h((size_t)pitch * rows);
        SSSI_HIP(hipMemcpyAsync(h.data(), sp[which].p, h.size() * sizeof(float),
                                hipMemcpyDeviceToHost, stream));
        SSSI_HIP(hipStreamSynchronize(stream));
        std::vector<float> gf(std::max(nB, 1));
        SSSI_HIP(hipMemcpy(gf.data(), gtab.p, gtab.bytes(), hipMemcpyDeviceToHost));
        auto gfac = [&](int64_t i, int64_t j) -> float {
            if (i >= nz - nB) return gf[nz - 1 - i];
            if (j < nB) return gf[j];
            if (j >= nx - nB) return gf[nx - 1 - j];
            return 1.f;
        };
        if (!padded_model) {
            for (int64_t j = 0; j < nx; ++j)
                for (int64_t i = 0; i < nz; ++i)
                    out[j * nz + i] = h[(size_t)(i + kHZ) * pitch + j + kHX];
            return;
        }
        const int64_t mz = nz + 2, mx = nx + 2;
        for (int64_t q = 0; q < mz * mx; ++q) out[q] = 0.0;
        for (int64_t j = 0; j < nx; ++j)
            for (int64_t i = 0; i < nz; ++i)
                out[(j + 1) * mz + i + 1] = gfac(i, j) * h[(size_t)(i + kHZ) * pitch + j + kHX];
    }

    void get_field(double *field) override {
        if (g) g->download((int)(done & 1), field, stream);
        else download_sponge(cur, field, false);
    }

    void get_model(double *model) override {
        if (!g) {
            download_sponge(cur ^ 1, model, true);
            return;
        }
        // reference model layout: (nz+2l) x (nx+2l), l-wide zero halo
        const int64_t l = 2 * N - 1, mz = nz + 2 * l, mx = nx + 2 * l;
        std::vector<double> in((size_t)nz * nx);
        g->download((int)((done & 1) ^ 1), in.data(), stream);
        for (int64_t q = 0; q < mz * mx; ++q) model[q] = 0.0;
        for (int64_t j = 0; j < nx; ++j)
            for (int64_t i = 0; i < nz; ++i) model[(j + l) * mz + i + l] = in[j * nz + i];
    }

    void alg_bytes(double *bytes, double *ncells) override {
        *ncells = (double)nz * nx;
        *bytes = g ? g->alg_bytes_per_step() : 16.0 * (double)nz * nx + 4.0 * nx;
    }
};

void Prop::check_health() {
    const int h = read_health(health_word(), stream);
    if (h != kHealthy)
        fail(SSSI_EUNSTABLE,
             "non-finite wavefield (NaN/Inf) at step %d of %lld since the last reset: the time "
             "step is likely beyond the scheme's stability limit (v_max dt sqrt(sum 1/d^2) = %.3f)",
             h + 1, (long long)steps_done, courant);
}

Prop *make_prop2d(const double *v, int64_t nz, int64_t nx, int abc, int N, int nB, double dz,
                  double dx, double dt) {
    return new Prop2D(v, nz, nx, abc, N, nB, dz, dx, dt);
}

}  // namespace sssi
