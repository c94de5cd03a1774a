// rtm.hip -- 2-D reverse-time migration on gfx950, the imaging loop of
// applications/rtm_2d_acoustic/mainRtmTimeCpmlFor2dAw.m:147-315:
//   per shot   dataTrue   = fwd(V,  src)                     (:165, cached per shot)
//              dataSmooth = fwd(VS, src)  + source-wavefield checkpoints
//              rvs(V, dataTrue - dataSmooth)                  (:252, :256)
//              M += f .* r ; s2 += f.^2  at the checkpoint times (:270-273)
//   stack      Stacked += M ./ (eps + s2)                     (:315)
// Each shot runs on an extended sub-window of the model (extBoundary: nB
// edge-replicated columns left/right and rows below).  Checkpoints are FP32 or
// 16-bit (FP16 with a per-checkpoint scale, BF16 unscaled); the per-shot accumulators are
// FP32 and the stack FP64, accumulated in shot order (deterministic).
//
// Two execution modes over the same kernels: one shot per launch (each time
// step of each shot is its own launch) or batched (one launch per time step
// covers the tiles of every shot, strip tiles first).
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "cpml2d.h"

namespace sssi {

__global__ void rtm_stack_image(const float *num, const float *den, double eps, int nz_ext,
                                int nx_ext, int nz_phys, int nB, int x_begin, int x_count,
                                double *stack, int nz_stack, float *img_out) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nz_ext || j >= nx_ext) return;
    const size_t cell = (size_t)i * nx_ext + j;
    const float img = (float)((double)num[cell] / (eps + (double)den[cell]));
    if (img_out) img_out[cell] = img;
    // physical part of the window -> model columns x_begin .. x_begin+x_count-1
    const int jj = j - nB;
    if (stack && i < nz_phys && jj >= 0 && jj < x_count)
        stack[(size_t)(x_begin + jj) * nz_stack + i] += (double)img;  // column-major stack
}

// A cached per-shot image (FP32 values, as computed) into the stack.
__global__ void rtm_stack_cached(const float *img, int nz_ext, int nx_ext, int nz_phys, int nB,
                                 int x_begin, int x_count, double *stack, int nz_stack) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nz_ext || j >= nx_ext) return;
    const int jj = j - nB;
    if (i < nz_phys && jj >= 0 && jj < x_count)
        stack[(size_t)(x_begin + jj) * nz_stack + i] += (double)img[(size_t)i * nx_ext + j];
}

__global__ void fill_f32(float *p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct RtmShot {
    int64_t x_begin = 0, x_count = 0;  // window in the physical model
    int src_ix = 0;                    // source column within the window (physical)
    std::vector<double> wavelet;
    std::unique_ptr<Cpml2dGrid> gV, gS;
    DevSources src;
    DevBuf<float> dtrue, dsmooth;      // [t][j]
    DevBuf<float> ck32, stage;
    DevBuf<uint16_t> ck16;
    DevBuf<unsigned> ckmax;
    DevBuf<float> ckscale;
    DevBuf<float> num, den, img;
    bool have_true = false;
    bool img_cached = false;  // image loaded from the result cache (not migrated)
    uint64_t hV = 0, hVS = 0, hW = 0;  // checksums of the window models and wavelet
};

// FNV-1a over raw bytes: identifies the inputs a cached result was made from.
static uint64_t fnv1a(const void *p, size_t n, uint64_t h = 1469598103934665603ull) {
    const uint8_t *b = (const uint8_t *)p;
    for (size_t q = 0; q < n; ++q) {
        h ^= b[q];
        h *= 1099511628211ull;
    }
    return h;
}

class Rtm2D {
   public:
    OwnedStream stream;  // first member: destroyed after the device state
    int64_t nz = 0, nx = 0, nt = 0;
    int N = 4, nB = 0, stride = 1, fmt = kCkF32;
    double dz = 0, dx = 0, dt = 0, eps = 1e-3;
    std::vector<double> V, VS;  // physical models, column-major
    std::vector<std::unique_ptr<RtmShot>> shots;
    DevBuf<double> stack;       // nz x nx physical, column-major
    int64_t launches = 0;
    // per-shot result cache (MAT-files in cache_dir, empty: off)
    std::string cache_dir;
    int loaded_true = 0, loaded_img = 0, saved_files = 0;

    Rtm2D(const double *V_, const double *VS_, int64_t nz_, int64_t nx_, int N_, int nB_,
          double dz_, double dx_, double dt_, int64_t nt_, int stride_, int fmt_, double eps_)
        : nz(nz_), nx(nx_), nt(nt_), N(N_), nB(nB_), stride(stride_), fmt(fmt_), dz(dz_), dx(dx_),
          dt(dt_), eps(eps_) {
        V.assign(V_, V_ + nz * nx);
        VS.assign(VS_, VS_ + nz * nx);
        stream.create();
        stack.alloc((size_t)nz * nx);
        SSSI_HIP(hipMemsetAsync(stack.p, 0, stack.bytes(), stream));
    }

    int64_t nck() const { return (nt + stride - 1) / stride; }

    // extBoundary of the window [x0, x0+n) (src/extBoundary.m:22-39)
    std::vector<double> window_ext(const std::vector<double> &M, int64_t x0, int64_t n) const {
        const int64_t ze = nz + nB, xe = n + 2 * nB;
        std::vector<double> out((size_t)ze * xe);
        for (int64_t j = 0; j < xe; ++j) {
            const int64_t src_col = x0 + std::min(std::max<int64_t>(j - nB, 0), n - 1);
            for (int64_t i = 0; i < ze; ++i) {
                const int64_t si = std::min(i, nz - 1);
                out[(size_t)j * ze + i] = M[(size_t)src_col * nz + si];
            }
        }
        return out;
    }

    int add_shot(int64_t x_begin, int64_t x_count, int src_ix, const double *wavelet) {
        if (x_begin < 0 || x_count < 1 || x_begin + x_count > nx)
            fail(SSSI_EINVAL, "shot window [%lld, %lld) outside the model", (long long)x_begin,
                 (long long)(x_begin + x_count));
        if (src_ix < 0 || src_ix >= x_count) fail(SSSI_EINVAL, "source column outside the window");
        if (!wavelet) fail(SSSI_EINVAL, "wavelet must not be NULL");
        auto s = std::make_unique<RtmShot>();
        s->x_begin = x_begin;
        s->x_count = x_count;
        s->src_ix = src_ix;
        s->wavelet.assign(wavelet, wavelet + nt);
        const int64_t ze = nz + nB, xe = x_count + 2 * nB;
        std::vector<double> wv = window_ext(V, x_begin, x_count);
        std::vector<double> ws = window_ext(VS, x_begin, x_count);
        s->gV.reset(new Cpml2dGrid(wv.data(), ze, xe, N, nB, dz, dx, dt));
        s->gS.reset(new Cpml2dGrid(ws.data(), ze, xe, N, nB, dz, dx, dt));
        s->hV = fnv1a(wv.data(), wv.size() * sizeof(double));
        s->hVS = fnv1a(ws.data(), ws.size() * sizeof(double));
        s->hW = fnv1a(s->wavelet.data(), s->wavelet.size() * sizeof(double));
        const int iz = 0, ix = src_ix + nB;
        s->src.set(*s->gS, 1, &iz, &ix, s->wavelet.data(), nt);
        const size_t ncell = (size_t)ze * xe;
        s->dtrue.alloc((size_t)nt * xe);
        s->dsmooth.alloc((size_t)nt * xe);
        if (fmt == kCkF32) {
            s->ck32.alloc((size_t)nck() * ncell);
        } else {
            if (fmt == kCkF16) s->stage.alloc(2 * ncell);
            s->ck16.alloc((size_t)nck() * ncell);
        }
        s->ckmax.alloc((size_t)nck() * kCkSlots);
        s->ckscale.alloc((size_t)nck());
        s->num.alloc(ncell);
        s->den.alloc(ncell);
        s->img.alloc(ncell);
        shots.push_back(std::move(s));
        return (int)shots.size() - 1;
    }

    // --- job descriptors of the three phases
    Job2D job_true(RtmShot &s) {
        Job2D J = s.gV->job();
        J.start_step = 0; J.nt = (int)nt; J.reverse = 0;
        J.src_ptr = s.src.ptr.p; J.src_cell = s.src.cell.p; J.src_amp = s.src.amp.p;
        J.nsrc = s.src.n; J.src_nt = s.src.nt;
        J.rec = s.dtrue.p;
        return J;
    }
    Job2D job_fwd(RtmShot &s) {
        Job2D J = s.gS->job();
        J.start_step = 0; J.nt = (int)nt; J.reverse = 0;
        J.src_ptr = s.src.ptr.p; J.src_cell = s.src.cell.p; J.src_amp = s.src.amp.p;
        J.nsrc = s.src.n; J.src_nt = s.src.nt;
        J.rec = s.dsmooth.p;
        J.ck_stride = stride;
        J.ck_fmt = fmt;
        J.ck32 = s.ck32.p; J.ck_stage = s.stage.p; J.ck16 = s.ck16.p;
        J.ck_max = s.ckmax.p; J.ck_scale = s.ckscale.p;
        return J;
    }
    Job2D job_rvs(RtmShot &s) {
        Job2D J = s.gV->job();
        J.start_step = 0; J.nt = (int)nt; J.reverse = 1;
        J.surf = s.dtrue.p; J.surf2 = s.dsmooth.p; J.surf_nt = (int)nt;
        J.im_stride = stride; J.im_fmt = fmt;
        J.im_ck32 = s.ck32.p; J.im_ck16 = s.ck16.p; J.im_scale = s.ckscale.p;
        J.im_num = s.num.p; J.im_den = s.den.p;
        return J;
    }

    void prepare_fwd(RtmShot &s) {
        s.gS->zero_state(stream);
        SSSI_HIP(hipMemsetAsync(s.ckmax.p, 0, s.ckmax.bytes(), stream));
    }
    void prepare_rvs(RtmShot &s) {
        s.gV->zero_state(stream);
        SSSI_HIP(hipMemsetAsync(s.num.p, 0, s.num.bytes(), stream));
        SSSI_HIP(hipMemsetAsync(s.den.p, 0, s.den.bytes(), stream));
    }
    void finish_fwd(RtmShot &s, const Job2D &J) {
        if (fmt == kCkF16) {
            launch_ck_pack(J, (int)((nt - 1) / stride), stream);
            launches++;
        }
    }
    void finish_shot(RtmShot &s) {
        const int ze = (int)(nz + nB), xe = (int)(s.x_count + 2 * nB);
        dim3 grid((unsigned)((xe + 63) / 64), (unsigned)((ze + 3) / 4));
        hipLaunchKernelGGL(rtm_stack_image, grid, dim3(256), 0, stream, s.num.p, s.den.p, eps, ze,
                           xe, (int)nz, nB, (int)s.x_begin, (int)s.x_count, stack.p, (int)nz,
                           s.img.p);
        SSSI_HIP(hipGetLastError());
        launches++;
    }

    // --- one shot per launch
    void run_single_phase(Cpml2dGrid &g, const Job2D &J) {
        std::vector<int2> ord;
        build_block_order({&g}, ord);
        DevBuf<int2> o(ord.size());
        SSSI_HIP(hipMemcpyAsync(o.p, ord.data(), o.bytes(), hipMemcpyHostToDevice, stream));
        {
            StepGraph graph;
            graph.run(&J, sizeof(J), 0, nt, stream, [&](int step, const int *base, hipStream_t s) {
                launch_cpml2d_single(g, J, o.p, (int)ord.size(), step, s, base);
            });
            launches += nt;
            SSSI_HIP(hipStreamSynchronize(stream));  // keep `o` alive until the launches ran
        }
    }

    // --- many shots per launch
    void run_batch_phase(std::vector<Cpml2dGrid *> grids, const std::vector<Job2D> &jobs) {
        std::vector<const Cpml2dGrid *> cg(grids.begin(), grids.end());
        for (const Cpml2dGrid *g : cg)
            if (g->rz != grids[0]->rz || g->ng != grids[0]->ng)
                fail(SSSI_ESTATE, "batched shots need one tile layout (rz %d/%d, groups %d/%d)",
                     g->rz, grids[0]->rz, g->ng, grids[0]->ng);
        std::vector<int2> ord;
        build_block_order(cg, ord);
        DevBuf<int2> o(ord.size());
        DevBuf<Job2D> dj(jobs.size());
        SSSI_HIP(hipMemcpyAsync(o.p, ord.data(), o.bytes(), hipMemcpyHostToDevice, stream));
        SSSI_HIP(hipMemcpyAsync(dj.p, jobs.data(), dj.bytes(), hipMemcpyHostToDevice, stream));
        {
            const void *key = dj.p;  // one capture per phase
            StepGraph graph;
            graph.run(&key, sizeof(key), 0, nt, stream,
                      [&](int step, const int *base, hipStream_t s) {
                          launch_cpml2d_batch(N, grids[0]->rz, grids[0]->ng, dj.p, o.p,
                                              (int)ord.size(), step, s, base);
                      });
            launches += nt;
            SSSI_HIP(hipStreamSynchronize(stream));
        }
    }

    // ---- per-shot result cache.  The reference's apps name their files by the
    // shot's source column xs (1-based) in the model: dataTrue<xs>.mat holds
    // dataTrue (nx_ext x nt), image<xs>.mat the shot's image M./(eps+s2) on its
    // extended window, StackedRTM.mat the stack; a file is written only if it
    // does not exist yet (mainRtmTimeCpmlFor2dAw.m:170-184, :259-264,
    // mainRtmTimeSpongeFor2dAw.m:343-346).  Each file also holds sssiShot, the
    // parameters and input checksums it was made from; a file made for other
    // inputs is refused rather than used.
    int xs_of(const RtmShot &s) const { return (int)(s.x_begin + s.src_ix + 1); }
    std::string cache_path(const char *what, const RtmShot &s) const {
        return cache_dir + "/" + what + std::to_string(xs_of(s)) + ".mat";
    }
    std::vector<double> key_of(const RtmShot &s, bool image) const {
        auto split = [](uint64_t h, std::vector<double> &k) {
            k.push_back((double)(h >> 32));
            k.push_back((double)(h & 0xffffffffu));
        };
        std::vector<double> k = {(double)s.x_begin, (double)s.x_count, (double)s.src_ix,
                                 (double)nz, (double)nt, (double)N, (double)nB, dz, dx, dt};
        split(s.hV, k);
        split(s.hW, k);
        if (image) {
            k.push_back((double)stride);
            k.push_back((double)fmt);
            k.push_back(eps);
            split(s.hVS, k);
        }
        return k;
    }
    // the variable `name` of a cached file if it exists and was made from the
    // same inputs (n values)
    bool cache_load(const std::string &path, const char *name, const std::vector<double> &key,
                    std::vector<double> &val, size_t n) {
        if (cache_dir.empty() || !file_exists(path.c_str())) return false;
        int nd = 0;
        int64_t dims[8] = {0};
        std::vector<double> k;
        bool same = mat_read(path.c_str(), "sssiShot", &nd, dims, nullptr, 0) && nd == 2 &&
                    (size_t)(dims[0] * dims[1]) == key.size();
        if (same) {
            k.resize(key.size());
            mat_read(path.c_str(), "sssiShot", nullptr, nullptr, k.data(), (int64_t)k.size());
            same = k == key;
        }
        if (!same)
            fail(SSSI_ESTATE, "cache file %s was made for other inputs (shot parameters or "
                 "models differ): remove it or use another cache directory", path.c_str());
        val.resize(n);
        if (!mat_read(path.c_str(), name, &nd, dims, nullptr, 0) ||
            (size_t)(nd == 2 ? dims[0] * dims[1] : 0) != n)
            fail(SSSI_ESTATE, "cache file %s holds no %s of %zu values", path.c_str(), name, n);
        mat_read(path.c_str(), name, nullptr, nullptr, val.data(), (int64_t)n);
        return true;
    }
    void cache_save(const std::string &path, const char *name, const double *val, int64_t d0,
                    int64_t d1, const std::vector<double> &key) {
        if (cache_dir.empty() || file_exists(path.c_str())) return;
        const char *names[2] = {name, "sssiShot"};
        const double *data[2] = {val, key.data()};
        const int nd[2] = {2, 2};
        const int64_t dims[4] = {d0, d1, 1, (int64_t)key.size()};
        mat_write(path.c_str(), 2, names, data, nd, dims);
        ++saved_files;
    }
    void check_cache_names() const {
        for (size_t a = 0; a < shots.size(); ++a)
            for (size_t b = a + 1; b < shots.size(); ++b)
                if (xs_of(*shots[a]) == xs_of(*shots[b]))
                    fail(SSSI_EINVAL, "shots %zu and %zu share source column %d, which names "
                         "their cache files", a, b, xs_of(*shots[a]));
    }
    // device FP32 array (n values) <-> host doubles
    void to_host(const float *d, size_t n, std::vector<double> &out) {
        std::vector<float> h(n);
        SSSI_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(float), hipMemcpyDeviceToHost, stream));
        SSSI_HIP(hipStreamSynchronize(stream));
        out.assign(h.begin(), h.end());
    }
    void to_device(const std::vector<double> &v, float *d) {
        std::vector<float> h(v.begin(), v.end());  // values are FP32 results: exact
        SSSI_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice,
                                stream));
        SSSI_HIP(hipStreamSynchronize(stream));
    }

    void precompute_true(bool batched) {
        loaded_true = 0;
        if (!cache_dir.empty()) {
            check_cache_names();
            std::vector<double> val;
            for (auto &s : shots) {
                if (s->have_true) continue;
                const size_t n = (size_t)nt * (s->x_count + 2 * nB);
                if (cache_load(cache_path("dataTrue", *s), "dataTrue", key_of(*s, false), val, n)) {
                    to_device(val, s->dtrue.p);  // data(nx_ext, nt) is [t][j]
                    s->have_true = true;
                    ++loaded_true;
                }
            }
        }
        std::vector<RtmShot *> todo;
        for (auto &s : shots)
            if (!s->have_true) todo.push_back(s.get());
        if (todo.empty()) return;
        if (batched) {
            std::vector<Cpml2dGrid *> g;
            std::vector<Job2D> j;
            for (auto *s : todo) {
                s->gV->zero_state(stream);
                g.push_back(s->gV.get());
                j.push_back(job_true(*s));
            }
            run_batch_phase(g, j);
        } else {
            for (auto *s : todo) {
                s->gV->zero_state(stream);
                run_single_phase(*s->gV, job_true(*s));
            }
        }
        SSSI_HIP(hipStreamSynchronize(stream));
        for (auto *s : todo) check_shot(*s, *s->gV, "V");
        for (auto *s : todo) s->have_true = true;
        if (!cache_dir.empty()) {
            std::vector<double> val;
            for (auto *s : todo) {
                const int64_t xe = s->x_count + 2 * nB;
                to_host(s->dtrue.p, (size_t)nt * xe, val);
                cache_save(cache_path("dataTrue", *s), "dataTrue", val.data(), xe, nt,
                           key_of(*s, false));
            }
        }
    }

    // watchdog of a shot's grid after its phase ran (the stream is idle)
    void check_shot(const RtmShot &s, const Cpml2dGrid &g, const char *model) {
        g.scan_nonfinite(0, (int)nt, stream);
        g.scan_nonfinite(1, (int)nt, stream);
        const int h = read_health(g.health.p, stream);
        if (h != kHealthy) {
            int k = 0;
            while (k < (int)shots.size() && shots[k].get() != &s) ++k;
            fail(SSSI_EUNSTABLE,
                 "shot %d: non-finite wavefield (NaN/Inf) at step %d of the propagation in %s: "
                 "the time step is likely beyond the scheme's stability limit", k, h + 1, model);
        }
    }

    // the imaging job of every shot: fwd(VS) + checkpoints, rvs(V) + imaging,
    // stack (in shot order); shots with a cached image are stacked from it
    void migrate(bool batched) {
        SSSI_HIP(hipMemsetAsync(stack.p, 0, stack.bytes(), stream));
        loaded_img = 0;
        std::vector<RtmShot *> run;
        {
            std::vector<double> val;
            if (!cache_dir.empty()) check_cache_names();
            for (auto &s : shots) {
                const int64_t ze = nz + nB, xe = s->x_count + 2 * nB;
                s->img_cached = false;
                if (cache_load(cache_path("image", *s), "image", key_of(*s, true), val,
                               (size_t)ze * xe)) {
                    // column-major (ze, xe) -> the device's [i][j]
                    std::vector<double> rows((size_t)ze * xe);
                    for (int64_t j = 0; j < xe; ++j)
                        for (int64_t i = 0; i < ze; ++i) rows[(size_t)i * xe + j] = val[j * ze + i];
                    to_device(rows, s->img.p);
                    s->img_cached = true;
                    ++loaded_img;
                } else {
                    run.push_back(s.get());
                }
            }
        }
        if (batched && !run.empty()) {
            std::vector<Cpml2dGrid *> gs, gv;
            std::vector<Job2D> jf, jr;
            for (RtmShot *s : run) {
                prepare_fwd(*s);
                gs.push_back(s->gS.get());
                jf.push_back(job_fwd(*s));
            }
            run_batch_phase(gs, jf);
            for (size_t k = 0; k < run.size(); ++k) finish_fwd(*run[k], jf[k]);
            for (RtmShot *s : run) {
                prepare_rvs(*s);
                gv.push_back(s->gV.get());
                jr.push_back(job_rvs(*s));
            }
            run_batch_phase(gv, jr);
        } else {
            for (RtmShot *s : run) {
                prepare_fwd(*s);
                const Job2D jf = job_fwd(*s);
                run_single_phase(*s->gS, jf);
                finish_fwd(*s, jf);
                prepare_rvs(*s);
                run_single_phase(*s->gV, job_rvs(*s));
            }
        }
        for (auto &s : shots) {
            if (s->img_cached) stack_cached(*s);
            else finish_shot(*s);
        }
        SSSI_HIP(hipStreamSynchronize(stream));
        for (RtmShot *s : run) {
            check_shot(*s, *s->gS, "VS");
            check_shot(*s, *s->gV, "V");
        }
        if (!cache_dir.empty()) {
            std::vector<double> val;
            for (RtmShot *s : run) {
                const int64_t ze = nz + nB, xe = s->x_count + 2 * nB;
                to_host(s->img.p, (size_t)ze * xe, val);
                std::vector<double> cm((size_t)ze * xe);
                for (int64_t j = 0; j < xe; ++j)
                    for (int64_t i = 0; i < ze; ++i) cm[j * ze + i] = val[(size_t)i * xe + j];
                cache_save(cache_path("image", *s), "image", cm.data(), ze, xe, key_of(*s, true));
            }
            const std::string sp = cache_dir + "/StackedRTM.mat";
            if (!file_exists(sp.c_str())) {
                std::vector<double> st((size_t)nz * nx);
                SSSI_HIP(hipMemcpyAsync(st.data(), stack.p, stack.bytes(), hipMemcpyDeviceToHost,
                                        stream));
                SSSI_HIP(hipStreamSynchronize(stream));
                const char *names[1] = {"StackedRTM"};
                const double *data[1] = {st.data()};
                const int nd[1] = {2};
                const int64_t dims[2] = {nz, nx};
                mat_write(sp.c_str(), 1, names, data, nd, dims);
                ++saved_files;
            }
        }
    }

    void stack_cached(RtmShot &s) {
        const int ze = (int)(nz + nB), xe = (int)(s.x_count + 2 * nB);
        dim3 grid((unsigned)((xe + 63) / 64), (unsigned)((ze + 3) / 4));
        hipLaunchKernelGGL(rtm_stack_cached, grid, dim3(256), 0, stream, s.img.p, ze, xe, (int)nz,
                           nB, (int)s.x_begin, (int)s.x_count, stack.p, (int)nz);
        SSSI_HIP(hipGetLastError());
        launches++;
    }
};

}  // namespace sssi

using namespace sssi;

struct sssi_rtm {
    std::unique_ptr<Rtm2D> impl;
};

namespace {
template <class F>
int rtm_guard(F &&f) {
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
Rtm2D &R(sssi_rtm *r) {
    if (!r || !r->impl) fail(SSSI_EINVAL, "NULL RTM handle");
    return *r->impl;
}
}  // namespace

extern "C" {

int sssi_rtm2d_create(const double *V, const double *VS, int64_t nz, int64_t nx, int nDiffOrder,
                      int nBoundary, double dz, double dx, double dt, int64_t nt, int img_stride,
                      int ck_format, double eps, sssi_rtm **out) {
    return rtm_guard([&] {
        if (!V || !VS || !out) fail(SSSI_EINVAL, "V, VS and out must not be NULL");
        validate_common(2, nz + nBoundary, 2 * (int64_t)nBoundary + 1, 1, nt, nDiffOrder, nBoundary,
                        dz, dx, 1.0, dt, true);
        if (nz < 1 || nx < 1 || nz * nx > ((int64_t)1 << 30))
            fail(SSSI_EINVAL, "model grid %lldx%lld out of range", (long long)nz, (long long)nx);
        validate_velocity(V, nz * nx, "V");
        validate_velocity(VS, nz * nx, "VS");
        if (img_stride < 1) fail(SSSI_EINVAL, "img_stride must be >= 1");
        if (ck_format < SSSI_CK_F32 || ck_format > SSSI_CK_BF16)
            fail(SSSI_EINVAL, "unknown checkpoint format %d", ck_format);
        if (!(eps >= 0) || !std::isfinite(eps)) fail(SSSI_EINVAL, "eps must be finite, >= 0");
        auto h = std::make_unique<sssi_rtm>();
        h->impl.reset(new Rtm2D(V, VS, nz, nx, nDiffOrder, nBoundary, dz, dx, dt, nt, img_stride,
                                ck_format, eps));
        *out = h.release();
    });
}

int sssi_rtm2d_add_shot(sssi_rtm *r, int64_t x_begin, int64_t x_count, int src_ix,
                        const double *wavelet, int *shot_id) {
    return rtm_guard([&] {
        const int id = R(r).add_shot(x_begin, x_count, src_ix, wavelet);
        if (shot_id) *shot_id = id;
    });
}

int sssi_rtm2d_precompute_data(sssi_rtm *r, int batched) {
    return rtm_guard([&] { R(r).precompute_true(batched != 0); });
}

int sssi_rtm2d_migrate(sssi_rtm *r, int batched, double *ms, int64_t *launches) {
    return rtm_guard([&] {
        Rtm2D &rt = R(r);
        rt.precompute_true(batched != 0);
        hipEvent_t e0, e1;
        SSSI_HIP(hipEventCreate(&e0));
        SSSI_HIP(hipEventCreate(&e1));
        const int64_t l0 = rt.launches;
        SSSI_HIP(hipEventRecord(e0, rt.stream));
        rt.migrate(batched != 0);
        SSSI_HIP(hipEventRecord(e1, rt.stream));
        SSSI_HIP(hipEventSynchronize(e1));
        float f = 0.f;
        SSSI_HIP(hipEventElapsedTime(&f, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms) *ms = f;
        if (launches) *launches = rt.launches - l0;
    });
}

int sssi_rtm2d_get_stack(sssi_rtm *r, double *stack) {
    return rtm_guard([&] {
        if (!stack) fail(SSSI_EINVAL, "stack must not be NULL");
        Rtm2D &rt = R(r);
        SSSI_HIP(hipMemcpyAsync(stack, rt.stack.p, rt.stack.bytes(), hipMemcpyDeviceToHost,
                                rt.stream));
        SSSI_HIP(hipStreamSynchronize(rt.stream));
    });
}

int sssi_rtm2d_get_shot(sssi_rtm *r, int shot, double *image, double *num, double *den,
                        double *data_true) {
    return rtm_guard([&] {
        Rtm2D &rt = R(r);
        if (shot < 0 || shot >= (int)rt.shots.size()) fail(SSSI_EINVAL, "no shot %d", shot);
        RtmShot &s = *rt.shots[shot];
        const int64_t ze = rt.nz + rt.nB, xe = s.x_count + 2 * rt.nB;
        const size_t n = (size_t)ze * xe;
        std::vector<float> h(n);
        auto get = [&](const DevBuf<float> &b, double *out) {
            if (!out) return;
            SSSI_HIP(hipMemcpyAsync(h.data(), b.p, n * sizeof(float), hipMemcpyDeviceToHost,
                                    rt.stream));
            SSSI_HIP(hipStreamSynchronize(rt.stream));
            for (int64_t j = 0; j < xe; ++j)
                for (int64_t i = 0; i < ze; ++i) out[j * ze + i] = h[(size_t)i * xe + j];
        };
        get(s.img, image);
        get(s.num, num);
        get(s.den, den);
        if (data_true) {
            std::vector<float> d((size_t)rt.nt * xe);
            SSSI_HIP(hipMemcpyAsync(d.data(), s.dtrue.p, d.size() * sizeof(float),
                                    hipMemcpyDeviceToHost, rt.stream));
            SSSI_HIP(hipStreamSynchronize(rt.stream));
            for (size_t q = 0; q < d.size(); ++q) data_true[q] = d[q];
        }
    });
}

int sssi_rtm2d_set_cache(sssi_rtm *r, const char *dir) {
    return rtm_guard([&] {
        Rtm2D &rt = R(r);
        if (!dir || !*dir) {
            rt.cache_dir.clear();
            return;
        }
        struct stat st;
        if (stat(dir, &st) != 0 && mkdir(dir, 0755) != 0)
            fail(SSSI_EINVAL, "cannot create cache directory %s", dir);
        if (stat(dir, &st) != 0 || !S_ISDIR(st.st_mode))
            fail(SSSI_EINVAL, "%s is not a directory", dir);
        rt.cache_dir = dir;
    });
}

int sssi_rtm2d_cache_stats(sssi_rtm *r, int *loaded_data, int *loaded_images, int *saved) {
    return rtm_guard([&] {
        Rtm2D &rt = R(r);
        if (loaded_data) *loaded_data = rt.loaded_true;
        if (loaded_images) *loaded_images = rt.loaded_img;
        if (saved) *saved = rt.saved_files;
    });
}

int sssi_rtm2d_device_bytes(sssi_rtm *r, double *bytes) {
    return rtm_guard([&] {
        if (!bytes) fail(SSSI_EINVAL, "bytes must not be NULL");
        Rtm2D &rt = R(r);
        double b = (double)rt.stack.bytes();
        for (auto &s : rt.shots) {
            for (const Cpml2dGrid *g : {s->gV.get(), s->gS.get()})
                b += 3.0 * g->p[0].bytes() + 2.0 * (g->zphi[0].bytes() + g->xphi[0].bytes()) * 2;
            b += s->dtrue.bytes() + s->dsmooth.bytes() + s->ck32.bytes() + s->ck16.bytes() +
                 s->stage.bytes() + 3.0 * s->num.bytes();
        }
        *bytes = b;
    });
}

void sssi_rtm2d_destroy(sssi_rtm *r) { delete r; }

}  // extern "C"
