// serve.cpp -- request loop for forward-modelling and RTM-shot jobs with
// continuous batching on one GPU.
//
// Jobs (2-D CPML forward, 2-D CPML RTM shot, 3-D CPML forward) arrive on a
// queue at any time.  A worker thread advances all resident jobs one time
// step per tick: the 2-D jobs of each nDiffOrder share one launch (the
// batched step kernel, every job at its own time index), 3-D jobs get one
// launch each.  A job is admitted when a slot is free and its device
// footprint fits the HBM budget; it joins the next tick, its wavefields,
// memory variables and checkpoints stay resident until it finishes, and its
// memory is released as soon as its results are copied out.  A job's
// arithmetic does not depend on its co-residents (per-job state, no
// cross-job reductions), so a result is identical alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <thread>

#include "cpml2d.h"

namespace sssi {

__global__ void serve_image(const float *num, const float *den, double eps, size_t n, float *img) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) img[i] = (float)((double)num[i] / (eps + (double)den[i]));
}

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum JobKind { kFwd2d = 0, kRtm2d = 1, kFwd3d = 2 };
enum JobState { kQueued = 0, kActive = 1, kDraining = 2, kDone = 3, kFailed = 4 };

struct ServeJob {
    int64_t id = 0;
    int kind = kFwd2d;
    std::atomic<int> state{kQueued};
    double t_arrive = 0, t_admit = 0, t_done = 0;
    // inputs
    std::vector<double> v, vs, wavelet;
    int64_t nz = 0, nx = 0, ny = 1, nt = 0;
    int N = 1, nB = 0, src_iz = 0, src_ix = 0, src_iy = 0, img_stride = 1, ck_fmt = kCkF32;
    double dz = 0, dx = 0, dy = 0, dt = 0, eps = 1e-3;
    size_t bytes = 0;
    // device state (2-D)
    std::unique_ptr<Cpml2dGrid> g, gs;
    DevSources src, src_s;
    DevBuf<float> rec, rec2, ck32, stage, num, den, img;
    DevBuf<uint16_t> ck16;
    DevBuf<unsigned> ckmax;
    DevBuf<float> ckscale;
    // device state (3-D)
    std::unique_ptr<Prop> p3;
    // schedule
    int phase = 0;          // RTM: 0 dataTrue, 1 fwd(VS)+checkpoints, 2 rvs(V)+imaging
    int64_t phase_start = 0;
    Job2D job{};
    // results
    std::vector<float> out;
    int health_h[2] = {kHealthy, kHealthy};  // watchdog words, copied with the result
    hipEvent_t done_ev = nullptr;
    std::string error;
    int code = SSSI_EDEVICE;  // status of a failed job
};

class Server {
   public:
    double budget = 0;
    int max_active = 1;
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeJob *> queue;       // submitted, waiting for a slot and memory
    std::deque<ServeJob *> ready;       // device state built, waiting for the next tick
    std::deque<ServeJob *> finished;    // results copied out, device state to release
    std::map<int64_t, std::unique_ptr<ServeJob>> jobs;
    std::vector<ServeJob *> active;
    int reserved_slots = 0;             // jobs admitted (being built, ready or active)
    std::condition_variable cv_helper;
    std::thread worker, helper;
    std::atomic<bool> stop{false};
    int64_t next_id = 1;
    int64_t tick = 0;
    size_t used = 0;
    double cell_updates = 0, busy_s = 0;
    int64_t launches = 0;
    bool groups_dirty = true;
    std::string fatal;

    struct Group {
        int N, rz, ng;
        std::vector<ServeJob *> members;
        DevBuf<Job2D> desc;
        DevBuf<int2> order;
        int nblk = 0;
    };
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<std::unique_ptr<Group>> retired;  // freed after the next sync

    Server(double budget_bytes, int max_active_) : budget(budget_bytes), max_active(max_active_) {
        SSSI_HIP(hipGetDevice(&device));
        SSSI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        worker = std::thread([this] { loop(); });
        helper = std::thread([this] { helper_loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        cv_helper.notify_all();
        if (worker.joinable()) worker.join();
        if (helper.joinable()) helper.join();
        for (auto &kv : jobs)
            if (kv.second->done_ev) (void)hipEventDestroy(kv.second->done_ev);
        active.clear();
        groups.clear();
        retired.clear();
        jobs.clear();
        if (stream) (void)hipStreamDestroy(stream);
    }

    int64_t submit(std::unique_ptr<ServeJob> j) {
        std::lock_guard<std::mutex> lk(mu);
        if (!fatal.empty()) fail(SSSI_ESTATE, "server stopped: %s", fatal.c_str());
        j->id = next_id++;
        j->t_arrive = now_s();
        j->bytes = footprint(*j);
        if ((double)j->bytes > budget)
            fail(SSSI_EBUDGET, "job needs %.3g GB, more than the %.3g GB budget", j->bytes / 1e9,
                 budget / 1e9);
        ServeJob *raw = j.get();
        const int64_t id = j->id;
        jobs[id] = std::move(j);
        queue.push_back(raw);
        cv_helper.notify_all();
        return id;
    }

    // device bytes of a job (what admission reserves)
    static size_t footprint(const ServeJob &j) {
        auto grid2 = [](int64_t nz, int64_t nx, int nB) {
            const size_t pitch = round_up(kHX + round_up(nx, kTX) + 32, 64);
            const size_t rows = kHZ + round_up(nz, kTZmax) + kHZbot;
            const size_t strips = (size_t)nB * nx + (size_t)nz * 2 * nB;
            return 3 * pitch * rows * 4 + 6 * strips * 4;
        };
        if (j.kind == kFwd2d) return grid2(j.nz, j.nx, j.nB) + (size_t)j.nt * j.nx * 4;
        if (j.kind == kRtm2d) {
            const size_t cells = (size_t)j.nz * j.nx;
            const size_t nck = (j.nt + j.img_stride - 1) / j.img_stride;
            const size_t ck = nck * cells * (j.ck_fmt == kCkF32 ? 4 : 2) +
                              (j.ck_fmt == kCkF32 ? 0 : 2 * cells * 4);
            return 2 * grid2(j.nz, j.nx, j.nB) + 2 * (size_t)j.nt * j.nx * 4 + ck +
                   3 * cells * 4 + nck * (kCkSlots + 1) * 4;
        }
        const size_t pitch = round_up(kHX + round_up(j.nx, kTX) + 32, 64);
        const size_t plane = pitch * (16 + round_up(j.ny, 16) + 16);
        const size_t np = plane * (kHZ + j.nz + kHZbot);
        const size_t strips = (size_t)j.nB * j.nx * j.ny + (size_t)j.nz * j.ny * 2 * j.nB +
                              (size_t)j.nz * 2 * j.nB * j.nx;
        return 3 * np * 4 + 3 * strips * 4 + (size_t)j.nt * j.nx * j.ny * 4;
    }

    // ---- helper thread: builds device state of admitted jobs (host-side
    // preparation and uploads) and releases the state of finished ones, so the
    // worker's launch stream never waits on them.
    void helper_loop() {
        try {
            SSSI_HIP(hipSetDevice(device));
            while (true) {
                ServeJob *build = nullptr, *drop = nullptr;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv_helper.wait(lk, [&] {
                        return stop || !finished.empty() ||
                               (!queue.empty() && reserved_slots < max_active &&
                                (double)(used + queue.front()->bytes) <= budget);
                    });
                    if (stop) break;
                    if (!finished.empty()) {
                        drop = finished.front();
                        finished.pop_front();
                    } else {
                        build = queue.front();  // FIFO: later jobs wait for memory
                        queue.pop_front();
                        used += build->bytes;
                        reserved_slots++;
                    }
                }
                if (drop) {
                    release(*drop);
                    std::lock_guard<std::mutex> lk(mu);
                    used -= drop->bytes;
                    reserved_slots--;
                    drop->state = drop->error.empty() ? kDone : kFailed;
                    cv_done.notify_all();
                    cv_helper.notify_all();
                    continue;
                }
                try {
                    setup(*build);
                    std::lock_guard<std::mutex> lk(mu);
                    build->t_admit = now_s();
                    ready.push_back(build);
                    cv_work.notify_all();
                } catch (const std::exception &e) {
                    release(*build);
                    std::lock_guard<std::mutex> lk(mu);
                    build->state = kFailed;
                    build->error = e.what();
                    used -= build->bytes;
                    reserved_slots--;
                    cv_done.notify_all();
                }
            }
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lk(mu);
            fatal = e.what();
            cv_done.notify_all();
        }
    }

    void release(ServeJob &j) {
        j.g.reset();
        j.gs.reset();
        j.p3.reset();
        for (auto *b : {&j.rec, &j.rec2, &j.ck32, &j.stage, &j.num, &j.den, &j.img, &j.ckscale})
            b->release();
        j.ck16.release();
        j.ckmax.release();
        j.src = DevSources();
        j.src_s = DevSources();
        if (j.done_ev) {
            (void)hipEventDestroy(j.done_ev);
            j.done_ev = nullptr;
        }
    }

    // ---- worker: one tick = one time step of every resident job
    void loop() {
        try {
            SSSI_HIP(hipSetDevice(device));
            while (true) {
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv_work.wait(lk, [&] { return stop || !ready.empty() || !active.empty(); });
                    if (stop && active.empty()) break;
                    while (!ready.empty()) {
                        ServeJob *j = ready.front();
                        ready.pop_front();
                        j->state = kActive;
                        active.push_back(j);
                        groups_dirty = true;
                    }
                }
                for (ServeJob *j : active)
                    if (j->state == kActive && j->phase_start < 0) start_phase(*j);
                if (active.empty()) continue;
                const double t0 = now_s();
                step_all();
                finish_phases();
                busy_s += now_s() - t0;
            }
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lk(mu);
            fatal = e.what();
            for (auto &kv : jobs)
                if (kv.second->state != kDone) {
                    kv.second->state = kFailed;
                    kv.second->error = fatal;
                }
            cv_done.notify_all();
        }
    }

    void setup(ServeJob &j) {
        if (j.kind == kFwd3d) {
            j.p3.reset(make_prop3d(j.v.data(), j.nz, j.nx, j.ny, j.N, j.nB, j.dz, j.dx, j.dy, j.dt));
            j.p3->use_stream(stream);
            j.p3->set_points(1, &j.src_iz, &j.src_ix, &j.src_iy, j.wavelet.data(), j.nt);
            j.p3->ensure_rec(j.nt, j.nx * j.ny);
            j.p3->sync();
            j.phase_start = -1;
            return;
        }
        j.g.reset(new Cpml2dGrid(j.v.data(), j.nz, j.nx, j.N, j.nB, j.dz, j.dx, j.dt));
        j.src.set(*j.g, 1, &j.src_iz, &j.src_ix, j.wavelet.data(), j.nt);
        j.rec.alloc((size_t)j.nt * j.nx);
        if (j.kind == kRtm2d) {
            j.gs.reset(new Cpml2dGrid(j.vs.data(), j.nz, j.nx, j.N, j.nB, j.dz, j.dx, j.dt));
            j.src_s.set(*j.gs, 1, &j.src_iz, &j.src_ix, j.wavelet.data(), j.nt);
            j.rec2.alloc((size_t)j.nt * j.nx);
            const size_t cells = (size_t)j.nz * j.nx;
            const size_t nck = (j.nt + j.img_stride - 1) / j.img_stride;
            if (j.ck_fmt == kCkF32) {
                j.ck32.alloc(nck * cells);
            } else {
                j.stage.alloc(2 * cells);
                j.ck16.alloc(nck * cells);
            }
            j.ckmax.alloc(nck * kCkSlots);
            j.ckscale.alloc(nck);
            j.num.alloc(cells);
            j.den.alloc(cells);
            j.img.alloc(cells);
        }
        j.phase = 0;
        j.phase_start = -1;  // uploads above are synchronous; the worker zeroes state
    }

    void start_phase(ServeJob &j) {
        j.phase_start = tick;
        if (j.kind == kFwd3d) return;
        Job2D J;
        if (j.kind == kFwd2d || (j.kind == kRtm2d && j.phase == 0)) {
            j.g->zero_state(stream);
            J = j.g->job();
            J.reverse = 0;
            J.src_ptr = j.src.ptr.p; J.src_cell = j.src.cell.p; J.src_amp = j.src.amp.p;
            J.nsrc = j.src.n; J.src_nt = j.src.nt;
            J.rec = j.rec.p;
        } else if (j.phase == 1) {
            j.gs->zero_state(stream);
            SSSI_HIP(hipMemsetAsync(j.ckmax.p, 0, j.ckmax.bytes(), stream));
            J = j.gs->job();
            J.reverse = 0;
            J.src_ptr = j.src_s.ptr.p; J.src_cell = j.src_s.cell.p; J.src_amp = j.src_s.amp.p;
            J.nsrc = j.src_s.n; J.src_nt = j.src_s.nt;
            J.rec = j.rec2.p;
            J.ck_stride = j.img_stride; J.ck_fmt = j.ck_fmt;
            J.ck32 = j.ck32.p; J.ck_stage = j.stage.p; J.ck16 = j.ck16.p;
            J.ck_max = j.ckmax.p; J.ck_scale = j.ckscale.p;
        } else {
            j.g->zero_state(stream);
            SSSI_HIP(hipMemsetAsync(j.num.p, 0, j.num.bytes(), stream));
            SSSI_HIP(hipMemsetAsync(j.den.p, 0, j.den.bytes(), stream));
            J = j.g->job();
            J.reverse = 1;
            J.surf = j.rec.p; J.surf2 = j.rec2.p; J.surf_nt = (int)j.nt;
            J.im_stride = j.img_stride; J.im_fmt = j.ck_fmt;
            J.im_ck32 = j.ck32.p; J.im_ck16 = j.ck16.p; J.im_scale = j.ckscale.p;
            J.im_num = j.num.p; J.im_den = j.den.p;
        }
        J.start_step = (int)tick;
        J.nt = (int)j.nt;
        j.job = J;
        groups_dirty = true;
    }

    void rebuild_groups() {
        for (auto &g : groups) retired.push_back(std::move(g));
        groups.clear();
        std::map<std::tuple<int, int, int>, Group *> by;
        for (ServeJob *j : active) {
            if (j->kind == kFwd3d || j->state != kActive) continue;
            auto key = std::make_tuple(j->g->N, j->g->rz, j->g->ng);
            if (!by.count(key)) {
                groups.emplace_back(new Group{std::get<0>(key), std::get<1>(key),
                                              std::get<2>(key), {}, {}, {}, 0});
                by[key] = groups.back().get();
            }
            by[key]->members.push_back(j);
        }
        for (auto &g : groups) {
            std::vector<const Cpml2dGrid *> grids;
            std::vector<Job2D> desc;
            for (ServeJob *j : g->members) {
                const Cpml2dGrid *gg = (j->kind == kRtm2d && j->phase == 1) ? j->gs.get()
                                                                            : j->g.get();
                grids.push_back(gg);
                desc.push_back(j->job);
            }
            std::vector<int2> ord;
            build_block_order(grids, ord);
            g->nblk = (int)ord.size();
            g->desc.alloc(desc.size());
            g->order.alloc(ord.size());
            // pageable source: the copy is staged before the call returns
            SSSI_HIP(hipMemcpyAsync(g->desc.p, desc.data(), g->desc.bytes(), hipMemcpyHostToDevice,
                                    stream));
            SSSI_HIP(hipMemcpyAsync(g->order.p, ord.data(), g->order.bytes(),
                                    hipMemcpyHostToDevice, stream));
        }
        groups_dirty = false;
    }

    void step_all() {
        if (groups_dirty) rebuild_groups();
        for (auto &g : groups) {
            launch_cpml2d_batch(g->N, g->rz, g->ng, g->desc.p, g->order.p, g->nblk,
                                (int)tick, stream);
            launches++;
            for (ServeJob *j : g->members)
                cell_updates += (double)j->nz * j->nx;
        }
        for (ServeJob *j : active)
            if (j->kind == kFwd3d && j->state == kActive) {
                j->p3->run(tick - j->phase_start, 1, false, true, nullptr, nullptr);
                launches++;
                cell_updates += (double)j->nz * j->nx * j->ny;
            }
        ++tick;
        // bound how far the host runs ahead of the device
        if (tick % 64 == 0) {
            SSSI_HIP(hipStreamSynchronize(stream));
            retired.clear();
        }
    }

    void finish_phases() {
        bool changed = false;
        for (ServeJob *j : active) {
            if (j->state != kActive || tick - j->phase_start < j->nt) continue;
            if (j->kind == kRtm2d && j->phase < 2) {
                if (j->phase == 1 && j->ck_fmt != kCkF32)
                    launch_ck_pack(j->job, (int)((j->nt - 1) / j->img_stride), stream);
                j->phase++;
                start_phase(*j);
                changed = true;
                continue;
            }
            // results out
            size_t n;
            const float *src;
            if (j->kind == kRtm2d) {
                n = (size_t)j->nz * j->nx;
                hipLaunchKernelGGL(serve_image, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                                   stream, j->num.p, j->den.p, j->eps, n, j->img.p);
                SSSI_HIP(hipGetLastError());
                src = j->img.p;
            } else if (j->kind == kFwd2d) {
                n = (size_t)j->nt * j->nx;
                src = j->rec.p;
            } else {
                n = (size_t)j->nt * j->nx * j->ny;
                src = j->p3->rec.p;
            }
            j->out.resize(n);
            SSSI_HIP(hipMemcpyAsync(j->out.data(), src, n * sizeof(float), hipMemcpyDeviceToHost,
                                    stream));
            // watchdog words of the job's grids (every phase reports into them)
            const int *hw[2] = {j->kind == kFwd3d ? j->p3->health_word() : j->g->health.p,
                                j->gs ? j->gs->health.p : nullptr};
            for (int q = 0; q < 2; ++q)
                if (hw[q])
                    SSSI_HIP(hipMemcpyAsync(&j->health_h[q], hw[q], sizeof(int),
                                            hipMemcpyDeviceToHost, stream));
            SSSI_HIP(hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming));
            SSSI_HIP(hipEventRecord(j->done_ev, stream));
            j->state = kDraining;
            changed = true;
        }
        if (changed) groups_dirty = true;
        // completed jobs (results on the host): the helper releases their memory
        std::vector<ServeJob *> keep;
        for (ServeJob *j : active) {
            if (j->state == kDraining && hipEventQuery(j->done_ev) == hipSuccess) {
                std::lock_guard<std::mutex> lk(mu);
                j->t_done = now_s();  // result is on the host
                if (j->health_h[0] != kHealthy || j->health_h[1] != kHealthy) {
                    const int h = std::min(j->health_h[0], j->health_h[1]);
                    char msg[256];
                    snprintf(msg, sizeof(msg),
                             "non-finite wavefield (NaN/Inf) at tick %d: the time step is likely "
                             "beyond the scheme's stability limit", h);
                    j->error = msg;
                    j->code = SSSI_EUNSTABLE;
                }
                finished.push_back(j);
                cv_helper.notify_all();
                continue;
            }
            keep.push_back(j);
        }
        if (keep.size() != active.size()) groups_dirty = true;
        active.swap(keep);
        // draining jobs only: wait for their copies instead of spinning
        bool any_running = false;
        for (ServeJob *j : active) any_running |= (j->state == kActive);
        if (!any_running && !active.empty()) {
            SSSI_HIP(hipStreamSynchronize(stream));
            retired.clear();
        }
    }

    ServeJob &get(int64_t id) {
        auto it = jobs.find(id);
        if (it == jobs.end()) fail(SSSI_EINVAL, "no job %lld", (long long)id);
        return *it->second;
    }

    bool wait(int64_t id, double timeout) {
        std::unique_lock<std::mutex> lk(mu);
        ServeJob &j = get(id);
        auto pred = [&] { return j.state == kDone || j.state == kFailed; };
        if (timeout < 0) cv_done.wait(lk, pred);
        else cv_done.wait_for(lk, std::chrono::duration<double>(timeout), pred);
        if (j.state == kFailed) fail(j.code, "job %lld failed: %s", (long long)id,
                                     j.error.c_str());
        return j.state == kDone;
    }

    void drain() {
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] {
            for (auto &kv : jobs)
                if (kv.second->state != kDone && kv.second->state != kFailed) return false;
            return true;
        });
        if (!fatal.empty()) fail(SSSI_EDEVICE, "server stopped: %s", fatal.c_str());
    }
};

}  // namespace sssi

using namespace sssi;

struct sssi_server {
    std::unique_ptr<Server> impl;
};

namespace {
template <class F>
int serve_guard(F &&f) {
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
Server &S(sssi_server *s) {
    if (!s || !s->impl) fail(SSSI_EINVAL, "NULL server handle");
    return *s->impl;
}
void check_src(int64_t n, int idx, const char *what) {
    if (idx < 0 || idx >= n) fail(SSSI_EINVAL, "source %s index %d outside [0, %lld)", what, idx,
                                  (long long)n);
}
}  // namespace

extern "C" {

int sssi_server_create(double hbm_budget_bytes, int max_active, sssi_server **out) {
    return serve_guard([&] {
        if (!out) fail(SSSI_EINVAL, "out must not be NULL");
        if (!(hbm_budget_bytes > 0)) fail(SSSI_EINVAL, "budget must be positive");
        if (max_active < 1) fail(SSSI_EINVAL, "max_active must be >= 1");
        auto h = std::make_unique<sssi_server>();
        h->impl.reset(new Server(hbm_budget_bytes, max_active));
        *out = h.release();
    });
}

int sssi_server_submit_fwd2d(sssi_server *s, const double *v, int64_t nz, int64_t nx, int64_t nt,
                             int nDiffOrder, int nBoundary, double dz, double dx, double dt,
                             int src_iz, int src_ix, const double *wavelet, int64_t *job_id) {
    return serve_guard([&] {
        if (!v || !wavelet) fail(SSSI_EINVAL, "v and wavelet must not be NULL");
        validate_common(2, nz, nx, 1, nt, nDiffOrder, nBoundary, dz, dx, 1.0, dt, true);
        validate_velocity(v, nz * nx, "v");
        check_src(nz, src_iz, "z");
        check_src(nx, src_ix, "x");
        auto j = std::make_unique<ServeJob>();
        j->kind = kFwd2d;
        j->v.assign(v, v + nz * nx);
        j->wavelet.assign(wavelet, wavelet + nt);
        j->nz = nz; j->nx = nx; j->nt = nt; j->N = nDiffOrder; j->nB = nBoundary;
        j->dz = dz; j->dx = dx; j->dt = dt; j->src_iz = src_iz; j->src_ix = src_ix;
        const int64_t id = S(s).submit(std::move(j));
        if (job_id) *job_id = id;
    });
}

int sssi_server_submit_rtm2d(sssi_server *s, const double *V, const double *VS, int64_t nz,
                             int64_t nx, int64_t nt, int nDiffOrder, int nBoundary, double dz,
                             double dx, double dt, int src_ix, const double *wavelet,
                             int img_stride, int ck_format, double eps, int64_t *job_id) {
    return serve_guard([&] {
        if (!V || !VS || !wavelet) fail(SSSI_EINVAL, "V, VS and wavelet must not be NULL");
        validate_common(2, nz, nx, 1, nt, nDiffOrder, nBoundary, dz, dx, 1.0, dt, true);
        validate_velocity(V, nz * nx, "V");
        validate_velocity(VS, nz * nx, "VS");
        check_src(nx, src_ix, "x");
        if (img_stride < 1) fail(SSSI_EINVAL, "img_stride must be >= 1");
        if (ck_format < SSSI_CK_F32 || ck_format > SSSI_CK_BF16)
            fail(SSSI_EINVAL, "unknown checkpoint format %d", ck_format);
        auto j = std::make_unique<ServeJob>();
        j->kind = kRtm2d;
        j->v.assign(V, V + nz * nx);
        j->vs.assign(VS, VS + nz * nx);
        j->wavelet.assign(wavelet, wavelet + nt);
        j->nz = nz; j->nx = nx; j->nt = nt; j->N = nDiffOrder; j->nB = nBoundary;
        j->dz = dz; j->dx = dx; j->dt = dt; j->src_iz = 0; j->src_ix = src_ix;
        j->img_stride = img_stride; j->ck_fmt = ck_format; j->eps = eps;
        const int64_t id = S(s).submit(std::move(j));
        if (job_id) *job_id = id;
    });
}

int sssi_server_submit_fwd3d(sssi_server *s, const double *v, int64_t nz, int64_t nx, int64_t ny,
                             int64_t nt, int nDiffOrder, int nBoundary, double dz, double dx,
                             double dy, double dt, int src_iz, int src_ix, int src_iy,
                             const double *wavelet, int64_t *job_id) {
    return serve_guard([&] {
        if (!v || !wavelet) fail(SSSI_EINVAL, "v and wavelet must not be NULL");
        validate_common(3, nz, nx, ny, nt, nDiffOrder, nBoundary, dz, dx, dy, dt, true);
        validate_velocity(v, nz * nx * ny, "v");
        check_src(nz, src_iz, "z");
        check_src(nx, src_ix, "x");
        check_src(ny, src_iy, "y");
        auto j = std::make_unique<ServeJob>();
        j->kind = kFwd3d;
        j->v.assign(v, v + nz * nx * ny);
        j->wavelet.assign(wavelet, wavelet + nt);
        j->nz = nz; j->nx = nx; j->ny = ny; j->nt = nt; j->N = nDiffOrder; j->nB = nBoundary;
        j->dz = dz; j->dx = dx; j->dy = dy; j->dt = dt;
        j->src_iz = src_iz; j->src_ix = src_ix; j->src_iy = src_iy;
        const int64_t id = S(s).submit(std::move(j));
        if (job_id) *job_id = id;
    });
}

int sssi_server_wait(sssi_server *s, int64_t job_id, double timeout_s, int *done) {
    return serve_guard([&] {
        const bool d = S(s).wait(job_id, timeout_s);
        if (done) *done = d ? 1 : 0;
    });
}

int sssi_server_drain(sssi_server *s) {
    return serve_guard([&] { S(s).drain(); });
}

int sssi_server_result(sssi_server *s, int64_t job_id, double *out, int64_t n) {
    return serve_guard([&] {
        Server &sv = S(s);
        std::lock_guard<std::mutex> lk(sv.mu);
        ServeJob &j = sv.get(job_id);
        if (j.state == kFailed)
            fail(j.code, "job %lld failed: %s", (long long)job_id, j.error.c_str());
        if (j.state != kDone) fail(SSSI_ESTATE, "job %lld not finished", (long long)job_id);
        if (!out || n != (int64_t)j.out.size())
            fail(SSSI_EINVAL, "result of job %lld has %zu values", (long long)job_id,
                 j.out.size());
        // data: [t][j] (2-D) or [t][k][j] (3-D) == column-major (nx[,ny],nt);
        // image: [i][j] -> column-major (nz,nx)
        if (j.kind == kRtm2d) {
            for (int64_t c = 0; c < j.nx; ++c)
                for (int64_t r = 0; r < j.nz; ++r) out[c * j.nz + r] = j.out[r * j.nx + c];
        } else {
            for (int64_t q = 0; q < n; ++q) out[q] = j.out[q];
        }
    });
}

int sssi_server_job_info(sssi_server *s, int64_t job_id, double *t_arrive, double *t_admit,
                         double *t_done, double *bytes) {
    return serve_guard([&] {
        Server &sv = S(s);
        std::lock_guard<std::mutex> lk(sv.mu);
        ServeJob &j = sv.get(job_id);
        if (t_arrive) *t_arrive = j.t_arrive;
        if (t_admit) *t_admit = j.t_admit;
        if (t_done) *t_done = j.t_done;
        if (bytes) *bytes = (double)j.bytes;
    });
}

int sssi_server_stats(sssi_server *s, double *cell_updates, double *busy_s, int64_t *launches,
                      int64_t *ticks) {
    return serve_guard([&] {
        Server &sv = S(s);
        std::lock_guard<std::mutex> lk(sv.mu);
        if (cell_updates) *cell_updates = sv.cell_updates;
        if (busy_s) *busy_s = sv.busy_s;
        if (launches) *launches = sv.launches;
        if (ticks) *ticks = sv.tick;
    });
}

void sssi_server_destroy(sssi_server *s) { delete s; }

}  // extern "C"
