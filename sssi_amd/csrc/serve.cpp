// serve.cpp -- request loop for forward-modelling and RTM-shot jobs with
// continuous batching on one GPU.
//
// Jobs (2-D CPML forward, 2-D CPML RTM shot, 3-D CPML forward) arrive on a
// queue at any time.  Helper threads build an admitted job's device state
// (uploads and conversions on their own streams, stream-ordered allocation, so
// no call waits for the whole device) while a worker thread keeps the resident
// jobs moving in ticks.  A tick issues, on each of four equal-priority streams,
// a chunk of time steps: the 2-D jobs of one layout (nDiffOrder, tile height,
// march groups) share one batched launch per step, every job at its own time
// index, on that layout's stream; each 3-D job steps on one of two 3-D
// streams.  The streams run concurrently on the device, so a large 3-D job
// does not set the step rate of the 2-D jobs, and the worker keeps at most two
// ticks in flight, so a job that becomes ready joins within two ticks.  A job
// is admitted when a slot is free and its device footprint fits the HBM
// budget; its wavefields, memory variables and checkpoints stay resident until
// it finishes, and a helper copies its result out and releases its memory.  A
// job's arithmetic does not depend on its co-residents (per-job state, no
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
    std::vector<float> dtrue;  // RTM: cached dataTrue [t][j], or empty (computed)
    // borrowed inputs (sssi_server_submit_rtm2d_borrowed): read when the job is
    // set up instead of copied at submission; null once it is
    const double *bv = nullptr, *bvs = nullptr, *bdtrue = nullptr;
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
    // schedule (worker)
    bool have_true = false;  // RTM: dataTrue given at submission
    int lane = 0;           // work stream
    int phase = 0;          // RTM: 0 dataTrue, 1 fwd(VS)+checkpoints, 2 rvs(V)+imaging
    int64_t steps = 0;      // steps issued in the current phase
    bool started = false;
    Job2D job{};
    // results
    const float *res_dev = nullptr;  // device result the helper copies out
    size_t res_n = 0;
    std::vector<float> out;
    int health_h[2] = {kHealthy, kHealthy};  // watchdog words, copied with the result
    hipEvent_t done_ev = nullptr;
    std::string error;
    int code = SSSI_EDEVICE;  // status of a failed job
};

class Server {
   public:
    static constexpr int kLanes = 4;    // work streams: 0-1 2-D layouts, 2-3 3-D jobs
    static constexpr int kHelpers = 8;  // threads building and releasing jobs
    static constexpr int kChunk2 = 16;  // time steps per tick of a 2-D layout
    static constexpr int kChunk3 = 4;   // time steps per tick of a 3-D job
    static constexpr int kAhead = 2;    // ticks in flight

    double budget = 0;
    int max_active = 1;
    int device = 0;
    hipStream_t lanes[kLanes] = {};
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_helper;
    std::deque<ServeJob *> queue;       // submitted, waiting for a slot and memory
    std::deque<ServeJob *> ready;       // device state built, waiting for the worker
    std::deque<ServeJob *> finished;    // GPU work complete: copy the result, release
    std::map<int64_t, std::unique_ptr<ServeJob>> jobs;
    std::vector<ServeJob *> active;     // owned by the worker
    int reserved_slots = 0;             // jobs admitted (being built, ready or active)
    std::vector<std::thread> helpers;
    std::thread worker;
    std::atomic<bool> stop{false};
    int64_t next_id = 1;
    int64_t tick = 0;
    size_t used = 0;
    double cell_updates = 0, busy_s = 0;
    int64_t launches = 0;
    std::string fatal;

    // 2-D jobs of one layout share a batched launch per step on the layout's
    // lane; `step` counts the layout's launches (a member's time index is
    // step - its phase's start step)
    struct Group {
        int N, rz, ng, lane;
        int step = 0;
        bool dirty = true;
        std::vector<ServeJob *> members;
        DevBuf<Job2D> desc;
        DevBuf<int2> order;
        int nblk = 0;
    };
    std::map<std::tuple<int, int, int>, std::unique_ptr<Group>> groups;
    std::deque<std::vector<hipEvent_t>> inflight;  // per tick: one event per used lane
    std::vector<hipEvent_t> spare_events;

    Server(double budget_bytes, int max_active_) : budget(budget_bytes), max_active(max_active_) {
        SSSI_HIP(hipGetDevice(&device));
        for (auto &l : lanes) l = stream_acquire();
        // keep freed job memory in the device pool for the next jobs instead of
        // returning it at every synchronisation
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t thr = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
        worker = std::thread([this] { loop(); });
        for (int h = 0; h < kHelpers; ++h) helpers.emplace_back([this] { helper_loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        cv_helper.notify_all();
        if (worker.joinable()) worker.join();
        for (auto &h : helpers)
            if (h.joinable()) h.join();
        for (auto &l : lanes)
            if (l) (void)hipStreamSynchronize(l);
        // the helpers' streams are gone: free in order on a lane instead
        alloc_stream() = lanes[0];
        for (auto &kv : jobs) release(*kv.second);
        groups.clear();
        active.clear();
        jobs.clear();
        alloc_stream() = nullptr;
        for (auto &t : inflight)
            for (hipEvent_t e : t) (void)hipEventDestroy(e);
        for (hipEvent_t e : spare_events) (void)hipEventDestroy(e);
        (void)hipDeviceSynchronize();  // stream-ordered frees above
        for (auto &l : lanes) stream_release(l);
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
                              (j.ck_fmt == kCkF16 ? 2 * cells * 4 : 0);
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

    // ---- helpers: build the device state of admitted jobs; copy the results
    // of completed ones out and release their memory.  Allocations and frees
    // go through the stream-ordered pool on the helper's own stream.
    void helper_loop() {
        hipStream_t hs = nullptr;
        try {
            SSSI_HIP(hipSetDevice(device));
            hs = stream_acquire();
            alloc_stream() = hs;
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
                    deliver(*drop, hs);
                    release(*drop);
                    SSSI_HIP(hipStreamSynchronize(hs));
                    std::lock_guard<std::mutex> lk(mu);
                    drop->t_done = now_s();  // the result is on the host
                    used -= drop->bytes;
                    reserved_slots--;
                    drop->state = drop->error.empty() ? kDone : kFailed;
                    cv_done.notify_all();
                    cv_helper.notify_all();
                    continue;
                }
                try {
                    setup(*build, hs);
                    SSSI_HIP(hipStreamSynchronize(hs));
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
                    cv_helper.notify_all();
                }
            }
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lk(mu);
            fatal = e.what();
            cv_done.notify_all();
        }
        alloc_stream() = nullptr;
        stream_release(hs);  // waits for the stream, then back to the pool
    }

    // result and watchdog words to the host (the job's GPU work is complete)
    void deliver(ServeJob &j, hipStream_t hs) {
        if (j.res_dev) {
            j.out.resize(j.res_n);
            SSSI_HIP(hipMemcpyAsync(j.out.data(), j.res_dev, j.res_n * sizeof(float),
                                    hipMemcpyDeviceToHost, hs));
        }
        const int *hw[2] = {j.kind == kFwd3d ? (j.p3 ? j.p3->health_word() : nullptr)
                                             : (j.g ? j.g->health.p : nullptr),
                            j.gs ? j.gs->health.p : nullptr};
        for (int q = 0; q < 2; ++q)
            if (hw[q])
                SSSI_HIP(hipMemcpyAsync(&j.health_h[q], hw[q], sizeof(int), hipMemcpyDeviceToHost,
                                        hs));
        SSSI_HIP(hipStreamSynchronize(hs));
        if (j.health_h[0] != kHealthy || j.health_h[1] != kHealthy) {
            char msg[256];
            snprintf(msg, sizeof(msg),
                     "non-finite wavefield (NaN/Inf) at step %d: the time step is likely beyond "
                     "the scheme's stability limit",
                     std::min(j.health_h[0], j.health_h[1]) + 1);
            j.error = msg;
            j.code = SSSI_EUNSTABLE;
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
        j.res_dev = nullptr;
        if (j.done_ev) {
            (void)hipEventDestroy(j.done_ev);
            j.done_ev = nullptr;
        }
    }

    void setup(ServeJob &j, hipStream_t hs) {
        if (j.kind == kFwd3d) {
            j.p3.reset(make_prop3d(j.v.data(), j.nz, j.nx, j.ny, j.N, j.nB, j.dz, j.dx, j.dy, j.dt));
            j.p3->set_points(1, &j.src_iz, &j.src_ix, &j.src_iy, j.wavelet.data(), j.nt);
            j.p3->ensure_rec(j.nt, j.nx * j.ny);
            j.p3->sync();
            j.lane = 2 + (int)(j.id & 1);
            return;
        }
        j.g.reset(new Cpml2dGrid(j.bv ? j.bv : j.v.data(), j.nz, j.nx, j.N, j.nB, j.dz, j.dx,
                                 j.dt));
        j.src.set(*j.g, 1, &j.src_iz, &j.src_ix, j.wavelet.data(), j.nt);
        j.rec.alloc((size_t)j.nt * j.nx);
        if (j.kind == kRtm2d) {
            j.gs.reset(new Cpml2dGrid(j.bvs ? j.bvs : j.vs.data(), j.nz, j.nx, j.N, j.nB, j.dz,
                                      j.dx, j.dt));
            j.src_s.set(*j.gs, 1, &j.src_iz, &j.src_ix, j.wavelet.data(), j.nt);
            j.rec2.alloc((size_t)j.nt * j.nx);
            const size_t cells = (size_t)j.nz * j.nx;
            const size_t nck = (j.nt + j.img_stride - 1) / j.img_stride;
            if (j.ck_fmt == kCkF32) {
                j.ck32.alloc(nck * cells);
            } else {
                if (j.ck_fmt == kCkF16) j.stage.alloc(2 * cells);
                j.ck16.alloc(nck * cells);
            }
            j.ckmax.alloc(nck * kCkSlots);
            j.ckscale.alloc(nck);
            j.num.alloc(cells);
            j.den.alloc(cells);
            j.img.alloc(cells);
            if (j.bdtrue) {
                // data(nx, nt) column-major is [t][j]: rounded once to FP32
                j.dtrue.resize((size_t)j.nt * j.nx);
                for (size_t q = 0; q < j.dtrue.size(); ++q) j.dtrue[q] = (float)j.bdtrue[q];
            }
            if (!j.dtrue.empty()) {
                SSSI_HIP(hipMemcpyAsync(j.rec.p, j.dtrue.data(), j.rec.bytes(),
                                        hipMemcpyHostToDevice, hs));
                SSSI_HIP(hipStreamSynchronize(hs));
                std::vector<float>().swap(j.dtrue);
            }
        }
        j.bv = j.bvs = j.bdtrue = nullptr;  // the caller's arrays are no longer read
        j.lane = group_lane(j.g->N);
        // an RTM shot with cached dataTrue starts at the fwd(VS) phase
        j.phase = j.kind == kRtm2d && j.have_true ? 1 : 0;
    }

    static int group_lane(int N) { return N & 1; }

    // ---- worker
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
                    }
                }
                const double t0 = now_s();
                for (ServeJob *j : active)
                    if (!j->started) start_phase(*j);
                run_tick();
                busy_s += now_s() - t0;
                if (stop) continue;
            }
            for (auto &l : lanes) SSSI_HIP(hipStreamSynchronize(l));
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

    Group &group_of(ServeJob &j) {
        const auto key = std::make_tuple(j.g->N, j.g->rz, j.g->ng);
        auto it = groups.find(key);
        if (it == groups.end()) {
            auto g = std::make_unique<Group>();
            g->N = j.g->N; g->rz = j.g->rz; g->ng = j.g->ng; g->lane = group_lane(j.g->N);
            it = groups.emplace(key, std::move(g)).first;
        }
        return *it->second;
    }

    // descriptor of the job's current phase, starting at its group's next step
    void start_phase(ServeJob &j) {
        j.started = true;
        j.steps = 0;
        hipStream_t s = lanes[j.lane];
        if (j.kind == kFwd3d) {
            j.p3->use_stream(s);  // state zeroed at construction
            return;
        }
        Job2D J;
        if (j.kind == kFwd2d || (j.kind == kRtm2d && j.phase == 0)) {
            j.g->zero_state(s);
            J = j.g->job();
            J.reverse = 0;
            J.src_ptr = j.src.ptr.p; J.src_cell = j.src.cell.p; J.src_amp = j.src.amp.p;
            J.nsrc = j.src.n; J.src_nt = j.src.nt;
            J.rec = j.rec.p;
        } else if (j.phase == 1) {
            j.gs->zero_state(s);
            SSSI_HIP(hipMemsetAsync(j.ckmax.p, 0, j.ckmax.bytes(), s));
            J = j.gs->job();
            J.reverse = 0;
            J.src_ptr = j.src_s.ptr.p; J.src_cell = j.src_s.cell.p; J.src_amp = j.src_s.amp.p;
            J.nsrc = j.src_s.n; J.src_nt = j.src_s.nt;
            J.rec = j.rec2.p;
            J.ck_stride = j.img_stride; J.ck_fmt = j.ck_fmt;
            J.ck32 = j.ck32.p; J.ck_stage = j.stage.p; J.ck16 = j.ck16.p;
            J.ck_max = j.ckmax.p; J.ck_scale = j.ckscale.p;
        } else {
            j.g->zero_state(s, false);  // keep the dataTrue phase's watchdog report
            SSSI_HIP(hipMemsetAsync(j.num.p, 0, j.num.bytes(), s));
            SSSI_HIP(hipMemsetAsync(j.den.p, 0, j.den.bytes(), s));
            J = j.g->job();
            J.reverse = 1;
            J.surf = j.rec.p; J.surf2 = j.rec2.p; J.surf_nt = (int)j.nt;
            J.im_stride = j.img_stride; J.im_fmt = j.ck_fmt;
            J.im_ck32 = j.ck32.p; J.im_ck16 = j.ck16.p; J.im_scale = j.ckscale.p;
            J.im_num = j.num.p; J.im_den = j.den.p;
        }
        Group &G = group_of(j);
        J.start_step = G.step;
        J.nt = (int)j.nt;
        j.job = J;
        G.dirty = true;
    }

    void rebuild(Group &G) {
        G.members.clear();
        for (ServeJob *j : active)
            if (j->kind != kFwd3d && j->state == kActive && j->started &&
                std::make_tuple(j->g->N, j->g->rz, j->g->ng) ==
                    std::make_tuple(G.N, G.rz, G.ng))
                G.members.push_back(j);
        G.dirty = false;
        if (G.members.empty()) return;
        std::vector<const Cpml2dGrid *> grids;
        std::vector<Job2D> desc;
        for (ServeJob *j : G.members) {
            grids.push_back((j->kind == kRtm2d && j->phase == 1) ? j->gs.get() : j->g.get());
            desc.push_back(j->job);
        }
        std::vector<int2> ord;
        build_block_order(grids, ord);
        G.nblk = (int)ord.size();
        // reallocated on the lane's stream: the old buffers are freed in stream
        // order, after the launches that still read them
        hipStream_t s = lanes[G.lane];
        alloc_stream() = s;
        G.desc.alloc(desc.size());
        G.order.alloc(ord.size());
        alloc_stream() = nullptr;
        // pageable sources: the copies are staged before the calls return
        SSSI_HIP(hipMemcpyAsync(G.desc.p, desc.data(), G.desc.bytes(), hipMemcpyHostToDevice, s));
        SSSI_HIP(hipMemcpyAsync(G.order.p, ord.data(), G.order.bytes(), hipMemcpyHostToDevice, s));
    }

    hipEvent_t take_event() {
        if (!spare_events.empty()) {
            hipEvent_t e = spare_events.back();
            spare_events.pop_back();
            return e;
        }
        hipEvent_t e;
        SSSI_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return e;
    }

    // a job's GPU work is all issued: its result is ready when done_ev is
    void end_job(ServeJob &j) {
        hipStream_t s = lanes[j.lane];
        if (j.kind == kRtm2d) {
            const size_t n = (size_t)j.nz * j.nx;
            hipLaunchKernelGGL(serve_image, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                               j.num.p, j.den.p, j.eps, n, j.img.p);
            SSSI_HIP(hipGetLastError());
            j.res_dev = j.img.p;
            j.res_n = n;
        } else if (j.kind == kFwd2d) {
            j.res_dev = j.rec.p;
            j.res_n = (size_t)j.nt * j.nx;
        } else {
            j.res_dev = j.p3->rec.p;
            j.res_n = (size_t)j.nt * j.nx * j.ny;
        }
        SSSI_HIP(hipEventCreateWithFlags(&j.done_ev, hipEventDisableTiming));
        SSSI_HIP(hipEventRecord(j.done_ev, s));
        j.state = kDraining;
    }

    void run_tick() {
        bool used_lane[kLanes] = {false, false, false, false};
        // 2-D layouts: a chunk of batched steps, cut where a member's phase ends
        for (auto &kv : groups) {
            Group &G = *kv.second;
            if (G.dirty) rebuild(G);
            if (G.members.empty()) continue;
            int64_t k = kChunk2;
            for (ServeJob *j : G.members) k = std::min<int64_t>(k, j->nt - j->steps);
            hipStream_t s = lanes[G.lane];
            for (int64_t q = 0; q < k; ++q) {
                launch_cpml2d_batch(G.N, G.rz, G.ng, G.desc.p, G.order.p, G.nblk, G.step, s);
                ++G.step;
                ++launches;
            }
            used_lane[G.lane] = true;
            for (ServeJob *j : G.members) {
                j->steps += k;
                cell_updates += (double)k * j->nz * j->nx;
                if (j->steps < j->nt) continue;
                G.dirty = true;
                if (j->kind == kRtm2d && j->phase < 2) {
                    if (j->phase == 1) launch_ck_pack(j->job, (int)((j->nt - 1) / j->img_stride), s);
                    j->phase++;
                    start_phase(*j);
                } else {
                    end_job(*j);
                }
            }
        }
        // 3-D jobs: a chunk of steps each on the 3-D lanes
        for (ServeJob *j : active) {
            if (j->kind != kFwd3d || j->state != kActive) continue;
            const int64_t k = std::min<int64_t>(kChunk3, j->nt - j->steps);
            j->p3->run(j->steps, k, false, true, nullptr, nullptr);
            j->steps += k;
            launches += k;
            cell_updates += (double)k * j->nz * j->nx * j->ny;
            used_lane[j->lane] = true;
            if (j->steps >= j->nt) {
                j->p3->scan_field();
                end_job(*j);
            }
        }
        ++tick;
        // bound the work in flight to kAhead ticks
        std::vector<hipEvent_t> ev;
        for (int l = 0; l < kLanes; ++l)
            if (used_lane[l]) {
                hipEvent_t e = take_event();
                SSSI_HIP(hipEventRecord(e, lanes[l]));
                ev.push_back(e);
            }
        inflight.push_back(std::move(ev));
        while ((int)inflight.size() > kAhead || (!inflight.empty() && !any_running())) {
            for (hipEvent_t e : inflight.front()) {
                SSSI_HIP(hipEventSynchronize(e));
                spare_events.push_back(e);
            }
            inflight.pop_front();
        }
        // jobs whose GPU work has completed go to the helpers
        std::vector<ServeJob *> keep;
        bool changed = false;
        for (ServeJob *j : active) {
            if (j->state == kDraining && hipEventQuery(j->done_ev) == hipSuccess) {
                std::lock_guard<std::mutex> lk(mu);
                finished.push_back(j);
                cv_helper.notify_all();
                changed = true;
                continue;
            }
            keep.push_back(j);
        }
        active.swap(keep);
        if (changed)
            for (auto &kv : groups) kv.second->dirty = true;
        // only draining jobs left: wait for them instead of spinning
        if (!any_running() && !active.empty()) {
            for (auto &l : lanes) SSSI_HIP(hipStreamSynchronize(l));
        }
    }

    bool any_running() const {
        for (ServeJob *j : active)
            if (j->state == kActive) return true;
        return false;
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
            if (!fatal.empty()) return true;
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
// Copy of a caller's array into a job (inputs are copied at submission), split
// over host threads for the large ones.
template <class T, class U>
void copy_in(std::vector<T> &dst, const U *src, size_t n) {
    dst.resize(n);
    T *d = dst.data();
    parallel_ranges((int64_t)n, [&](int64_t a, int64_t b) {
        for (int64_t q = a; q < b; ++q) d[q] = (T)src[q];
    }, 1 << 20);
}

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
        copy_in(j->v, v, (size_t)(nz * nx));
        j->wavelet.assign(wavelet, wavelet + nt);
        j->nz = nz; j->nx = nx; j->nt = nt; j->N = nDiffOrder; j->nB = nBoundary;
        j->dz = dz; j->dx = dx; j->dt = dt; j->src_iz = src_iz; j->src_ix = src_ix;
        const int64_t id = S(s).submit(std::move(j));
        if (job_id) *job_id = id;
    });
}

}  // extern "C"

namespace {
// RTM shot submission; with `borrow` the models and dataTrue are read when
// the job is set up (the caller keeps them until the job is done).
int submit_rtm(sssi_server *s, const double *V, const double *VS, int64_t nz, int64_t nx,
               int64_t nt, int nDiffOrder, int nBoundary, double dz, double dx, double dt,
               int src_ix, const double *wavelet, const double *data_true, int img_stride,
               int ck_format, double eps, int64_t *job_id, bool borrow) {
    return serve_guard([&] {
        if (!V || !VS || !wavelet) fail(SSSI_EINVAL, "V, VS and wavelet must not be NULL");
        validate_common(2, nz, nx, 1, nt, nDiffOrder, nBoundary, dz, dx, 1.0, dt, true);
        validate_velocity(V, nz * nx, "V");
        validate_velocity(VS, nz * nx, "VS");
        check_src(nx, src_ix, "x");
        if (img_stride < 1) fail(SSSI_EINVAL, "img_stride must be >= 1");
        if (ck_format < SSSI_CK_F32 || ck_format > SSSI_CK_BF16)
            fail(SSSI_EINVAL, "unknown checkpoint format %d", ck_format);
        if (!(eps >= 0) || !std::isfinite(eps)) fail(SSSI_EINVAL, "eps must be finite, >= 0");
        auto j = std::make_unique<ServeJob>();
        j->kind = kRtm2d;
        if (borrow) {
            j->bv = V;
            j->bvs = VS;
            j->bdtrue = data_true;
        } else {
            copy_in(j->v, V, (size_t)(nz * nx));
            copy_in(j->vs, VS, (size_t)(nz * nx));
            // data(nx, nt) column-major is [t][j]: rounded once to the device's FP32
            if (data_true) copy_in(j->dtrue, data_true, (size_t)(nt * nx));
        }
        j->have_true = data_true != nullptr;
        j->wavelet.assign(wavelet, wavelet + nt);
        j->nz = nz; j->nx = nx; j->nt = nt; j->N = nDiffOrder; j->nB = nBoundary;
        j->dz = dz; j->dx = dx; j->dt = dt; j->src_iz = 0; j->src_ix = src_ix;
        j->img_stride = img_stride; j->ck_fmt = ck_format; j->eps = eps;
        const int64_t id = S(s).submit(std::move(j));
        if (job_id) *job_id = id;
    });
}
}  // namespace

extern "C" {

int sssi_server_submit_rtm2d_data(sssi_server *s, const double *V, const double *VS, int64_t nz,
                                  int64_t nx, int64_t nt, int nDiffOrder, int nBoundary,
                                  double dz, double dx, double dt, int src_ix,
                                  const double *wavelet, const double *data_true, int img_stride,
                                  int ck_format, double eps, int64_t *job_id) {
    return submit_rtm(s, V, VS, nz, nx, nt, nDiffOrder, nBoundary, dz, dx, dt, src_ix, wavelet,
                      data_true, img_stride, ck_format, eps, job_id, false);
}

int sssi_server_submit_rtm2d_borrowed(sssi_server *s, const double *V, const double *VS,
                                      int64_t nz, int64_t nx, int64_t nt, int nDiffOrder,
                                      int nBoundary, double dz, double dx, double dt, int src_ix,
                                      const double *wavelet, const double *data_true,
                                      int img_stride, int ck_format, double eps,
                                      int64_t *job_id) {
    return submit_rtm(s, V, VS, nz, nx, nt, nDiffOrder, nBoundary, dz, dx, dt, src_ix, wavelet,
                      data_true, img_stride, ck_format, eps, job_id, true);
}

int sssi_server_submit_rtm2d(sssi_server *s, const double *V, const double *VS, int64_t nz,
                             int64_t nx, int64_t nt, int nDiffOrder, int nBoundary, double dz,
                             double dx, double dt, int src_ix, const double *wavelet,
                             int img_stride, int ck_format, double eps, int64_t *job_id) {
    return sssi_server_submit_rtm2d_data(s, V, VS, nz, nx, nt, nDiffOrder, nBoundary, dz, dx, dt,
                                         src_ix, wavelet, nullptr, img_stride, ck_format, eps,
                                         job_id);
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
        copy_in(j->v, v, (size_t)(nz * nx * ny));
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
