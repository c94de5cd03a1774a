// cpml2d.h -- 2-D CPML propagation jobs: device descriptor, per-model device
// state and launch helpers shared by the single-model handles (prop2d.hip) and
// the batched multi-shot RTM driver (rtm.hip).
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "health.h"
#include "prop.h"

namespace sssi {

// Half-precision checkpoint formats.
enum CkFmt : int { kCkNone = 0, kCkF32 = 1, kCkF16 = 2, kCkBF16 = 3 };
constexpr int kCkSlots = 64;  // per-checkpoint |max| slots (spread the atomics)

// Everything a time step of one 2-D CPML propagation needs, resident in
// device memory (batched launches) or passed by value (single launches).
// Per-step quantities (time index, ping-pong parity) are derived in-kernel
// from the launch's step counter, so the descriptor stays fixed for a phase.
struct Job2D {
    int nz, nx, pitch, rows, ntx, nwz, nB;
    int start_step, nt, reverse;  // t = start..: forward t = rel, reverse t = nt-1-rel
    // interior march regions: rows [0, mz_rows) x column tiles [mta, mta+mntx),
    // blocks of one column tile x 4 wavefronts of (march groups * rz) rows
    int mz_rows, mta, mntx;
    float cz[8], cx[8];           // c_k/dz, c_k/dx (two-stage form)
    float w0, w0lo, wz[16], wx[16];  // composite D-D+ weights / d^2 (w0 + w0lo: centre)
    float w0z, w0zlo, w0x, w0xlo;    // centre weights of the z / x stencils alone
    float *p[2];                  // wavefield levels (padded); parity picks current
    const float *vdt2;            // v^2 dt^2 (padded)
    float *zphi[2], *zpsi;        // z strip [(i-(nz-nB))*nx + j]
    const float *zbm1;
    float *xphi[2], *xpsi;        // x strips [i*2nB + xs]
    const float *xbm1;
    // forward sources: CSR over wave tiles (64 x rz), amplitudes [t][k]
    const int *src_ptr, *src_cell;
    const float *src_amp;
    int nsrc, src_nt;
    // reverse surface injection: s(0,j) = surf[t*nx+j] - surf2[t*nx+j] (surf2 may be null)
    const float *surf, *surf2;
    int surf_nt;
    float *rec;                   // receivers [t*nx + j] or null
    float *snap;                  // FP32 snapshots at t % snap_stride == 0, slot t/stride - slot0
    int snap_stride, snap_slot0;
    // checkpoints of the forward field at t % ck_stride == 0 (slot c = t/ck_stride)
    int ck_stride, ck_fmt;
    float *ck32;                  // kCkF32: [c][nz*nx]
    float *ck_stage;              // kCkF16: FP32 staging [c&1][nz*nx]
    uint16_t *ck16;               // [c][nz*nx], value = half(p / scale[c])
    unsigned *ck_max;             // [c][kCkSlots] bits of max |p| (per wave, atomicMax)
    float *ck_scale;              // [c]
    // imaging against the checkpoints (reverse): num += f*p, den += f*f
    const float *im_ck32;
    const uint16_t *im_ck16;
    const float *im_scale;
    int im_stride, im_fmt;
    float *im_num, *im_den;
    // non-finite watchdog word (kHealthy, or the first launch step whose top
    // row held a NaN/Inf); null: no check
    int *health;
};

// Stencil weights of a 2-D CPML job (Job2D fields cz .. w0xlo): staggered
// coefficients c_k / d for the two-stage form and the composite D-D+ weights
// / d^2, each centre split as w + wlo so that the FP32 weights sum to zero.
void stencil_2d(int N, double dz, double dx, Job2D &J);

// Per-model device state of a 2-D CPML propagation (velocity-dependent
// coefficients and damping, wavefield levels, memory variables).
class Cpml2dGrid {
   public:
    int64_t nz = 0, nx = 0, pitch = 0, rows = 0;
    int N = 1, nB = 0, rz = 4;
    // interior march: ng groups of rz rows per wave (0: tile path only), over
    // rows [0, mz_rows) and column tiles [mta, mta + mntx)
    int ng = 0, mz_rows = 0, mta = 0, mntx = 0;
    double dz = 0, dx = 0, dt = 0;
    DevBuf<float> p[2], vdt2;
    DevBuf<float> zphi[2], zpsi, zbm1, xphi[2], xpsi, xbm1;
    DevBuf<int> health;  // non-finite watchdog word of the jobs on this grid
    Job2D coef{};  // stencil weights (stencil_2d); job() copies them

    // rz 0 / ng -1: the SSSI_RZ / SSSI_NG environment variables or the defaults
    Cpml2dGrid(const double *v, int64_t nz, int64_t nx, int N, int nB, double dz, double dx,
               double dt, int rz = 0, int ng = -1);
    // tile height (4 or 8) and interior march groups (0, 2, 4; 3 and 8 with
    // rz 4): the
    // block geometry only; arrays and results at one rz do not depend on ng
    void set_layout(int rz, int ng);
    int ntx() const { return (int)((nx + kTX - 1) / kTX); }
    int nwz() const { return (int)((nz + rz - 1) / rz); }
    int tz() const { return 4 * rz; }
    int ntz() const { return (int)((nz + tz() - 1) / tz()); }
    int nblocks() const { return ntx() * ntz(); }
    bool block_is_strip(int tile) const;
    int strip_kind(int tile) const;  // bit 0: x strip in reach, bit 1: z strip
    // tile block covered by the interior march (not launched as a tile block)
    bool block_in_march(int tile) const;
    int n_march_blocks() const { return mntx > 0 ? (mz_rows / (4 * ng * rz)) * mntx : 0; }
    // zero the wavefield and memory variables and (arm) re-arm the watchdog
    void zero_state(hipStream_t s, bool arm = true);
    // scan level `which` for NaN/Inf (reported into health as `tag`)
    void scan_nonfinite(int which, int tag, hipStream_t s) const;
    // descriptor with geometry, coefficients and state pointers filled in
    Job2D job() const;
    // interior (nz x nx, x fastest) of level `which` -> column-major doubles
    void download(int which, double *out, hipStream_t s) const;
    double alg_bytes_per_step() const;
};

// Point sources sorted into the per-wave-tile CSR the kernel scans.
struct DevSources {
    DevBuf<int> ptr, cell;
    DevBuf<float> amp;
    int n = 0, nt = 0;
    void set(const Cpml2dGrid &g, int n, const int *iz, const int *ix, const double *wavelet,
             int64_t nt);
};

// Interior march groups (rz rows each) per wave for a tile height rz, from
// the environment variable SSSI_NG (default 3 at rz 4, 0 -- tile path only --
// at rz 8).
int march_groups(int rz);

// Block launch order for a set of jobs: (job, block) pairs; block >= 0 is a
// 4-tile block of the tile path, block < 0 the interior march block -1-block.
// Strip tiles of every job come first (they run longest), then the interior
// march blocks, then the remaining interior tiles; each of the three lists is
// split into 8 contiguous bands (block b is dispatched to XCD b % 8), so
// neighbouring tiles share an L2.
void build_block_order(const std::vector<const Cpml2dGrid *> &grids, std::vector<int2> &order);

// Launch one step of a single job (descriptor by value).  With `base`, the
// kernel's step is step + *base (a launch inside a captured time loop).
void launch_cpml2d_single(const Cpml2dGrid &g, const Job2D &job, const int2 *order, int nblk,
                          int step, hipStream_t s, const int *base = nullptr);
void launch_cpml2d_single(int N, int rz, int ng, const Job2D &job, const int2 *order, int nblk,
                          int step, hipStream_t s, const int *base = nullptr);
// Launch one step of many jobs sharing nDiffOrder, rz and march layout
// (descriptors in device memory).
void launch_cpml2d_batch(int N, int rz, int ng, const Job2D *jobs, const int2 *order, int nblk,
                         int step, hipStream_t s, const int *base = nullptr);
// hipGraph-captured time loop: kGraphSteps step launches that read their step
// base from device memory, then a one-thread kernel that advances the base, so
// one captured graph replays for any run of the same launch configuration.
// Replays launch the same kernels with the same step indices as direct
// launches.  Capture is on with SSSI_GRAPH=1 (read at every run); direct
// back-to-back launches are as fast on this kernel, so it is off by default.
constexpr int kGraphSteps = 64;
bool graphs_enabled();

class StepGraph {
   public:
    using Launch = std::function<void(int step, const int *base, hipStream_t s)>;
    StepGraph() = default;
    StepGraph(const StepGraph &) = delete;
    StepGraph &operator=(const StepGraph &) = delete;
    ~StepGraph();
    // Steps step0 .. step0+count-1.  `key` (compared bytewise) identifies the
    // launch configuration; a different key re-captures.  clear() (and the
    // destructor) wait for the last replays' stream, which must still exist.
    void run(const void *key, size_t key_bytes, int step0, int64_t count, hipStream_t s,
             const Launch &launch);
    void clear();

   private:
    std::vector<char> key_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipStream_t stream_ = nullptr;  // stream of the last replays
    DevBuf<int> base_;
};

// Pack the staged FP32 checkpoint `c` of a job into FP16 (after its last step;
// nothing to do for FP32 and BF16 checkpoints, which are written directly).
void launch_ck_pack(const Job2D &job, int c, hipStream_t s);

}  // namespace sssi
