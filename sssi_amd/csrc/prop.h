// prop.h -- device-resident propagation handles (2-D CPML/sponge, 3-D CPML).
#pragma once

#include "common.h"

struct sssi_prop;  // opaque ABI handle; defined as a subclass of sssi::Prop

namespace sssi {

// Padded device layout of a 2-D wavefield: interior cell (i, j) lives at
// (i + kHZ) * pitch + (j + kHX).  The halo is zero and never written, so the
// stencil reads past the interior without bounds checks (the reference's
// zero padding of width l = 2N-1, N <= 8).
constexpr int kHZ = 16;
constexpr int kHZbot = 32;  // rows below the (rounded-up) interior: the z sweep
                            // prefetches up to RZ + L - 1 rows past the last row
constexpr int kHX = 32;
constexpr int kTX = 64;     // tile width (one wavefront across x)
constexpr int kTZmax = 64;  // tallest tile the kernels use

// Bytes of one padded 2-D level in the layout above.
inline int64_t padded_bytes_2d(int64_t nz, int64_t nx) {
    const int64_t pitch = round_up(kHX + round_up(nx, kTX) + 32, 64);
    const int64_t rows = kHZ + round_up(nz, kTZmax) + kHZbot;
    return pitch * rows * 4;
}
// The sponge step drops the stores of lanes outside the model by giving them
// the out-of-range byte offset 2^31, which is out of range only for buffers
// below 2 GiB: its padded levels are capped there.
inline void validate_sponge_grid(int64_t nz, int64_t nx) {
    if (padded_bytes_2d(nz, nx) >= ((int64_t)1 << 31))
        fail(SSSI_EINVAL, "sponge grid %lldx%lld too large: a padded level needs %lld bytes "
             "(limit 2 GiB)", (long long)nz, (long long)nx, (long long)padded_bytes_2d(nz, nx));
}

struct LaunchStats {
    int64_t launches = 0;
};

// Where snapshots go during a run: device buffer of compact FP32 slices
// (x fastest), one per step t with t % stride == 0, at slot t/stride - slot0.
struct SnapSink {
    float *dev = nullptr;
    int64_t stride = 1;
    int64_t slot0 = 0;
};

class Prop {
   public:
    // the stream is destroyed here, so that a derived constructor that throws
    // after creating it does not leak it
    virtual ~Prop() {
        if (own_stream) stream_release(stream);
    }
    int dim = 2;
    int abc = SSSI_ABC_CPML;
    int64_t nz = 0, nx = 0, ny = 1;
    int N = 1, nB = 0;
    double dz = 0, dx = 0, dy = 0, dt = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false when the stream is lent by a server
    int device = 0;

    // point sources, sorted by tile; amplitudes [t][k] in sorted order
    int nsrc = 0;
    int64_t src_nt = 0;
    DevBuf<int> src_ptr, src_cell;
    DevBuf<float> src_amp;
    // surface data for reverse runs: [t][j]
    int64_t surf_nt = 0;
    DevBuf<float> surf;
    // receivers: [t][j] (2-D) or [t][k][j] (3-D)
    int64_t rec_nt = 0;
    DevBuf<float> rec;

    virtual void reset() = 0;
    virtual void set_points(int n, const int *iz, const int *ix, const int *iy,
                            const double *wavelet, int64_t nt) = 0;
    virtual void set_surface(const double *data, int64_t nt) = 0;
    // Enqueue `count` steps starting at time index t_first: forward runs use
    // t = t_first, t_first+1, ...; reverse runs t = t_first, t_first-1, ...
    // Point sources drive forward runs, surface data reverse runs.
    virtual void run(int64_t t_first, int64_t count, bool reverse, bool record,
                     const SnapSink *snap, LaunchStats *st) = 0;
    virtual int64_t cells() const = 0;
    virtual void get_data(double *data, int64_t nt) = 0;
    virtual void get_field(double *field) = 0;
    virtual void get_model(double *model) = 0;
    virtual void alg_bytes(double *bytes, double *cells) = 0;

    void sync() { SSSI_HIP(hipStreamSynchronize(stream)); }

    // Non-finite watchdog.  The kernels report a NaN/Inf on the top (receiver)
    // row into a device word every step; scan_field() also scans the whole
    // current level.  check_health() waits for the stream and fails with
    // SSSI_EUNSTABLE if either saw one since the last reset().
    virtual int *health_word() const = 0;
    virtual void scan_field() = 0;
    int64_t steps_done = 0;  // steps since the last reset()
    double courant = 0.0;    // vmax dt sqrt(sum 1/d^2), for the message
    void check_health();

    // Grow the receiver buffer to nt steps of `per_step` values, keeping what
    // was recorded so far.
    void ensure_rec(int64_t nt, int64_t per_step) {
        const size_t need = (size_t)nt * per_step;
        if (rec.n < need) {
            DevBuf<float> nr(need);
            SSSI_HIP(hipMemsetAsync(nr.p, 0, nr.bytes(), stream));
            if (rec.p)
                SSSI_HIP(hipMemcpyAsync(nr.p, rec.p, rec.bytes(), hipMemcpyDeviceToDevice, stream));
            SSSI_HIP(hipStreamSynchronize(stream));
            rec = std::move(nr);
        }
        if (rec_nt < nt) rec_nt = nt;
    }
    // Run on a stream owned by someone else (the serving loop).
    // layout in use: 2-D (tile height, march groups), 3-D (block y extent, z chunk)
    virtual void layout(int *a, int *b) const = 0;
    virtual void use_stream(hipStream_t s) {
        if (own_stream) stream_release(stream);
        stream = s;
        own_stream = false;
    }
};

Prop *make_prop2d(const double *v, int64_t nz, int64_t nx, int abc, int N, int nB, double dz,
                  double dx, double dt);
Prop *make_prop3d(const double *v, int64_t nz, int64_t nx, int64_t ny, int N, int nB,
                  double dz, double dx, double dy, double dt);

}  // namespace sssi
