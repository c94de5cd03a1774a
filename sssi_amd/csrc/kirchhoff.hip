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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "common.h"

namespace sssi {

constexpr double kT0 = 65535;  // initial traveltime (eikonal2d_mex.cpp:31)

static unsigned grid_blocks(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256));
}

// The reference's cell update (eikonal2d_mex.cpp:80-105 and its three
// repeats): a = min of the z neighbours, b = min of the x neighbours (the
// single inner neighbour at a grid edge), Godunov update, min with the old
// value.  Contraction is off at every caller so the operation order, and with
// it every bit, is the reference's.
__device__ __forceinline__ double eik_update(double me, double a, double b, double s, double dx) {
#pragma clang fp contract(off)
    double TT;
    if (fabs(a - b) < s * dx)
        TT = 0.5 * (a + b + sqrt((2 * s * s * dx * dx) - (a - b) * (a - b)));
    else
        TT = fmin(a, b) + (s * dx);
    return fmin(me, TT);
}

// One Jacobi pass for nsrc tables [k][j][i] (column-major per table): the
// plain form, one launch per pass.
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
        double a, b;
        if (i == 0) a = c[1];
        else if (i == nz - 1) a = c[nz - 2];
        else a = fmin(c[i - 1], c[i + 1]);
        if (j == 0) b = T[i + nz];
        else if (j == nx - 1) b = T[i + (nx - 2) * nz];
        else b = fmin(T[i + (j - 1) * nz], T[i + (j + 1) * nz]);
        const double t = eik_update(c[i], a, b, slow[q], dx);
        tnew[g] = q == src_cell[k] ? 0.0 : t;
    }
}

// 64-bit wave shifts by one lane (DPP wave_shr:1 / wave_shl:1 on both
// dwords): lane l receives lane l-1's (shr) or lane l+1's (shl) value; the
// lane with no source lane reads 0 (bound_ctrl).
__device__ __forceinline__ double wave_shr1(double v) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x138, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_shl1(double v) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x130, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Temporally blocked passes.  A block holds a 64 x TW tile of one table
// (rows z0-H .. z0-H+63, columns x0-H .. x0-H+TW-1) and runs up to H Jacobi
// passes on it, then writes the centre (64-2H) x (TW-2H) cells.  A pass at a
// tile edge that is not a grid edge reads a clamped neighbour and goes wrong,
// and each pass moves that error one cell inwards, so after npass <= H passes
// the centre is exactly the plain form's; a pass skips the columns that can
// no longer reach the centre (H >= npass keeps every read column in the
// tile).  Grid edges use the reference's inner-neighbour rule on in-tile
// cells; cells outside the grid are computed but never read by a cell inside
// it and never stored.
//   lane = tile row (z neighbours by DPP wave shifts of the lane's own
//   register), wave w owns tile columns w, w+4, ... (x neighbours from the
//   LDS copy of the previous level); two barriers per pass.
// The launch is 1-D over ntiles * nsrc blocks, XCD-banded: block b runs on XCD
// b % 8 and takes the b/8-th tile of that XCD's contiguous band, so
// neighbouring tiles of a table share one L2 for their halos.
template <int H, int TW>
__global__ void __launch_bounds__(256) eik_tile(const double *__restrict__ slow,
                                                const double *__restrict__ told,
                                                double *__restrict__ tnew, int nz, int nx,
                                                int ntz, int ntiles,
                                                const int64_t *__restrict__ src_cell, double dx,
                                                int npass) {
#pragma clang fp contract(off)
    constexpr int C = TW / 4;  // columns per wave
    constexpr int OZ = 64 - 2 * H, OX = TW - 2 * H;
    __shared__ double L[TW * 64];
    const int lz = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int qb = nb / 8, rem = nb % 8, xc = b % 8;
    const int t = xc * qb + min(xc, rem) + b / 8;
    const int k = t / ntiles, tile = t - k * ntiles;
    const int tx = tile / ntz, tz = tile - tx * ntz;
    const int gi = tz * OZ - H + lz;
    const int x0 = tx * OX - H;
    const int nn = nz * nx;
    const double *T = told + (int64_t)k * nn;
    double *TN = tnew + (int64_t)k * nn;
    const int src = (int)src_cell[k];
    const bool zin = gi >= 0 && gi < nz;

    double tv[C], sv[C];
    unsigned srcmask = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int lx = w + 4 * c, gj = x0 + lx;
        const bool in = zin && gj >= 0 && gj < nx;
        const int q = gj * nz + gi;
        tv[c] = in ? T[q] : kT0;
        sv[c] = in ? slow[q] : 0.0;
        if (in && q == src) srcmask |= 1u << c;
        L[lx * 64 + lz] = tv[c];
    }
    __syncthreads();
    for (int p = 0; p < npass; ++p) {
        // columns whose value can still reach the centre by the last pass
        const int r = npass - 1 - p;
        double tn[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int lx = w + 4 * c, gj = x0 + lx;
            tn[c] = tv[c];
            if (lx < H - r || lx >= H + OX + r) continue;  // wave-uniform
            const double me = tv[c];
            // edge rule as neighbour substitution: min(x, x) = x
            const double up0 = wave_shr1(me), dn0 = wave_shl1(me);
            const double up = gi == 0 ? dn0 : up0, dn = gi == nz - 1 ? up0 : dn0;
            const int cl = gj == 0 ? lx + 1 : lx - 1, cr = gj == nx - 1 ? lx - 1 : lx + 1;
            const double a = fmin(up, dn);
            const double bb = fmin(L[cl * 64 + lz], L[cr * 64 + lz]);
            const double v = eik_update(me, a, bb, sv[c], dx);
            tn[c] = (srcmask >> c) & 1u ? 0.0 : v;
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < C; ++c) {
            tv[c] = tn[c];
            L[(w + 4 * c) * 64 + lz] = tn[c];
        }
        __syncthreads();
    }
    if (lz < H || lz >= H + OZ || !zin) return;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int lx = w + 4 * c, gj = x0 + lx;
        if (lx >= H && lx < H + OX && gj < nx) TN[gj * nz + gi] = tv[c];
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
// receivers r in order (table tr[r], record column xr[r]).  A thread owns a
// pixel and takes the shots S at a time, so each receiver's index is loaded
// once per S shots; every shot keeps its own sum over the receivers in order
// and the stack adds the shots in order -- the reference's sums exactly.
template <int S>
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
        for (int s0 = 0; s0 < nshot; s0 += S) {
            const int ns = min(S, nshot - s0);
            int ks[S];
            double m[S];
#pragma unroll
            for (int u = 0; u < S; ++u) {
                ks[u] = u < ns ? kidx[(int64_t)ts[s0 + u] * nn + q] : 0;
                m[u] = 0.0;
            }
            for (int r = 0; r < nrec; ++r) {
                const int64_t kr = kidx[(int64_t)tr[r] * nn + q];
                const double *col = shots + (int64_t)s0 * nt * ncol + (int64_t)xr[r] * nt - 1;
#pragma unroll
                for (int u = 0; u < S; ++u) {
                    if (u >= ns) break;
                    const int64_t it = ks[u] + kr;  // 1-based sample
                    if (it < 1 || it > nt) {
                        oob = true;
                        continue;
                    }
                    m[u] = m[u] + col[(int64_t)u * nt * ncol + it];
                }
            }
#pragma unroll
            for (int u = 0; u < S; ++u) {
                if (u >= ns) break;
                if (images) images[(int64_t)(s0 + u) * nn + q] = m[u];
                st = st + m[u];
            }
        }
        stack[q] = st;
        if (oob) atomicOr(bad, 1);
    }
}
// Shot groups in parallel: workgroup row blockIdx.y migrates shots
// S*y .. S*y+S-1 (the loop of kirch_stack for one group) and writes their
// images; kirch_sum then stacks the images in shot order, the same additions
// in the same order as kirch_stack's running sum.  Four times the workgroups
// of kirch_stack at 16 shots, for the latency of the record gathers.
template <int S>
__global__ void __launch_bounds__(256) kirch_group(const int *__restrict__ kidx, int64_t nn,
                                                   const int *__restrict__ ts, int nshot,
                                                   const int *__restrict__ tr,
                                                   const int *__restrict__ xr, int nrec,
                                                   const double *__restrict__ shots, int64_t nt,
                                                   int64_t ncol, double *__restrict__ images,
                                                   int *__restrict__ bad) {
    const int s0 = (int)blockIdx.y * S;
    const int ns = min(S, nshot - s0);
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nn;
         q += (int64_t)gridDim.x * blockDim.x) {
        bool oob = false;
        int ks[S];
        double m[S];
#pragma unroll
        for (int u = 0; u < S; ++u) {
            ks[u] = u < ns ? kidx[(int64_t)ts[s0 + u] * nn + q] : 0;
            m[u] = 0.0;
        }
        for (int r = 0; r < nrec; ++r) {
            const int64_t kr = kidx[(int64_t)tr[r] * nn + q];
            const double *col = shots + (int64_t)s0 * nt * ncol + (int64_t)xr[r] * nt - 1;
#pragma unroll
            for (int u = 0; u < S; ++u) {
                if (u >= ns) break;
                const int64_t it = ks[u] + kr;  // 1-based sample
                if (it < 1 || it > nt) {
                    oob = true;
                    continue;
                }
                m[u] = m[u] + col[(int64_t)u * nt * ncol + it];
            }
        }
#pragma unroll
        for (int u = 0; u < S; ++u) {
            if (u >= ns) break;
            images[(int64_t)(s0 + u) * nn + q] = m[u];
        }
        if (oob) atomicOr(bad, 1);
    }
}

__global__ void __launch_bounds__(256) kirch_sum(const double *__restrict__ images, int64_t nn,
                                                 int nshot, double *__restrict__ stack) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nn;
         q += (int64_t)gridDim.x * blockDim.x) {
        double st = 0.0;
        for (int s = 0; s < nshot; ++s) st = st + images[(int64_t)s * nn + q];
        stack[q] = st;
    }
}

constexpr int kShotBatch = 4;

// Pass kernel choice: H passes per launch on 64 x TW tiles, or (h = 0) the
// plain one-pass kernel.  SSSI_EIK = "plain" | "<H>x<TW>" (4x32, 4x64, 8x48,
// 8x64) overrides the default 4x32.  Cell-passes per second on 600 tables of
// a 300 x 600 grid, 1800 passes: plain 148 G, 4x32 369 G, 8x48 315 G,
// 4x64 273 G, 8x64 261 G (the wider tiles hold 16 columns per wave in
// registers, 176+ VGPRs, and run 2 waves per SIMD).
struct EikLayout {
    int h = 4, tw = 32;
    void launch(dim3 grid, hipStream_t s, const double *slow, const double *told, double *tnew,
                int nz, int nx, int ntz, int ntiles, const int64_t *src, double dx,
                int np) const {
#define SSSI_EIK_CASE(H_, TW_)                                                               \
    if (h == H_ && tw == TW_) {                                                              \
        hipLaunchKernelGGL((eik_tile<H_, TW_>), grid, dim3(256), 0, s, slow, told, tnew, nz, nx, \
                           ntz, ntiles, src, dx, np);                                        \
        return;                                                                              \
    }
        SSSI_EIK_CASE(4, 32)
        SSSI_EIK_CASE(4, 64)
        SSSI_EIK_CASE(8, 48)
        SSSI_EIK_CASE(8, 64)
#undef SSSI_EIK_CASE
        fail(SSSI_EINVAL, "eikonal layout %dx%d not built", h, tw);
    }
};

static EikLayout eik_layout() {
    EikLayout l;
    const char *e = getenv("SSSI_EIK");
    if (!e || !*e) return l;
    if (!strcmp(e, "plain")) {
        l.h = 0;
        return l;
    }
    int h = 0, tw = 0;
    if (sscanf(e, "%dx%d", &h, &tw) == 2) {
        l.h = h;
        l.tw = tw;
        if ((h == 4 && (tw == 32 || tw == 64)) || (h == 8 && (tw == 48 || tw == 64))) return l;
    }
    fail(SSSI_EINVAL, "SSSI_EIK must be plain, 4x32, 4x64, 8x48 or 8x64 (got \"%s\")", e);
}

// ---------------------------------------------------------------------------
// Traveltime tables of sources (1-based (sz, sx)) on one model, resident.
class Eikonal2D {
   public:
    int64_t nz = 0, nx = 0;
    int nsrc = 0;
    OwnedStream stream;  // before the buffers: destroyed after them
    DevBuf<double> slow, t[2];
    DevBuf<int64_t> src;
    int cur = 0;
    int64_t passes = 0, launches = 0;
    float pass_ms = 0;  // kernel time of the passes (HIP events on `stream`)

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
        stream.create();
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
        hipEvent_t ev[2];
        SSSI_HIP(hipEventCreate(&ev[0]));
        SSSI_HIP(hipEventCreate(&ev[1]));
        SSSI_HIP(hipEventRecord(ev[0], stream));
        const EikLayout lay = eik_layout();
        if (lay.h == 0) {
            for (int p = 0; p < 4 * iMax; ++p) {
                hipLaunchKernelGGL(eik_pass, dim3(grid_blocks(total)), dim3(256), 0, stream,
                                   slow.p, t[cur].p, t[cur ^ 1].p, nz, nx, nsrc, src.p, dxi);
                SSSI_HIP(hipGetLastError());
                cur ^= 1;
            }
            launches = 4 * (int64_t)iMax;
        } else {
            const int oz = 64 - 2 * lay.h, ox = lay.tw - 2 * lay.h;
            const int ntz = (int)((nz + oz - 1) / oz), ntx = (int)((nx + ox - 1) / ox);
            const int64_t nblk = (int64_t)ntz * ntx * nsrc;
            if (nblk > INT32_MAX) fail(SSSI_EINVAL, "too many traveltime tables for one launch");
            launches = 0;
            for (int p = 0; p < 4 * iMax; p += lay.h) {
                const int np = std::min(lay.h, 4 * iMax - p);
                lay.launch(dim3((unsigned)nblk), stream, slow.p, t[cur].p, t[cur ^ 1].p, (int)nz,
                           (int)nx, ntz, ntz * ntx, src.p, dxi, np);
                SSSI_HIP(hipGetLastError());
                cur ^= 1;
                ++launches;
            }
        }
        passes = 4 * (int64_t)iMax;
        SSSI_HIP(hipEventRecord(ev[1], stream));
        SSSI_HIP(hipStreamSynchronize(stream));
        SSSI_HIP(hipEventElapsedTime(&pass_ms, ev[0], ev[1]));
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
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

// Device buffers of a migration call: the stack and the per-shot images (or
// their scratch); a resident handle keeps one across calls, so repeated
// migrations allocate nothing.
struct KirchWork {
    DevBuf<double> stack, img;
    DevBuf<int> ts, tr, xr, bad;
    HostBuf<double> hstack;  // pinned (resident handles): the stack comes back at link speed
    HostBuf<int> hbad;
    hipEvent_t ev[2] = {nullptr, nullptr};  // kernel timing, created once
    KirchWork() = default;
    KirchWork(const KirchWork &) = delete;
    KirchWork &operator=(const KirchWork &) = delete;
    ~KirchWork() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    void ensure(int64_t nn, size_t nimg, int nshot, int nrec, bool pinned) {
        for (hipEvent_t &e : ev)
            if (!e) SSSI_HIP(hipEventCreate(&e));
        if (stack.n < (size_t)nn) stack.alloc(nn);
        if (img.n < nimg) img.alloc(nimg);
        if (ts.n < (size_t)nshot) ts.alloc(nshot);
        if (tr.n < (size_t)std::max(nrec, 1)) tr.alloc(std::max(nrec, 1));
        if (xr.n < (size_t)std::max(nrec, 1)) xr.alloc(std::max(nrec, 1));
        if (bad.n < 1) bad.alloc(1);
        if (pinned && hstack.n < (size_t)nn) hstack.alloc(nn);
        if (pinned && hbad.n < 1) hbad.alloc(1);
    }
};

// Migration of nshot device-resident records against resident index tables;
// returns the kernel time (HIP events on stream s).  `work` may be null (the
// call's own buffers).
static float kirchhoff_run(const int *kidx_dev, int64_t nz, int64_t nx, int64_t ntab,
                           const int *xrec, int nrec, const int *xsr, int nsr,
                           const double *dshots, int64_t nt, int nshot, const int *ixs,
                           double *stack, double *images, hipStream_t s, KirchWork *work) {
    if (nt < 1 || nshot < 1 || !dshots || !ixs || !stack || nrec < 0 || (nrec > 0 && !xrec))
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
    // shot groups in parallel (then the stack in shot order) when there is
    // more than one group and their images fit a scratch buffer
    const int ngroups = (nshot + kShotBatch - 1) / kShotBatch;
    const bool grouped = ngroups > 1 && ngroups <= 65535 && (double)nn * nshot * 8 <= 4e9;
    KirchWork local;
    KirchWork &w = work ? *work : local;
    const bool pinned = work != nullptr;  // a one-shot call copies straight to the caller
    w.ensure(nn, images || grouped ? (size_t)nn * nshot : 0, nshot, nrec, pinned);
    const int *dts = w.ts.p, *dtr = w.tr.p, *dxr = w.xr.p;
    SSSI_HIP(hipMemcpyAsync(w.ts.p, ts.data(), (size_t)nshot * 4, hipMemcpyHostToDevice, s));
    if (nrec > 0) {
        SSSI_HIP(hipMemcpyAsync(w.tr.p, tr.data(), (size_t)nrec * 4, hipMemcpyHostToDevice, s));
        SSSI_HIP(hipMemcpyAsync(w.xr.p, xr0.data(), (size_t)nrec * 4, hipMemcpyHostToDevice, s));
    }
    SSSI_HIP(hipMemsetAsync(w.bad.p, 0, 4, s));
    hipEvent_t *ev = w.ev;
    SSSI_HIP(hipEventRecord(ev[0], s));
    if (grouped) {
        hipLaunchKernelGGL(kirch_group<kShotBatch>, dim3(grid_blocks(nn), ngroups), dim3(256), 0,
                           s, kidx_dev, nn, dts, nshot, dtr, dxr, nrec, dshots, nt, nx,
                           w.img.p, w.bad.p);
        hipLaunchKernelGGL(kirch_sum, dim3(grid_blocks(nn)), dim3(256), 0, s, w.img.p, nn, nshot,
                           w.stack.p);
    } else {
        hipLaunchKernelGGL(kirch_stack<kShotBatch>, dim3(grid_blocks(nn)), dim3(256), 0, s,
                           kidx_dev, nn, dts, nshot, dtr, dxr, nrec, dshots, nt, nx,
                           w.stack.p, images ? w.img.p : nullptr, w.bad.p);
    }
    SSSI_HIP(hipGetLastError());
    SSSI_HIP(hipEventRecord(ev[1], s));
    int hbad_local = 0;
    int *hbad_dst = pinned ? w.hbad.p : &hbad_local;
    SSSI_HIP(hipMemcpyAsync(hbad_dst, w.bad.p, 4, hipMemcpyDeviceToHost, s));
    SSSI_HIP(hipMemcpyAsync(pinned ? w.hstack.p : stack, w.stack.p, (size_t)nn * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    if (images)
        SSSI_HIP(hipMemcpyAsync(images, w.img.p, (size_t)nn * nshot * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    SSSI_HIP(hipStreamSynchronize(s));
    if (pinned) std::memcpy(stack, w.hstack.p, (size_t)nn * sizeof(double));
    const int hbad = *hbad_dst;
    float ms = 0;
    SSSI_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    if (hbad)
        fail(SSSI_EINVAL, "a traveltime sample index falls outside the %lld-sample record "
             "(the reference's migrate indexes past the shot there)", (long long)nt);
    return ms;
}

// Traveltimes of surface positions and their sample indices, resident.
class Kirchhoff2D {
   public:
    int64_t nz = 0, nx = 0;
    std::vector<int> xsr;
    double dt = 0;
    std::unique_ptr<Eikonal2D> eik;
    DevBuf<int> kidx;
    DevBuf<double> rec;  // loaded records (nt, nx, nrec_shots), column-major
    int64_t rec_nt = 0;
    int rec_shots = 0;
    float migrate_ms = 0;  // kernel time of the last migration
    KirchWork work;        // stack and image buffers, kept across migrations

    void load(const double *shots, int64_t nt, int nshot) {
        if (!shots || nt < 1 || nshot < 1) fail(SSSI_EINVAL, "kirchhoff: no records to load");
        const size_t n = (size_t)nt * (size_t)nx * (size_t)nshot;
        if (rec.n != n) {
            rec_shots = 0;
            rec.alloc(n);  // frees the previous records first
        }
        SSSI_HIP(hipMemcpyAsync(rec.p, shots, rec.bytes(), hipMemcpyHostToDevice, eik->stream));
        SSSI_HIP(hipStreamSynchronize(eik->stream));
        rec_nt = nt;
        rec_shots = nshot;
    }
    void migrate(const int *xrec, int nrec, const int *ixs, double *stack, double *images) {
        if (rec_shots < 1) fail(SSSI_ESTATE, "kirchhoff: no records loaded");
        migrate_ms = kirchhoff_run(kidx.p, nz, nx, (int64_t)xsr.size(), xrec, nrec, xsr.data(),
                                   (int)xsr.size(), rec.p, rec_nt, rec_shots, ixs, stack,
                                   images, eik->stream, &work);
    }

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
        OwnedStream s;
        s.create();
        {
            const int64_t n = nz * nx * ntab;
            DevBuf<double> tt(n);
            DevBuf<int> kidx(n);
            SSSI_HIP(hipMemcpyAsync(tt.p, travelTime, tt.bytes(), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(eik_index, dim3(grid_blocks(n)), dim3(256), 0, s, tt.p, n, dt,
                               kidx.p);
            SSSI_HIP(hipGetLastError());
            if (!shot || nt < 1) fail(SSSI_EINVAL, "shot must not be empty");
        DevBuf<double> rec((size_t)nt * nx);
        SSSI_HIP(hipMemcpyAsync(rec.p, shot, rec.bytes(), hipMemcpyHostToDevice, s));
        kirchhoff_run(kidx.p, nz, nx, ntab, xRecGrid, nrec, xShotAndRecGrid, nsr, rec.p, nt, 1,
                      &ixs, m, nullptr, s, nullptr);
        }
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
        h->impl->load(shots, nt, nshot);
        h->impl->migrate(xRecGrid, nrec, ixs, stack, images);
    });
}

int sssi_kirchhoff_load_shots(sssi_kirchhoff *h, const double *shots, int64_t nt, int nshot) {
    return kguard([&] {
        if (!h || !h->impl) fail(SSSI_EINVAL, "NULL handle");
        h->impl->load(shots, nt, nshot);
    });
}

int sssi_kirchhoff_migrate_loaded(sssi_kirchhoff *h, const int *xRecGrid, int nrec,
                                  const int *ixs, double *stack, double *images) {
    return kguard([&] {
        if (!h || !h->impl) fail(SSSI_EINVAL, "NULL handle");
        h->impl->migrate(xRecGrid, nrec, ixs, stack, images);
    });
}

int sssi_kirchhoff_timing(sssi_kirchhoff *h, double *eikonal_ms, int64_t *passes,
                          int64_t *launches, double *migrate_ms) {
    return kguard([&] {
        if (!h || !h->impl) fail(SSSI_EINVAL, "NULL handle");
        if (eikonal_ms) *eikonal_ms = h->impl->eik->pass_ms;
        if (passes) *passes = h->impl->eik->passes;
        if (launches) *launches = h->impl->eik->launches;
        if (migrate_ms) *migrate_ms = h->impl->migrate_ms;
    });
}

void sssi_kirchhoff_destroy(sssi_kirchhoff *h) { delete h; }

}  // extern "C"
