// prop2d.hip -- 2-D acoustic propagators for gfx950: CPML (fused time step of
// fwdTimeCpmlFor2dAw_mex.c:183-290 / rvsTimeCpmlFor2dAw_mex.c:172-281) and the
// Cerjan sponge (src/fwdTimeSpongeFor2dAw.m:69-106, src/rvsTimeSpongeFor2dAw.m).
//
// One kernel launch per time step does everything the reference does in six
// derivative passes, eight elementwise passes and two copies:
//   D+p, Phi update, A = D+p + Phi, D-A, Psi update, P = D-A + Psi (z and x),
//   p_new = v^2 dt^2 (P_z + P_x + s) + 2 p - p_old, source injection,
//   receiver gather (top row), optional snapshot.
// p_new overwrites p_old in place (each cell reads p_old only at itself), so
// the time rotation is a pointer swap.  Phi is ping-ponged (a tile also
// computes the N half points on each side that its neighbours own), Psi is
// updated in place.  Phi/Psi and b-1 exist only for the damping strips.
//
// Tiling: 256 threads = 4 wavefronts; a tile is 64 columns x TZ rows, lane =
// column, each wavefront owns RZ = TZ/4 consecutive rows.  The tile plus an
// l-wide halo is staged once in LDS; the z derivative runs from registers
// (column window), the x derivative via a second LDS array of A_x at half
// points.  Tiles are numbered so that each of the 8 XCDs gets a contiguous
// band (halo rows shared by neighbouring tiles stay in that XCD's L2).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "prop.h"

namespace sssi {

struct Cpml2dArgs {
    const float *pcur;
    float *pold;  // read p_old, write p_new
    const float *vdt2;
    int nz, nx, pitch, rows, ntx, nblk, nB;
    const int *order;           // block -> tile (strip tiles first, XCD-banded interior)
    float cz[8], cx[8];         // c_k/dz, c_k/dx (two-stage form)
    float w0, wz[16], wx[16];   // composite D-D+ weights / d^2 (w0: both axes)
    const float *zphi_in;
    float *zphi_out, *zpsi;
    const float *zbm1;
    const float *xphi_in;
    float *xphi_out, *xpsi;
    const float *xbm1;
    const int *src_ptr;   // per wave tile [ptr, ptr+1) into src_cell/src_amp, or null
    const int *src_cell;  // i * nx + j
    const float *src_amp; // amplitudes of this step (sorted order)
    const float *surf;    // this step's surface row (reverse runs) or null
    float *rec;           // this step's receiver row or null
    float *snap;          // this step's snapshot slice (nz*nx) or null
};

__device__ __forceinline__ int xcd_tile(int b, int nblk) {
    // blocks are dispatched round-robin over 8 XCDs: block b runs on XCD b%8
    const int q = nblk >> 3, r = nblk & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// LDS written by some lanes of a wavefront and read by others: wavefront-scope
// release/acquire (no workgroup barrier; waves of a block are independent).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Raw buffer access: the wave-uniform row offset goes in the scalar offset,
// the lane's column in the vector offset, so a load costs no vector address
// arithmetic.  Out-of-range offsets (beyond num_records) read 0 and drop stores.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                       uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, 0);
}

// One time step.  Each wavefront owns a 64 x RZ tile (lane = column); the four
// wavefronts of a block take four consecutive tiles in z and never
// synchronise with each other.  All global loads of a tile (z window, x halo,
// p_old, v^2dt^2) are issued before the first use.  Tiles whose stencil reach
// touches no damping strip use the composite (D-D+) stencil of radius
// L = 2N-1, which equals the two-stage form there (Phi = Psi = 0); strip
// tiles run the two-stage CPML update.
template <int N, int RZ>
__global__ void __launch_bounds__(256) cpml2d_step(Cpml2dArgs a) {
    constexpr int L = 2 * N - 1;
    constexpr int PW = kTX + 2 * L;  // p rows: columns x0-L .. x0+63+L
    constexpr int AW = kTX + L;      // A_x rows: half points x0-N .. x0+63+N-1
    constexpr int HN = 2 * L * RZ;   // x-halo values per tile
    constexpr int HPL = (HN + 63) / 64;
    __shared__ float s_p[4][RZ * PW];
    __shared__ float s_a[4][RZ * AW];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tile = a.order[blockIdx.x];
    const int tzi = tile / a.ntx, txi = tile - tzi * a.ntx;
    const int x0 = txi * kTX;
    const int nz = a.nz, nx = a.nx, nB = a.nB;
    const int zs = (tzi * 4 + w) * RZ;
    if (zs >= nz) return;
    const uint32_t rowb = (uint32_t)a.pitch * 4u;
    const uint32_t nbytes = rowb * (uint32_t)a.rows;
    const __amdgpu_buffer_rsrc_t rP = buf_rsrc(a.pcur, nbytes);
    const __amdgpu_buffer_rsrc_t rO = buf_rsrc(a.pold, nbytes);
    const __amdgpu_buffer_rsrc_t rV = buf_rsrc(a.vdt2, nbytes);
    const int j = x0 + lane;
    const bool jin = j < nx;
    const uint32_t vcol = (uint32_t)(j + kHX) * 4u;
    const uint32_t srow = (uint32_t)(zs + kHZ) * rowb;  // first row of the tile

    // ---- issue every global load of the tile
    float col[RZ + 2 * L], po[RZ], v2[RZ], hal[HPL];
#pragma unroll
    for (int r = 0; r < RZ + 2 * L; ++r) col[r] = buf_ld(rP, vcol, srow + (r - L) * rowb);
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        po[r] = buf_ld(rO, vcol, srow + r * rowb);
        v2[r] = buf_ld(rV, vcol, srow + r * rowb);
    }
    int h_lds[HPL];
#pragma unroll
    for (int u = 0; u < HPL; ++u) {
        const int e = lane + 64 * u;
        const int r = e / (2 * L), c = e - r * (2 * L);
        const int xc = c < L ? x0 - L + c : x0 + kTX - L + c;
        h_lds[u] = e < HN ? r * PW + (c < L ? c : kTX + c) : -1;
        hal[u] = buf_ld(rP, (uint32_t)(r * a.pitch + xc + kHX) * 4u, srow);
    }

    // ---- the tile's rows with an L-wide x halo into the wave's LDS rows
    float *sp = s_p[w];
#pragma unroll
    for (int r = 0; r < RZ; ++r) sp[r * PW + L + lane] = col[L + r];
#pragma unroll
    for (int u = 0; u < HPL; ++u)
        if (h_lds[u] >= 0) sp[h_lds[u]] = hal[u];
    wave_lds_sync();

    const int zb0 = nz - nB;
    const bool zstrip = nB > 0 && (zs + RZ + N - 1 > zb0);
    const bool xstrip = nB > 0 && (x0 - N < nB || x0 + kTX + N - 1 > nx - nB);
    float lap[RZ];
    if (!zstrip && !xstrip) {
        // ---- interior: composite stencil
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            const float *row = sp + r * PW + L + lane;
            float acc = a.w0 * col[L + r];
#pragma unroll
            for (int m = 1; m <= L; ++m) acc = fmaf(a.wz[m], col[L + r + m] + col[L + r - m], acc);
#pragma unroll
            for (int m = 1; m <= L; ++m) acc = fmaf(a.wx[m], row[m] + row[-m], acc);
            lap[r] = acc;
        }
    } else {
        // ---- damping strips: two-stage form with Phi/Psi
        // A_z at half points h = zs-N+q (q < RZ+L): D+p(h) = sum c_k (p(h+1+k) - p(h-k))
        float Az[RZ + L];
#pragma unroll
        for (int q = 0; q < RZ + L; ++q) {
            float d = 0.f;
#pragma unroll
            for (int k = 0; k < N; ++k) d = fmaf(a.cz[k], col[q + N + k] - col[q + N - 1 - k], d);
            Az[q] = d;
        }
        if (zstrip && jin) {
            // all z-strip memory variables of the tile in one batch of loads
            // (b-1) at half point h and at row i = h share one index
            float phz[RZ + L], bmh[RZ + L], psz[RZ];
#pragma unroll
            for (int q = 0; q < RZ + L; ++q) {
                const int h = zs - N + q;
                const bool in = h >= zb0 && h < nz;
                const size_t o = in ? (size_t)(h - zb0) * nx + j : 0;
                phz[q] = in ? a.zphi_in[o] : 0.f;
                bmh[q] = in ? a.zbm1[o] : 0.f;
            }
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                const int i = zs + r;
                const bool in = i >= zb0 && i < nz;
                psz[r] = in ? a.zpsi[(size_t)(i - zb0) * nx + j] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < RZ + L; ++q) {
                const int h = zs - N + q;
                if (h >= zb0 && h < nz) {
                    const float phi = fmaf(bmh[q], phz[q] + Az[q], phz[q]);  // b*phi + (b-1)*D
                    Az[q] += phi;
                    if (h >= zs && h < zs + RZ) a.zphi_out[(size_t)(h - zb0) * nx + j] = phi;
                }
            }
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                float d = 0.f;
#pragma unroll
                for (int k = 0; k < N; ++k) d = fmaf(a.cz[k], Az[r + N + k] - Az[r + N - 1 - k], d);
                const int i = zs + r;
                if (i >= zb0 && i < nz) {
                    const float psi = fmaf(bmh[r + N], psz[r] + d, psz[r]);
                    a.zpsi[(size_t)(i - zb0) * nx + j] = psi;
                    d += psi;
                }
                lap[r] = d;
            }
        } else {
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                float d = 0.f;
#pragma unroll
                for (int k = 0; k < N; ++k) d = fmaf(a.cz[k], Az[r + N + k] - Az[r + N - 1 - k], d);
                lap[r] = d;
            }
        }
        // A_x at half points h = x0-N+q, q < 64+L: lane q, lanes < L also q = 64+lane
        float *sa = s_a[w];
        const int nb2 = 2 * nB;
        const bool xown = xstrip && jin && (j < nB || j >= nx - nB);
        const int xs = j < nB ? j : j - (nx - nb2);
        // all x-strip memory variables of the tile in one batch of loads
        float phx[2][RZ], bmx[2][RZ], psx[RZ], bmo[RZ];
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            const int i = zs + r;
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int q = lane + half * kTX;
                const int h = x0 - N + q;
                const bool in = xstrip && (half == 0 || lane < L) && i < nz && h >= 0 && h < nx &&
                                (h < nB || h >= nx - nB);
                const size_t o = in ? (size_t)i * nb2 + (h < nB ? h : h - (nx - nb2)) : 0;
                phx[half][r] = in ? a.xphi_in[o] : 0.f;
                bmx[half][r] = in ? a.xbm1[o] : 0.f;
            }
            const bool own = xown && i < nz;
            const size_t o = own ? (size_t)i * nb2 + xs : 0;
            psx[r] = own ? a.xpsi[o] : 0.f;
            bmo[r] = own ? a.xbm1[o] : 0.f;
        }
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            const float *row = sp + r * PW;
            const int i = zs + r;
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int q = lane + half * kTX;
                if (half == 1 && lane >= L) break;
                float d = 0.f;
#pragma unroll
                for (int k = 0; k < N; ++k)
                    d = fmaf(a.cx[k], row[q + N + k] - row[q + N - 1 - k], d);
                const int h = x0 - N + q;
                if (xstrip && i < nz && h >= 0 && h < nx && (h < nB || h >= nx - nB)) {
                    const float phi = fmaf(bmx[half][r], phx[half][r] + d, phx[half][r]);
                    d += phi;
                    if (h >= x0 && h < x0 + kTX)
                        a.xphi_out[(size_t)i * nb2 + (h < nB ? h : h - (nx - nb2))] = phi;
                }
                sa[r * AW + q] = d;
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            const float *ar = sa + r * AW + lane + N;
            float d = 0.f;
#pragma unroll
            for (int k = 0; k < N; ++k) d = fmaf(a.cx[k], ar[k] - ar[-1 - k], d);
            const int i = zs + r;
            if (xown && i < nz) {
                const float psi = fmaf(bmo[r], psx[r] + d, psx[r]);
                a.xpsi[(size_t)i * nb2 + xs] = psi;
                d += psi;
            }
            lap[r] += d;
        }
    }

    // ---- sources
    float s[RZ];
#pragma unroll
    for (int r = 0; r < RZ; ++r) s[r] = 0.f;
    if (a.src_ptr) {
        const int wt = (zs / RZ) * a.ntx + txi;
        const int k0 = a.src_ptr[wt], k1 = a.src_ptr[wt + 1];
        for (int k = k0; k < k1; ++k) {
            const int cell = a.src_cell[k];
            const int ci = cell / nx, cj = cell - ci * nx;
            if (cj == j) {
#pragma unroll
                for (int r = 0; r < RZ; ++r)
                    if (ci == zs + r) s[r] += a.src_amp[k];
            }
        }
    }
    if (a.surf && zs == 0 && jin) s[0] += a.surf[j];

    // ---- update: p_new = v^2dt^2 (P_z + P_x + s) + 2p - p_old
    if (!jin) return;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        const int i = zs + r;
        if (i >= nz) break;
        const float pc = col[r + L];
        // pc + (v2*(...) + (pc - po)): pc - po is exact when the two levels are
        // within a factor 2, which keeps the rounding of the update at the
        // scale of the increment rather than of the field.
        const float pn = pc + fmaf(v2[r], lap[r] + s[r], pc - po[r]);
        buf_st(pn, rO, vcol, srow + r * rowb);
        if (a.rec && i == 0) a.rec[j] = pn;
        if (a.snap) a.snap[(size_t)i * nx + j] = pn;
    }
}

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
    int nz, nx, pitch, nB;
    float idx2, idz2;  // 1/dx^2, 1/dz^2
    const int *src_cell;
    const float *src_amp;
    int nsrc;
    const float *surf;
    float *rec;
    float *snap;
};

__global__ void __launch_bounds__(256) sponge2d_step(Sponge2dArgs a) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * 4;
    if (j >= a.nx) return;
    const ptrdiff_t pitch = a.pitch;
    for (int r = 0; r < 4; ++r) {
        const int i = i0 + r;
        if (i >= a.nz) return;
        const size_t g0 = (size_t)(i + kHZ) * pitch + (j + kHX);
        const float vd = a.vdt[g0];
        const float v2 = vd * vd;
        const float ax = v2 * a.idx2, az = v2 * a.idz2;
        float gf = 1.f;
        if (i >= a.nz - a.nB) gf = a.g[a.nz - 1 - i];
        else if (j < a.nB) gf = a.g[j];
        else if (j >= a.nx - a.nB) gf = a.g[a.nx - 1 - j];
        const float *c = a.pcur + g0;
        float s = 0.f;
        for (int k = 0; k < a.nsrc; ++k)
            if (a.src_cell[k] == i * a.nx + j) s += a.src_amp[k];
        if (a.surf && i == 0) s += a.surf[j];
        // b p - p_old + ax (pR+pL) + az (pD+pU) regrouped as
        // p + (p - p_old) + ax((pR-p)+(pL-p)) + az((pD-p)+(pU-p)): the FP32
        // rounding then scales with the increments instead of with 2p.
        const float pc = c[0];
        const float lap = ax * ((c[1] - pc) + (c[-1] - pc)) + az * ((c[pitch] - pc) + (c[-pitch] - pc));
        float pn = pc + ((pc - gf * a.pold[g0]) + (lap - v2 * s));
        pn *= gf;
        a.pold[g0] = pn;
        if (a.rec && i == 0) a.rec[j] = pn;
        if (a.snap) a.snap[(size_t)i * a.nx + j] = pn;
    }
}

// Compact interior (nz x nx, x fastest) <-> padded layout.
__global__ void unpad2d(const float *src, float *dst, int nz, int nx, int pitch, float scale_g,
                        const float *g, int nB) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nz || j >= nx) return;
    float f = 1.f;
    if (g) {
        if (i >= nz - nB) f = g[nz - 1 - i];
        else if (j < nB) f = g[j];
        else if (j >= nx - nB) f = g[nx - 1 - j];
    }
    (void)scale_g;
    dst[(size_t)i * nx + j] = f * src[(size_t)(i + kHZ) * pitch + j + kHX];
}

// Composite weights of D-D+ (before division by d^2): w[L+m], m = -L..L, with
// D+p(h) = sum_q c_q (p(h+1+q) - p(h-q)) and D-A(i) = sum_k c_k (A(i+k) - A(i-1-k)).
static void composite_weights(int N, const double *c, double *w) {
    const int L = 2 * N - 1;
    for (int m = 0; m <= 2 * L; ++m) w[m] = 0.0;
    for (int k = 0; k < N; ++k)
        for (int q = 0; q < N; ++q) {
            const double cc = c[k] * c[q];
            w[L + k + 1 + q] += cc;   // A(i+k) <- p(i+k+1+q)
            w[L + k - q] -= cc;       // A(i+k) <- p(i+k-q)
            w[L - k + q] -= cc;       // A(i-1-k) <- p(i-k+q)
            w[L - 1 - k - q] += cc;   // A(i-1-k) <- p(i-1-k-q)
        }
}

static int default_rz(int N) { return N >= 6 ? 4 : 4; }

// ---------------------------------------------------------------------------
class Prop2D final : public Prop {
   public:
    int64_t pitch = 0, rows = 0;
    DevBuf<float> p[2], vfield;  // vfield: v^2dt^2 (CPML) or v*dt (sponge), padded
    DevBuf<float> zphi[2], zpsi, zbm1, xphi[2], xpsi, xbm1, gtab;
    int cur = 0, phase = 0;
    float cz[8] = {0}, cx[8] = {0};
    int rz = 8;  // rows per wavefront tile (64 x rz); a block holds 4 tiles stacked in z
    float w0 = 0.f, wz[16] = {0}, wx[16] = {0};
    std::vector<int> h_src_cell;  // sponge: unsorted cell list
    DevBuf<float> snap_tmp;
    DevBuf<int> order;  // launch position -> block tile

    Prop2D(const double *v, int64_t nz_, int64_t nx_, int abc_, int N_, int nB_, double dz_,
           double dx_, double dt_) {
        dim = 2;
        abc = abc_;
        nz = nz_; nx = nx_; ny = 1; N = N_; nB = nB_; dz = dz_; dx = dx_; dt = dt_;
        SSSI_HIP(hipGetDevice(&device));
        SSSI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        pitch = round_up(kHX + round_up(nx, kTX) + 32, 64);
        rows = kHZ + round_up(nz, kTZmax) + kHZbot;
        const size_t np = (size_t)pitch * rows;
        p[0].alloc(np);
        p[1].alloc(np);
        vfield.alloc(np);
        std::vector<float> hv(np, 0.f);
        for (int64_t j = 0; j < nx; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                const double vv = v[j * nz + i];
                const double f = (abc == SSSI_ABC_CPML) ? (vv * dt) * (vv * dt) : vv * dt;
                hv[(size_t)(i + kHZ) * pitch + j + kHX] = (float)f;
            }
        SSSI_HIP(hipMemcpy(vfield.p, hv.data(), np * sizeof(float), hipMemcpyHostToDevice));
        if (abc == SSSI_ABC_CPML) {
            double c[8];
            staggered_coefficients(N, c);
            for (int k = 0; k < N; ++k) {
                cz[k] = (float)(c[k] / dz);
                cx[k] = (float)(c[k] / dx);
            }
            double w[32];
            composite_weights(N, c, w);
            const int L = 2 * N - 1;
            w0 = (float)(w[L] / (dz * dz) + w[L] / (dx * dx));
            for (int m = 1; m <= L; ++m) {
                wz[m] = (float)(w[L + m] / (dz * dz));
                wx[m] = (float)(w[L + m] / (dx * dx));
            }
            if (nB > 0) {
                std::vector<float> zb, xb;
                cpml2d_bm1(v, nz, nx, nB, dz, dx, dt, zb, xb);
                zbm1.alloc(zb.size());
                xbm1.alloc(xb.size());
                SSSI_HIP(hipMemcpy(zbm1.p, zb.data(), zbm1.bytes(), hipMemcpyHostToDevice));
                SSSI_HIP(hipMemcpy(xbm1.p, xb.data(), xbm1.bytes(), hipMemcpyHostToDevice));
                for (int q = 0; q < 2; ++q) {
                    zphi[q].alloc(zb.size());
                    xphi[q].alloc(xb.size());
                }
                zpsi.alloc(zb.size());
                xpsi.alloc(xb.size());
            }
            rz = default_rz(N);
            if (const char *e = std::getenv("SSSI_RZ")) {
                const int r = std::atoi(e);
                if (r == 4 || r == 8) rz = r;
            }
            build_order();
        } else {
            std::vector<float> g(std::max(nB, 1), 1.f);
            const double alpha = 0.015;
            for (int m = 0; m < nB; ++m) {
                const double e = std::exp(-((alpha * (nB - m)) * (alpha * (nB - m))));
                g[m] = (float)std::pow(e, 10);
            }
            gtab.alloc(g.size());
            SSSI_HIP(hipMemcpy(gtab.p, g.data(), gtab.bytes(), hipMemcpyHostToDevice));
        }
        reset();
    }

    ~Prop2D() override {
        if (stream) (void)hipStreamDestroy(stream);
    }

    int64_t cells() const override { return nz * nx; }

    int tz() const { return 4 * rz; }
    int ntx() const { return (int)((nx + kTX - 1) / kTX); }
    int ntz() const { return (int)((nz + tz() - 1) / tz()); }
    int nwz() const { return (int)((nz + rz - 1) / rz); }  // wave tiles along z

    // Launch order of the block tiles (64 x 4*rz).  Tiles that run the
    // two-stage CPML update (damping strips within the stencil reach) are
    // launched first so their longer dependency chains overlap the interior
    // tiles instead of forming a tail.  Interior tiles follow, split into 8
    // contiguous bands: block b is dispatched to XCD b % 8, so each XCD works
    // through one band and the halo rows shared by neighbouring tiles stay in
    // its L2.
    void build_order() {
        const int ntx_ = ntx(), ntz_ = ntz();
        const int nblk = ntx_ * ntz_;
        std::vector<int> strip, inner;
        for (int tzi = 0; tzi < ntz_; ++tzi)
            for (int txi = 0; txi < ntx_; ++txi) {
                const int64_t z0 = (int64_t)tzi * tz(), x0 = (int64_t)txi * kTX;
                const bool zs = nB > 0 && z0 + tz() + N - 1 > nz - nB;
                const bool xs = nB > 0 && (x0 - N < nB || x0 + kTX + N - 1 > nx - nB);
                (zs || xs ? strip : inner).push_back(tzi * ntx_ + txi);
            }
        std::vector<int> ord(strip);
        const size_t ni = inner.size();
        std::vector<size_t> next(8), end(8);
        for (int x = 0; x < 8; ++x) {
            next[x] = ni * x / 8;
            end[x] = ni * (x + 1) / 8;
        }
        for (size_t b = ord.size(); b < (size_t)nblk; ++b) {
            int x = (int)(b % 8);
            if (next[x] == end[x])
                for (int y = 0; y < 8; ++y)
                    if (next[y] < end[y]) { x = y; break; }
            ord.push_back(inner[next[x]++]);
        }
        order.alloc(ord.size());
        SSSI_HIP(hipMemcpy(order.p, ord.data(), order.bytes(), hipMemcpyHostToDevice));
    }

    void reset() override {
        SSSI_HIP(hipMemsetAsync(p[0].p, 0, p[0].bytes(), stream));
        SSSI_HIP(hipMemsetAsync(p[1].p, 0, p[1].bytes(), stream));
        for (auto *b : {&zphi[0], &zphi[1], &zpsi, &xphi[0], &xphi[1], &xpsi})
            if (b->p) SSSI_HIP(hipMemsetAsync(b->p, 0, b->bytes(), stream));
        cur = 0;
        phase = 0;
        SSSI_HIP(hipStreamSynchronize(stream));
    }

    void set_points(int n, const int *iz, const int *ix, const int *, const double *wavelet,
                    int64_t nt) override {
        if (n < 0 || nt < 1 || (n > 0 && (!iz || !ix || !wavelet)))
            fail(SSSI_EINVAL, "set_points: bad arguments");
        for (int k = 0; k < n; ++k)
            if (iz[k] < 0 || iz[k] >= nz || ix[k] < 0 || ix[k] >= nx)
                fail(SSSI_EINVAL, "set_points: source %d at (%d,%d) outside the %lldx%lld grid",
                     k, iz[k], ix[k], (long long)nz, (long long)nx);
        nsrc = n;
        src_nt = nt;
        if (n == 0) {
            src_ptr.release(); src_cell.release(); src_amp.release();
            return;
        }
        // sort by tile (CPML) so each tile scans only its own points
        const int ntx_ = ntx();
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        // per wavefront tile (64 x rz), the unit the kernel scans
        auto tile_of = [&](int k) { return (iz[k] / rz) * ntx_ + ix[k] / kTX; };
        std::stable_sort(order.begin(), order.end(),
                         [&](int a_, int b_) { return tile_of(a_) < tile_of(b_); });
        const int ntiles = ntx_ * nwz();
        std::vector<int> ptr(ntiles + 1, 0), cell(n);
        for (int k = 0; k < n; ++k) {
            ptr[tile_of(order[k]) + 1]++;
            cell[k] = iz[order[k]] * (int)nx + ix[order[k]];
        }
        for (int t = 0; t < ntiles; ++t) ptr[t + 1] += ptr[t];
        std::vector<float> amp((size_t)nt * n);
        for (int64_t t = 0; t < nt; ++t)
            for (int k = 0; k < n; ++k) amp[(size_t)t * n + k] = (float)wavelet[t * n + order[k]];
        src_ptr.alloc(ptr.size());
        src_cell.alloc(n);
        src_amp.alloc(amp.size());
        SSSI_HIP(hipMemcpy(src_ptr.p, ptr.data(), src_ptr.bytes(), hipMemcpyHostToDevice));
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

    void ensure_rec(int64_t nt) {
        if (rec.n < (size_t)(nt * nx)) rec.alloc((size_t)nt * nx);
        rec_nt = nt;
    }

    template <int NN, int RZ>
    void launch_cpml(const Cpml2dArgs &args) {
        hipLaunchKernelGGL((cpml2d_step<NN, RZ>), dim3(args.nblk), dim3(256), 0, stream, args);
    }

    void step_cpml(int64_t t, bool reverse, bool record, float *snap) {
        Cpml2dArgs a{};
        a.pcur = p[cur].p;
        a.pold = p[cur ^ 1].p;
        a.vdt2 = vfield.p;
        a.nz = (int)nz; a.nx = (int)nx; a.pitch = (int)pitch; a.rows = (int)rows;
        a.ntx = ntx();
        a.nblk = ntx() * ntz();
        a.order = order.p;
        a.nB = nB;
        for (int k = 0; k < 8; ++k) { a.cz[k] = cz[k]; a.cx[k] = cx[k]; }
        a.w0 = w0;
        for (int m = 0; m < 16; ++m) { a.wz[m] = wz[m]; a.wx[m] = wx[m]; }
        if (nB > 0) {
            a.zphi_in = zphi[phase].p; a.zphi_out = zphi[phase ^ 1].p; a.zpsi = zpsi.p;
            a.zbm1 = zbm1.p;
            a.xphi_in = xphi[phase].p; a.xphi_out = xphi[phase ^ 1].p; a.xpsi = xpsi.p;
            a.xbm1 = xbm1.p;
        }
        if (!reverse && nsrc > 0 && t < src_nt) {
            a.src_ptr = src_ptr.p;
            a.src_cell = src_cell.p;
            a.src_amp = src_amp.p + (size_t)t * nsrc;
        }
        if (reverse && surf.p && t < surf_nt) a.surf = surf.p + (size_t)t * nx;
        if (record) a.rec = rec.p + (size_t)t * nx;
        a.snap = snap;
        const int key = N * 16 + rz;
        switch (key) {
#define SSSI_CASE(NN, R) case NN * 16 + R: launch_cpml<NN, R>(a); break;
            SSSI_CASE(1, 4) SSSI_CASE(1, 8) SSSI_CASE(2, 4) SSSI_CASE(2, 8)
            SSSI_CASE(3, 4) SSSI_CASE(3, 8) SSSI_CASE(4, 4) SSSI_CASE(4, 8)
            SSSI_CASE(5, 4) SSSI_CASE(5, 8) SSSI_CASE(6, 4) SSSI_CASE(6, 8)
            SSSI_CASE(7, 4) SSSI_CASE(7, 8) SSSI_CASE(8, 4) SSSI_CASE(8, 8)
#undef SSSI_CASE
            default: fail(SSSI_EINVAL, "no kernel for nDiffOrder=%d rz=%d", N, rz);
        }
        SSSI_HIP(hipGetLastError());
        cur ^= 1;
        phase ^= 1;
    }

    void step_sponge(int64_t t, bool reverse, bool record, float *snap) {
        Sponge2dArgs a{};
        a.pcur = p[cur].p;
        a.pold = p[cur ^ 1].p;
        a.vdt = vfield.p;
        a.g = gtab.p;
        a.nz = (int)nz; a.nx = (int)nx; a.pitch = (int)pitch; a.nB = nB;
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
        dim3 grid((unsigned)((nx + 63) / 64), (unsigned)((nz + 15) / 16));
        hipLaunchKernelGGL(sponge2d_step, grid, dim3(256), 0, stream, a);
        SSSI_HIP(hipGetLastError());
        cur ^= 1;
    }

    void run(int64_t t_first, int64_t count, bool reverse, bool record, const SnapSink *sink,
             LaunchStats *st) override {
        if (count < 0) fail(SSSI_EINVAL, "run: negative step count");
        if (record) {
            const int64_t need = reverse ? t_first + 1 : t_first + count;
            if (rec_nt < need) ensure_rec(need);
        }
        for (int64_t q = 0; q < count; ++q) {
            const int64_t t = reverse ? t_first - q : t_first + q;
            float *snap = nullptr;
            if (sink && sink->dev && t % sink->stride == 0)
                snap = sink->dev + (size_t)(t / sink->stride - sink->slot0) * nz * nx;
            if (abc == SSSI_ABC_CPML) step_cpml(t, reverse, record, snap);
            else step_sponge(t, reverse, record, snap);
            if (st) st->launches++;
        }
    }

    void get_data(double *data, int64_t nt) override {
        if (nt > rec_nt || !rec.p) fail(SSSI_ESTATE, "get_data: no recorded traces for %lld steps",
                                        (long long)nt);
        std::vector<float> h((size_t)nt * nx);
        SSSI_HIP(hipMemcpyAsync(h.data(), rec.p, h.size() * sizeof(float), hipMemcpyDeviceToHost,
                                stream));
        SSSI_HIP(hipStreamSynchronize(stream));
        for (size_t q = 0; q < h.size(); ++q) data[q] = h[q];  // [t][j] == data(nx,nt)
    }

    // interior of the padded level `which` as column-major (nz, nx) doubles,
    // optionally with the sponge's second damping factor applied
    void download_level(int which, double *out, bool padded_model, bool sponge_g) {
        std::vector<float> h((size_t)pitch * rows);
        SSSI_HIP(hipMemcpyAsync(h.data(), p[which].p, h.size() * sizeof(float),
                                hipMemcpyDeviceToHost, stream));
        SSSI_HIP(hipStreamSynchronize(stream));
        std::vector<double> g;
        if (sponge_g) {
            g.assign(nB, 1.0);
            std::vector<float> gf(std::max(nB, 1));
            SSSI_HIP(hipMemcpy(gf.data(), gtab.p, gtab.bytes(), hipMemcpyDeviceToHost));
            for (int m = 0; m < nB; ++m) g[m] = gf[m];
        }
        auto gfac = [&](int64_t i, int64_t j) -> double {
            if (!sponge_g) return 1.0;
            if (i >= nz - nB) return g[nz - 1 - i];
            if (j < nB) return g[j];
            if (j >= nx - nB) return g[nx - 1 - j];
            return 1.0;
        };
        if (!padded_model) {
            for (int64_t j = 0; j < nx; ++j)
                for (int64_t i = 0; i < nz; ++i)
                    out[j * nz + i] = h[(size_t)(i + kHZ) * pitch + j + kHX] * gfac(i, j);
            return;
        }
        // reference model layout: (nz+2l) x (nx+2l) with l-wide zero halo (CPML)
        // or (nz+2) x (nx+2) (sponge)
        const int64_t l = (abc == SSSI_ABC_CPML) ? 2 * N - 1 : 1;
        const int64_t mz = nz + 2 * l, mx = nx + 2 * l;
        for (int64_t q = 0; q < mz * mx; ++q) out[q] = 0.0;
        for (int64_t j = 0; j < nx; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                double val = h[(size_t)(i + kHZ) * pitch + j + kHX];
                if (sponge_g) {
                    float f = (float)gfac(i, j);
                    val = (double)(f * (float)val);
                }
                out[(j + l) * mz + i + l] = val;
            }
    }

    void get_field(double *field) override { download_level(cur, field, false, false); }
    void get_model(double *model) override {
        download_level(cur ^ 1, model, true, abc == SSSI_ABC_SPONGE);
    }

    void alg_bytes(double *bytes, double *ncells) override {
        // per step: read p, read p_old, write p_new, read v-field (16 B/cell);
        // per damped direction and strip cell: Phi read+write, Psi read+write,
        // b-1 read (20 B); receivers 4 B per column.
        const double c = (double)nz * nx;
        double strips = 0;
        if (abc == SSSI_ABC_CPML) strips = (double)nB * nx + (double)nz * 2 * nB;
        *bytes = 16.0 * c + 20.0 * strips + 4.0 * nx;
        *ncells = c;
    }
};

Prop *make_prop2d(const double *v, int64_t nz, int64_t nx, int abc, int N, int nB, double dz,
                  double dx, double dt) {
    return new Prop2D(v, nz, nx, abc, N, nB, dz, dx, dt);
}

}  // namespace sssi
