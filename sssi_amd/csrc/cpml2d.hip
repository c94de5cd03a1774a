// cpml2d.hip -- 2-D CPML time step for gfx950: the fused form of
// fwdTimeCpmlFor2dAw_mex.c:183-290 / rvsTimeCpmlFor2dAw_mex.c:172-281.
//
// One kernel launch per time step does everything the reference does in six
// derivative passes, eight elementwise passes and two copies:
//   D+p, Phi update, A = D+p + Phi, D-A, Psi update, P = D-A + Psi (z and x),
//   p_new = v^2 dt^2 (P_z + P_x + s) + 2 p - p_old, source injection,
//   receiver gather (top row), and optionally a snapshot, a source-wavefield
//   checkpoint (FP32 or 16-bit with a per-checkpoint scale) or the RTM
//   cross-correlation imaging condition.
// p_new overwrites p_old in place (each cell reads p_old only at itself), so
// the time rotation is a parity flip.  Phi is ping-ponged (a tile also
// computes the N half points on each side that its neighbours own), Psi is
// updated in place.  Phi/Psi and b-1 exist only for the damping strips.
//
// Tiling: a wavefront owns a 64 x RZ tile (lane = column) and never
// synchronises with the other wavefronts of its block (4 tiles stacked in
// z).  All global loads of a tile (z window, x halo, p_old, v^2dt^2) go out
// before the first use, as raw buffer loads with the wave-uniform row offset
// in the scalar offset.  Tiles whose stencil reach touches no damping strip use
// the composite D-D+ stencil of radius L = 2N-1 (equal to the two-stage form
// there, where Phi = Psi = 0); strip tiles run the two-stage CPML update.
// The same tile code runs one job per launch or many jobs (shots) per launch.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "cpml2d.h"

namespace sssi {

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

// LDS written by some lanes of a wavefront and read by others: wavefront-scope
// release/acquire (no workgroup barrier; waves of a block are independent).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint16_t f32_to_16(float x, int fmt) {
    if (fmt == kCkF16) return __half_as_ushort(__float2half_rn(x));
    uint32_t u = __float_as_uint(x);  // bf16, round to nearest even
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float f16_to_32(uint16_t h, int fmt) {
    if (fmt == kCkF16) return __half2float(__ushort_as_half(h));
    return __uint_as_float((uint32_t)h << 16);
}

template <int N, int RZ>
__device__ __forceinline__ void cpml2d_tile(const Job2D &J, int tile, int step) {
    constexpr int L = 2 * N - 1;
    constexpr int PW = kTX + 2 * L;  // p rows: columns x0-L .. x0+63+L
    constexpr int AW = kTX + L;      // A_x rows: half points x0-N .. x0+63+N-1
    constexpr int HN = 2 * L * RZ;   // x-halo values per tile
    constexpr int HPL = (HN + 63) / 64;
    __shared__ float s_p[4][RZ * PW];
    __shared__ float s_a[4][RZ * AW];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntx = J.ntx;
    const int tzi = tile / ntx, txi = tile - tzi * ntx;
    const int x0 = txi * kTX;
    const int nz = J.nz, nx = J.nx, nB = J.nB;
    const int zs = (tzi * 4 + w) * RZ;
    if (zs >= nz) return;
    const int rel = step - J.start_step;
    const int t = J.reverse ? J.nt - 1 - rel : rel;
    const int par = rel & 1;
    const uint32_t rowb = (uint32_t)J.pitch * 4u;
    const uint32_t nbytes = rowb * (uint32_t)J.rows;
    const __amdgpu_buffer_rsrc_t rP = buf_rsrc(J.p[par], nbytes);
    const __amdgpu_buffer_rsrc_t rO = buf_rsrc(J.p[par ^ 1], nbytes);
    const __amdgpu_buffer_rsrc_t rV = buf_rsrc(J.vdt2, nbytes);
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
        hal[u] = buf_ld(rP, (uint32_t)(r * J.pitch + xc + kHX) * 4u, srow);
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
            float acc = J.w0 * col[L + r];
#pragma unroll
            for (int m = 1; m <= L; ++m) acc = fmaf(J.wz[m], col[L + r + m] + col[L + r - m], acc);
#pragma unroll
            for (int m = 1; m <= L; ++m) acc = fmaf(J.wx[m], row[m] + row[-m], acc);
            // the low part of the centre weight, added once the large terms have
            // cancelled, makes the weights sum to zero to FP32 resolution of lap
            lap[r] = fmaf(J.w0lo, col[L + r], acc);
        }
    } else {
        // ---- damping strips: two-stage form with Phi/Psi
        const float *zphi_in = J.zphi[par];
        float *zphi_out = J.zphi[par ^ 1];
        // A_z at half points h = zs-N+q (q < RZ+L): D+p(h) = sum c_k (p(h+1+k) - p(h-k))
        float Az[RZ + L];
#pragma unroll
        for (int q = 0; q < RZ + L; ++q) {
            float d = 0.f;
#pragma unroll
            for (int k = 0; k < N; ++k) d = fmaf(J.cz[k], col[q + N + k] - col[q + N - 1 - k], d);
            Az[q] = d;
        }
        if (zstrip && jin) {
            // all z-strip memory variables of the tile in one batch of loads;
            // (b-1) at half point h and at row i = h share one index
            float phz[RZ + L], bmh[RZ + L], psz[RZ];
#pragma unroll
            for (int q = 0; q < RZ + L; ++q) {
                const int h = zs - N + q;
                const bool in = h >= zb0 && h < nz;
                const size_t o = in ? (size_t)(h - zb0) * nx + j : 0;
                phz[q] = in ? zphi_in[o] : 0.f;
                bmh[q] = in ? J.zbm1[o] : 0.f;
            }
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                const int i = zs + r;
                const bool in = i >= zb0 && i < nz;
                psz[r] = in ? J.zpsi[(size_t)(i - zb0) * nx + j] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < RZ + L; ++q) {
                const int h = zs - N + q;
                if (h >= zb0 && h < nz) {
                    const float phi = fmaf(bmh[q], phz[q] + Az[q], phz[q]);  // b*phi + (b-1)*D
                    Az[q] += phi;
                    if (h >= zs && h < zs + RZ) zphi_out[(size_t)(h - zb0) * nx + j] = phi;
                }
            }
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                float d = 0.f;
#pragma unroll
                for (int k = 0; k < N; ++k) d = fmaf(J.cz[k], Az[r + N + k] - Az[r + N - 1 - k], d);
                const int i = zs + r;
                if (i >= zb0 && i < nz) {
                    const float psi = fmaf(bmh[r + N], psz[r] + d, psz[r]);
                    J.zpsi[(size_t)(i - zb0) * nx + j] = psi;
                    d += psi;
                }
                lap[r] = d;
            }
        } else {
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                float d = 0.f;
#pragma unroll
                for (int k = 0; k < N; ++k) d = fmaf(J.cz[k], Az[r + N + k] - Az[r + N - 1 - k], d);
                lap[r] = d;
            }
        }
        // A_x at half points h = x0-N+q, q < 64+L: lane q, lanes < L also q = 64+lane
        const float *xphi_in = J.xphi[par];
        float *xphi_out = J.xphi[par ^ 1];
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
                phx[half][r] = in ? xphi_in[o] : 0.f;
                bmx[half][r] = in ? J.xbm1[o] : 0.f;
            }
            const bool own = xown && i < nz;
            const size_t o = own ? (size_t)i * nb2 + xs : 0;
            psx[r] = own ? J.xpsi[o] : 0.f;
            bmo[r] = own ? J.xbm1[o] : 0.f;
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
                    d = fmaf(J.cx[k], row[q + N + k] - row[q + N - 1 - k], d);
                const int h = x0 - N + q;
                if (xstrip && i < nz && h >= 0 && h < nx && (h < nB || h >= nx - nB)) {
                    const float phi = fmaf(bmx[half][r], phx[half][r] + d, phx[half][r]);
                    d += phi;
                    if (h >= x0 && h < x0 + kTX)
                        xphi_out[(size_t)i * nb2 + (h < nB ? h : h - (nx - nb2))] = phi;
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
            for (int k = 0; k < N; ++k) d = fmaf(J.cx[k], ar[k] - ar[-1 - k], d);
            const int i = zs + r;
            if (xown && i < nz) {
                const float psi = fmaf(bmo[r], psx[r] + d, psx[r]);
                J.xpsi[(size_t)i * nb2 + xs] = psi;
                d += psi;
            }
            lap[r] += d;
        }
    }

    // ---- sources
    float s[RZ];
#pragma unroll
    for (int r = 0; r < RZ; ++r) s[r] = 0.f;
    if (!J.reverse && J.src_ptr && t < J.src_nt) {
        const int wt = (zs / RZ) * ntx + txi;
        const int k0 = J.src_ptr[wt], k1 = J.src_ptr[wt + 1];
        const float *amp = J.src_amp + (size_t)t * J.nsrc;
        for (int k = k0; k < k1; ++k) {
            const int cell = J.src_cell[k];
            const int ci = cell / nx, cj = cell - ci * nx;
            if (cj == j) {
#pragma unroll
                for (int r = 0; r < RZ; ++r)
                    if (ci == zs + r) s[r] += amp[k];
            }
        }
    }
    if (J.reverse && J.surf && zs == 0 && jin && t < J.surf_nt) {
        const size_t o = (size_t)t * nx + j;
        s[0] += J.surf2 ? J.surf[o] - J.surf2[o] : J.surf[o];
    }

    // ---- update p_new = v^2dt^2 (P_z + P_x + s) + 2p - p_old, and epilogues
    const size_t ncell = (size_t)nz * nx;
    const bool snap_now = J.snap && (t % J.snap_stride == 0);
    const bool ck_now = J.ck_fmt != kCkNone && (t % J.ck_stride == 0);
    const int c = ck_now ? t / J.ck_stride : 0;
    const bool stage_now = ck_now && J.ck_fmt != kCkF32;
    const bool pack_now = stage_now && c > 0;
    const bool im_now = J.im_num && (t % J.im_stride == 0);
    const int ci_im = im_now ? t / J.im_stride : 0;
    float inv_scale = 1.f, im_scale = 1.f;
    if (pack_now) {
        // scale of checkpoint c-1 from its per-wave maxima
        const unsigned *mx = J.ck_max + (size_t)(c - 1) * kCkSlots;
        unsigned m = 0;
        for (int q = 0; q < kCkSlots; ++q) m = max(m, mx[q]);
        const float amax = __uint_as_float(m);
        const float scale = amax > 0.f ? amax * (1.0f / 32768.0f) : 1.0f;
        inv_scale = 1.0f / scale;
        if (tile == 0 && w == 0 && lane == 0) J.ck_scale[c - 1] = scale;
    }
    if (im_now && J.im_fmt != kCkF32) im_scale = J.im_scale[ci_im];
    float wmax = 0.f;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        const int i = zs + r;
        if (jin && i < nz) {
            const float pc = col[r + L];
            // pc + (v2*(...) + (pc - po)): pc - po is exact when the two levels are
            // within a factor 2, which keeps the rounding of the update at the
            // scale of the increment rather than of the field.
            const float pn = pc + fmaf(v2[r], lap[r] + s[r], pc - po[r]);
            buf_st(pn, rO, vcol, srow + r * rowb);
            const size_t cell = (size_t)i * nx + j;
            if (J.rec && i == 0) J.rec[(size_t)t * nx + j] = pn;
            if (snap_now) J.snap[(size_t)(t / J.snap_stride - J.snap_slot0) * ncell + cell] = pn;
            if (ck_now) {
                if (J.ck_fmt == kCkF32) {
                    J.ck32[(size_t)c * ncell + cell] = pn;
                } else {
                    J.ck_stage[(size_t)(c & 1) * ncell + cell] = pn;
                    wmax = fmaxf(wmax, fabsf(pn));
                }
            }
            if (pack_now) {
                const float prev = J.ck_stage[(size_t)((c - 1) & 1) * ncell + cell];
                J.ck16[(size_t)(c - 1) * ncell + cell] = f32_to_16(prev * inv_scale, J.ck_fmt);
            }
            if (im_now) {
                const size_t o = (size_t)ci_im * ncell + cell;
                const float f = J.im_fmt == kCkF32 ? J.im_ck32[o]
                                                   : f16_to_32(J.im_ck16[o], J.im_fmt) * im_scale;
                J.im_num[cell] = fmaf(f, pn, J.im_num[cell]);
                J.im_den[cell] = fmaf(f, f, J.im_den[cell]);
            }
        }
    }
    if (stage_now) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, off, 64));
        if (lane == 0)
            atomicMax(J.ck_max + (size_t)c * kCkSlots + ((tile * 4 + w) & (kCkSlots - 1)),
                      __float_as_uint(wmax));
    }
}

template <int N, int RZ>
__global__ void __launch_bounds__(256) cpml2d_single(Job2D J, const int2 *__restrict__ order,
                                                     int step) {
    cpml2d_tile<N, RZ>(J, order[blockIdx.x].y, step);
}

template <int N, int RZ>
__global__ void __launch_bounds__(256) cpml2d_batch(const Job2D *__restrict__ jobs,
                                                    const int2 *__restrict__ order, int step) {
    const int2 jt = order[blockIdx.x];
    cpml2d_tile<N, RZ>(jobs[jt.x], jt.y, step);
}

// Pack a staged FP32 checkpoint into 16 bits after the job's last step.
__global__ void ck_pack(float *stage, uint16_t *out, const unsigned *mx, float *scale_out,
                        size_t n, int fmt) {
    unsigned m = 0;
    for (int q = 0; q < kCkSlots; ++q) m = max(m, mx[q]);
    const float amax = __uint_as_float(m);
    const float scale = amax > 0.f ? amax * (1.0f / 32768.0f) : 1.0f;
    const float inv = 1.0f / scale;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *scale_out = scale;
    if (i < n) out[i] = f32_to_16(stage[i] * inv, fmt);
}

// ---------------------------------------------------------------------------
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

Cpml2dGrid::Cpml2dGrid(const double *v, int64_t nz_, int64_t nx_, int N_, int nB_, double dz_,
                       double dx_, double dt_, int rz_)
    : nz(nz_), nx(nx_), N(N_), nB(nB_), dz(dz_), dx(dx_), dt(dt_) {
    rz = 4;
    if (rz_ == 4 || rz_ == 8) rz = rz_;
    else if (const char *e = std::getenv("SSSI_RZ")) {
        const int r = std::atoi(e);
        if (r == 4 || r == 8) rz = r;
    }
    pitch = round_up(kHX + round_up(nx, kTX) + 32, 64);
    rows = kHZ + round_up(nz, kTZmax) + kHZbot;
    const size_t np = (size_t)pitch * rows;
    if (np * 4 >= ((size_t)1 << 32)) fail(SSSI_EINVAL, "2-D padded grid exceeds 4 GiB");
    p[0].alloc(np);
    p[1].alloc(np);
    vdt2.alloc(np);
    std::vector<float> hv(np, 0.f);
    for (int64_t j = 0; j < nx; ++j)
        for (int64_t i = 0; i < nz; ++i) {
            const double vv = v[j * nz + i];
            hv[(size_t)(i + kHZ) * pitch + j + kHX] = (float)((vv * dt) * (vv * dt));
        }
    SSSI_HIP(hipMemcpy(vdt2.p, hv.data(), np * sizeof(float), hipMemcpyHostToDevice));
    double c[8];
    staggered_coefficients(N, c);
    for (int k = 0; k < N; ++k) {
        cz[k] = (float)(c[k] / dz);
        cx[k] = (float)(c[k] / dx);
    }
    double w[32];
    composite_weights(N, c, w);
    const int L = 2 * N - 1;
    double side = 0.0;  // exact in double: a sum of 4L floats
    for (int m = 1; m <= L; ++m) {
        wz[m] = (float)(w[L + m] / (dz * dz));
        wx[m] = (float)(w[L + m] / (dx * dx));
        side += 2.0 * ((double)wz[m] + (double)wx[m]);
    }
    // The centre weight is chosen so that the FP32 weights sum to zero (the
    // Laplacian of a constant vanishes, as in the two-stage form where
    // D+ of a constant is exactly 0); it is carried as w0 + w0lo.
    w0 = (float)(-side);
    w0lo = (float)(-side - (double)w0);
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
}

bool Cpml2dGrid::block_is_strip(int tile) const {
    if (nB == 0) return false;
    const int tzi = tile / ntx(), txi = tile % ntx();
    const int64_t z0 = (int64_t)tzi * tz(), x0 = (int64_t)txi * kTX;
    const bool zs = z0 + tz() + N - 1 > nz - nB;
    const bool xs = x0 - N < nB || x0 + kTX + N - 1 > nx - nB;
    return zs || xs;
}

void Cpml2dGrid::zero_state(hipStream_t s) {
    for (auto *b : {&p[0], &p[1], &zphi[0], &zphi[1], &zpsi, &xphi[0], &xphi[1], &xpsi})
        if (b->p) SSSI_HIP(hipMemsetAsync(b->p, 0, b->bytes(), s));
}

Job2D Cpml2dGrid::job() const {
    Job2D J;
    std::memset(&J, 0, sizeof(J));
    J.nz = (int)nz; J.nx = (int)nx; J.pitch = (int)pitch; J.rows = (int)rows;
    J.ntx = ntx(); J.nwz = nwz(); J.nB = nB;
    for (int k = 0; k < 8; ++k) { J.cz[k] = cz[k]; J.cx[k] = cx[k]; }
    J.w0 = w0;
    J.w0lo = w0lo;
    for (int m = 0; m < 16; ++m) { J.wz[m] = wz[m]; J.wx[m] = wx[m]; }
    J.p[0] = p[0].p; J.p[1] = p[1].p;
    J.vdt2 = vdt2.p;
    J.zphi[0] = zphi[0].p; J.zphi[1] = zphi[1].p; J.zpsi = zpsi.p; J.zbm1 = zbm1.p;
    J.xphi[0] = xphi[0].p; J.xphi[1] = xphi[1].p; J.xpsi = xpsi.p; J.xbm1 = xbm1.p;
    J.snap_stride = 1;
    J.ck_stride = 1;
    J.im_stride = 1;
    return J;
}

void Cpml2dGrid::download(int which, double *out, hipStream_t s) const {
    std::vector<float> h((size_t)pitch * rows);
    SSSI_HIP(hipMemcpyAsync(h.data(), p[which].p, h.size() * sizeof(float), hipMemcpyDeviceToHost,
                            s));
    SSSI_HIP(hipStreamSynchronize(s));
    for (int64_t j = 0; j < nx; ++j)
        for (int64_t i = 0; i < nz; ++i) out[j * nz + i] = h[(size_t)(i + kHZ) * pitch + j + kHX];
}

double Cpml2dGrid::alg_bytes_per_step() const {
    // read p, p_old, v^2dt^2, write p_new (16 B/cell); per damped-direction strip
    // cell Phi read+write, Psi read+write, b-1 read (20 B); receivers 4 B/column
    const double strips = (double)nB * nx + (double)nz * 2 * nB;
    return 16.0 * (double)nz * nx + 20.0 * strips + 4.0 * nx;
}

void DevSources::set(const Cpml2dGrid &g, int n_, const int *iz, const int *ix,
                     const double *wavelet, int64_t nt_) {
    if (n_ < 0 || nt_ < 1 || (n_ > 0 && (!iz || !ix || !wavelet)))
        fail(SSSI_EINVAL, "sources: bad arguments");
    for (int k = 0; k < n_; ++k)
        if (iz[k] < 0 || iz[k] >= g.nz || ix[k] < 0 || ix[k] >= g.nx)
            fail(SSSI_EINVAL, "source %d at (%d,%d) outside the %lldx%lld grid", k, iz[k], ix[k],
                 (long long)g.nz, (long long)g.nx);
    n = n_;
    nt = (int)nt_;
    if (n == 0) {
        ptr.release(); cell.release(); amp.release();
        return;
    }
    const int ntx = g.ntx(), rz = g.rz;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    auto tile_of = [&](int k) { return (iz[k] / rz) * ntx + ix[k] / kTX; };
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return tile_of(a) < tile_of(b); });
    const int ntiles = ntx * g.nwz();
    std::vector<int> hp(ntiles + 1, 0), hc(n);
    for (int k = 0; k < n; ++k) {
        hp[tile_of(order[k]) + 1]++;
        hc[k] = iz[order[k]] * (int)g.nx + ix[order[k]];
    }
    for (int t = 0; t < ntiles; ++t) hp[t + 1] += hp[t];
    std::vector<float> ha((size_t)nt * n);
    for (int64_t t = 0; t < nt; ++t)
        for (int k = 0; k < n; ++k) ha[(size_t)t * n + k] = (float)wavelet[t * n + order[k]];
    ptr.alloc(hp.size());
    cell.alloc(n);
    amp.alloc(ha.size());
    SSSI_HIP(hipMemcpy(ptr.p, hp.data(), ptr.bytes(), hipMemcpyHostToDevice));
    SSSI_HIP(hipMemcpy(cell.p, hc.data(), cell.bytes(), hipMemcpyHostToDevice));
    SSSI_HIP(hipMemcpy(amp.p, ha.data(), amp.bytes(), hipMemcpyHostToDevice));
}

void build_block_order(const std::vector<const Cpml2dGrid *> &grids, std::vector<int2> &order) {
    std::vector<int2> strip, inner;
    for (int g = 0; g < (int)grids.size(); ++g)
        for (int t = 0; t < grids[g]->nblocks(); ++t)
            (grids[g]->block_is_strip(t) ? strip : inner).push_back(make_int2(g, t));
    order = strip;
    const size_t ni = inner.size(), total = strip.size() + inner.size();
    size_t next[8], end[8];
    for (int x = 0; x < 8; ++x) {
        next[x] = ni * x / 8;
        end[x] = ni * (x + 1) / 8;
    }
    for (size_t b = order.size(); b < total; ++b) {
        int x = (int)(b % 8);
        if (next[x] == end[x])
            for (int y = 0; y < 8; ++y)
                if (next[y] < end[y]) { x = y; break; }
        order.push_back(inner[next[x]++]);
    }
}

#define SSSI_NRZ_SWITCH(N, RZ, CALL)                                                       \
    switch ((N) * 16 + (RZ)) {                                                            \
        case 1 * 16 + 4: CALL(1, 4); break;  case 1 * 16 + 8: CALL(1, 8); break;            \
        case 2 * 16 + 4: CALL(2, 4); break;  case 2 * 16 + 8: CALL(2, 8); break;            \
        case 3 * 16 + 4: CALL(3, 4); break;  case 3 * 16 + 8: CALL(3, 8); break;            \
        case 4 * 16 + 4: CALL(4, 4); break;  case 4 * 16 + 8: CALL(4, 8); break;            \
        case 5 * 16 + 4: CALL(5, 4); break;  case 5 * 16 + 8: CALL(5, 8); break;            \
        case 6 * 16 + 4: CALL(6, 4); break;  case 6 * 16 + 8: CALL(6, 8); break;            \
        case 7 * 16 + 4: CALL(7, 4); break;  case 7 * 16 + 8: CALL(7, 8); break;            \
        case 8 * 16 + 4: CALL(8, 4); break;  case 8 * 16 + 8: CALL(8, 8); break;            \
        default: fail(SSSI_EINVAL, "no 2-D kernel for nDiffOrder=%d rz=%d", (N), (RZ));    \
    }

void launch_cpml2d_single(const Cpml2dGrid &g, const Job2D &job, const int2 *order, int nblk,
                          int step, hipStream_t s) {
#define SSSI_L1(NN, R)                                                                    \
    hipLaunchKernelGGL((cpml2d_single<NN, R>), dim3(nblk), dim3(256), 0, s, job, order, step)
    SSSI_NRZ_SWITCH(g.N, g.rz, SSSI_L1)
#undef SSSI_L1
    SSSI_HIP(hipGetLastError());
}

void launch_cpml2d_batch(int N, int rz, const Job2D *jobs, const int2 *order, int nblk, int step,
                         hipStream_t s) {
#define SSSI_LB(NN, R)                                                                    \
    hipLaunchKernelGGL((cpml2d_batch<NN, R>), dim3(nblk), dim3(256), 0, s, jobs, order, step)
    SSSI_NRZ_SWITCH(N, rz, SSSI_LB)
#undef SSSI_LB
    SSSI_HIP(hipGetLastError());
}

void launch_ck_pack(const Job2D &job, int c, hipStream_t s) {
    const size_t n = (size_t)job.nz * job.nx;
    hipLaunchKernelGGL(ck_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       job.ck_stage + (size_t)(c & 1) * n, job.ck16 + (size_t)c * n,
                       job.ck_max + (size_t)c * kCkSlots, job.ck_scale + c, n, job.ck_fmt);
    SSSI_HIP(hipGetLastError());
}

}  // namespace sssi
