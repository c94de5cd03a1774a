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
#include "devbuf.h"

namespace sssi {

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

// Wave-uniform quantities of one time step of one job.
//
// 16-bit checkpoints carry one FP32 scale each.  BF16 has FP32's exponent
// range, so its checkpoint c is written directly in its own step, unscaled
// (scale 1: a scale would not change BF16's relative precision) -- one 2-byte
// store per cell and no |p| maximum.  FP16 would overflow a field that grows
// past twice the previous maximum, so its checkpoint c is staged in FP32 and
// packed during the next step (or after the last one) with its own maximum.
struct StepCtx {
    int t, par;
    uint32_t rowb, nbytes;
    bool snap_now, ck_now, direct_now, stage_now, pack_now, im_now;
    int c, ci_im;
    float scale, inv_scale, im_scale;
};

__device__ __forceinline__ float ck_scale_of(const Job2D &J, int c) {
    const unsigned *mx = J.ck_max + (size_t)c * kCkSlots;
    unsigned m = 0;
    for (int q = 0; q < kCkSlots; ++q) m = max(m, ld_uniform(mx + q));
    const float amax = __uint_as_float(m);
    return amax > 0.f ? amax * (1.0f / 32768.0f) : 1.0f;
}

__device__ __forceinline__ StepCtx step_ctx(const Job2D &J, int step) {
    StepCtx S;
    const int rel = step - J.start_step;
    S.t = J.reverse ? J.nt - 1 - rel : rel;
    S.par = rel & 1;
    S.rowb = (uint32_t)J.pitch * 4u;
    S.nbytes = S.rowb * (uint32_t)J.rows;
    S.snap_now = J.snap && (S.t % J.snap_stride == 0);
    S.ck_now = J.ck_fmt != kCkNone && (S.t % J.ck_stride == 0);
    S.c = S.ck_now ? S.t / J.ck_stride : 0;
    S.direct_now = S.ck_now && J.ck_fmt == kCkBF16;
    S.stage_now = S.ck_now && J.ck_fmt == kCkF16;
    S.pack_now = S.stage_now && S.c > 0;
    S.im_now = J.im_num && (S.t % J.im_stride == 0);
    S.ci_im = S.im_now ? S.t / J.im_stride : 0;
    S.scale = S.inv_scale = S.im_scale = 1.f;
    if (S.pack_now) {
        // checkpoint c-1's own scale
        S.scale = ck_scale_of(J, S.c - 1);
        S.inv_scale = 1.0f / S.scale;
    }
    if (S.im_now && J.im_fmt != kCkF32) S.im_scale = ld_uniform(J.im_scale + S.ci_im);
    return S;
}

// Source terms of the RZ rows zs.. of column j: forward point sources from the
// per-wave-tile CSR, reverse surface injection at row 0.
template <int RZ>
__device__ __forceinline__ void tile_sources(const Job2D &J, int t, int zs, int txi, int j,
                                             bool jin, float (&s)[RZ]) {
#pragma unroll
    for (int r = 0; r < RZ; ++r) s[r] = 0.f;
    if (!J.reverse && J.src_ptr && t < J.src_nt) {
        const int wt = (zs / RZ) * J.ntx + txi;
        const int k0 = ld_uniform(J.src_ptr + wt), k1 = ld_uniform(J.src_ptr + wt + 1);
        const float *amp = J.src_amp + (size_t)t * J.nsrc;
        for (int k = k0; k < k1; ++k) {
            const int cell = ld_uniform(J.src_cell + k);
            const int ci = cell / J.nx, cj = cell - ci * J.nx;
            const float a = ld_uniform(amp + k);
            if (cj == j) {
#pragma unroll
                for (int r = 0; r < RZ; ++r)
                    if (ci == zs + r) s[r] += a;
            }
        }
    }
    if (J.reverse && J.surf && zs == 0 && jin && t < J.surf_nt) {
        const size_t o = (size_t)t * J.nx + j;
        s[0] += J.surf2 ? J.surf[o] - J.surf2[o] : J.surf[o];
    }
}

// p_new = v^2dt^2 (P_z + P_x + s) + 2p - p_old for the RZ rows zs.. of column
// j, stored over p_old, then receivers, snapshot, checkpoint and imaging.
template <int RZ>
__device__ __forceinline__ void tile_update(const Job2D &J, const StepCtx &S,
                                            __amdgpu_buffer_rsrc_t rO, uint32_t vcol,
                                            uint32_t srow, int zs, int j, bool jin,
                                            const float (&pc)[RZ], const float (&po)[RZ],
                                            const float (&v2)[RZ], const float (&lap)[RZ],
                                            const float (&s)[RZ], float &wmax, float &pn0) {
    const size_t ncell = (size_t)J.nz * J.nx;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        const int i = zs + r;
        if (jin && i < J.nz) {
            // pc + (v2*(...) + (pc - po)): pc - po is exact when the two levels are
            // within a factor 2, which keeps the rounding of the update at the
            // scale of the increment rather than of the field.
            const float pn = pc[r] + fmaf(v2[r], lap[r] + s[r], pc[r] - po[r]);
            buf_st(pn, rO, vcol, srow + r * S.rowb);
            if (r == 0) pn0 = pn;
            const size_t cell = (size_t)i * J.nx + j;
            if (J.rec && i == 0) J.rec[(size_t)S.t * J.nx + j] = pn;
            if (S.snap_now)
                J.snap[(size_t)(S.t / J.snap_stride - J.snap_slot0) * ncell + cell] = pn;
            if (S.ck_now) {
                if (J.ck_fmt == kCkF32) {
                    J.ck32[(size_t)S.c * ncell + cell] = pn;
                } else if (S.direct_now) {
                    J.ck16[(size_t)S.c * ncell + cell] = f32_to_16(pn, J.ck_fmt);
                } else {
                    J.ck_stage[(size_t)(S.c & 1) * ncell + cell] = pn;
                    wmax = fmaxf(wmax, fabsf(pn));
                }
            }
            if (S.pack_now) {
                const float prev = J.ck_stage[(size_t)((S.c - 1) & 1) * ncell + cell];
                J.ck16[(size_t)(S.c - 1) * ncell + cell] = f32_to_16(prev * S.inv_scale, J.ck_fmt);
            }
            if (S.im_now) {
                const size_t o = (size_t)S.ci_im * ncell + cell;
                const float f = J.im_fmt == kCkF32
                                    ? J.im_ck32[o]
                                    : f16_to_32(J.im_ck16[o], J.im_fmt) * S.im_scale;
                J.im_num[cell] = fmaf(f, pn, J.im_num[cell]);
                J.im_den[cell] = fmaf(f, f, J.im_den[cell]);
            }
        }
    }
}

// Materialise values here: the compiler would otherwise sink their arithmetic
// past the branches of the source and epilogue code and keep every operand
// (the LDS neighbourhood of all rows) live across them.
template <int R>
__device__ __forceinline__ void pin(float (&v)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r) asm volatile("" : "+v"(v[r]));
}

// Per-wave |p| maximum of a staged FP16 checkpoint into one of the kCkSlots slots.
__device__ __forceinline__ void ck_max_commit(const Job2D &J, const StepCtx &S, float wmax,
                                              int slot, int lane) {
    if (!S.stage_now) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, off, 64));
    if (lane == 0)
        atomicMax(J.ck_max + (size_t)S.c * kCkSlots + (slot & (kCkSlots - 1)),
                  __float_as_uint(wmax));
}

// Composite D-D+ Laplacian of RZ rows: the z window win[OFF + r] holds row
// zs - L + r (r < RZ + 2L), the wave's LDS rows sp[r * PW + L + lane] the x
// neighbourhood of row zs + r.  OFF is a constant after unrolling.
template <int N, int RZ, int W>
__device__ __forceinline__ void interior_lap(const Job2D &J, const float (&win)[W], int OFF,
                                             const float *sp, int lane, float (&lap)[RZ]) {
    constexpr int L = 2 * N - 1;
    constexpr int PW = kTX + 2 * L;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        const float *row = sp + r * PW + L + lane;
        float acc = J.w0 * win[OFF + L + r];
#pragma unroll
        for (int m = 1; m <= L; ++m)
            acc = fmaf(J.wz[m], win[OFF + L + r + m] + win[OFF + L + r - m], acc);
#pragma unroll
        for (int m = 1; m <= L; ++m) acc = fmaf(J.wx[m], row[m] + row[-m], acc);
        // the low part of the centre weight, added once the large terms have
        // cancelled, makes the weights sum to zero to FP32 resolution of lap
        lap[r] = fmaf(J.w0lo, win[OFF + L + r], acc);
    }
    pin(lap);
}

// The same Laplacian with the x neighbourhood exchanged inside the wavefront
// instead of through LDS: a chain of DPP wave shifts by one lane (wave_shr:1
// / wave_shl:1) walks each row's value m = 1 .. L columns left and right, and
// the lane with no source lane (lane 0 / lane 63) keeps the DPP's old operand,
// the x-halo value of that column broadcast from the lane of the halo
// registers that loaded it (readlane).  Same operands in the same FMA order
// as interior_lap, so the results are identical bit for bit.
__device__ __forceinline__ float shr1_in(float v, float from, int src_lane) {
    const int h = __builtin_amdgcn_readlane(__float_as_int(from), src_lane);
    return __int_as_float(
        __builtin_amdgcn_update_dpp(h, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float shl1_in(float v, float from, int src_lane) {
    const int h = __builtin_amdgcn_readlane(__float_as_int(from), src_lane);
    return __int_as_float(
        __builtin_amdgcn_update_dpp(h, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
template <int N, int RZ, int W, int HPL>
__device__ __forceinline__ void interior_lap_dpp(const Job2D &J, const float (&win)[W], int OFF,
                                                 const float (&hal)[HPL], float (&lap)[RZ]) {
    constexpr int L = 2 * N - 1;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        const float v = win[OFF + L + r];
        float acc = J.w0 * v;
#pragma unroll
        for (int m = 1; m <= L; ++m)
            acc = fmaf(J.wz[m], win[OFF + L + r + m] + win[OFF + L + r - m], acc);
        float sl = v, sr = v;
#pragma unroll
        for (int m = 1; m <= L; ++m) {
            // halo element e = lane + 64u holds row e / 2L, slot e % 2L: slot L-m is
            // column x0-m, slot L+m-1 column x0+63+m
            const int el = r * 2 * L + L - m, er = r * 2 * L + L + m - 1;
            sl = shr1_in(sl, hal[el / 64], el % 64);
            sr = shl1_in(sr, hal[er / 64], er % 64);
            acc = fmaf(J.wx[m], sr + sl, acc);
        }
        lap[r] = fmaf(J.w0lo, v, acc);
    }
    pin(lap);
}

// The same Laplacian with the x neighbourhood read as loads shifted by m
// columns (the sponge step's layout: no LDS, no halo exchange); the row zs+r
// is at byte row offset srow + r * rowb of the padded field, whose 32 padding
// columns each side keep every shifted load in bounds.  Same operands in the
// same FMA order as interior_lap.
template <int N, int RZ, int W>
__device__ __forceinline__ void interior_lap_glb(const Job2D &J, const float (&win)[W], int OFF,
                                                 __amdgpu_buffer_rsrc_t rP, uint32_t vcol,
                                                 uint32_t srow, uint32_t rowb,
                                                 float (&lap)[RZ]) {
    constexpr int L = 2 * N - 1;
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        const uint32_t ro = srow + r * rowb;
        float xl[L + 1], xr[L + 1];
#pragma unroll
        for (int m = 1; m <= L; ++m) {
            xl[m] = buf_ld(rP, vcol - 4u * m, ro);
            xr[m] = buf_ld(rP, vcol + 4u * m, ro);
        }
        float acc = J.w0 * win[OFF + L + r];
#pragma unroll
        for (int m = 1; m <= L; ++m)
            acc = fmaf(J.wz[m], win[OFF + L + r + m] + win[OFF + L + r - m], acc);
#pragma unroll
        for (int m = 1; m <= L; ++m) acc = fmaf(J.wx[m], xr[m] + xl[m], acc);
        lap[r] = fmaf(J.w0lo, win[OFF + L + r], acc);
    }
    pin(lap);
}

// x halo of RZ rows (L columns each side of the 64-wide tile): lane e of
// load u fetches value e + 64u; h_lds is its LDS slot, or for lanes without a
// value the lane's word of the wave's spare LDS row (index dummy + lane), so
// the LDS write is unconditional and the compiler cannot sink the load into a
// branch, where waiting for it would drain every load issued before.
template <int N, int RZ>
__device__ __forceinline__ void halo_slots(int lane, int (&h_lds)[(2 * (2 * N - 1) * RZ + 63) / 64],
                                           uint32_t (&h_off)[(2 * (2 * N - 1) * RZ + 63) / 64],
                                           int x0, int pitch, int dummy) {
    constexpr int L = 2 * N - 1;
    constexpr int PW = kTX + 2 * L;
    constexpr int HN = 2 * L * RZ;
    constexpr int HPL = (HN + 63) / 64;
#pragma unroll
    for (int u = 0; u < HPL; ++u) {
        const int e = lane + 64 * u;
        const int r = e / (2 * L), c = e - r * (2 * L);
        const int xc = c < L ? x0 - L + c : x0 + kTX - L + c;
        h_lds[u] = e < HN ? r * PW + (c < L ? c : kTX + c) : dummy + lane;
        h_off[u] = e < HN ? (uint32_t)(r * pitch + xc + kHX) * 4u : (uint32_t)(x0 + kHX) * 4u;
    }
}

// LDS of a 4-wave block: RZ rows of PW floats per wave, then one spare row of
// 64 floats per wave (halo_slots' sink for lanes without a halo value).
template <int N, int RZ>
constexpr int lds_p_floats() { return 4 * RZ * (kTX + 2 * (2 * N - 1)) + 4 * kTX; }
// index of wave w's spare row relative to its rows at lds_p + w * RZ * PW
template <int N, int RZ>
__device__ __forceinline__ int spare_row(int w) {
    constexpr int PW = kTX + 2 * (2 * N - 1);
    return (4 - w) * RZ * PW + w * kTX;
}

template <int N, int RZ>
__device__ __forceinline__ void cpml2d_tile(const Job2D &J, int tile, int step, float *lds_p,
                                            float *lds_a) {
    constexpr int L = 2 * N - 1;
    constexpr int PW = kTX + 2 * L;  // p rows: columns x0-L .. x0+63+L
    constexpr int AW = kTX + L;      // A_x rows: half points x0-N .. x0+63+N-1
    constexpr int HN = 2 * L * RZ;   // x-halo values per tile
    constexpr int HPL = (HN + 63) / 64;

    const int lane = threadIdx.x & 63, w = wave_id();
    const int ntx = J.ntx;
    const int tzi = tile / ntx, txi = tile - tzi * ntx;
    const int x0 = txi * kTX;
    const int nz = J.nz, nx = J.nx, nB = J.nB;
    const int zs = (tzi * 4 + w) * RZ;
    if (zs >= nz) return;
    const StepCtx S = step_ctx(J, step);
    const int t = S.t, par = S.par;
    const uint32_t rowb = S.rowb;
    const __amdgpu_buffer_rsrc_t rP = buf_rsrc(J.p[par], S.nbytes);
    const __amdgpu_buffer_rsrc_t rO = buf_rsrc(J.p[par ^ 1], S.nbytes);
    const __amdgpu_buffer_rsrc_t rV = buf_rsrc(J.vdt2, S.nbytes);
    const int j = x0 + lane;
    const bool jin = j < nx;
    const uint32_t vcol = (uint32_t)(j + kHX) * 4u;
    const uint32_t srow = (uint32_t)(zs + kHZ) * rowb;  // first row of the tile

    // ---- issue every global load of the tile: the x halo first (vmcnt counts
    // in issue order, so its LDS exchange then waits for it alone), then the
    // z window, then p_old and v^2dt^2
    float col[RZ + 2 * L], po[RZ], v2[RZ], hal[HPL];
    int h_lds[HPL];
    uint32_t h_off[HPL];
    halo_slots<N, RZ>(lane, h_lds, h_off, x0, J.pitch, spare_row<N, RZ>(w));
#pragma unroll
    for (int u = 0; u < HPL; ++u) hal[u] = buf_ld(rP, h_off[u], srow);
#pragma unroll
    for (int r = 0; r < RZ + 2 * L; ++r) col[r] = buf_ld(rP, vcol, srow + (r - L) * rowb);
#pragma unroll
    for (int r = 0; r < RZ; ++r) {
        po[r] = buf_ld(rO, vcol, srow + r * rowb);
        v2[r] = buf_ld(rV, vcol, srow + r * rowb);
    }

    // ---- the tile's rows with an L-wide x halo into the wave's LDS rows
    float *sp = lds_p + w * (RZ * PW);
#pragma unroll
    for (int u = 0; u < HPL; ++u) sp[h_lds[u]] = hal[u];
#pragma unroll
    for (int r = 0; r < RZ; ++r) sp[r * PW + L + lane] = col[L + r];
    wave_lds_sync();

    const int zb0 = nz - nB;
    const bool zstrip = nB > 0 && (zs + RZ + N - 1 > zb0);
    const bool xstrip = nB > 0 && (x0 - N < nB || x0 + kTX + N - 1 > nx - nB);
    float lap[RZ];
    if (!zstrip && !xstrip) {
        interior_lap<N, RZ>(J, col, 0, sp, lane, lap);
    } else {
        // ---- damping strips.  Each direction takes the two-stage form with its
        // memory variables where its strip is within stencil reach, and its
        // composite stencil otherwise (equal there, where Phi = Psi = 0).
        // ---- z: two-stage CPML form where a z strip is within reach
        if (zstrip) {
            const uint32_t zbytes = (uint32_t)nB * (uint32_t)nx * 4u;
            const __amdgpu_buffer_rsrc_t rZi = buf_rsrc(J.zphi[par], zbytes);
            const __amdgpu_buffer_rsrc_t rZo = buf_rsrc(J.zphi[par ^ 1], zbytes);
            const __amdgpu_buffer_rsrc_t rZb = buf_rsrc(J.zbm1, zbytes);
            const __amdgpu_buffer_rsrc_t rZs = buf_rsrc(J.zpsi, zbytes);
            // byte offset of this lane's value in the strip row of depth h
            // (wave-uniform test), kOOB where there is none
            const uint32_t zcol = jin ? (uint32_t)j * 4u : kOOB;
            const uint32_t nxb = (uint32_t)nx * 4u;
            auto zoff = [&](int h) -> uint32_t {
                return h >= zb0 && h < nz ? zcol + (uint32_t)(h - zb0) * nxb : kOOB;
            };
            // A_z at half points h = zs-N+q (q < RZ+L): D+p(h) = sum c_k (p(h+1+k) - p(h-k))
            float Az[RZ + L];
#pragma unroll
            for (int q = 0; q < RZ + L; ++q) {
                float d = 0.f;
#pragma unroll
                for (int k = 0; k < N; ++k) d = fmaf(J.cz[k], col[q + N + k] - col[q + N - 1 - k], d);
                Az[q] = d;
            }
            // all z-strip memory variables of the tile in one batch of loads
            // (zero for lanes outside the model); (b-1) at half point h and at
            // row i = h share one index
            float phz[RZ + L], bmh[RZ + L], psz[RZ];
#pragma unroll
            for (int q = 0; q < RZ + L; ++q) {
                const uint32_t o = zoff(zs - N + q);
                phz[q] = buf_ld(rZi, o, 0);
                bmh[q] = buf_ld(rZb, o, 0);
            }
#pragma unroll
            for (int r = 0; r < RZ; ++r) psz[r] = buf_ld(rZs, zoff(zs + r), 0);
#pragma unroll
            for (int q = 0; q < RZ + L; ++q) {
                const int h = zs - N + q;
                if (h >= zb0 && h < nz) {
                    const float phi = fmaf(bmh[q], phz[q] + Az[q], phz[q]);  // b*phi + (b-1)*D
                    Az[q] += phi;
                    if (h >= zs && h < zs + RZ) buf_st(phi, rZo, zoff(h), 0);
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
                    buf_st(psi, rZs, zoff(i), 0);
                    d += psi;
                }
                lap[r] = d;
            }
        } else {
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                float acc = J.w0z * col[L + r];
#pragma unroll
                for (int m = 1; m <= L; ++m)
                    acc = fmaf(J.wz[m], col[L + r + m] + col[L + r - m], acc);
                lap[r] = fmaf(J.w0zlo, col[L + r], acc);
            }
        }
        // ---- x: two-stage CPML form where an x strip is within reach
        if (xstrip) {
            // A_x at half points h = x0-N+q, q < 64+L: lane q, lanes < L also q = 64+lane
            const int nb2 = 2 * nB;
            const uint32_t xbytes = (uint32_t)nz * (uint32_t)nb2 * 4u;
            const __amdgpu_buffer_rsrc_t rXi = buf_rsrc(J.xphi[par], xbytes);
            const __amdgpu_buffer_rsrc_t rXo = buf_rsrc(J.xphi[par ^ 1], xbytes);
            const __amdgpu_buffer_rsrc_t rXb = buf_rsrc(J.xbm1, xbytes);
            const __amdgpu_buffer_rsrc_t rXs = buf_rsrc(J.xpsi, xbytes);
            float *sa = lds_a + w * (RZ * AW);
            // byte offset within a strip row of column h, kOOB if h lies in no
            // x strip; a row adds i * nb2 * 4 (< 2^31), which keeps kOOB out of range
            auto xcol = [&](int h) -> uint32_t {
                return h >= 0 && h < nx && (h < nB || h >= nx - nB)
                           ? (uint32_t)(h < nB ? h : h - (nx - nb2)) * 4u
                           : kOOB;
            };
            uint32_t xc[2];
            xc[0] = xcol(x0 - N + lane);
            xc[1] = lane < L ? xcol(x0 - N + kTX + lane) : kOOB;
            // half points this lane owns (Phi written by this tile): x0 .. x0+63
            const bool hown[2] = {lane >= N, lane < N};
            const bool xown = jin && (j < nB || j >= nx - nB);
            const uint32_t xo = xown ? xcol(j) : kOOB;
            const uint32_t rb2 = (uint32_t)nb2 * 4u;
            // all x-strip memory variables of the tile in one batch of loads
            float phx[2][RZ], bmx[2][RZ], psx[RZ], bmo[RZ];
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                const int i = zs + r;
                const bool rin = i < nz;
                const uint32_t ro = (uint32_t)i * rb2;
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const uint32_t o = rin ? xc[half] + ro : kOOB;
                    phx[half][r] = buf_ld(rXi, o, 0);
                    bmx[half][r] = buf_ld(rXb, o, 0);
                }
                const uint32_t o = rin ? xo + ro : kOOB;
                psx[r] = buf_ld(rXs, o, 0);
                bmo[r] = buf_ld(rXb, o, 0);
            }
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                const float *row = sp + r * PW;
                const int i = zs + r;
                const uint32_t ro = (uint32_t)i * rb2;
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const int q = lane + half * kTX;
                    if (half == 1 && lane >= L) break;
                    float d = 0.f;
#pragma unroll
                    for (int k = 0; k < N; ++k)
                        d = fmaf(J.cx[k], row[q + N + k] - row[q + N - 1 - k], d);
                    if (i < nz && xc[half] != kOOB) {
                        const float phi = fmaf(bmx[half][r], phx[half][r] + d, phx[half][r]);
                        d += phi;
                        if (hown[half]) buf_st(phi, rXo, xc[half] + ro, 0);
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
                    buf_st(psi, rXs, xo + (uint32_t)i * rb2, 0);
                    d += psi;
                }
                lap[r] += d;
            }
        } else {
#pragma unroll
            for (int r = 0; r < RZ; ++r) {
                const float *row = sp + r * PW + L + lane;
                float acc = J.w0x * col[L + r];
#pragma unroll
                for (int m = 1; m <= L; ++m) acc = fmaf(J.wx[m], row[m] + row[-m], acc);
                lap[r] += fmaf(J.w0xlo, col[L + r], acc);
            }
        }
    }

    pin(lap);
    float s[RZ], pc[RZ];
    tile_sources<RZ>(J, t, zs, txi, j, jin, s);
#pragma unroll
    for (int r = 0; r < RZ; ++r) pc[r] = col[L + r];
    if (tile == 0 && w == 0 && lane == 0) {
        if (S.pack_now) J.ck_scale[S.c - 1] = S.scale;
        if (S.direct_now) J.ck_scale[S.c] = S.scale;
    }
    float wmax = 0.f, pn0 = 0.f;
    tile_update<RZ>(J, S, rO, vcol, srow, zs, j, jin, pc, po, v2, lap, s, wmax, pn0);
    ck_max_commit(J, S, wmax, tile * 4 + w, lane);
    // watchdog on the top row (the receivers' row)
    if (J.health && zs == 0) report_nonfinite(J.health, jin && nonfinite(pn0), step);
}

// Interior march: a wavefront owns 64 columns x NG*RZ rows of the interior
// (no damping strip within stencil reach, every cell inside the model) and
// walks down them RZ rows at a time.  The z window stays in registers from
// one group to the next (each p row is loaded once per wave instead of
// (RZ + 2L) / RZ times), and the loads of group g+1 are issued before group g
// is computed, so every wave keeps memory requests in flight while it
// computes.  Arithmetic is that of the tile path's interior branch.  The 4
// wavefronts of a block march the same 64 columns, stacked in z, so the rows
// their windows share are read by one CU.
// NGC = groups, plus 10 for the DPP x exchange (interior_lap_dpp, no LDS) or
// 20 for shifted global loads (interior_lap_glb, no LDS).
template <int N, int RZ, int NGC>
__device__ __forceinline__ void cpml2d_march(const Job2D &J, int mb, int step, float *lds_p) {
    constexpr int NG = NGC % 10;
    constexpr int XMODE = NGC / 10;
    constexpr int L = 2 * N - 1;
    constexpr int PW = kTX + 2 * L;
    constexpr int HPL = (2 * L * RZ + 63) / 64;
    constexpr int ZB = NG * RZ;
    constexpr int WN = ZB + 2 * L;  // window rows z0-L .. z0+ZB-1+L

    const int lane = threadIdx.x & 63, w = wave_id();
    // block mb: column tile mt of the march region, its 4 wavefronts stacked
    // in z (wave w walks rows z0 .. z0 + ZB - 1 of the block's 4 ZB rows)
    const int mzr = mb / J.mntx, mt = mb - mzr * J.mntx;
    const int txi = J.mta + mt;
    const int x0 = txi * kTX, z0 = (mzr * 4 + w) * ZB;
    const StepCtx S = step_ctx(J, step);
    const uint32_t rowb = S.rowb;
    const __amdgpu_buffer_rsrc_t rP = buf_rsrc(J.p[S.par], S.nbytes);
    const __amdgpu_buffer_rsrc_t rO = buf_rsrc(J.p[S.par ^ 1], S.nbytes);
    const __amdgpu_buffer_rsrc_t rV = buf_rsrc(J.vdt2, S.nbytes);
    const int j = x0 + lane;
    const uint32_t vcol = (uint32_t)(j + kHX) * 4u;
    const uint32_t srow0 = (uint32_t)(z0 + kHZ) * rowb;
    int h_lds[HPL];
    uint32_t h_off[HPL];
    halo_slots<N, RZ>(lane, h_lds, h_off, x0, J.pitch, spare_row<N, RZ>(w));

    float win[WN], po[ZB], v2[ZB], hal[NG][HPL];
    float *sp = lds_p + w * (RZ * PW);
    float wmax = 0.f;
    // group q's window loads: its RZ new window rows (group 0: all RZ + 2L)
    // and x halo, prefetched one group ahead
    auto issue = [&](int q) {
        const uint32_t sn = srow0 + q * RZ * rowb;
        if (q == 0) {
#pragma unroll
            for (int r = 0; r < 2 * L; ++r) win[r] = buf_ld(rP, vcol, srow0 + (r - L) * rowb);
        }
#pragma unroll
        for (int r = 0; r < RZ; ++r) win[q * RZ + 2 * L + r] = buf_ld(rP, vcol, sn + (r + L) * rowb);
#pragma unroll
        for (int u = 0; u < HPL; ++u) hal[q][u] = buf_ld(rP, h_off[u], sn);
    };
    // group q's p_old and v^2dt^2, needed only by its update: issued when the
    // group starts, ahead of the next group's prefetch (vmcnt counts in issue
    // order, so waiting for them leaves the prefetch in flight).  Holding them
    // one group instead of two keeps the march at 72 VGPRs (occupancy 7 at
    // N = 5; 76 and occupancy 6 with them prefetched: 12.31 against 11.63 us
    // per step on the 1040 x 3080 workload at NG = 3).
    auto issue_pv = [&](int q) {
        const uint32_t sn = srow0 + q * RZ * rowb;
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            po[q * RZ + r] = buf_ld(rO, vcol, sn + r * rowb);
            v2[q * RZ + r] = buf_ld(rV, vcol, sn + r * rowb);
        }
    };
    issue(0);
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        // keep the prefetch one group deep: nothing below moves above this point
        __builtin_amdgcn_sched_barrier(0);
        issue_pv(g);
        __builtin_amdgcn_sched_barrier(0);
        if (g + 1 < NG) issue(g + 1);
        const int zs = z0 + g * RZ;
        float lap[RZ], s[RZ], pc[RZ], pog[RZ], v2g[RZ];
        if constexpr (XMODE == 1) {
            interior_lap_dpp<N, RZ>(J, win, g * RZ, hal[g], lap);
        } else if constexpr (XMODE == 2) {
            interior_lap_glb<N, RZ>(J, win, g * RZ, rP, vcol, srow0 + g * RZ * rowb, rowb, lap);
        } else {
#pragma unroll
            for (int r = 0; r < RZ; ++r) sp[r * PW + L + lane] = win[g * RZ + L + r];
#pragma unroll
            for (int u = 0; u < HPL; ++u) sp[h_lds[u]] = hal[g][u];
            wave_lds_sync();
            interior_lap<N, RZ>(J, win, g * RZ, sp, lane, lap);
        }
        tile_sources<RZ>(J, S.t, zs, txi, j, true, s);
#pragma unroll
        for (int r = 0; r < RZ; ++r) {
            pc[r] = win[g * RZ + L + r];
            pog[r] = po[g * RZ + r];
            v2g[r] = v2[g * RZ + r];
        }
        float pn0 = 0.f;
        tile_update<RZ>(J, S, rO, vcol, srow0 + g * RZ * rowb, zs, j, true, pc, pog, v2g, lap, s,
                        wmax, pn0);
        if (J.health && zs == 0) report_nonfinite(J.health, nonfinite(pn0), step);
    }
    if (mb == 0 && w == 0 && lane == 0) {
        if (S.pack_now) J.ck_scale[S.c - 1] = S.scale;
        if (S.direct_now) J.ck_scale[S.c] = S.scale;
    }
    ck_max_commit(J, S, wmax, mb * 4 + w, lane);
}

// Block entry: order entries >= 0 are 4-tile blocks (tile path), entries < 0
// interior march blocks (-1 - index).
template <int N, int RZ, int NG>
__device__ __forceinline__ void cpml2d_block(const Job2D &J, int tile, int step) {
    __shared__ float s_p[lds_p_floats<N, RZ>()];
    __shared__ float s_a[4 * RZ * (kTX + 2 * N - 1)];
    if constexpr (NG > 0) {
        if (tile < 0) {
            cpml2d_march<N, RZ, NG>(J, -1 - tile, step, s_p);
            return;
        }
    }
    cpml2d_tile<N, RZ>(J, tile, step, s_p, s_a);
}

// step = the launch's step argument, plus a device-resident base when the
// launch is replayed from a captured time loop (StepGraph).
// Occupancy targets (minimum waves per SIMD) for 4-row tiles up to nDiffOrder
// 5: 7 for the tile path alone (72 VGPRs at N = 5), 6 with the interior march
// (which compiles to 72 as well).  Without them the compiler scheduled the tile
// path at 71 VGPRs to a slower order (13.42 against 12.25 us per step on the
// 1040 x 3080, N = 5 workload).  Taller tiles and higher orders need more
// registers and keep the compiler's own choice (1: no target), as a target
// would spill.
template <int N, int RZ, int NG>
constexpr int occ_target() { return RZ == 4 && N <= 5 ? (NG == 0 ? 7 : 6) : 1; }

template <int N, int RZ, int NG>
__global__ void __launch_bounds__(256, (occ_target<N, RZ, NG>())) cpml2d_single(Job2D J, const int2 *__restrict__ order,
                                                     int step, const int *__restrict__ base) {
    if (base) step += ld_uniform(base);
    cpml2d_block<N, RZ, NG>(J, ld_uniform(&order[blockIdx.x].y), step);
}

template <int N, int RZ, int NG>
__global__ void __launch_bounds__(256, (occ_target<N, RZ, NG>())) cpml2d_batch(const Job2D *__restrict__ jobs,
                                                    const int2 *__restrict__ order, int step,
                                                    const int *__restrict__ base) {
    if (base) step += ld_uniform(base);
    const int jx = ld_uniform(&order[blockIdx.x].x), jy = ld_uniform(&order[blockIdx.x].y);
    cpml2d_block<N, RZ, NG>(jobs[jx], jy, step);
}

__global__ void step_base_set(int *base, int v) {
    if (threadIdx.x == 0) *base = v;
}
__global__ void step_base_add(int *base, int k) {
    if (threadIdx.x == 0) *base += k;
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

void stencil_2d(int N, double dz, double dx, Job2D &J) {
    double c[8];
    staggered_coefficients(N, c);
    for (int k = 0; k < 8; ++k) J.cz[k] = J.cx[k] = 0.f;
    for (int m = 0; m < 16; ++m) J.wz[m] = J.wx[m] = 0.f;
    for (int k = 0; k < N; ++k) {
        J.cz[k] = (float)(c[k] / dz);
        J.cx[k] = (float)(c[k] / dx);
    }
    double w[32];
    composite_weights(N, c, w);
    const int L = 2 * N - 1;
    double side = 0.0, sidez = 0.0, sidex = 0.0;  // exact in double: sums of floats
    for (int m = 1; m <= L; ++m) {
        J.wz[m] = (float)(w[L + m] / (dz * dz));
        J.wx[m] = (float)(w[L + m] / (dx * dx));
        side += 2.0 * ((double)J.wz[m] + (double)J.wx[m]);
        sidez += 2.0 * (double)J.wz[m];
        sidex += 2.0 * (double)J.wx[m];
    }
    // The centre weight is chosen so that the FP32 weights sum to zero (the
    // Laplacian of a constant vanishes, as in the two-stage form where
    // D+ of a constant is exactly 0); it is carried as w0 + w0lo.  The
    // per-direction centres serve strip tiles that use the composite stencil
    // in one direction only.
    J.w0 = (float)(-side);
    J.w0lo = (float)(-side - (double)J.w0);
    J.w0z = (float)(-sidez);
    J.w0zlo = (float)(-sidez - (double)J.w0z);
    J.w0x = (float)(-sidex);
    J.w0xlo = (float)(-sidex - (double)J.w0x);
}

Cpml2dGrid::Cpml2dGrid(const double *v, int64_t nz_, int64_t nx_, int N_, int nB_, double dz_,
                       double dx_, double dt_, int rz_, int ng_)
    : nz(nz_), nx(nx_), N(N_), nB(nB_), dz(dz_), dx(dx_), dt(dt_) {
    // strip state is addressed with 32-bit buffer offsets below 2^31 (the
    // kernel marks lanes without a value with an offset of 2^31)
    if ((int64_t)nB * nx * 4 >= ((int64_t)1 << 31) || nz * 2 * nB * 4 >= ((int64_t)1 << 31))
        fail(SSSI_EINVAL, "2-D damping-strip state exceeds 2 GiB per array");
    int r = 4;
    if (rz_ == 4 || rz_ == 8) r = rz_;
    else if (const char *e = std::getenv("SSSI_RZ")) {
        const int q = std::atoi(e);
        if (q == 4 || q == 8) r = q;
    }
    set_layout(r, ng_ >= 0 ? ng_ : march_groups(r));
    pitch = round_up(kHX + round_up(nx, kTX) + 32, 64);
    rows = kHZ + round_up(nz, kTZmax) + kHZbot;
    const size_t np = (size_t)pitch * rows;
    if (np * 4 >= ((size_t)1 << 32)) fail(SSSI_EINVAL, "2-D padded grid exceeds 4 GiB");
    p[0].alloc(np);
    p[1].alloc(np);
    vdt2.alloc(np);
    health.alloc(1);
    SSSI_HIP(hipMemset(health.p, 0x7f, sizeof(int)));
    // v^2 dt^2 in the padded layout, converted on the device from the FP64 model
    upload_model(v, nz, nx, 1, dt, true, vdt2.p, vdt2.n, pitch, 0, kHZ, 0, kHX, alloc_stream());
    stencil_2d(N, dz, dx, coef);
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

bool Cpml2dGrid::block_is_strip(int tile) const { return strip_kind(tile) != 0; }

int Cpml2dGrid::strip_kind(int tile) const {
    if (nB == 0) return 0;
    const int tzi = tile / ntx(), txi = tile % ntx();
    const int64_t z0 = (int64_t)tzi * tz(), x0 = (int64_t)txi * kTX;
    const bool zs = z0 + tz() + N - 1 > nz - nB;
    const bool xs = x0 - N < nB || x0 + kTX + N - 1 > nx - nB;
    return (xs ? 1 : 0) | (zs ? 2 : 0);
}

void Cpml2dGrid::set_layout(int rz_, int ng_) {
    if (!(rz_ == 4 || rz_ == 8)) fail(SSSI_EINVAL, "tile height %d (4 or 8)", rz_);
    if (!(ng_ == 0 || ng_ == 2 || ng_ == 4 || ((ng_ == 3 || ng_ == 8) && rz_ == 4)))
        fail(SSSI_EINVAL, "march groups %d with tile height %d", ng_, rz_);
    rz = rz_;
    ng = ng_;
    mz_rows = mta = mntx = 0;
    if (ng == 0) return;
    // column tiles with every lane inside the model and no x strip within
    // stencil reach (the tile path's interior test); a contiguous range
    int ta = -1, tb = -1;
    for (int t = 0; t < ntx(); ++t) {
        const int64_t x0 = (int64_t)t * kTX;
        bool ok = x0 + kTX <= nx;
        if (nB > 0) ok = ok && x0 - N >= nB && x0 + kTX + N - 1 <= nx - nB;
        if (ok) {
            if (ta < 0) ta = t;
            tb = t;
        }
    }
    // rows with no z strip within reach, in whole tile blocks and march bands
    const int64_t zlim = nB > 0 ? nz - nB - N + 1 : nz;
    const int64_t unit = (int64_t)4 * rz * ng;
    const int64_t rows_m = zlim >= unit ? (zlim / unit) * unit : 0;
    if (ta >= 0 && rows_m > 0) {
        mz_rows = (int)rows_m;
        mta = ta;
        mntx = tb - ta + 1;
    }
}

bool Cpml2dGrid::block_in_march(int tile) const {
    if (mntx == 0) return false;
    const int tzi = tile / ntx(), txi = tile % ntx();
    return txi >= mta && txi < mta + mntx && (int64_t)(tzi + 1) * tz() <= mz_rows;
}

static int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

int march_groups(int rz) {
    int ng = env_int("SSSI_NG", rz == 4 ? 3 : 0);
    if (rz == 8 && ng > 4) ng = 4;
    if (rz == 8 && ng == 3) ng = 2;
    return (ng >= 2 && ng <= 4) || ng == 8 ? ng : 0;
}

void Cpml2dGrid::zero_state(hipStream_t s, bool arm) {
    for (auto *b : {&p[0], &p[1], &zphi[0], &zphi[1], &zpsi, &xphi[0], &xphi[1], &xpsi})
        if (b->p) SSSI_HIP(hipMemsetAsync(b->p, 0, b->bytes(), s));
    if (arm) arm_health(health.p, s);
}

void Cpml2dGrid::scan_nonfinite(int which, int tag, hipStream_t s) const {
    launch_nonfinite_scan(p[which].p, p[which].n, health.p, tag, s);
}

// Grid-stride scan: one atomic per wave that holds a NaN/Inf.
__global__ void __launch_bounds__(256) nonfinite_scan(const float *__restrict__ p, size_t n,
                                                      int *health, int tag) {
    bool bad = false;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        bad |= nonfinite(p[q]);
    report_nonfinite(health, bad, tag);
}

void launch_nonfinite_scan(const float *p, size_t n, int *health, int tag, hipStream_t s) {
    if (!p || !health || n == 0) return;
    const size_t blocks = std::min<size_t>(2048, (n + 255) / 256);
    hipLaunchKernelGGL(nonfinite_scan, dim3((unsigned)blocks), dim3(256), 0, s, p, n, health, tag);
    SSSI_HIP(hipGetLastError());
}

int read_health(const int *health, hipStream_t s) {
    if (!health) return kHealthy;
    int h = kHealthy;
    SSSI_HIP(hipMemcpyAsync(&h, health, sizeof(int), hipMemcpyDeviceToHost, s));
    SSSI_HIP(hipStreamSynchronize(s));
    return h;
}

void arm_health(int *health, hipStream_t s) {
    if (health) SSSI_HIP(hipMemsetAsync(health, 0x7f, sizeof(int), s));
}

Job2D Cpml2dGrid::job() const {
    Job2D J;
    std::memset(&J, 0, sizeof(J));
    J.nz = (int)nz; J.nx = (int)nx; J.pitch = (int)pitch; J.rows = (int)rows;
    J.ntx = ntx(); J.nwz = nwz(); J.nB = nB;
    J.mz_rows = mz_rows; J.mta = mta; J.mntx = mntx;
    for (int k = 0; k < 8; ++k) { J.cz[k] = coef.cz[k]; J.cx[k] = coef.cx[k]; }
    J.w0 = coef.w0;
    J.w0lo = coef.w0lo;
    J.w0z = coef.w0z;
    J.w0zlo = coef.w0zlo;
    J.w0x = coef.w0x;
    J.w0xlo = coef.w0xlo;
    for (int m = 0; m < 16; ++m) { J.wz[m] = coef.wz[m]; J.wx[m] = coef.wx[m]; }
    J.p[0] = p[0].p; J.p[1] = p[1].p;
    J.vdt2 = vdt2.p;
    J.zphi[0] = zphi[0].p; J.zphi[1] = zphi[1].p; J.zpsi = zpsi.p; J.zbm1 = zbm1.p;
    J.xphi[0] = xphi[0].p; J.xphi[1] = xphi[1].p; J.xpsi = xpsi.p; J.xbm1 = xbm1.p;
    J.health = health.p;
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

// Append `list` so that the k-th appended entry lands on XCD (base + k) % 8
// with each XCD taking a contiguous band of the list.
static void append_banded(std::vector<int2> &order, const std::vector<int2> &list) {
    const size_t n = list.size(), base = order.size();
    size_t next[8], end[8];
    for (int x = 0; x < 8; ++x) {
        next[x] = n * x / 8;
        end[x] = n * (x + 1) / 8;
    }
    for (size_t b = base; b < base + n; ++b) {
        int x = (int)(b % 8);
        if (next[x] == end[x])
            for (int y = 0; y < 8; ++y)
                if (next[y] < end[y]) { x = y; break; }
        order.push_back(list[next[x]++]);
    }
}

// Strip tiles first (banded per kind of strip in reach when SSSI_STRIPBAND2 is
// 2, as one list when 1), then the interior march blocks, then the remaining
// interior tiles, each banded over the 8 XCDs.
void build_block_order(const std::vector<const Cpml2dGrid *> &grids, std::vector<int2> &order) {
    std::vector<int2> march, strip, inner, kind[4];
    for (int g = 0; g < (int)grids.size(); ++g) {
        const Cpml2dGrid &G = *grids[g];
        for (int b = 0; b < G.n_march_blocks(); ++b) march.push_back(make_int2(g, -1 - b));
        for (int t = 0; t < G.nblocks(); ++t) {
            if (G.block_in_march(t)) continue;
            const int k = G.strip_kind(t);
            (k ? strip : inner).push_back(make_int2(g, t));
            if (k) kind[k].push_back(make_int2(g, t));
        }
    }
    order.clear();
    const char *e = std::getenv("SSSI_STRIPBAND2");
    if (e && std::atoi(e) == 2) {
        for (int k = 3; k >= 1; --k) append_banded(order, kind[k]);
    } else {
        append_banded(order, strip);
    }
    append_banded(order, march);
    append_banded(order, inner);
}

// Instantiated layouts: (rz, march groups); groups 0 is the tile path alone.
#define SSSI_NRZ_CASES(NN, CALL)                                                          \
    case NN * 100 + 4 * 10 + 0: CALL(NN, 4, 0); break;                                    \
    case NN * 100 + 4 * 10 + 2: CALL(NN, 4, 2); break;                                    \
    case NN * 100 + 4 * 10 + 3: CALL(NN, 4, 3); break;                                    \
    case NN * 100 + 4 * 10 + 4: CALL(NN, 4, 4); break;                                    \
    case NN * 100 + 4 * 10 + 8: CALL(NN, 4, 8); break;                                    \
    case NN * 100 + 8 * 10 + 0: CALL(NN, 8, 0); break;                                    \
    case NN * 100 + 8 * 10 + 2: CALL(NN, 8, 2); break;                                    \
    case NN * 100 + 8 * 10 + 4: CALL(NN, 8, 4); break;
#define SSSI_NRZ_SWITCH(N, RZ, NG, CALL)                                                   \
    switch ((N) * 100 + (RZ) * 10 + (NG)) {                                               \
        SSSI_NRZ_CASES(1, CALL) SSSI_NRZ_CASES(2, CALL) SSSI_NRZ_CASES(3, CALL)           \
        SSSI_NRZ_CASES(4, CALL) SSSI_NRZ_CASES(5, CALL) SSSI_NRZ_CASES(6, CALL)           \
        SSSI_NRZ_CASES(7, CALL) SSSI_NRZ_CASES(8, CALL)                                   \
        case 5 * 100 + 4 * 10 + 13: CALL(5, 4, 13); break;                                 \
        case 5 * 100 + 4 * 10 + 23: CALL(5, 4, 23); break;                                 \
        default: fail(SSSI_EINVAL, "no 2-D kernel for nDiffOrder=%d rz=%d groups=%d", (N), \
                      (RZ), (NG));                                                        \
    }

void launch_cpml2d_single(int N, int rz, int ng, const Job2D &job, const int2 *order, int nblk,
                          int step, hipStream_t s, const int *base) {
#define SSSI_L1(NN, R, G)                                                                 \
    hipLaunchKernelGGL((cpml2d_single<NN, R, G>), dim3(nblk), dim3(256), 0, s, job, order, step, \
                       base)
    SSSI_NRZ_SWITCH(N, rz, ng, SSSI_L1)
#undef SSSI_L1
    SSSI_HIP(hipGetLastError());
}

// SSSI_XDPP=1: the interior march exchanges x neighbours by DPP; =2: reads
// them as shifted global loads (built for nDiffOrder 5, 4-row tiles, 3 groups:
// the benchmark layout).  Returns the layout code's offset (0, 10 or 20).
static int xdpp_enabled(int N, int rz, int ng) {
    const char *e = std::getenv("SSSI_XDPP");
    const int m = e ? std::atoi(e) : 0;
    return (m == 1 || m == 2) && N == 5 && rz == 4 && ng == 3 ? 10 * m : 0;
}

void launch_cpml2d_single(const Cpml2dGrid &g, const Job2D &job, const int2 *order, int nblk,
                          int step, hipStream_t s, const int *base) {
    launch_cpml2d_single(g.N, g.rz, g.ng + xdpp_enabled(g.N, g.rz, g.ng), job, order, nblk, step,
                         s, base);
}

void launch_cpml2d_batch(int N, int rz, int ng, const Job2D *jobs, const int2 *order, int nblk,
                         int step, hipStream_t s, const int *base) {
#define SSSI_LB(NN, R, G)                                                                 \
    hipLaunchKernelGGL((cpml2d_batch<NN, R, G>), dim3(nblk), dim3(256), 0, s, jobs, order, step, \
                       base)
    SSSI_NRZ_SWITCH(N, rz, ng, SSSI_LB)
#undef SSSI_LB
    SSSI_HIP(hipGetLastError());
}

void launch_ck_pack(const Job2D &job, int c, hipStream_t s) {
    if (job.ck_fmt != kCkF16) return;  // only FP16 checkpoints are staged
    const size_t n = (size_t)job.nz * job.nx;
    hipLaunchKernelGGL(ck_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       job.ck_stage + (size_t)(c & 1) * n, job.ck16 + (size_t)c * n,
                       job.ck_max + (size_t)c * kCkSlots, job.ck_scale + c, n, job.ck_fmt);
    SSSI_HIP(hipGetLastError());
}

bool graphs_enabled() { return env_int("SSSI_GRAPH", 0) != 0; }

// Columns c0 .. c0+ncol-1 (column c = k*nx + j) of the FP64 model in the
// staging buffer -> the padded FP32 field.
__global__ void model_to_padded(const double *__restrict__ vc, int64_t nz, int64_t nx,
                                int64_t c0, int64_t ncol, double dt, int square,
                                float *__restrict__ dst, int64_t plane, int64_t pitch, int hz,
                                int hy, int hx) {
    const size_t n = (size_t)nz * ncol;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        const int64_t c = c0 + (int64_t)(q / nz), i = (int64_t)(q % nz);
        const int64_t k = c / nx, j = c - k * nx;
        const double t = vc[q] * dt;
        dst[(size_t)(i + hz) * plane + (size_t)(k + hy) * pitch + (j + hx)] =
            (float)(square ? t * t : t);
    }
}

__global__ void fill_zero_f32(float *p, size_t n) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        p[q] = 0.f;
}

void upload_model(const double *v, int64_t nz, int64_t nx, int64_t ny, double dt, bool square,
                  float *dst, size_t dst_n, int64_t plane, int64_t pitch, int hz, int hy, int hx,
                  hipStream_t s) {
    auto blocks = [](size_t n) {
        return (unsigned)std::max<size_t>(1, std::min<size_t>(4096, (n + 255) / 256));
    };
    hipLaunchKernelGGL(fill_zero_f32, dim3(blocks(dst_n)), dim3(256), 0, s, dst, dst_n);
    SSSI_HIP(hipGetLastError());
    const int64_t ncols = nx * ny;
    const int64_t per = std::max<int64_t>(1, ((int64_t)4 << 20) / nz);  // <= 32 MB chunks
    const int64_t chunk = std::min(per, ncols);
    DevBuf<double> stage((size_t)chunk * nz);
    for (int64_t c0 = 0; c0 < ncols; c0 += chunk) {
        const int64_t nc = std::min(chunk, ncols - c0);
        // pageable source: the copy has left the host buffer when the call returns
        SSSI_HIP(hipMemcpyAsync(stage.p, v + c0 * nz, (size_t)nc * nz * sizeof(double),
                                hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(model_to_padded, dim3(blocks((size_t)nc * nz)), dim3(256), 0, s,
                           stage.p, nz, nx, c0, nc, dt, square ? 1 : 0, dst, plane, pitch, hz, hy,
                           hx);
        SSSI_HIP(hipGetLastError());
    }
    SSSI_HIP(hipStreamSynchronize(s));  // the staging buffer is released on return
}

StepGraph::~StepGraph() { clear(); }

void StepGraph::clear() {
    if (exec_ && stream_) (void)hipStreamSynchronize(stream_);  // replays still queued
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    key_.clear();
}

void StepGraph::run(const void *key, size_t key_bytes, int step0, int64_t count, hipStream_t s,
                    const Launch &launch) {
    const int64_t reps = count / kGraphSteps;
    if (!graphs_enabled() || reps < 2) {
        for (int64_t q = 0; q < count; ++q) launch((int)(step0 + q), nullptr, s);
        return;
    }
    const char *kb = static_cast<const char *>(key);
    if (!exec_ || key_.size() != key_bytes || std::memcmp(key_.data(), kb, key_bytes) != 0) {
        clear();
        if (!base_.p) base_.alloc(1);
        SSSI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            for (int q = 0; q < kGraphSteps; ++q) launch(q, base_.p, s);
            hipLaunchKernelGGL(step_base_add, dim3(1), dim3(64), 0, s, base_.p, kGraphSteps);
            SSSI_HIP(hipGetLastError());
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SSSI_HIP(hipStreamEndCapture(s, &graph_));
        SSSI_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        key_.assign(kb, kb + key_bytes);
    }
    stream_ = s;
    hipLaunchKernelGGL(step_base_set, dim3(1), dim3(64), 0, s, base_.p, step0);
    SSSI_HIP(hipGetLastError());
    for (int64_t r = 0; r < reps; ++r) SSSI_HIP(hipGraphLaunch(exec_, s));
    for (int64_t q = reps * kGraphSteps; q < count; ++q) launch((int)(step0 + q), nullptr, s);
}

}  // namespace sssi
