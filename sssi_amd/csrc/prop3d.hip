// prop3d.hip -- 3-D acoustic CPML propagator for gfx950 (fwdTimeCpmlFor3dAw,
// src/fwdTimeCpmlFor3dAw.m:37-160), including its damping conventions
// (fwdTimeCpmlFor3dAw.m:46-70: u in grid-index units, L = nBoundary; the x and
// z profiles take v from the first y slice, the y profile from the top row).
//
// Device layout: [z][y][x], x fastest; interior cell (i, k, j) = (z, y, x) at
// ((i + kHZ) * npy + (k + kHY)) * pitch + (j + kHX).  Halos are zero.
//
// One launch per time step.  A block of 64 x TY threads owns a 64 (x) by TY (y)
// column tile and sweeps a chunk of ZC planes downwards (2.5-D): the z
// derivative runs from registers (a window of 2L+1 p values and a queue of
// 2N A_z half-point values per column, so each D+ is computed once), x and y
// from a double-buffered LDS plane tile with an L-wide halo.  The planes'
// global loads run PD planes ahead in a register ring.
//
// The z chunks are split at the depth where the z stencil first reaches the
// bottom damping strip, so that a block either touches no strip at all (the
// interior sweep: composite D-D+ stencil in x, y and z, nothing but the ring
// in flight) or is a strip block (x/y strip tile or bottom chunk: the two-stage
// CPML form in the directions whose strip is within reach, A_x and A_y at half
// points through LDS; Phi ping-ponged, neighbouring tiles and chunks
// recomputing the half points at their edges; Psi in place).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "health.h"
#include "prop.h"
#include "tune.h"

namespace sssi {

constexpr int kHY = 16;  // y halo (>= L for N <= 8)

struct Cpml3dArgs {
    const float *pcur;
    float *pold;  // read p_old, write p_new
    const float *vdt2;
    int nz, nx, ny, pitch, npy, nB;
    int ntx, nty, zc;  // tile counts in x, y; planes per chunk
    // z chunks: nzA chunks tile [0, zsplit), whose z reach stays above the
    // bottom damping strip; the rest tile [zsplit, nz)
    int zsplit, nzA;
    uint32_t rowb, planeb, nbytes;
    float cz[8], cx[8], cy[8];
    float w0, w0lo, wx[16], wy[16];  // composite x/y weights / d^2 (w0 + w0lo: centre)
    float w0x, w0xlo, w0y, w0ylo;    // centres of the x / y stencils alone
    // Damping-strip arrays, byte offsets into one pool of pool_bytes < 2 GiB:
    // z strip: half points / rows i in [nz-nB, nz): [(i-(nz-nB))*ny + k]*nx + j,
    //   b-1 at [(i-(nz-nB))*nx + j] (replicated over y);
    // x strip: [(i*ny + k)*2nB + xs], b-1 at [i*2nB + xs] (replicated over y);
    // y strip: [(i*2nB + ys)*nx + j], b-1 at [ys*nx + j] (replicated over z).
    float *pool;
    uint32_t pool_bytes;
    uint32_t o_zphi_in, o_zphi_out, o_zpsi, o_zbm1;
    uint32_t o_xphi_in, o_xphi_out, o_xpsi, o_xbm1;
    uint32_t o_yphi_in, o_yphi_out, o_ypsi, o_ybm1;
    const int *order;     // launch order of the blocks (strip blocks first) or null
    // point sources bucketed by (plane i, row k, 64-column x tile): bucket
    // (i*ny + k)*ntx + tx holds entries src_ptr[b] .. src_ptr[b+1]-1, in the
    // caller's order within a bucket; entry e is cell src_cell[e] with this
    // step's amplitude src_amp[e]
    const int *src_ptr;
    const int *src_zr;    // [(k*ntx + tx)*2]: first and last plane with a source (lo > hi: none)
    const int *src_cell;  // (i*ny + k)*nx + j
    const float *src_amp;
    int nsrc;
    float *rec;   // this step's receivers [k*nx + j] or null
    float *snap;  // this step's snapshot [(i*ny + k)*nx + j] or null
    int *health;  // non-finite watchdog word, or null
    int step;     // launch step reported to it
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc3(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}
// Scalar load of wave-uniform, launch-invariant data (source tables): waited
// for with lgkmcnt, so the in-order vmcnt of the prefetched planes is left alone.
template <class T>
__device__ __forceinline__ T ld_uniform3(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}
__device__ __forceinline__ float ld3(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
// Streams touched once per step (p_old, v^2dt^2 reads; the p_new store) carry
// the non-temporal cache policy so that they do not evict the p planes that
// neighbouring blocks re-read as their x/y halo from L2.
constexpr int kNonTemporal = 2;  // cache-policy NT bit of gfx950 buffer operations
__device__ __forceinline__ float ld3_nt(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(
        __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, kNonTemporal));
}
__device__ __forceinline__ void st3_nt(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                       uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff,
                                          kNonTemporal);
}
__device__ __forceinline__ void st3(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                    uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, 0);
}

constexpr int kPD3max = 4;  // deepest prefetch ring of the 3-D sweep (planes)
static_assert(2 * kPD3max + 15 <= kHZbot, "bottom padding must hold the sweep's overrun");

// Strip state lives in one pool behind one buffer resource; a lane without a
// value in a strip array takes this byte offset, past the end of any pool
// (pools stay below 2 GiB): its buffer load returns 0 and its store is dropped,
// so the strip code runs without per-lane branches.
constexpr uint32_t kOOB = 0x80000000u;
__device__ __forceinline__ float ldS(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, 0, 0));
}
__device__ __forceinline__ void stS(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, 0, 0);
}

// Tile geometry shared by both sweeps.
template <int N, int TY>
struct Tile3 {
    static constexpr int L = 2 * N - 1;
    static constexpr int NT = 64 * TY;
    static constexpr int PX = kTX + 2 * L;  // LDS plane tile: x0-L .. x0+63+L
    static constexpr int PY = TY + 2 * L;   //                 y0-L .. y0+TY-1+L
    // halo values per plane: the y halo at the tile's own 64 columns and the x
    // halo at its own rows.  The corners (y halo x x halo) are never read: the
    // y stencil runs down owned columns, the x stencil along owned rows, in
    // both sweeps.  The y-halo rows are then 64 aligned floats, two cache lines.
    static constexpr int HN = 2 * L * kTX + TY * 2 * L;
    static constexpr int HPT = (HN + NT - 1) / NT;
    static constexpr int AXW = kTX + L;  // A_x half points x0-N .. x0+63+N-1 per row
    static constexpr int AYH = TY + L;   // A_y half points y0-N .. y0+TY-1+N-1 per column
    static constexpr int QY = (AYH + TY - 1) / TY;  // A_y half points per thread
    // Opt-in LDS plane ring of L+2 tiles, where it fits the LDS share of a
    // block at 4 waves per SIMD (10 KiB per wave): the halo of plane i+L is
    // loaded with the z window's front, plane i+L, so a cache line of p is
    // requested by its owner and by the neighbouring tiles' halos at one time,
    // and waits in LDS until plane i is computed.  At 420 x 440 x 440, N=4 it
    // fetches 4 % less (L2 hits 39 -> 42 %) and runs 1 % slower, so the default
    // is 2 tiles, the halo loaded at plane i.
    static constexpr int RING_BYTES = (L + 2) * PX * PY * 4 + (TY * AXW + AYH * kTX) * 4;
    static constexpr int RING = RING_BYTES <= 10240 * TY ? L + 2 : 2;
};

// Halo slots of this thread in the LDS plane tile (fixed for the sweep), the
// cells of Tile3::HN.  Threads without a slot load a valid cell that is never
// written to LDS (an unconditional load keeps the prefetched registers out of
// branches).
template <int N, int TY>
__device__ __forceinline__ void halo_slots3(const Cpml3dArgs &a, int tid, int x0, int y0,
                                            int (&h_lds)[Tile3<N, TY>::HPT],
                                            uint32_t (&h_voff)[Tile3<N, TY>::HPT]) {
    using T = Tile3<N, TY>;
    constexpr int L = T::L, PX = T::PX;
#pragma unroll
    for (int u = 0; u < T::HPT; ++u) {
        const int e = tid + T::NT * u;
        int lidx = -1, gy = 0, gx = 0;
        if (e < T::HN) {
            // the y halo above (rows y0-L .. y0-1, owned columns), the x halo
            // of the owned rows, the y halo below (rows y0+TY .. y0+TY+L-1)
            const int top = L * kTX;
            const int mid = TY * 2 * L;
            if (e < top) {
                const int r = e / kTX, c = e - r * kTX;
                lidx = r * PX + L + c; gy = y0 - L + r; gx = x0 + c;
            } else if (e < top + mid) {
                const int q = e - top, r = q / (2 * L), c = q - r * (2 * L);
                const int cc = c < L ? c : kTX + c;
                lidx = (L + r) * PX + cc; gy = y0 + r; gx = x0 - L + cc;
            } else {
                const int q = e - top - mid, r = q / kTX, c = q - r * kTX;
                lidx = (L + TY + r) * PX + L + c; gy = y0 + TY + r; gx = x0 + c;
            }
        }
        h_lds[u] = lidx;
        h_voff[u] = ((uint32_t)(gy + kHY) * a.pitch + (uint32_t)(gx + kHX)) * 4u;
    }
}

// Sweep of a column tile and z chunk whose stencil reach touches no damping
// strip: the composite D-D+ stencil in x, y and z, no CPML state, and the only
// global loads in the plane loop are the prefetch ring's, so the in-order
// vmcnt waits never drain the ring.
template <int N, int TY, int PD, int RG>
__device__ __forceinline__ void cpml3d_sweep_inner(const Cpml3dArgs &a, float *sp0, int tx, int tyi,
                                                   int zbeg, int zend) {
    using T = Tile3<N, TY>;
    constexpr int L = T::L, PX = T::PX, PY = T::PY, HPT = T::HPT;
    static_assert(RG == 2 || RG >= L + 2, "plane ring: 2 tiles, or L+2 and more");
    constexpr int HL = RG >= L + 2 ? L : 0;  // planes the halo loads run ahead
    // the wave's row of the tile as a scalar: the compiler cannot prove
    // tid >> 6 uniform, and every row-dependent offset and test would be
    // vector arithmetic
    const int tid = threadIdx.x, lane = tid & 63, ty = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x0 = tx * kTX, y0 = tyi * TY;
    const int nz = a.nz, nx = a.nx, ny = a.ny;
    const int j = x0 + lane, k = y0 + ty;
    const bool own = j < nx && k < ny;
    const __amdgpu_buffer_rsrc_t rP = buf_rsrc3(a.pcur, a.nbytes);
    const __amdgpu_buffer_rsrc_t rO = buf_rsrc3(a.pold, a.nbytes);
    const __amdgpu_buffer_rsrc_t rV = buf_rsrc3(a.vdt2, a.nbytes);
    const uint32_t vcol = ((uint32_t)(k + kHY) * a.pitch + (uint32_t)(j + kHX)) * 4u;
    // planes holding a source of this wave's row and x tile (wave-uniform)
    int s_zlo = 1, s_zhi = 0;
    if (a.nsrc > 0 && k < ny) {
        s_zlo = ld_uniform3(a.src_zr + ((size_t)k * a.ntx + tx) * 2);
        s_zhi = ld_uniform3(a.src_zr + ((size_t)k * a.ntx + tx) * 2 + 1);
    }
    auto plane = [&](int i) -> uint32_t { return (uint32_t)(i + kHZ) * a.planeb; };
    int h_lds[HPT];
    uint32_t h_voff[HPT];
    halo_slots3<N, TY>(a, tid, x0, y0, h_lds, h_voff);

    // z window p(i-L .. i+L) and A_z queue (half points i-N .. i+N-1)
    float zq[2 * L + 1], aq[2 * N];
#pragma unroll
    for (int r = 0; r < 2 * L; ++r) zq[r + 1] = ld3(rP, vcol, plane(zbeg - L + r));
#pragma unroll
    for (int q = 0; q < 2 * N - 1; ++q) {
        // half point h = zbeg-N+q uses p(h-N+1 .. h+N) = zq[1+q .. 1+q+2N-1]
        float d = 0.f;
#pragma unroll
        for (int m = 0; m < N; ++m) d = fmaf(a.cz[m], zq[1 + q + N + m] - zq[q + N - m], d);
        aq[q + 1] = d;
    }

    // prefetch ring: planes zbeg .. zbeg+PD-1 (slot d holds plane i with
    // (i - zbeg) % PD == d); a slot is refilled right after it is consumed, so
    // each plane's loads have PD plane-iterations to arrive.  The slots are
    // issued in slot order (scheduling barriers): vmcnt counts in issue order,
    // and the loop's waits must also hold on entry from here, so a slot-0 load
    // scheduled after slot 1's would make every iteration wait for all loads.
    // halo of planes zbeg .. zbeg+HL-1 into ring slots 0 .. HL-1, all loads
    // issued before the first write (the loop's first barrier orders these
    // writes before their reads)
    if constexpr (HL > 0) {
        float hv[HL][HPT];
#pragma unroll
        for (int q = 0; q < HL; ++q)
#pragma unroll
            for (int u = 0; u < HPT; ++u) hv[q][u] = ld3(rP, h_voff[u], plane(zbeg + q));
#pragma unroll
        for (int q = 0; q < HL; ++q) {
            float *Q = sp0 + q * (PX * PY);
#pragma unroll
            for (int u = 0; u < HPT; ++u)
                if (h_lds[u] >= 0) Q[h_lds[u]] = hv[q][u];
        }
    }
    float f_front[PD], f_po[PD], f_v2[PD], f_h[PD][HPT];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int d = 0; d < PD; ++d) {
        const int ip = min(zbeg + d, nz - 1);  // in bounds; unused past zend
        f_po[d] = ld3_nt(rO, vcol, plane(ip));
        f_v2[d] = ld3_nt(rV, vcol, plane(ip));
        f_front[d] = ld3(rP, vcol, plane(ip + L));
#pragma unroll
        for (int u = 0; u < HPT; ++u) f_h[d][u] = ld3(rP, h_voff[u], plane(ip + HL));
        __builtin_amdgcn_sched_barrier(0);
    }

    // ring slots of plane i and of plane i+HL (equal when HL is 0)
    int sI = 0, sH = HL % RG;
    for (int i0 = zbeg; i0 < zend; i0 += PD) {
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            // planes past the chunk end (last group of a chunk) run on the zero
            // padding below the model and store nothing
            const int i = i0 + d;
            const bool live = i < zend;
            // advance the z window; the ring's front value enters it through a
            // move, so that the window owns a register of its own and the
            // refill below loads into the slot's register (no copy of an
            // in-flight load at the loop back-edge)
#pragma unroll
            for (int r = 0; r < 2 * L; ++r) zq[r] = zq[r + 1];
            asm volatile("v_mov_b32 %0, %1" : "=v"(zq[2 * L]) : "v"(f_front[d]));
#pragma unroll
            for (int q = 0; q < 2 * N - 1; ++q) aq[q] = aq[q + 1];
            {
                float dd = 0.f;  // new half point i+N-1: p(i .. i+2N-1) = zq[L .. L+2N-1]
#pragma unroll
                for (int m = 0; m < N; ++m) dd = fmaf(a.cz[m], zq[L + N + m] - zq[L + N - 1 - m], dd);
                aq[2 * N - 1] = dd;
            }
            float pz = 0.f;
#pragma unroll
            for (int m = 0; m < N; ++m) pz = fmaf(a.cz[m], aq[N + m] - aq[N - 1 - m], pz);

            float *P = sp0 + sI * (PX * PY);
            P[(ty + L) * PX + lane + L] = zq[L];
            float *PH = sp0 + sH * (PX * PY);
#pragma unroll
            for (int u = 0; u < HPT; ++u)
                if (h_lds[u] >= 0) PH[h_lds[u]] = f_h[d][u];
            __syncthreads();
            const float *c = P + (ty + L) * PX + lane + L;
            float acc = a.w0 * c[0];
#pragma unroll
            for (int m = 1; m <= L; ++m) acc = fmaf(a.wx[m], c[m] + c[-m], acc);
#pragma unroll
            for (int m = 1; m <= L; ++m) acc = fmaf(a.wy[m], c[m * PX] + c[-m * PX], acc);
            const float pxy = fmaf(a.w0lo, c[0], acc);  // low part of the centre weight, after cancellation

            float pn = 0.f;
            if (own && live) {
                float s = 0.f;
                if (a.nsrc > 0 && i >= s_zlo && i <= s_zhi) {
                    // the wave's bucket: one plane, one row, its 64-column tile
                    const int cell = (i * ny + k) * nx + j;
                    const int bk = (i * ny + k) * a.ntx + tx;
                    const int e0 = ld_uniform3(a.src_ptr + bk), e1 = ld_uniform3(a.src_ptr + bk + 1);
                    for (int e = e0; e < e1; ++e) {
                        const int cq = ld_uniform3(a.src_cell + e);
                        const float amp = ld_uniform3(a.src_amp + e);
                        if (cq == cell) s += amp;
                    }
                }
                const float pc = zq[L];
                pn = pc + fmaf(f_v2[d], (pz + pxy) + s, pc - f_po[d]);
                st3_nt(pn, rO, vcol, plane(i));
                if (a.rec && i == 0) a.rec[(size_t)k * nx + j] = pn;
                if (a.snap) a.snap[((size_t)i * ny + k) * nx + j] = pn;
            }
            // watchdog on the top (receiver) plane
            if (a.health && i == 0) report_nonfinite(a.health, own && live && nonfinite(pn), a.step);
            // refill the slot with plane i+PD; past the chunk the loads read the
            // bottom padding (2*PD + L <= kHZbot)
            f_front[d] = ld3(rP, vcol, plane(i + PD + L));
            f_po[d] = ld3_nt(rO, vcol, plane(i + PD));
            f_v2[d] = ld3_nt(rV, vcol, plane(i + PD));
#pragma unroll
            for (int u = 0; u < HPT; ++u) f_h[d][u] = ld3(rP, h_voff[u], plane(i + PD + HL));
            sI = sI + 1 == RG ? 0 : sI + 1;
            sH = sH + 1 == RG ? 0 : sH + 1;
        }
    }
}

// Sweep of a tile or chunk within reach of a damping strip: each direction
// takes the two-stage CPML form where its strip is within reach (A_x, A_y at
// half points through LDS; Phi ping-ponged, neighbouring tiles and chunks
// recomputing the half points at their edges; Psi in place), its composite
// stencil otherwise.  A plane's strip values are loaded, branch-free (pool
// offsets, kOOB for lanes without a value), at the end of the previous plane
// and before that plane's ring refill, so waiting for them leaves the refill
// in flight.
template <int N, int TY, int PD, int RG>
__device__ __forceinline__ void cpml3d_sweep_strip(const Cpml3dArgs &a, float *sp0, float *sax,
                                                   float *say, int tx, int tyi, int zbeg,
                                                   int zend) {
    using T = Tile3<N, TY>;
    constexpr int L = T::L, PX = T::PX, PY = T::PY, HPT = T::HPT, AXW = T::AXW, AYH = T::AYH,
                  QY = T::QY;
    static_assert(RG == 2 || RG >= L + 2, "plane ring: 2 tiles, or L+2 and more");
    constexpr int HL = RG >= L + 2 ? L : 0;  // planes the halo loads run ahead
    // the wave's row of the tile as a scalar: the compiler cannot prove
    // tid >> 6 uniform, and every row-dependent offset and test would be
    // vector arithmetic
    const int tid = threadIdx.x, lane = tid & 63, ty = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x0 = tx * kTX, y0 = tyi * TY;
    const int nz = a.nz, nx = a.nx, ny = a.ny, nB = a.nB;
    const int j = x0 + lane, k = y0 + ty;
    const bool own = j < nx && k < ny;
    const __amdgpu_buffer_rsrc_t rP = buf_rsrc3(a.pcur, a.nbytes);
    const __amdgpu_buffer_rsrc_t rO = buf_rsrc3(a.pold, a.nbytes);
    const __amdgpu_buffer_rsrc_t rV = buf_rsrc3(a.vdt2, a.nbytes);
    const __amdgpu_buffer_rsrc_t rS = buf_rsrc3(a.pool, a.pool_bytes);
    const uint32_t vcol = ((uint32_t)(k + kHY) * a.pitch + (uint32_t)(j + kHX)) * 4u;
    // planes holding a source of this wave's row and x tile (wave-uniform)
    int s_zlo = 1, s_zhi = 0;
    if (a.nsrc > 0 && k < ny) {
        s_zlo = ld_uniform3(a.src_zr + ((size_t)k * a.ntx + tx) * 2);
        s_zhi = ld_uniform3(a.src_zr + ((size_t)k * a.ntx + tx) * 2 + 1);
    }
    auto plane = [&](int i) -> uint32_t { return (uint32_t)(i + kHZ) * a.planeb; };
    int h_lds[HPT];
    uint32_t h_voff[HPT];
    halo_slots3<N, TY>(a, tid, x0, y0, h_lds, h_voff);

    const bool xstrip = x0 - N < nB || x0 + kTX + N - 1 > nx - nB;
    const bool ystrip = y0 - N < nB || y0 + TY + N - 1 > ny - nB;
    const bool zgroup = zbeg >= a.zsplit;  // chunk of the bottom z strip
    const int zb0 = nz - nB;
    const int nb2 = 2 * nB;
    auto xsi = [&](int x) { return x < nB ? x : x - (nx - nb2); };  // x strip index
    auto ysi = [&](int y) { return y < nB ? y : y - (ny - nb2); };  // y strip index
    auto in_xs = [&](int x) { return x >= 0 && x < nx && (x < nB || x >= nx - nB); };
    auto in_ys = [&](int y) { return y >= 0 && y < ny && (y < nB || y >= ny - nB); };
    // Byte offsets of this lane's strip values for plane i, as a lane part
    // (kOOB for a lane without a value, fixed for the sweep) plus a
    // wave-uniform part (kNone for a plane without values): kOOB + part,
    // lane + kNone and kOOB + kNone all stay past the pool's end (pools stay
    // below 2 GiB) without wrapping, so an offset costs one vector add.
    constexpr uint32_t kNone = 0x7fffffffu;
    const uint32_t l_zc = own ? ((uint32_t)k * nx + j) * 4u : kOOB;  // z cell arrays
    const uint32_t l_j = own ? (uint32_t)j * 4u : kOOB;              // z profile, y own cells
    const uint32_t l_jx = j < nx ? (uint32_t)j * 4u : kOOB;          // y half points
    uint32_t l_xh[2];                                                  // x half points
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const int hx = x0 - N + lane + half * kTX;
        l_xh[half] = (half == 0 || lane < L) && in_xs(hx) ? (uint32_t)xsi(hx) * 4u : kOOB;
    }
    const uint32_t l_xo = own && in_xs(j) ? (uint32_t)xsi(j) * 4u : kOOB;  // x own cells
    auto o_zh = [&](int i, uint32_t base, bool cell) -> uint32_t {  // half point i+N-1
        const int hz = i + N - 1;
        const uint32_t r = (uint32_t)(hz - zb0);
        const uint32_t u = hz >= zb0 && hz < nz ? base + r * (cell ? ny * nx : nx) * 4u : kNone;
        return (cell ? l_zc : l_j) + u;
    };
    auto o_zi = [&](int i, uint32_t base, bool cell) -> uint32_t {  // cell i
        const uint32_t r = (uint32_t)(i - zb0);
        const uint32_t u = i < zend && i >= zb0 ? base + r * (cell ? ny * nx : nx) * 4u : kNone;
        return (cell ? l_zc : l_j) + u;
    };
    auto o_xu = [&](int i, uint32_t base, bool cell) -> uint32_t {  // uniform part, x arrays
        return i < zend && k < ny
                   ? base + (cell ? ((uint32_t)i * ny + k) * nb2 : (uint32_t)i * nb2) * 4u
                   : kNone;
    };
    auto o_xh = [&](int i, int half, uint32_t base, bool cell) -> uint32_t {
        return l_xh[half] + o_xu(i, base, cell);
    };
    auto o_xo = [&](int i, uint32_t base, bool cell) -> uint32_t {
        return l_xo + o_xu(i, base, cell);
    };
    auto o_yh = [&](int i, int u, uint32_t base) -> uint32_t {
        const int q = ty + u * TY, hy = y0 - N + q;
        return l_jx + (i < zend && q < AYH && in_ys(hy)
                           ? base + ((uint32_t)i * nb2 + ysi(hy)) * nx * 4u : kNone);
    };
    auto o_yo = [&](int i, uint32_t base) -> uint32_t {
        return l_j + (i < zend && in_ys(k) ? base + ((uint32_t)i * nb2 + ysi(k)) * nx * 4u
                                           : kNone);
    };

    // this plane's strip values (loaded during the previous plane)
    float zph, zbh, zps, zbi, xph[2], xbh[2], xps, xbo, yph[QY], yps;
    auto load_strip = [&](int i) {
        if (zgroup) {
            zph = ldS(rS, o_zh(i, a.o_zphi_in, true));
            zbh = ldS(rS, o_zh(i, a.o_zbm1, false));
            zps = ldS(rS, o_zi(i, a.o_zpsi, true));
            zbi = ldS(rS, o_zi(i, a.o_zbm1, false));
        }
        if (xstrip) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                xph[half] = ldS(rS, o_xh(i, half, a.o_xphi_in, true));
                xbh[half] = ldS(rS, o_xh(i, half, a.o_xbm1, false));
            }
            xps = ldS(rS, o_xo(i, a.o_xpsi, true));
            xbo = ldS(rS, o_xo(i, a.o_xbm1, false));
        }
        if (ystrip) {
#pragma unroll
            for (int u = 0; u < QY; ++u) yph[u] = ldS(rS, o_yh(i, u, a.o_yphi_in));
            yps = ldS(rS, o_yo(i, a.o_ypsi));
        }
    };
    // y-strip b-1 values do not change along z: loaded once per block
    float ybh[QY], ybo = 0.f;
    if (ystrip) {
#pragma unroll
        for (int u = 0; u < QY; ++u) {
            const int q = ty + u * TY, hy = y0 - N + q;
            ybh[u] = ldS(rS, j < nx && q < AYH && in_ys(hy)
                                 ? a.o_ybm1 + ((uint32_t)ysi(hy) * nx + j) * 4u : kOOB);
        }
        ybo = ldS(rS, own && in_ys(k) ? a.o_ybm1 + ((uint32_t)ysi(k) * nx + j) * 4u : kOOB);
    }

    // z window p(i-L .. i+L) and A_z queue (half points i-N .. i+N-1)
    float zq[2 * L + 1], aq[2 * N];
#pragma unroll
    for (int r = 0; r < 2 * L; ++r) zq[r + 1] = ld3(rP, vcol, plane(zbeg - L + r));
#pragma unroll
    for (int q = 0; q < 2 * N - 1; ++q) {
        // half point h = zbeg-N+q uses p(h-N+1 .. h+N) = zq[1+q .. 1+q+2N-1]
        float d = 0.f;
#pragma unroll
        for (int m = 0; m < N; ++m) d = fmaf(a.cz[m], zq[1 + q + N + m] - zq[q + N - m], d);
        if (zgroup) {
            const int h = zbeg - N + q;
            const bool ok = own && h >= zb0 && h < nz;
            const uint32_t o = ok ? (((uint32_t)(h - zb0) * ny + k) * nx + j) * 4u : kOOB;
            const float ph = ldS(rS, ok ? a.o_zphi_in + o : kOOB);
            const float bh = ldS(rS, ok ? a.o_zbm1 + ((uint32_t)(h - zb0) * nx + j) * 4u : kOOB);
            const float phi = fmaf(bh, ph + d, ph);
            d += phi;
            stS(phi, rS, ok && h >= zbeg && h < zend ? a.o_zphi_out + o : kOOB);
        }
        aq[q + 1] = d;
    }

    // strip values of the first plane, then the prefetch ring in slot order
    load_strip(zbeg);
    // halo of planes zbeg .. zbeg+HL-1 into ring slots 0 .. HL-1, all loads
    // issued before the first write (the loop's first barrier orders these
    // writes before their reads)
    if constexpr (HL > 0) {
        float hv[HL][HPT];
#pragma unroll
        for (int q = 0; q < HL; ++q)
#pragma unroll
            for (int u = 0; u < HPT; ++u) hv[q][u] = ld3(rP, h_voff[u], plane(zbeg + q));
#pragma unroll
        for (int q = 0; q < HL; ++q) {
            float *Q = sp0 + q * (PX * PY);
#pragma unroll
            for (int u = 0; u < HPT; ++u)
                if (h_lds[u] >= 0) Q[h_lds[u]] = hv[q][u];
        }
    }
    float f_front[PD], f_po[PD], f_v2[PD], f_h[PD][HPT];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int d = 0; d < PD; ++d) {
        const int ip = min(zbeg + d, nz - 1);  // in bounds; unused past zend
        f_po[d] = ld3_nt(rO, vcol, plane(ip));
        f_v2[d] = ld3_nt(rV, vcol, plane(ip));
        f_front[d] = ld3(rP, vcol, plane(ip + L));
#pragma unroll
        for (int u = 0; u < HPT; ++u) f_h[d][u] = ld3(rP, h_voff[u], plane(ip + HL));
        __builtin_amdgcn_sched_barrier(0);
    }

    // ring slots of plane i and of plane i+HL (equal when HL is 0)
    int sI = 0, sH = HL % RG;
    for (int i0 = zbeg; i0 < zend; i0 += PD) {
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            const int i = i0 + d;
            const bool live = i < zend;
#pragma unroll
            for (int r = 0; r < 2 * L; ++r) zq[r] = zq[r + 1];
            asm volatile("v_mov_b32 %0, %1" : "=v"(zq[2 * L]) : "v"(f_front[d]));
#pragma unroll
            for (int q = 0; q < 2 * N - 1; ++q) aq[q] = aq[q + 1];

            // plane i into LDS
            float *P = sp0 + sI * (PX * PY);
            P[(ty + L) * PX + lane + L] = zq[L];
            float *PH = sp0 + sH * (PX * PY);
#pragma unroll
            for (int u = 0; u < HPT; ++u)
                if (h_lds[u] >= 0) PH[h_lds[u]] = f_h[d][u];
            __syncthreads();

            // z: new A_z half point i+N-1 from p(i .. i+2N-1) = zq[L .. L+2N-1],
            // then P_z = D-A_z + Psi_z
            {
                float dd = 0.f;
#pragma unroll
                for (int m = 0; m < N; ++m) dd = fmaf(a.cz[m], zq[L + N + m] - zq[L + N - 1 - m], dd);
                if (zgroup) {
                    const float phi = fmaf(zbh, zph + dd, zph);
                    dd += phi;
                    const int hz = i + N - 1;
                    stS(phi, rS, hz >= zbeg && hz < zend ? o_zh(i, a.o_zphi_out, true) : kOOB);
                }
                aq[2 * N - 1] = dd;
            }
            float pz = 0.f;
#pragma unroll
            for (int m = 0; m < N; ++m) pz = fmaf(a.cz[m], aq[N + m] - aq[N - 1 - m], pz);
            if (zgroup) {
                const float psi = fmaf(zbi, zps + pz, zps);
                stS(psi, rS, o_zi(i, a.o_zpsi, true));
                pz += psi;
            }

            const float *c = P + (ty + L) * PX + lane + L;
            float pxy;
            if (!xstrip && !ystrip) {
                // a z-strip chunk of an interior column tile: composite x and y
                float acc = a.w0 * c[0];
#pragma unroll
                for (int m = 1; m <= L; ++m) acc = fmaf(a.wx[m], c[m] + c[-m], acc);
#pragma unroll
                for (int m = 1; m <= L; ++m) acc = fmaf(a.wy[m], c[m * PX] + c[-m * PX], acc);
                pxy = fmaf(a.w0lo, c[0], acc);
            } else {
            float px, py;
            if (xstrip) {
                // A_x at half points hx = x0-N+q (q < 64+L) for row ty
                const float *row = P + (ty + L) * PX;
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const int q = lane + half * kTX;
                    if (half == 1 && lane >= L) break;
                    float dx = 0.f;
#pragma unroll
                    for (int m = 0; m < N; ++m) dx = fmaf(a.cx[m], row[q + N + m] - row[q + N - 1 - m], dx);
                    const float phi = fmaf(xbh[half], xph[half] + dx, xph[half]);
                    dx += phi;
                    const int hx = x0 - N + q;
                    stS(phi, rS, hx >= x0 && hx < x0 + kTX ? o_xh(i, half, a.o_xphi_out, true) : kOOB);
                    sax[ty * AXW + q] = dx;
                }
            }
            if (ystrip) {
                // A_y at half points hy = y0-N+q (q < TY+L) for column lane
#pragma unroll
                for (int u = 0; u < QY; ++u) {
                    const int q = ty + u * TY;
                    if (q >= AYH) break;
                    const float *cl = P + lane + L;
                    float dy = 0.f;
#pragma unroll
                    for (int m = 0; m < N; ++m)
                        dy = fmaf(a.cy[m], cl[(q + N + m) * PX] - cl[(q + N - 1 - m) * PX], dy);
                    const float phi = fmaf(ybh[u], yph[u] + dy, yph[u]);
                    dy += phi;
                    const int hy = y0 - N + q;
                    stS(phi, rS, hy >= y0 && hy < y0 + TY ? o_yh(i, u, a.o_yphi_out) : kOOB);
                    say[q * kTX + lane] = dy;
                }
            }
            // A_x rows are written and read by the same wavefront (row ty); A_y
            // columns cross wavefronts and need the block barrier
            if (ystrip) {
                __syncthreads();
            } else {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            if (xstrip) {
                px = 0.f;
                const float *ax = sax + ty * AXW + lane + N;
#pragma unroll
                for (int m = 0; m < N; ++m) px = fmaf(a.cx[m], ax[m] - ax[-1 - m], px);
                const float psi = fmaf(xbo, xps + px, xps);
                stS(psi, rS, o_xo(i, a.o_xpsi, true));
                px += psi;
            } else {
                float acc = a.w0x * c[0];
#pragma unroll
                for (int m = 1; m <= L; ++m) acc = fmaf(a.wx[m], c[m] + c[-m], acc);
                px = fmaf(a.w0xlo, c[0], acc);
            }
            if (ystrip) {
                py = 0.f;
                const float *ay = say + (ty + N) * kTX + lane;
#pragma unroll
                for (int m = 0; m < N; ++m) py = fmaf(a.cy[m], ay[m * kTX] - ay[(-1 - m) * kTX], py);
                const float psi = fmaf(ybo, yps + py, yps);
                stS(psi, rS, o_yo(i, a.o_ypsi));
                py += psi;
            } else {
                float acc = a.w0y * c[0];
#pragma unroll
                for (int m = 1; m <= L; ++m) acc = fmaf(a.wy[m], c[m * PX] + c[-m * PX], acc);
                py = fmaf(a.w0ylo, c[0], acc);
            }
            pxy = px + py;
            }

            float pn = 0.f;
            if (own && live) {
                float s = 0.f;
                if (a.nsrc > 0 && i >= s_zlo && i <= s_zhi) {
                    // the wave's bucket: one plane, one row, its 64-column tile
                    const int cell = (i * ny + k) * nx + j;
                    const int bk = (i * ny + k) * a.ntx + tx;
                    const int e0 = ld_uniform3(a.src_ptr + bk), e1 = ld_uniform3(a.src_ptr + bk + 1);
                    for (int e = e0; e < e1; ++e) {
                        const int cq = ld_uniform3(a.src_cell + e);
                        const float amp = ld_uniform3(a.src_amp + e);
                        if (cq == cell) s += amp;
                    }
                }
                const float pc = zq[L];
                pn = pc + fmaf(f_v2[d], (pz + pxy) + s, pc - f_po[d]);
                st3_nt(pn, rO, vcol, plane(i));
                if (a.rec && i == 0) a.rec[(size_t)k * nx + j] = pn;
                if (a.snap) a.snap[((size_t)i * ny + k) * nx + j] = pn;
            }
            // watchdog on the top (receiver) plane
            if (a.health && i == 0) report_nonfinite(a.health, own && live && nonfinite(pn), a.step);
            // next plane's strip values, then the ring refill with plane i+PD
            load_strip(i + 1);
            f_front[d] = ld3(rP, vcol, plane(i + PD + L));
            f_po[d] = ld3_nt(rO, vcol, plane(i + PD));
            f_v2[d] = ld3_nt(rV, vcol, plane(i + PD));
#pragma unroll
            for (int u = 0; u < HPT; ++u) f_h[d][u] = ld3(rP, h_voff[u], plane(i + PD + HL));
            sI = sI + 1 == RG ? 0 : sI + 1;
            sH = sH + 1 == RG ? 0 : sH + 1;
        }
    }
}

// Chunk tz of the z split: chunks of zc planes over [0, zsplit), then over
// [zsplit, nz).  Host and device share it.
__host__ __device__ inline void chunk_bounds3(int tz, int zc, int zsplit, int nzA, int nz,
                                              int *zbeg, int *zend) {
    if (tz < nzA) {
        *zbeg = tz * zc;
        *zend = *zbeg + zc < zsplit ? *zbeg + zc : zsplit;
    } else {
        *zbeg = zsplit + (tz - nzA) * zc;
        *zend = *zbeg + zc < nz ? *zbeg + zc : nz;
    }
}

// One time step.  Each block decodes its (x tile, y tile, z chunk) from the
// launch order and runs the strip or the interior sweep (a block-uniform
// branch: the two loops are compiled apart, so the interior loop carries none
// of the strip code's loads).
template <int N, int TY, int PD, int RG>
__global__ void __launch_bounds__(64 * TY) cpml3d_step(Cpml3dArgs a) {
    static_assert(PD >= 1 && PD <= kPD3max, "prefetch ring depth");
    using T = Tile3<N, TY>;
    __shared__ float sp[RG * T::PX * T::PY];
    __shared__ float sax[TY * T::AXW];
    __shared__ float say[T::AYH * kTX];
    int b = a.order ? ld_uniform3(a.order + blockIdx.x) : (int)blockIdx.x;
    const int tx = b % a.ntx;
    b /= a.ntx;
    const int tyi = b % a.nty;
    const int tz = b / a.nty;
    int zbeg, zend;
    chunk_bounds3(tz, a.zc, a.zsplit, a.nzA, a.nz, &zbeg, &zend);
    const int x0 = tx * kTX, y0 = tyi * TY, nB = a.nB;
    const bool strip = nB > 0 && (x0 - N < nB || x0 + kTX + N - 1 > a.nx - nB ||
                                  y0 - N < nB || y0 + TY + N - 1 > a.ny - nB || zbeg >= a.zsplit);
    if (strip)
        cpml3d_sweep_strip<N, TY, PD, RG>(a, sp, sax, say, tx, tyi, zbeg, zend);
    else
        cpml3d_sweep_inner<N, TY, PD, RG>(a, sp, tx, tyi, zbeg, zend);
}

// ---------------------------------------------------------------------------
static void composite_weights3(int N, const double *c, double *w) {
    const int L = 2 * N - 1;
    for (int m = 0; m <= 2 * L; ++m) w[m] = 0.0;
    for (int k = 0; k < N; ++k)
        for (int q = 0; q < N; ++q) {
            const double cc = c[k] * c[q];
            w[L + k + 1 + q] += cc;
            w[L + k - q] -= cc;
            w[L - k + q] -= cc;
            w[L - 1 - k - q] += cc;
        }
}

class Prop3D final : public Prop {
   public:
    int64_t pitch = 0, npy = 0, nzp = 0;
    DevBuf<float> p[2], vfield;
    DevBuf<int> src_zr;  // per (row, x tile) source plane range
    int pd_override = 0;  // SSSI_PD3
    bool ring_on = false;  // SSSI_RING3=1: LDS plane ring (Tile3::RING tiles)
    // damping strips: one pool; b-1 tables first, then the Phi/Psi state
    DevBuf<float> pool;
    size_t o_zbm1 = 0, o_xbm1 = 0, o_ybm1 = 0, o_zphi[2] = {0, 0}, o_zpsi = 0, o_xphi[2] = {0, 0},
           o_xpsi = 0, o_yphi[2] = {0, 0}, o_ypsi = 0, o_state = 0;  // in floats
    int cur = 0, phase = 0;
    float cz[8] = {0}, cx[8] = {0}, cy[8] = {0}, w0 = 0.f, w0lo = 0.f, wx[16] = {0}, wy[16] = {0};
    float w0x = 0.f, w0xlo = 0.f, w0y = 0.f, w0ylo = 0.f;
    int ty = 8, zc = 32;
    DevBuf<int> order;  // block launch order for (order_ty, order_zc)
    int order_ty = 0, order_zc = 0;
    DevBuf<int> health;

    Prop3D(const double *v, int64_t nz_, int64_t nx_, int64_t ny_, int N_, int nB_, double dz_,
           double dx_, double dy_, double dt_) {
        dim = 3;
        abc = SSSI_ABC_CPML;
        nz = nz_; nx = nx_; ny = ny_; N = N_; nB = nB_;
        dz = dz_; dx = dx_; dy = dy_; dt = dt_;
        if (nz * nx * ny > ((int64_t)1 << 31) / 2)
            fail(SSSI_EINVAL, "3-D grid too large (%lld cells)", (long long)(nz * nx * ny));
        SSSI_HIP(hipGetDevice(&device));
        stream = stream_acquire();
        {
            double vmax = 0.0;
            for (int64_t q = 0; q < nz * nx * ny; ++q) vmax = std::max(vmax, v[q]);
            courant = vmax * dt * std::sqrt(1.0 / (dz * dz) + 1.0 / (dx * dx) + 1.0 / (dy * dy));
        }
        health.alloc(1);
        if (const char *e = std::getenv("SSSI_TY3")) {
            const int t = std::atoi(e);
            if (t == 4 || t == 8 || t == 16) ty = t;
        }
        if (const char *e = std::getenv("SSSI_PD3")) pd_override = std::atoi(e);
        if (const char *e = std::getenv("SSSI_RING3")) ring_on = std::atoi(e) != 0;
        if (const char *e = std::getenv("SSSI_ZC3")) {
            const int z = std::atoi(e);
            if (z >= 8 && z <= 4096) zc = z;
        }
        pitch = round_up(kHX + round_up(nx, kTX) + 32, 64);
        npy = kHY + round_up(ny, 16) + kHY;
        nzp = kHZ + nz + kHZbot;
        const size_t np = (size_t)pitch * npy * nzp;
        if (np * 4 >= ((size_t)1 << 32)) fail(SSSI_EINVAL, "3-D padded grid exceeds 4 GiB");
        p[0].alloc(np);
        p[1].alloc(np);
        vfield.alloc(np);
        // v^2 dt^2 in the padded [z][y][x] layout, converted on the device
        upload_model(v, nz, nx, ny, dt, true, vfield.p, vfield.n, pitch * npy, pitch, kHZ, kHY,
                     kHX, stream);
        double c[8];
        staggered_coefficients(N, c);
        for (int q = 0; q < N; ++q) {
            cz[q] = (float)(c[q] / dz);
            cx[q] = (float)(c[q] / dx);
            cy[q] = (float)(c[q] / dy);
        }
        double w[32];
        composite_weights3(N, c, w);
        const int L = 2 * N - 1;
        double side = 0.0, sidex = 0.0, sidey = 0.0;  // exact in double
        for (int m = 1; m <= L; ++m) {
            wx[m] = (float)(w[L + m] / (dx * dx));
            wy[m] = (float)(w[L + m] / (dy * dy));
            side += 2.0 * ((double)wx[m] + (double)wy[m]);
            sidex += 2.0 * (double)wx[m];
            sidey += 2.0 * (double)wy[m];
        }
        // centre weight makes the FP32 weights sum to zero, carried as w0 + w0lo
        // (per direction for strip blocks that use one composite direction)
        w0 = (float)(-side);
        w0lo = (float)(-side - (double)w0);
        w0x = (float)(-sidex);
        w0xlo = (float)(-sidex - (double)w0x);
        w0y = (float)(-sidey);
        w0ylo = (float)(-sidey - (double)w0y);
        if (nB > 0) {
            // fwdTimeCpmlFor3dAw.m:46-70 with dampPml.m: u in samples, L = nB
            const double logR = std::log(1e-6), Lb = (double)nB;
            auto bm1 = [&](double vv, double u) {
                const double d0 = -(3 * vv) / (2 * Lb) * logR;
                const double q = (u / Lb) * (u / Lb);
                return (float)(std::exp(-(d0 * q) * dt) - 1.0);
            };
            auto V = [&](int64_t i, int64_t j, int64_t k) { return v[(k * nx + j) * nz + i]; };
            std::vector<float> hx((size_t)nz * 2 * nB), hy((size_t)2 * nB * nx), hz((size_t)nB * nx);
            for (int64_t i = 0; i < nz; ++i)
                for (int q = 0; q < nB; ++q) {
                    hx[(size_t)i * 2 * nB + q] = bm1(V(i, q, 0), (double)(nB - q));
                    hx[(size_t)i * 2 * nB + nB + q] = bm1(V(i, nx - nB + q, 0), (double)(q + 1));
                }
            for (int q = 0; q < nB; ++q)
                for (int64_t j = 0; j < nx; ++j) {
                    hy[(size_t)q * nx + j] = bm1(V(0, j, q), (double)(nB - q));
                    hy[(size_t)(nB + q) * nx + j] = bm1(V(0, j, ny - nB + q), (double)(q + 1));
                }
            for (int q = 0; q < nB; ++q)
                for (int64_t j = 0; j < nx; ++j)
                    hz[(size_t)q * nx + j] = bm1(V(nz - nB + q, j, 0), (double)(q + 1));
            const size_t zn = (size_t)nB * ny * nx, xn = (size_t)nz * ny * 2 * nB,
                         yn = (size_t)nz * 2 * nB * nx;
            size_t top = 0;
            auto take = [&](size_t n) { const size_t o = top; top += round_up((int64_t)n, 64); return o; };
            o_zbm1 = take(hz.size());
            o_xbm1 = take(hx.size());
            o_ybm1 = take(hy.size());
            o_state = top;
            for (int q = 0; q < 2; ++q) {
                o_zphi[q] = take(zn);
                o_xphi[q] = take(xn);
                o_yphi[q] = take(yn);
            }
            o_zpsi = take(zn);
            o_xpsi = take(xn);
            o_ypsi = take(yn);
            // byte offsets are 32-bit and must stay below the out-of-range marker
            if (top * 4 >= ((size_t)1 << 31))
                fail(SSSI_EINVAL, "3-D damping-strip state exceeds 2 GiB (nBoundary %d)", nB);
            pool.alloc(top);
            SSSI_HIP(hipMemcpy(pool.p + o_zbm1, hz.data(), hz.size() * 4, hipMemcpyHostToDevice));
            SSSI_HIP(hipMemcpy(pool.p + o_xbm1, hx.data(), hx.size() * 4, hipMemcpyHostToDevice));
            SSSI_HIP(hipMemcpy(pool.p + o_ybm1, hy.data(), hy.size() * 4, hipMemcpyHostToDevice));
        }
        reset();
        if (autotune_enabled()) {
            tune_layout();
            reset();
        }
    }

    // Autotuner: time every (ty, z chunk) candidate on this grid (a short
    // warm-up, then kTuneSteps steps timed with HIP events), keep the fastest,
    // cache it for the shape.  Runs before any source is set; the state is
    // zeroed afterwards.
    void tune_layout() {
        const TuneKey key{3, nz, nx, ny, N, nB};
        int a = 0, b = 0;
        // a cache entry is used only if it is a valid layout (as for SSSI_TY3 / SSSI_ZC3)
        if (tune_lookup(key, a, b) && (a == 4 || a == 8 || a == 16) && b >= 8 && b <= 4096) {
            ty = a;
            zc = b;
            return;
        }
        // z chunks: at 420 x 440 x 440, 8-row tiles, us per step at 48, 64, 80, 96,
        // 112 planes: 457.6, 461.7, 454.2, 472.8, 472.5 (one box)
        static const int cty[] = {4, 8, 16};
        static const int czc[] = {32, 48, 64, 80, 128};
        constexpr int kWarm = 2, kTuneSteps = 10;
        hipEvent_t e0, e1;
        SSSI_HIP(hipEventCreate(&e0));
        SSSI_HIP(hipEventCreate(&e1));
        float best_ms = 1e30f;
        int best_ty = ty, best_zc = zc;
        try {
            for (int t : cty)
                for (int z : czc) {
                    ty = t;
                    zc = z;
                    for (int q = 0; q < kWarm; ++q) step(q, false, nullptr);
                    SSSI_HIP(hipEventRecord(e0, stream));
                    for (int q = 0; q < kTuneSteps; ++q) step(kWarm + q, false, nullptr);
                    SSSI_HIP(hipEventRecord(e1, stream));
                    SSSI_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    SSSI_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best_ty = t;
                        best_zc = z;
                    }
                }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ty = best_ty;
        zc = best_zc;
        tune_store(key, ty, zc);
    }

    void layout(int *a, int *b) const override {
        *a = ty;
        *b = zc;
    }

    int64_t cells() const override { return nz * nx * ny; }

    void reset() override {
        for (auto *b : {&p[0], &p[1]}) SSSI_HIP(hipMemsetAsync(b->p, 0, b->bytes(), stream));
        if (pool.p)
            SSSI_HIP(hipMemsetAsync(pool.p + o_state, 0, (pool.n - o_state) * sizeof(float), stream));
        arm_health(health.p, stream);
        cur = 0;
        phase = 0;
        steps_done = 0;
        SSSI_HIP(hipStreamSynchronize(stream));
    }

    int *health_word() const override { return health.p; }

    void scan_field() override {
        launch_nonfinite_scan(p[cur].p, p[cur].n, health.p, (int)steps_done, stream);
    }

    void set_points(int n, const int *iz, const int *ix, const int *iy, const double *wavelet,
                    int64_t nt) override {
        if (n < 0 || nt < 1 || (n > 0 && (!iz || !ix || !iy || !wavelet)))
            fail(SSSI_EINVAL, "set_points: bad arguments");
        if ((double)n * (double)nt * 4 > 64e9)
            fail(SSSI_ENOMEM, "set_points: %d sources x %lld steps of amplitudes exceed 64 GB", n,
                 (long long)nt);
        for (int q = 0; q < n; ++q)
            if (iz[q] < 0 || iz[q] >= nz || ix[q] < 0 || ix[q] >= nx || iy[q] < 0 || iy[q] >= ny)
                fail(SSSI_EINVAL, "set_points: source %d outside the grid", q);
        nsrc = n;
        src_nt = nt;
        if (n == 0) {
            src_ptr.release();
            src_zr.release();
            src_cell.release();
            src_amp.release();
            return;
        }
        // bucket (i*ny + k)*ntx + tx by a stable counting sort (caller's order kept
        // within a bucket, so a cell's sources are summed in that order)
        const int64_t ntx = (nx + kTX - 1) / kTX, nbk = nz * ny * ntx;
        std::vector<int> ptr(nbk + 1, 0), bk(n), perm(n);
        for (int q = 0; q < n; ++q) {
            bk[q] = (int)(((int64_t)iz[q] * ny + iy[q]) * ntx + ix[q] / kTX);
            ++ptr[bk[q] + 1];
        }
        for (int64_t b = 0; b < nbk; ++b) ptr[b + 1] += ptr[b];
        {
            std::vector<int> fill(ptr.begin(), ptr.end() - 1);
            for (int q = 0; q < n; ++q) perm[fill[bk[q]]++] = q;
        }
        std::vector<int> cell(n);
        for (int e = 0; e < n; ++e) {
            const int q = perm[e];
            cell[e] = (int)(((int64_t)iz[q] * ny + iy[q]) * nx + ix[q]);
        }
        std::vector<float> amp((size_t)nt * n);
        for (int64_t t = 0; t < nt; ++t)
            for (int e = 0; e < n; ++e) amp[(size_t)t * n + e] = (float)wavelet[(size_t)t * n + perm[e]];
        // per (row, x tile): the first and last plane with a source
        std::vector<int> zr((size_t)ny * ntx * 2);
        for (size_t r = 0; r < zr.size(); r += 2) { zr[r] = INT32_MAX; zr[r + 1] = -1; }
        for (int q = 0; q < n; ++q) {
            int *z = &zr[((size_t)iy[q] * ntx + ix[q] / kTX) * 2];
            z[0] = std::min(z[0], iz[q]);
            z[1] = std::max(z[1], iz[q]);
        }
        src_zr.alloc(zr.size());
        SSSI_HIP(hipMemcpy(src_zr.p, zr.data(), src_zr.bytes(), hipMemcpyHostToDevice));
        src_ptr.alloc(ptr.size());
        src_cell.alloc(n);
        src_amp.alloc(amp.size());
        SSSI_HIP(hipMemcpy(src_ptr.p, ptr.data(), src_ptr.bytes(), hipMemcpyHostToDevice));
        SSSI_HIP(hipMemcpy(src_cell.p, cell.data(), src_cell.bytes(), hipMemcpyHostToDevice));
        SSSI_HIP(hipMemcpy(src_amp.p, amp.data(), src_amp.bytes(), hipMemcpyHostToDevice));
    }

    void set_surface(const double *, int64_t) override {
        fail(SSSI_ESTATE, "3-D handles run forward modelling only");
    }

    template <int NN, int TY, int RG>
    void launch_pd(const Cpml3dArgs &args, int nblk) {
        // ring depth: 4 planes where the registers allow it at occupancy 4;
        // SSSI_PD3=2 forces 2 (a shallower ring keeps fewer planes in flight)
        constexpr int PD = TY >= 8 && NN <= 4 ? 4 : 2;
        if constexpr (PD == 4) {
            if (pd_override == 2) {
                hipLaunchKernelGGL((cpml3d_step<NN, TY, 2, RG>), dim3(nblk), dim3(64 * TY), 0,
                                   stream, args);
                return;
            }
        }
        hipLaunchKernelGGL((cpml3d_step<NN, TY, PD, RG>), dim3(nblk), dim3(64 * TY), 0, stream,
                           args);
    }

    template <int NN, int TY>
    void launch(const Cpml3dArgs &args, int nblk) {
        // LDS plane buffer: 2 tiles, or Tile3::RING tiles with SSSI_RING3=1
        constexpr int RG = Tile3<NN, TY>::RING;
        if constexpr (RG > 2) {
            if (ring_on) {
                launch_pd<NN, TY, RG>(args, nblk);
                return;
            }
        }
        launch_pd<NN, TY, 2>(args, nblk);
    }

    // z split of the sweep: chunks above zsplit never reach the bottom strip
    int zsplit() const { return nB > 0 ? (int)std::max<int64_t>(0, nz - nB - N + 1) : (int)nz; }
    int nzA() const { return (zsplit() + zc - 1) / zc; }
    int nzc() const { return nzA() + (int)((nz - zsplit() + zc - 1) / zc); }

    // Blocks whose stencil reach touches a damping strip (x, y, or a chunk of
    // the bottom z strip) go first (they run the two-stage form and carry the
    // strip arrays' traffic); the others follow.  Both lists are split into 8
    // contiguous bands, block b landing on XCD b % 8.
    void build_order(int ntx, int nty, int nzc) {
        const int nblk = ntx * nty * nzc;
        const int zs = zsplit(), na = nzA();
        std::vector<int> strip, inner;
        std::vector<int> cat[8];  // strip blocks by which strips are in reach (x, y, z bits)
        for (int b = 0; b < nblk; ++b) {
            const int tx = b % ntx, tyi = (b / ntx) % nty, tz = b / (ntx * nty);
            const int64_t x0 = (int64_t)tx * kTX, y0 = (int64_t)tyi * ty;
            int zbeg, zend;
            chunk_bounds3(tz, zc, zs, na, (int)nz, &zbeg, &zend);
            const bool xs = nB > 0 && (x0 - N < nB || x0 + kTX + N - 1 > nx - nB);
            const bool ys = nB > 0 && (y0 - N < nB || y0 + ty + N - 1 > ny - nB);
            const bool zs_ = nB > 0 && zbeg >= zs;
            (xs || ys || zs_ ? strip : inner).push_back(b);
            if (xs || ys || zs_) cat[(xs ? 1 : 0) | (ys ? 2 : 0) | (zs_ ? 4 : 0)].push_back(b);
        }
        // each list split into 8 contiguous bands, block b of the launch taking
        // the next entry of XCD b % 8's band, so that the tiles one XCD runs at
        // the same time are neighbours and share their halo rows in its L2
        std::vector<int> ord;
        ord.reserve(nblk);
        auto append_banded = [&](const std::vector<int> &list) {
            const size_t n = list.size(), base = ord.size();
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
                ord.push_back(list[next[x]++]);
            }
        };
        // strip blocks: SSSI_STRIPBAND3 = 0 natural order (round robin over the
        // XCDs), 1 one banded list, 2 (default) banded per kind of strip in
        // reach, so each XCD gets its share of every kind (they differ in cost)
        const char *sb = std::getenv("SSSI_STRIPBAND3");
        const int mode = sb ? std::atoi(sb) : 2;
        if (mode == 0) {
            ord = strip;
        } else if (mode == 1) {
            append_banded(strip);
        } else {
            for (int c = 7; c >= 1; --c) append_banded(cat[c]);
        }
        append_banded(inner);
        SSSI_HIP(hipStreamSynchronize(stream));  // the previous order may be in use
        order.alloc(ord.size());
        SSSI_HIP(hipMemcpy(order.p, ord.data(), order.bytes(), hipMemcpyHostToDevice));
        order_ty = ty;
        order_zc = zc;
    }

    void step(int64_t t, bool record, float *snap) {
        Cpml3dArgs a{};
        a.pcur = p[cur].p;
        a.pold = p[cur ^ 1].p;
        a.vdt2 = vfield.p;
        a.nz = (int)nz; a.nx = (int)nx; a.ny = (int)ny;
        a.pitch = (int)pitch; a.npy = (int)npy; a.nB = nB;
        a.ntx = (int)((nx + kTX - 1) / kTX);
        a.nty = (int)((ny + ty - 1) / ty);
        a.zc = zc;
        a.zsplit = zsplit();
        a.nzA = nzA();
        a.rowb = (uint32_t)pitch * 4u;
        a.planeb = (uint32_t)(pitch * npy) * 4u;
        a.nbytes = (uint32_t)(p[0].bytes());
        for (int q = 0; q < 8; ++q) { a.cz[q] = cz[q]; a.cx[q] = cx[q]; a.cy[q] = cy[q]; }
        a.w0 = w0;
        a.w0lo = w0lo;
        a.w0x = w0x;
        a.w0xlo = w0xlo;
        a.w0y = w0y;
        a.w0ylo = w0ylo;
        for (int m = 0; m < 16; ++m) { a.wx[m] = wx[m]; a.wy[m] = wy[m]; }
        if (nB > 0) {
            auto ob = [](size_t o) { return (uint32_t)(o * 4); };
            a.pool = pool.p;
            a.pool_bytes = (uint32_t)pool.bytes();
            a.o_zphi_in = ob(o_zphi[phase]); a.o_zphi_out = ob(o_zphi[phase ^ 1]);
            a.o_zpsi = ob(o_zpsi); a.o_zbm1 = ob(o_zbm1);
            a.o_xphi_in = ob(o_xphi[phase]); a.o_xphi_out = ob(o_xphi[phase ^ 1]);
            a.o_xpsi = ob(o_xpsi); a.o_xbm1 = ob(o_xbm1);
            a.o_yphi_in = ob(o_yphi[phase]); a.o_yphi_out = ob(o_yphi[phase ^ 1]);
            a.o_ypsi = ob(o_ypsi); a.o_ybm1 = ob(o_ybm1);
        }
        if (nsrc > 0 && t < src_nt) {
            a.src_ptr = src_ptr.p;
            a.src_zr = src_zr.p;
            a.src_cell = src_cell.p;
            a.src_amp = src_amp.p + (size_t)t * nsrc;
            a.nsrc = nsrc;
        }
        if (record) a.rec = rec.p + (size_t)t * nx * ny;
        a.snap = snap;
        a.health = health.p;
        a.step = (int)steps_done;
        const int nblk = a.ntx * a.nty * nzc();
        if (order_ty != ty || order_zc != zc) build_order(a.ntx, a.nty, nzc());
        a.order = order.p;
        switch (N * 100 + ty) {
#define SSSI_C3(NN, T) case NN * 100 + T: launch<NN, T>(a, nblk); break;
            SSSI_C3(1, 4) SSSI_C3(1, 8) SSSI_C3(1, 16) SSSI_C3(2, 4) SSSI_C3(2, 8) SSSI_C3(2, 16)
            SSSI_C3(3, 4) SSSI_C3(3, 8) SSSI_C3(3, 16) SSSI_C3(4, 4) SSSI_C3(4, 8) SSSI_C3(4, 16)
            SSSI_C3(5, 4) SSSI_C3(5, 8) SSSI_C3(5, 16) SSSI_C3(6, 4) SSSI_C3(6, 8) SSSI_C3(6, 16)
            SSSI_C3(7, 4) SSSI_C3(7, 8) SSSI_C3(7, 16) SSSI_C3(8, 4) SSSI_C3(8, 8) SSSI_C3(8, 16)
#undef SSSI_C3
            default: fail(SSSI_EINVAL, "no 3-D kernel for nDiffOrder=%d ty=%d", N, ty);
        }
        SSSI_HIP(hipGetLastError());
        cur ^= 1;
        phase ^= 1;
    }

    void run(int64_t t_first, int64_t count, bool reverse, bool record, const SnapSink *sink,
             LaunchStats *st) override {
        if (reverse) fail(SSSI_ESTATE, "3-D handles run forward modelling only");
        if (record) ensure_rec(t_first + count, nx * ny);
        for (int64_t q = 0; q < count; ++q) {
            const int64_t t = t_first + q;
            float *snap = nullptr;
            if (sink && sink->dev && t % sink->stride == 0)
                snap = sink->dev + (size_t)(t / sink->stride - sink->slot0) * cells();
            step(t, record, snap);
            ++steps_done;
            if (st) st->launches++;
        }
    }

    void get_data(double *data, int64_t nt) override {
        if (nt > rec_nt || !rec.p) fail(SSSI_ESTATE, "get_data: no recorded traces");
        HostBuf<float> h((size_t)nt * nx * ny);
        SSSI_HIP(hipMemcpyAsync(h.p, rec.p, h.bytes(), hipMemcpyDeviceToHost, stream));
        SSSI_HIP(hipStreamSynchronize(stream));
        parallel_ranges((int64_t)h.n, [&](int64_t a, int64_t b) {
            for (int64_t q = a; q < b; ++q) data[q] = h.p[q];  // [t][k][j] == data(nx,ny,nt)
        }, 1 << 16);
    }

    void get_field(double *field) override {
        const size_t np = (size_t)pitch * npy * nzp;
        std::vector<float> h(np);
        SSSI_HIP(hipMemcpyAsync(h.data(), p[cur].p, np * sizeof(float), hipMemcpyDeviceToHost,
                                stream));
        SSSI_HIP(hipStreamSynchronize(stream));
        for (int64_t k = 0; k < ny; ++k)
            for (int64_t j = 0; j < nx; ++j)
                for (int64_t i = 0; i < nz; ++i)
                    field[(k * nx + j) * nz + i] =
                        h[((size_t)(i + kHZ) * npy + k + kHY) * pitch + j + kHX];
    }

    void get_model(double *) override { fail(SSSI_ESTATE, "3-D handles have no model output"); }

    void alg_bytes(double *bytes, double *ncells) override {
        const double c = (double)nz * nx * ny;
        const double strips = nB > 0 ? (double)nB * nx * ny + (double)nz * ny * 2 * nB +
                                           (double)nz * nx * 2 * nB
                                     : 0.0;
        *bytes = 16.0 * c + 20.0 * strips + 4.0 * nx * ny;
        *ncells = c;
    }
};

Prop *make_prop3d(const double *v, int64_t nz, int64_t nx, int64_t ny, int N, int nB, double dz,
                  double dx, double dy, double dt) {
    return new Prop3D(v, nz, nx, ny, N, nB, dz, dx, dy, dt);
}

}  // namespace sssi
