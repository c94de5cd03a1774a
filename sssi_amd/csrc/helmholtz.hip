// helmholtz.hip -- frequency-domain acoustic CPML Helmholtz solve on gfx950:
// freqCpmlFor2dAw (src/freqCpmlFor2dAw.m:43-133), A U = -S with A the
// (4k+1)-point impedance matrix (k = 2N-1) of squared slowness m, angular
// frequency w, CPML damping on the side and bottom strips.
//
// Matrix.  One thread per row cell computes the row's two complex damping
// factors (1j w / (s (1j w + d)))^2 in FP64 and writes the row into the band
// store (and, on request, CSR), in the reference's entry order.  Arithmetic
// is FP64 with contraction off, each operation rounded as the oracle rounds it.
//
// Solve.  A is banded in MATLAB's column-major cell order, kl = ku = k nz.
// The factorisation is LU with partial pivoting in LAPACK's band layout
// (A(i, j) at ab[kl + ku + i - j + j ldab], ldab = 2 kl + ku + 1; the top kl
// rows hold the fill of U), run as panels of kNB columns:
//   freq_panel   one workgroup factors the panel's columns in order: pivot =
//                first row of maximal |re| + |im| within kl below the
//                diagonal, row swap over the columns up to the running reach
//                of U, scaling by the reciprocal of the pivot, rank-1 update
//                of the panel's later columns;
//   freq_trail   one workgroup per later column within reach replays the
//                panel's swaps and updates on that column, in pivot order.
// Each element therefore receives the same operations in the same order as
// the unblocked algorithm, and the factors equal the oracle's bit for bit.
//   freq_solve   one workgroup per right-hand side: the forward sweep with
//                the interchanges, then the column-oriented backward sweep,
//                with the band window of the right-hand side kept in LDS.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"

namespace sssi {

struct cx {
    double re, im;
};

__device__ __forceinline__ cx cmul(cx a, cx b) {
#pragma clang fp contract(off)
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cx csub_mul(cx a, cx l, cx u) {  // a - l u
#pragma clang fp contract(off)
    return {a.re - (l.re * u.re - l.im * u.im), a.im - (l.re * u.im + l.im * u.re)};
}
__device__ __forceinline__ cx cdiv(cx a, cx b) {
#pragma clang fp contract(off)
    const double den = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
__device__ __forceinline__ cx crecip(cx b) {
#pragma clang fp contract(off)
    const double den = b.re * b.re + b.im * b.im;
    return {b.re / den, -b.im / den};
}

constexpr int kNB = 16;      // panel width (columns)
constexpr int kMaxOff = 31;  // 2k + 1 summed coefficients, k <= 15

struct BandDesc {
    cx *ab;
    int64_t n, kl, ku, ldab;
    __device__ __forceinline__ cx &at(int64_t i, int64_t j) const {
        return ab[(kl + ku + i - j) + j * ldab];
    }
};

// (1j w / (s (1j w + d)))^2: t = (s d, s w); q = (0, w) / t; q^2
__device__ __forceinline__ cx freq_factor(double w, double s, double d) {
#pragma clang fp contract(off)
    const double tr = s * d, ti = s * w;
    const double den = tr * tr + ti * ti;
    const double qr = w * ti / den, qi = w * tr / den;
    return {qr * qr - qi * qi, qr * qi + qi * qr};
}

struct FreqCoef {
    double c[kMaxOff];
};

// Row r = j nz + i of A: x neighbours left, z above, centre, z below, x right
// (ascending column order).  ab (band) and/or CSR (rowptr given) are written.
__global__ void __launch_bounds__(256) freq_assemble(const double *__restrict__ m,
                                                     const double *__restrict__ xd,
                                                     const double *__restrict__ zd, FreqCoef fc,
                                                     int nz, int nx, int k, double w, double dz,
                                                     double dx, BandDesc B,
                                                     const int64_t *__restrict__ rowptr,
                                                     int64_t *__restrict__ col,
                                                     cx *__restrict__ val) {
#pragma clang fp contract(off)
    const int64_t n = (int64_t)nz * nx;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(r / nz), i = (int)(r - (int64_t)j * nz);
        const cx fz = freq_factor(w, dz, zd[r]), fx = freq_factor(w, dx, xd[r]);
        int64_t e = rowptr ? rowptr[r] : 0;
        auto emit = [&](int64_t cidx, cx v) {
            if (B.ab) B.at(r, cidx) = v;
            if (rowptr) {
                col[e] = cidx;
                val[e] = v;
                ++e;
            }
        };
        for (int o = -k; o <= -1; ++o)
            if (j + o >= 0) emit(r + (int64_t)o * nz, cx{fc.c[o + k] * fx.re, fc.c[o + k] * fx.im});
        for (int o = -k; o <= -1; ++o)
            if (i + o >= 0) emit(r + o, cx{fc.c[o + k] * fz.re, fc.c[o + k] * fz.im});
        {
            const double mw = m[r] * (w * w);
            emit(r, cx{(mw + fc.c[k] * fz.re) + fc.c[k] * fx.re, fc.c[k] * fz.im + fc.c[k] * fx.im});
        }
        for (int o = 1; o <= k; ++o)
            if (i + o < nz) emit(r + o, cx{fc.c[o + k] * fz.re, fc.c[o + k] * fz.im});
        for (int o = 1; o <= k; ++o)
            if (j + o < nx) emit(r + (int64_t)o * nz, cx{fc.c[o + k] * fx.re, fc.c[o + k] * fx.im});
    }
}

// State of the factorisation carried between launches (device resident).
struct LuState {
    int64_t ju;        // running reach of U (last column touched)
    int singular;      // first zero pivot column + 1 (0: none)
};

// Factor panel columns j0 .. j0+kNB-1 (256 threads).  With LDS, the panel's
// rows j0 .. j0+kNB-1+kl (R of them) are staged in dynamic LDS, column-major,
// for the whole factorisation; entries outside the band are neither read nor
// written (none of the steps touches them).  Per column: pivot search (wave
// reductions), the decision by thread 0, the row swap, then each thread
// scales its rows of the column and applies them to the later columns (the
// multiplier stays in a register: one pass, no barrier between the two).
// Measured per panel of 16 columns at kl = 360: 62.7 us with 256 threads,
// 75.7 us with 1024 (rank-1 split over column groups), 75.1 us with the rows
// batched through registers.
template <bool LDS>
__global__ void __launch_bounds__(256) freq_panel(BandDesc B, int64_t j0, int64_t *__restrict__ piv,
                                                  int64_t *__restrict__ juv, LuState *st,
                                                  int64_t R) {
#pragma clang fp contract(off)
    extern __shared__ cx P[];
    __shared__ double s_best[4];
    __shared__ int s_arg[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t n = B.n, kl = B.kl, ku = B.ku;
    const int64_t jend = min(j0 + kNB, n);
    const int Ri = (int)R;
    auto inband = [&](int64_t i, int64_t j) {
        return i < n && j < n && i - j <= kl && j - i <= kl + ku;
    };
    auto A = [&](int64_t i, int64_t j) -> cx & {
        if constexpr (LDS) return P[(int)(j - j0) * Ri + (int)(i - j0)];
        else return B.at(i, j);
    };
    if constexpr (LDS) {
        // a row's kNB entries are loaded together, then stored to LDS: one
        // memory latency per row instead of one per entry
        for (int r = tid; r < Ri; r += 256) {
            cx v[kNB];
#pragma unroll
            for (int c = 0; c < kNB; ++c) {
                const int64_t i = j0 + r, j = j0 + c;
                v[c] = inband(i, j) ? B.at(i, j) : cx{0.0, 0.0};
            }
#pragma unroll
            for (int c = 0; c < kNB; ++c) P[c * Ri + r] = v[c];
        }
    }
    // furthest column any step so far reaches (kept by every thread alike)
    int64_t ju = st->ju;
    __syncthreads();
    for (int64_t j = j0; j < jend; ++j) {
        const int km = (int)min(kl, n - 1 - j);
        // pivot: first maximum of |re| + |im| over rows j .. j+km
        double best = -1.0;
        int arg = 0;
        for (int t = tid; t <= km; t += 256) {
            const cx a = A(j + t, j);
            const double s = fabs(a.re) + fabs(a.im);
            if (s > best) { best = s; arg = t; }
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int oa = __shfl_xor(arg, off);
            if (ob > best || (ob == best && oa < arg)) { best = ob; arg = oa; }
        }
        if (lane == 0) { s_best[wv] = best; s_arg[wv] = arg; }
        __syncthreads();
        // every thread combines the four waves' candidates (the same first
        // maximum), so no second barrier hands the pivot out
        double b = s_best[0];
        int a = s_arg[0];
        for (int q = 1; q < 4; ++q)
            if (s_best[q] > b || (s_best[q] == b && s_arg[q] < a)) { b = s_best[q]; a = s_arg[q]; }
        const int64_t reach = min(j + ku + a, n - 1);
        if (reach > ju) ju = reach;
        if (tid == 0) {
            if (b == 0.0 && st->singular == 0) st->singular = (int)min<int64_t>(j + 1, INT32_MAX);
            piv[j] = j + a;
            juv[j] = ju;
        }
        const int jp = a;
        const int64_t cend = min(jend - 1, ju);  // panel columns within reach
        if (jp != 0 && j + tid <= cend) {
            const int64_t c = j + tid;
            const cx t = A(j, c);
            A(j, c) = A(j + jp, c);
            A(j + jp, c) = t;
        }
        __syncthreads();
        if (km > 0) {
            const cx r = crecip(A(j, j));
            for (int t = 1 + tid; t <= km; t += 256) {
                const cx l = cmul(A(j + t, j), r);
                A(j + t, j) = l;
                // columns four at a time: their loads are independent and go
                // out together (each entry's update is the same operation)
                int64_t c = j + 1;
                for (; c + 3 <= cend; c += 4) {
                    cx a[4], u[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        a[q] = A(j + t, c + q);
                        u[q] = A(j, c + q);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) A(j + t, c + q) = csub_mul(a[q], l, u[q]);
                }
                for (; c <= cend; ++c) A(j + t, c) = csub_mul(A(j + t, c), l, A(j, c));
            }
        }
        __syncthreads();
    }
    if constexpr (LDS) {
        for (int r = tid; r < Ri; r += 256) {
            cx v[kNB];
#pragma unroll
            for (int c = 0; c < kNB; ++c) v[c] = P[c * Ri + r];
#pragma unroll
            for (int c = 0; c < kNB; ++c) {
                const int64_t i = j0 + r, j = j0 + c;
                if (inband(i, j)) B.at(i, j) = v[c];
            }
        }
    }
    if (tid == 0) st->ju = ju;
}

// Replay the panel's swaps and updates on column c = j0 + kNB + blockIdx.x,
// for the steps whose reach includes it (general form: global memory).
__global__ void __launch_bounds__(256) freq_trail(BandDesc B, int64_t j0,
                                                  const int64_t *__restrict__ piv,
                                                  const int64_t *__restrict__ juv) {
#pragma clang fp contract(off)
    const int64_t n = B.n, kl = B.kl;
    const int64_t c = j0 + kNB + blockIdx.x;
    const int64_t jend = min(j0 + kNB, n);
    if (c >= n || c > juv[jend - 1]) return;  // beyond every step's reach: untouched
    const int tid = threadIdx.x;
    for (int64_t p = j0; p < jend; ++p) {
        if (c > juv[p]) continue;  // block-uniform
        const int64_t jp = piv[p] - p;
        if (jp != 0 && tid == 0) {
            const cx t = B.at(p, c);
            B.at(p, c) = B.at(p + jp, c);
            B.at(p + jp, c) = t;
        }
        __syncthreads();
        const int64_t km = min(kl, n - 1 - p);
        const cx u = B.at(p, c);
        for (int64_t t = 1 + tid; t <= km; t += blockDim.x)
            B.at(p + t, c) = csub_mul(B.at(p + t, c), B.at(p + t, p), u);
        __syncthreads();
    }
}

// The same with the column's rows j0 .. j0+R-1 staged in LDS, and all of a
// thread's multipliers of the panel (rows tid + 256 q, q < RPT), the panel's
// pivots and reaches loaded up front, so the 16 steps touch only LDS and
// registers.
template <int RPT, int NC>
__global__ void __launch_bounds__(256) freq_trail_lds(BandDesc B, int64_t j0,
                                                      const int64_t *__restrict__ piv,
                                                      const int64_t *__restrict__ juv, int64_t R) {
#pragma clang fp contract(off)
    __shared__ cx Cl[NC][RPT * 256];
    __shared__ int64_t s_jv[kNB], s_pv[kNB];
    const int64_t n = B.n, kl = B.kl, ku = B.ku;
    const int64_t c0 = j0 + kNB + (int64_t)blockIdx.x * NC;  // columns c0 .. c0+NC-1
    const int64_t jend = min(j0 + kNB, n);
    const int tid = threadIdx.x;
    auto inband = [&](int64_t i, int64_t c) {
        return c < n && i < n && i - c <= kl && c - i <= kl + ku;
    };
    // The reaches, pivots and the columns' rows go out together; the panel's
    // multipliers are loaded one step ahead of their use (two register sets)
    // and serve all NC columns of the workgroup.
    const int64_t reach = juv[jend - 1];
    if (tid < kNB) {
        s_jv[tid] = j0 + tid < jend ? juv[j0 + tid] : -1;
        s_pv[tid] = j0 + tid < jend ? piv[j0 + tid] : j0 + tid;
    }
    cx col[NC][RPT];
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t r = tid + 256 * q;
            col[k][q] = r < R && inband(j0 + r, c0 + k) ? B.at(j0 + r, c0 + k) : cx{0.0, 0.0};
        }
    // multiplier of local row r at step p: L(j0 + r, p) when p < j0 + r <= p + km(p)
    auto lmul = [&](int pp, cx (&L)[RPT]) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t p = j0 + pp, i = j0 + tid + 256 * q;
            L[q] = (pp < kNB && p < jend && i > p && i <= p + min(kl, n - 1 - p)) ? B.at(i, p)
                                                                                : cx{0.0, 0.0};
        }
    };
    cx La[RPT], Lb[RPT];
    lmul(0, La);
    if (c0 >= n || c0 > reach) return;  // every column beyond every step's reach: untouched
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t r = tid + 256 * q;
            if (r < R) Cl[k][r] = col[k][q];
        }
    __syncthreads();
    auto step = [&](int pp, const cx (&L)[RPT], cx (&Ln)[RPT]) {
        lmul(pp + 1, Ln);
        const int64_t p = j0 + pp, jv = s_jv[pp];
        if (c0 > jv) return;  // block-uniform: no column of the block within reach
        const int64_t jp = s_pv[pp] - p;
        // column k within reach (block-uniform; the columns within reach are a prefix)
        if (jp != 0 && tid < NC && c0 + tid <= jv && c0 + tid < n) {
            cx *C = Cl[tid];
            const cx t = C[pp];
            C[pp] = C[pp + jp];
            C[pp + jp] = t;
        }
        __syncthreads();
        const int64_t km = min(kl, n - 1 - p);
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if (c0 + k > jv) break;
            const cx u = Cl[k][pp];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int64_t r = tid + 256 * q, i = j0 + r;
                if (i > p && i <= p + km) Cl[k][r] = csub_mul(Cl[k][r], L[q], u);
            }
        }
        __syncthreads();
    };
#pragma unroll 1
    for (int pp = 0; pp < kNB; pp += 2) {
        step(pp, La, Lb);
        step(pp + 1, Lb, La);
    }
#pragma unroll
    for (int k = 0; k < NC; ++k)
        for (int64_t r = tid; r < R; r += 256)
            if (inband(j0 + r, c0 + k)) B.at(j0 + r, c0 + k) = Cl[k][r];
}

// One right-hand side per workgroup of 1024: b = -s, forward sweep with
// interchanges (window b[j .. j+kl]), backward sweep (window b[j-kl-ku .. j]),
// the windows a power-of-two ring in LDS (slot of b[q] = q & mask).  The
// values entering the window come from global memory in chunks of 1024 per
// sweep into a double-buffered LDS stage, and each thread loads its
// multipliers of step j+1 while step j runs (two register sets, alternating).
constexpr int kStage = 1024;

template <int RPT>
__global__ void __launch_bounds__(1024) freq_solve(BandDesc B, const int64_t *__restrict__ piv,
                                                   const cx *__restrict__ S, cx *__restrict__ U,
                                                   int64_t mask) {
#pragma clang fp contract(off)
    extern __shared__ cx W[];
    cx *St = W + (mask + 1);  // 2 x kStage
    const int64_t n = B.n, kl = B.kl, ku = B.ku;
    const cx *s = S + (int64_t)blockIdx.x * n;
    cx *u = U + (int64_t)blockIdx.x * n;
    const int tid = threadIdx.x;
    cx la[RPT], lb[RPT];

    // ---- forward: step j's rows j+t, t = 1 + tid + 1024 q <= lm(j); entering b[j+kl+1]
    auto lf = [&](int64_t j, cx (&l)[RPT]) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t t = 1 + tid + 1024 * q;
            l[q] = (j < n && t <= min(kl, n - 1 - j)) ? B.at(j + t, j) : cx{0.0, 0.0};
        }
    };
    auto fwd = [&](int64_t j, const cx (&cur)[RPT], cx (&nxt)[RPT]) {
        if (j % kStage == 0) {  // stage b[j+kl+1 ..] of the next chunk
            const int64_t q = j + kl + 1 + tid;
            St[((j / kStage) & 1) * kStage + tid] = q < n ? cx{-s[q].re, -s[q].im} : cx{0.0, 0.0};
        }
        lf(j + 1, nxt);
        if (tid == 0) {
            const int64_t l = piv[j];
            if (l != j) {
                const cx t = W[l & mask];
                W[l & mask] = W[j & mask];
                W[j & mask] = t;
            }
        }
        __syncthreads();
        const cx bj = W[j & mask];
        const int64_t lm = min(kl, n - 1 - j);
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t t = 1 + tid + 1024 * q;
            if (t <= lm) W[(j + t) & mask] = csub_mul(W[(j + t) & mask], cur[q], bj);
        }
        __syncthreads();
        if (tid == 0) {
            u[j] = bj;
            const int64_t q = j + kl + 1;
            if (q < n) W[q & mask] = St[((j / kStage) & 1) * kStage + (j % kStage)];
        }
    };
    for (int64_t q = tid; q <= min(kl, n - 1); q += blockDim.x) W[q & mask] = cx{-s[q].re, -s[q].im};
    lf(0, la);
    __syncthreads();
    int64_t j = 0;
    for (; j + 1 < n; j += 2) {
        fwd(j, la, lb);
        fwd(j + 1, lb, la);
    }
    if (j < n) fwd(j, la, lb);
    __syncthreads();

    // ---- backward: step j's rows i = max(0, j-kl-ku) + tid + 1024 q < j; entering
    // b[j-kl-ku-1], staged in chunks counted from the top (k = n-1-j)
    auto ub = [&](int64_t jj, cx (&l)[RPT]) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t i = max<int64_t>(0, jj - kl - ku) + tid + 1024 * q;
            l[q] = (jj >= 0 && i < jj) ? B.at(i, jj) : cx{0.0, 0.0};
        }
    };
    auto bwd = [&](int64_t jj, const cx (&cur)[RPT], cx (&nxt)[RPT]) {
        const int64_t k = n - 1 - jj;
        if (k % kStage == 0) {
            const int64_t q = jj - kl - ku - 1 - tid;
            St[((k / kStage) & 1) * kStage + tid] = q >= 0 ? u[q] : cx{0.0, 0.0};
        }
        ub(jj - 1, nxt);
        if (tid == 0) W[jj & mask] = cdiv(W[jj & mask], B.at(jj, jj));
        __syncthreads();
        const cx bj = W[jj & mask];
        const int64_t i0 = max<int64_t>(0, jj - kl - ku);
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t i = i0 + tid + 1024 * q;
            if (i < jj) W[i & mask] = csub_mul(W[i & mask], cur[q], bj);
        }
        __syncthreads();
        if (tid == 0) {
            u[jj] = bj;
            const int64_t q = jj - kl - ku - 1;
            if (q >= 0) W[q & mask] = St[((k / kStage) & 1) * kStage + (k % kStage)];
        }
    };
    const int64_t w0 = max<int64_t>(0, n - 1 - kl - ku);
    for (int64_t q = w0 + tid; q < n; q += blockDim.x) W[q & mask] = u[q];
    ub(n - 1, la);
    __syncthreads();
    int64_t jj = n - 1;
    for (; jj - 1 >= 0; jj -= 2) {
        bwd(jj, la, lb);
        bwd(jj - 1, lb, la);
    }
    if (jj >= 0) bwd(jj, la, lb);
}

// ---------------------------------------------------------------------------
// Host side: damping and summed coefficients exactly as the oracle computes
// them (src/freqCpmlFor2dAw.m:52-89 with src/dampPml.m:16-18).
static void freq_host_setup(const double *m, int64_t nz, int64_t nx, int N, int nB, double dz,
                            double dx, std::vector<double> &xd, std::vector<double> &zd,
                            FreqCoef &fc) {
    std::vector<double> coeff(N);
    staggered_coefficients(N, coeff.data());
    const int k = 2 * N - 1;
    for (int o = 0; o < kMaxOff; ++o) fc.c[o] = 0;
    for (int ii = 1; ii <= N; ++ii)
        for (int jj = 1; jj <= N; ++jj) {
            const int o1 = ii + jj - 1, o2 = ii - jj;
            const double p = coeff[ii - 1] * coeff[jj - 1];
            fc.c[o1 + k] = fc.c[o1 + k] + p;
            fc.c[o2 + k] = fc.c[o2 + k] - p;
            fc.c[-o1 + k] = fc.c[-o1 + k] + p;
            fc.c[-o2 + k] = fc.c[-o2 + k] - p;
        }
    xd.assign(nz * nx, 0.0);
    zd.assign(nz * nx, 0.0);
    if (nB == 0) return;
    const double logR = std::log(1e-6);
    const double Lx = nB * dx, Lz = nB * dz;
    for (int j = 0; j < nB; ++j)
        for (int64_t i = 0; i < nz; ++i) {
            const int64_t jl = j, jr = nx - nB + j;
            const double ul = (nB - j) * dx, ur = (j + 1) * dx;
            const double vl = std::sqrt(1. / m[jl * nz + i]), vr = std::sqrt(1. / m[jr * nz + i]);
            const double d0l = -(3 * vl) / (2 * Lx) * logR, d0r = -(3 * vr) / (2 * Lx) * logR;
            const double sl = (ul / Lx) * (ul / Lx), sr = (ur / Lx) * (ur / Lx);
            xd[jl * nz + i] = d0l * sl;
            xd[jr * nz + i] = d0r * sr;
        }
    for (int64_t j = 0; j < nx; ++j)
        for (int i = 0; i < nB; ++i) {
            const int64_t ii = nz - nB + i;
            const double u = (i + 1) * dz;
            const double v = std::sqrt(1. / m[j * nz + ii]);
            const double d0 = -(3 * v) / (2 * Lz) * logR;
            zd[j * nz + ii] = d0 * ((u / Lz) * (u / Lz));
        }
}

static int64_t freq_nnz(int64_t nz, int64_t nx, int k, std::vector<int64_t> *rowptr) {
    std::vector<int64_t> cz(nz), cxn(nx);
    for (int64_t i = 0; i < nz; ++i) {
        int64_t t = 0;
        for (int o = -k; o <= k; ++o) t += (o != 0 && i + o >= 0 && i + o < nz);
        cz[i] = t;
    }
    for (int64_t j = 0; j < nx; ++j) {
        int64_t t = 0;
        for (int o = -k; o <= k; ++o) t += (o != 0 && j + o >= 0 && j + o < nx);
        cxn[j] = t;
    }
    int64_t e = 0;
    if (rowptr) rowptr->resize(nz * nx + 1);
    for (int64_t j = 0; j < nx; ++j)
        for (int64_t i = 0; i < nz; ++i) {
            if (rowptr) (*rowptr)[j * nz + i] = e;
            e += cz[i] + cxn[j] + 1;
        }
    if (rowptr) (*rowptr)[nz * nx] = e;
    return e;
}

static void freq_validate(const double *m, int64_t nz, int64_t nx, double w, int N, int nB,
                          double dz, double dx) {
    if (!m) fail(SSSI_EINVAL, "model must not be NULL");
    if (nz < 1 || nx < 1 || nz > (1 << 24) || nx > (1 << 24) || nz * nx > ((int64_t)1 << 28))
        fail(SSSI_EINVAL, "model %lldx%lld out of range", (long long)nz, (long long)nx);
    if (N < 1 || N > 8) fail(SSSI_EINVAL, "nDiffOrder must be in 1..8");
    if (nB < 0 || nB > nz || 2 * (int64_t)nB > nx)
        fail(SSSI_EINVAL, "nBoundary must satisfy 0 <= nBoundary <= nz and 2 nBoundary <= nx");
    if (!std::isfinite(w) || !(dz > 0) || !(dx > 0) || !std::isfinite(dz) || !std::isfinite(dx))
        fail(SSSI_EINVAL, "w must be finite and dz, dx positive");
    for (int64_t q = 0; q < nz * nx; ++q)
        if (!std::isfinite(m[q]) || !(m[q] > 0))
            fail(SSSI_EINVAL, "model (squared slowness) must be finite and positive (element %lld)",
                 (long long)q);
}

// Device matrix in band form (and CSR on request).
struct FreqSystem {
    int64_t nz, nx, n, kl, ku, ldab;
    int k;
    OwnedStream s;  // before the buffers: destroyed after them
    DevBuf<cx> ab;
    DevBuf<int64_t> piv, juv;
    DevBuf<LuState> st;

    FreqSystem(const double *m, int64_t nz_, int64_t nx_, double w, int N, int nB, double dz,
               double dx, bool band, int64_t *rowptr_out, int64_t *col_out, double *val_out)
        : nz(nz_), nx(nx_), n(nz_ * nx_), k(2 * N - 1) {
        kl = (int64_t)k * nz;
        ku = kl;
        ldab = 2 * kl + ku + 1;
        if (band && (double)ldab * n * sizeof(cx) > 200e9)
            fail(SSSI_ENOMEM, "band storage of %.1f GB exceeds the device budget",
                 (double)ldab * n * sizeof(cx) / 1e9);
        std::vector<double> xd, zd;
        FreqCoef fc;
        freq_host_setup(m, nz, nx, N, nB, dz, dx, xd, zd, fc);
        s.create();
        DevBuf<double> dm(n), dxd(n), dzd(n);
        SSSI_HIP(hipMemcpyAsync(dm.p, m, dm.bytes(), hipMemcpyHostToDevice, s));
        SSSI_HIP(hipMemcpyAsync(dxd.p, xd.data(), dxd.bytes(), hipMemcpyHostToDevice, s));
        SSSI_HIP(hipMemcpyAsync(dzd.p, zd.data(), dzd.bytes(), hipMemcpyHostToDevice, s));
        std::vector<int64_t> rp;
        DevBuf<int64_t> drp, dcol;
        DevBuf<cx> dval;
        if (rowptr_out) {
            const int64_t nnz = freq_nnz(nz, nx, k, &rp);
            drp.alloc(n + 1);
            dcol.alloc(nnz);
            dval.alloc(nnz);
            SSSI_HIP(hipMemcpyAsync(drp.p, rp.data(), drp.bytes(), hipMemcpyHostToDevice, s));
        }
        BandDesc B{nullptr, n, kl, ku, ldab};
        if (band) {
            ab.alloc((size_t)ldab * n);
            SSSI_HIP(hipMemsetAsync(ab.p, 0, ab.bytes(), s));
            B.ab = ab.p;
        }
        const unsigned nb = (unsigned)std::min<int64_t>(65535, (n + 255) / 256);
        hipLaunchKernelGGL(freq_assemble, dim3(nb), dim3(256), 0, s, dm.p, dxd.p, dzd.p, fc,
                           (int)nz, (int)nx, k, w, dz, dx, B, rowptr_out ? drp.p : nullptr,
                           dcol.p, dval.p);
        SSSI_HIP(hipGetLastError());
        if (rowptr_out) {
            std::memcpy(rowptr_out, rp.data(), rp.size() * sizeof(int64_t));
            SSSI_HIP(hipMemcpyAsync(col_out, dcol.p, dcol.bytes(), hipMemcpyDeviceToHost, s));
            SSSI_HIP(hipMemcpyAsync(val_out, dval.p, dval.bytes(), hipMemcpyDeviceToHost, s));
        }
        SSSI_HIP(hipStreamSynchronize(s));
    }

    // dynamic LDS a workgroup may use on the current device
    static size_t lds_limit() {
        int dev = 0;
        SSSI_HIP(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        SSSI_HIP(hipGetDeviceProperties(&prop, dev));
        return prop.sharedMemPerBlock;
    }

    float factor() {
        piv.alloc(n);
        juv.alloc(n);
        st.alloc(1);
        SSSI_HIP(hipMemsetAsync(st.p, 0, sizeof(LuState), s));
        BandDesc B{ab.p, n, kl, ku, ldab};
        hipEvent_t e0, e1;
        SSSI_HIP(hipEventCreate(&e0));
        SSSI_HIP(hipEventCreate(&e1));
        SSSI_HIP(hipEventRecord(e0, s));
        const int64_t R = std::min<int64_t>(n, kNB + kl);  // panel rows (and a column's)
        const size_t plds = (size_t)kNB * R * sizeof(cx);
        const bool panel_lds = plds + 1024 <= lds_limit();
        const int rpt = R <= 256 ? 1 : R <= 512 ? 2 : R <= 1024 ? 4 : 0;
        // columns per workgroup of the replay: two share one load of the
        // panel's multipliers (every workgroup reads the same panel, so fewer
        // workgroups contend less for its cache lines).  SSSI_FREQ_NC = 1 or 4
        // for the measured alternatives: 1050 replays take 67.2 ms at 2 columns,
        // 77.3 ms at 1 and 78.5 ms at 4 on the 120 x 140, nDiffOrder 2 workload.
        int trail_nc = 2;
        if (const char *e = std::getenv("SSSI_FREQ_NC")) {
            const int v = std::atoi(e);
            trail_nc = v == 4 ? 4 : v == 1 ? 1 : 2;
        }
        for (int64_t j0 = 0; j0 < n; j0 += kNB) {
            if (panel_lds)
                hipLaunchKernelGGL(freq_panel<true>, dim3(1), dim3(256), plds, s, B, j0, piv.p,
                                   juv.p, st.p, R);
            else
                hipLaunchKernelGGL(freq_panel<false>, dim3(1), dim3(256), 0, s, B, j0, piv.p,
                                   juv.p, st.p, R);
            const int64_t c0 = j0 + kNB, c1 = std::min(n - 1, j0 + kNB - 1 + kl + ku);
            if (c1 < c0) continue;
            const dim3 g((unsigned)(c1 - c0 + 1));
            const dim3 g2((unsigned)((c1 - c0 + 2) / 2)), g4((unsigned)((c1 - c0 + 4) / 4));
            switch (rpt * (trail_nc == 4 ? 100 : trail_nc == 2 ? 10 : 1)) {
                case 1: hipLaunchKernelGGL((freq_trail_lds<1, 1>), g, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                case 2: hipLaunchKernelGGL((freq_trail_lds<2, 1>), g, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                case 4: hipLaunchKernelGGL((freq_trail_lds<4, 1>), g, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                case 10: hipLaunchKernelGGL((freq_trail_lds<1, 2>), g2, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                case 20: hipLaunchKernelGGL((freq_trail_lds<2, 2>), g2, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                case 40: hipLaunchKernelGGL((freq_trail_lds<4, 2>), g2, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                case 100: hipLaunchKernelGGL((freq_trail_lds<1, 4>), g4, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                case 200: hipLaunchKernelGGL((freq_trail_lds<2, 4>), g4, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                case 400: hipLaunchKernelGGL((freq_trail_lds<4, 4>), g4, dim3(256), 0, s, B, j0, piv.p, juv.p, R); break;
                default: hipLaunchKernelGGL(freq_trail, g, dim3(256), 0, s, B, j0, piv.p, juv.p);
            }
        }
        SSSI_HIP(hipGetLastError());
        SSSI_HIP(hipEventRecord(e1, s));
        LuState h{};
        SSSI_HIP(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
        SSSI_HIP(hipStreamSynchronize(s));
        float ms = 0;
        SSSI_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (h.singular)
            fail(SSSI_EINVAL, "the impedance matrix is singular (zero pivot in column %d)",
                 h.singular);
        return ms;
    }

    float solve(const double *S, int nsrc, double *U) {
        int64_t ring = 1;
        while (ring < kl + ku + 1) ring <<= 1;
        const size_t sh = (ring + 2 * kStage) * sizeof(cx);
        if (sh > lds_limit())
            fail(SSSI_EINVAL, "band too wide for the solve's LDS window (kl = (2 nDiffOrder - 1) "
                 "nz = %lld; the window of 2 kl + 1 values, rounded up to a power of two, and "
                 "the staging must fit %zu bytes of LDS)", (long long)kl, lds_limit());
        const int64_t rows = kl + ku;  // per step, the backward sweep's is the wider
        const int rpt = rows <= 1024 ? 1 : rows <= 2048 ? 2 : rows <= 4096 ? 4 : 8;
        DevBuf<cx> dS((size_t)n * nsrc), dU((size_t)n * nsrc);
        SSSI_HIP(hipMemcpyAsync(dS.p, S, dS.bytes(), hipMemcpyHostToDevice, s));
        BandDesc B{ab.p, n, kl, ku, ldab};
        hipEvent_t e0, e1;
        SSSI_HIP(hipEventCreate(&e0));
        SSSI_HIP(hipEventCreate(&e1));
        SSSI_HIP(hipEventRecord(e0, s));
        const dim3 g((unsigned)nsrc), blk(1024);
        const int64_t mask = ring - 1;
        switch (rpt) {
            case 1: hipLaunchKernelGGL(freq_solve<1>, g, blk, sh, s, B, piv.p, dS.p, dU.p, mask); break;
            case 2: hipLaunchKernelGGL(freq_solve<2>, g, blk, sh, s, B, piv.p, dS.p, dU.p, mask); break;
            case 4: hipLaunchKernelGGL(freq_solve<4>, g, blk, sh, s, B, piv.p, dS.p, dU.p, mask); break;
            default: hipLaunchKernelGGL(freq_solve<8>, g, blk, sh, s, B, piv.p, dS.p, dU.p, mask);
        }
        SSSI_HIP(hipGetLastError());
        SSSI_HIP(hipEventRecord(e1, s));
        SSSI_HIP(hipMemcpyAsync(U, dU.p, dU.bytes(), hipMemcpyDeviceToHost, s));
        SSSI_HIP(hipStreamSynchronize(s));
        float ms = 0;
        SSSI_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return ms;
    }
};

}  // namespace sssi

using namespace sssi;

namespace {
template <class F>
int fguard(F &&f) {
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
thread_local double g_factor_ms = 0, g_solve_ms = 0;
}  // namespace

extern "C" {

int sssi_freq_cpml_2d_matrix(const double *model, int64_t nz, int64_t nx, double w,
                             int nDiffOrder, int nBoundary, double dz, double dx, int64_t *nnz,
                             int64_t *rowptr, int64_t *col, double *val) {
    return fguard([&] {
        freq_validate(model, nz, nx, w, nDiffOrder, nBoundary, dz, dx);
        if (!nnz) fail(SSSI_EINVAL, "nnz must not be NULL");
        *nnz = freq_nnz(nz, nx, 2 * nDiffOrder - 1, nullptr);
        if (!rowptr) return;
        if (!col || !val) fail(SSSI_EINVAL, "col and val must not be NULL with rowptr");
        FreqSystem fs(model, nz, nx, w, nDiffOrder, nBoundary, dz, dx, false, rowptr, col, val);
    });
}

int sssi_freqCpmlFor2dAw(const double *model, int64_t nz, int64_t nx, const double *source,
                         int nsrc, double w, int nDiffOrder, int nBoundary, double dz, double dx,
                         double *snapshot) {
    return fguard([&] {
        freq_validate(model, nz, nx, w, nDiffOrder, nBoundary, dz, dx);
        if (nsrc < 1 || !source || !snapshot)
            fail(SSSI_EINVAL, "source and snapshot must hold at least one column");
        FreqSystem fs(model, nz, nx, w, nDiffOrder, nBoundary, dz, dx, true, nullptr, nullptr,
                      nullptr);
        g_factor_ms = fs.factor();
        g_solve_ms = fs.solve(source, nsrc, snapshot);
    });
}

int sssi_freq_last_timing(double *factor_ms, double *solve_ms) {
    if (factor_ms) *factor_ms = g_factor_ms;
    if (solve_ms) *solve_ms = g_solve_ms;
    return SSSI_OK;
}

}  // extern "C"
