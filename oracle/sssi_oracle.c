/*
 * sssi_oracle.c -- FP64 CPU restatement of the SSSI acoustic propagators.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped).
 *
 * Every function restates the reference algorithm in the same floating-point
 * operation order, so that a build with -ffp-contract=off (no FMA) reproduces
 * the reference's FP64 arithmetic step for step.  Loops are fused per step
 * (the reference makes six full-grid derivative passes and two memcpy per
 * step; here each derived quantity is computed once, in the same order of
 * operations), and OpenMP splits independent columns; the per-cell result does
 * not depend on the thread count.
 *
 * Parity status (how each function is pinned; the checks live in
 * tests/test_oracle.py; each one fails under a planted index mistake):
 *   orc_dcoef            - pinned: coefficient table of UserGuide.tex:339-341
 *                          (N=1..3) and the exact rational solution of the
 *                          Vandermonde system for N=1..8.
 *   orc_cpml_profiles_2d - pinned: closed form of dampPml and its invariants.
 *   orc_fwd_cpml_2d      - pinned: interior = composite stencil from exact
 *                          rational weights, light cone, linearity, mirror
 *                          symmetry, analytic 2-D Green's function, CPML
 *                          absorption; strip conventions (b(h) at half point
 *                          h+1/2, b(i) at i, which rows/columns are damped) by
 *                          closed-form impulse responses in every strip and by
 *                          the exact step and row/column at which each strip
 *                          first changes the field.
 *   orc_rvs_cpml_2d      - pinned: equals orc_fwd_cpml_2d driven by the
 *                          time-reversed surface data (bit for bit).
 *   orc_fwd_cpml_3d      - pinned: bit-exact equality with the 2-D path for
 *                          y-invariant problems inside the y-strip light cone;
 *                          closed-form impulse responses in the x, y and z
 *                          strips with the index-unit profiles.
 *   orc_fwd_sponge_2d    - pinned: closed-form one-, two- and three-step
 *                          impulse responses in every region (side strips,
 *                          bottom strip, corners, interior).
 *   orc_rvs_sponge_2d    - pinned: equals orc_fwd_sponge_2d driven by the
 *                          time-reversed surface data (bit for bit), model =
 *                          damped previous level.
 *   orc_rtm_shot_cpml_2d / orc_xcorr_image - pinned: brute-force sums over
 *                          stored snapshots and composition of the above.
 *   Parity unpinned: the last-bit FP64 summation order of the reference
 *   C-MEX (it cannot be built here) and MATLAB's mldivide (replaced by the
 *   closed-form rationals).
 */
#include "sssi_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EINVAL (-1)
#define ORC_ENOMEM (-2)

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static int64_t gcd64(int64_t a, int64_t b)
{
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a;
}

/* dCoef(N,'s') -- finiteDifference.c:24-82 / src/dCoef.m:34-48 solve
 *   sum_j (2j-1)^(2i-1) c_j = delta_{i1}.
 * With a_j = (2j-1) c_j this is a Vandermonde system in x_j = (2j-1)^2 whose
 * solution is the Lagrange basis at 0:  a_j = prod_{m!=j} x_m / (x_m - x_j).
 * The rational is reduced exactly (numerator and denominator stay below 2^53
 * for N <= 8) and rounded once by the division.  The reference obtains the
 * same numbers through MATLAB's mldivide (LAPACK), whose last-bit rounding
 * is not reproducible here. */
int orc_dcoef(int N, double *c)
{
    if (N < 1 || N > 8 || !c) return ORC_EINVAL;
    for (int j = 1; j <= N; ++j) {
        int64_t num = 1, den = 2 * j - 1;
        int neg = 0;
        for (int m = 1; m <= N; ++m) {
            if (m == j) continue;
            int64_t a = (int64_t)(2 * m - 1) * (2 * m - 1);
            int64_t b = a - (int64_t)(2 * j - 1) * (2 * j - 1);
            if (b < 0) { neg ^= 1; b = -b; }
            int64_t g1 = gcd64(a, den), g2 = gcd64(num, b);
            num = (num / g2) * (a / g1);
            den = (den / g1) * (b / g2);
            int64_t g = gcd64(num, den);
            num /= g; den /= g;
        }
        c[j - 1] = (neg ? -1.0 : 1.0) * ((double)num / (double)den);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 2-D CPML damping: fwdTimeCpmlFor2dAw_mex.c:100-145 with dampPml
 * finiteDifference.c:253-280.  d0 = -(3v)/(2L)*log(R); d = d0*(u/L)*(u/L);
 * b = exp(-d*dt).  Left u=(nB-j)dx, right u=(j+1)dx over the last nB columns,
 * bottom u=(i+1)dz over the last nB rows; no top layer. */
int orc_cpml_profiles_2d(const double *v, int64_t nz, int64_t nx, int nB,
                         double dz, double dx, double dt, double *xb, double *zb)
{
    if (!v || !xb || !zb || nz <= 0 || nx <= 0 || nB < 0 || 2 * (int64_t)nB > nx || nB > nz)
        return ORC_EINVAL;
    const double logR = log(1e-6);
    for (int64_t q = 0; q < nz * nx; ++q) { xb[q] = exp(-0.0 * dt); zb[q] = exp(-0.0 * dt); }
    if (nB == 0) return 0;
    {
        const double L = nB * dx;
        for (int j = 0; j < nB; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                double u = (nB - j) * dx;
                double d0 = -(3.0 * v[(int64_t)j * nz + i]) / (2 * L) * logR;
                double d = d0 * (u / L) * (u / L);
                xb[(int64_t)j * nz + i] = exp(-d * dt);
            }
        for (int j = 0; j < nB; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                int64_t jj = nx - nB + j;
                double u = (j + 1) * dx;
                double d0 = -(3.0 * v[jj * nz + i]) / (2 * L) * logR;
                double d = d0 * (u / L) * (u / L);
                xb[jj * nz + i] = exp(-d * dt);
            }
    }
    {
        const double L = nB * dz;
        for (int64_t j = 0; j < nx; ++j)
            for (int i = 0; i < nB; ++i) {
                int64_t ii = nz - nB + i;
                double u = (i + 1) * dz;
                double d0 = -(3.0 * v[j * nz + ii]) / (2 * L) * logR;
                double d = d0 * (u / L) * (u / L);
                zb[j * nz + ii] = exp(-d * dt);
            }
    }
    return 0;
}

/* 3-D CPML damping, src/fwdTimeCpmlFor3dAw.m:46-70 with src/dampPml.m:16-18.
 * The reference passes u in grid-index units and L = nBoundary (not times the
 * spacing); the x and z profiles take v from the first y slice, the y profile
 * from the top z row, and each is replicated along the remaining axis.
 * dampPml.m computes d0 .* (u/L).^2 (square first, then multiply). */
int orc_cpml_profiles_3d(const double *v, int64_t nz, int64_t nx, int64_t ny, int nB,
                         double dt, double *xb, double *yb, double *zb)
{
    if (!v || !xb || !yb || !zb || nz <= 0 || nx <= 0 || ny <= 0 || nB < 0 ||
        2 * (int64_t)nB > nx || 2 * (int64_t)nB > ny || nB > nz)
        return ORC_EINVAL;
    const int64_t n = nz * nx * ny;
    for (int64_t q = 0; q < n; ++q) { xb[q] = exp(-0.0 * dt); yb[q] = xb[q]; zb[q] = xb[q]; }
    if (nB == 0) return 0;
    const double logR = log(1e-6);
    const double L = (double)nB;
#define V3(i, j, k) v[((int64_t)(k) * nx + (j)) * nz + (i)]
#define IDX3(i, j, k) (((int64_t)(k) * nx + (j)) * nz + (i))
    for (int64_t k = 0; k < ny; ++k)
        for (int j = 0; j < nB; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                double ul = (double)(nB - j), ur = (double)(j + 1);
                double dl = -(3 * V3(i, j, 0)) / (2 * L) * logR;
                double dr = -(3 * V3(i, nx - nB + j, 0)) / (2 * L) * logR;
                double ql = (ul / L) * (ul / L), qr = (ur / L) * (ur / L);
                xb[IDX3(i, j, k)] = exp(-(dl * ql) * dt);
                xb[IDX3(i, nx - nB + j, k)] = exp(-(dr * qr) * dt);
            }
    for (int kk = 0; kk < nB; ++kk)
        for (int64_t j = 0; j < nx; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                double uf = (double)(nB - kk), ur = (double)(kk + 1);
                double df = -(3 * V3(0, j, kk)) / (2 * L) * logR;
                double dr = -(3 * V3(0, j, ny - nB + kk)) / (2 * L) * logR;
                double qf = (uf / L) * (uf / L), qr = (ur / L) * (ur / L);
                yb[IDX3(i, j, kk)] = exp(-(df * qf) * dt);
                yb[IDX3(i, j, ny - nB + kk)] = exp(-(dr * qr) * dt);
            }
    for (int64_t k = 0; k < ny; ++k)
        for (int64_t j = 0; j < nx; ++j)
            for (int ii = 0; ii < nB; ++ii) {
                double u = (double)(ii + 1);
                double d = -(3 * V3(nz - nB + ii, j, 0)) / (2 * L) * logR;
                double q = (u / L) * (u / L);
                zb[IDX3(nz - nB + ii, j, k)] = exp(-(d * q) * dt);
            }
#undef V3
#undef IDX3
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 2-D CPML state.  Arrays have the shapes of fwdTimeCpmlFor2dAw_mex.c:152-175:
 *   p (old/cur/new)  (nz+2l) x (nx+2l)     p[j*npz + ip]
 *   zPhi, zA         (nz+2l) x nx          [j*npz + ip]
 *   xPhi, xA         nz x (nx+2l)          [jp*nz + i]
 *   zPsi             (nz+l) x nx           [j*(nz+l) + ip], rows l..nz+l-1 used
 *   xPsi             nz x (nx+l)           [jp*nz + i],  cols l..nx+l-1 used */
typedef struct {
    int64_t nz, nx, npz, npx;
    int N, l;
    double c[8], dz, dx;
    double *pold, *pcur, *pnew;
    double *zPhi, *zA, *xPhi, *xA, *zPsi, *xPsi;
    double *xb, *zb, *vdtSq, *src;
} cpml2d_t;

static void cpml2d_free(cpml2d_t *s)
{
    free(s->pold); free(s->pcur); free(s->pnew);
    free(s->zPhi); free(s->zA); free(s->xPhi); free(s->xA); free(s->zPsi); free(s->xPsi);
    free(s->xb); free(s->zb); free(s->vdtSq); free(s->src);
}

static int cpml2d_init(cpml2d_t *s, const double *v, int64_t nz, int64_t nx, int N, int nB,
                       double dz, double dx, double dt)
{
    memset(s, 0, sizeof(*s));
    if (!v || nz <= 0 || nx <= 0 || N < 1 || N > 8 || nB < 0 || 2 * (int64_t)nB > nx || nB > nz)
        return ORC_EINVAL;
    s->nz = nz; s->nx = nx; s->N = N; s->l = 2 * N - 1;
    s->npz = nz + 2 * s->l; s->npx = nx + 2 * s->l;
    s->dz = dz; s->dx = dx;
    orc_dcoef(N, s->c);
    const int64_t l = s->l;
    s->pold = calloc(s->npz * s->npx, sizeof(double));
    s->pcur = calloc(s->npz * s->npx, sizeof(double));
    s->pnew = calloc(s->npz * s->npx, sizeof(double));
    s->zPhi = calloc(s->npz * nx, sizeof(double));
    s->zA   = calloc(s->npz * nx, sizeof(double));
    s->xPhi = calloc(nz * s->npx, sizeof(double));
    s->xA   = calloc(nz * s->npx, sizeof(double));
    s->zPsi = calloc((nz + l) * nx, sizeof(double));
    s->xPsi = calloc(nz * (nx + l), sizeof(double));
    s->xb = malloc(nz * nx * sizeof(double));
    s->zb = malloc(nz * nx * sizeof(double));
    s->vdtSq = malloc(nz * nx * sizeof(double));
    s->src = calloc(nz * nx, sizeof(double));
    if (!s->pold || !s->pcur || !s->pnew || !s->zPhi || !s->zA || !s->xPhi || !s->xA ||
        !s->zPsi || !s->xPsi || !s->xb || !s->zb || !s->vdtSq || !s->src) {
        cpml2d_free(s);
        return ORC_ENOMEM;
    }
    orc_cpml_profiles_2d(v, nz, nx, nB, dz, dx, dt, s->xb, s->zb);
    /* fwdTimeCpmlFor2dAw_mex.c:165-168 */
    for (int64_t q = 0; q < nz * nx; ++q) s->vdtSq[q] = (v[q] * dt) * (v[q] * dt);
    return 0;
}

/* One time step, fwdTimeCpmlFor2dAw_mex.c:185-270 (rvs…_mex.c:180-265).
 * s->src holds source(:,:,t) as an nz x nx slice.  Each derivative is summed
 * over k = 0..N-1 as acc = acc + c[k]*(a-b)/d, as diffOperator2d does
 * (finiteDifference.c:100-147). */
static void cpml2d_step(cpml2d_t *s)
{
    const int64_t nz = s->nz, nx = s->nx, npz = s->npz, npx = s->npx;
    const int N = s->N;
    const int64_t l = s->l;
    const double *c = s->c, dz = s->dz, dx = s->dx;
    const double *P = s->pcur;

    /* zPhi (izi) and zA (izl): half points ip in [N-1, nz+2l-N-1] */
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nx; ++j) {
        const double *col = P + (j + l) * npz;
        for (int64_t ip = N - 1; ip <= npz - N - 1; ++ip) {
            double D = 0.0;
            for (int k = 0; k < N; ++k) D = D + c[k] * (col[ip + 1 + k] - col[ip - k]) / dz;
            double *phi = s->zPhi + j * npz + ip;
            if (ip >= l && ip < nz + l) {
                double b = s->zb[j * nz + (ip - l)];
                *phi = b * (*phi) + (b - 1) * D;
            }
            s->zA[j * npz + ip] = D + *phi;
        }
    }
    /* xPhi (ixi) and xA (ixl) */
    #pragma omp parallel for schedule(static)
    for (int64_t jp = N - 1; jp <= npx - N - 1; ++jp) {
        for (int64_t i = 0; i < nz; ++i) {
            double D = 0.0;
            for (int k = 0; k < N; ++k)
                D = D + c[k] * (P[(jp + 1 + k) * npz + i + l] - P[(jp - k) * npz + i + l]) / dx;
            double *phi = s->xPhi + jp * nz + i;
            if (jp >= l && jp < nx + l) {
                double b = s->xb[(jp - l) * nz + i];
                *phi = b * (*phi) + (b - 1) * D;
            }
            s->xA[jp * nz + i] = D + *phi;
        }
    }
    /* zPsi, xPsi, zP, xP and the update (fwd…_mex.c:225-263) */
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nx; ++j) {
        for (int64_t i = 0; i < nz; ++i) {
            const int64_t ip = i + l, jp = j + l;
            const double *zA = s->zA + j * npz;
            double Dz = 0.0;
            for (int k = 0; k < N; ++k) Dz = Dz + c[k] * (zA[ip + k] - zA[ip - 1 - k]) / dz;
            double bz = s->zb[j * nz + i];
            double *zpsi = s->zPsi + j * (nz + l) + ip;
            *zpsi = bz * (*zpsi) + (bz - 1) * Dz;
            double zP = Dz + *zpsi;

            double Dx = 0.0;
            for (int k = 0; k < N; ++k)
                Dx = Dx + c[k] * (s->xA[(jp + k) * nz + i] - s->xA[(jp - 1 - k) * nz + i]) / dx;
            double bx = s->xb[j * nz + i];
            double *xpsi = s->xPsi + jp * nz + i;
            *xpsi = bx * (*xpsi) + (bx - 1) * Dx;
            double xP = Dx + *xpsi;

            s->pnew[jp * npz + ip] = s->vdtSq[j * nz + i] * (zP + xP + s->src[j * nz + i]) +
                                     2 * P[jp * npz + ip] - s->pold[jp * npz + ip];
        }
    }
    /* rotate: old <- cur, cur <- new (fwd…_mex.c:267-270) */
    double *t = s->pold;
    s->pold = s->pcur;
    s->pcur = s->pnew;
    s->pnew = t;
}

static void fill_point_sources(double *slice, int64_t nz, int nsrc, const int *iz,
                               const int *ix, const double *wav_t, int64_t nz_nx)
{
    memset(slice, 0, nz_nx * sizeof(double));
    for (int q = 0; q < nsrc; ++q) slice[(int64_t)ix[q] * nz + iz[q]] += wav_t[q];
}

static int check_points(int nsrc, const int *iz, const int *ix, const int *iy,
                        int64_t nz, int64_t nx, int64_t ny, const double *wavelet)
{
    if (nsrc < 0) return ORC_EINVAL;
    if (nsrc == 0) return 0;
    if (!iz || !ix || !wavelet || (ny > 0 && !iy)) return ORC_EINVAL;
    for (int q = 0; q < nsrc; ++q) {
        if (iz[q] < 0 || iz[q] >= nz || ix[q] < 0 || ix[q] >= nx) return ORC_EINVAL;
        if (ny > 0 && (iy[q] < 0 || iy[q] >= ny)) return ORC_EINVAL;
    }
    return 0;
}

int orc_fwd_cpml_2d(const double *v, int64_t nz, int64_t nx, int64_t nt,
                    const double *dense,
                    int nsrc, const int *src_iz, const int *src_ix, const double *wavelet,
                    int N, int nB, double dz, double dx, double dt,
                    double *data, double *snap, int64_t snap_stride)
{
    if (nt < 1 || (!dense && check_points(nsrc, src_iz, src_ix, NULL, nz, nx, 0, wavelet)))
        return ORC_EINVAL;
    if (snap && snap_stride < 1) return ORC_EINVAL;
    cpml2d_t s;
    int rc = cpml2d_init(&s, v, nz, nx, N, nB, dz, dx, dt);
    if (rc) return rc;
    const int64_t l = s.l, nn = nz * nx;
    for (int64_t t = 0; t < nt; ++t) {
        if (dense) memcpy(s.src, dense + t * nn, nn * sizeof(double));
        else fill_point_sources(s.src, nz, nsrc, src_iz, src_ix, wavelet + t * nsrc, nn);
        cpml2d_step(&s);
        /* data(:, it) = fdm(l, ixi, 2)  (fwd…_mex.c:274-275) */
        if (data)
            for (int64_t j = 0; j < nx; ++j) data[t * nx + j] = s.pcur[(j + l) * s.npz + l];
        if (snap && t % snap_stride == 0) {
            double *o = snap + (t / snap_stride) * nn;
            for (int64_t j = 0; j < nx; ++j)
                memcpy(o + j * nz, s.pcur + (j + l) * s.npz + l, nz * sizeof(double));
        }
    }
    cpml2d_free(&s);
    return 0;
}

int orc_rvs_cpml_2d(const double *v, int64_t nz, int64_t nx, int64_t nt,
                    const double *data, int N, int nB, double dz, double dx, double dt,
                    double *model, double *snap, int64_t snap_stride)
{
    if (nt < 1 || !data) return ORC_EINVAL;
    if (snap && snap_stride < 1) return ORC_EINVAL;
    cpml2d_t s;
    int rc = cpml2d_init(&s, v, nz, nx, N, nB, dz, dx, dt);
    if (rc) return rc;
    const int64_t l = s.l, nn = nz * nx;
    /* rvsTimeCpmlFor2dAw_mex.c:172-281: t runs backward, source(0,:) = data(:,t) */
    for (int64_t t = nt - 1; t >= 0; --t) {
        memset(s.src, 0, nn * sizeof(double));
        for (int64_t j = 0; j < nx; ++j) s.src[j * nz] = data[t * nx + j];
        cpml2d_step(&s);
        if (snap && t % snap_stride == 0) {
            double *o = snap + (t / snap_stride) * nn;
            for (int64_t j = 0; j < nx; ++j)
                memcpy(o + j * nz, s.pcur + (j + l) * s.npz + l, nz * sizeof(double));
        }
    }
    /* model = rtm(:, :, 1)  (rvs…_mex.c:285-287) */
    if (model) memcpy(model, s.pold, s.npz * s.npx * sizeof(double));
    cpml2d_free(&s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3-D CPML, src/fwdTimeCpmlFor3dAw.m:37-160.  Column-major (z, x, y):
 *   p (nz+2k, nx+2k, ny+2k); zPhi/zA (nz+2k, nx, ny); xPhi/xA (nz, nx+2k, ny);
 *   yPhi/yA (nz, nx, ny+2k); zPsi (nz+k, nx, ny); xPsi (nz, nx+k, ny);
 *   yPsi (nz, nx, ny+k).  Update order as in the .m file: all Phi, all A,
 *   all Psi/P, then fdm3 = vdtSq.*(((zP+xP)+yP)+s) + 2*fdm2 - fdm1. */
typedef struct {
    int64_t nz, nx, ny, npz, npx, npy;
    int N, l;
    double c[8], dz, dx, dy;
    double *pold, *pcur, *pnew;
    double *zA, *xA, *yA, *zPhi, *xPhi, *yPhi, *zPsi, *xPsi, *yPsi;
    double *xb, *yb, *zb, *vdtSq, *src;
} cpml3d_t;

static void cpml3d_free(cpml3d_t *s)
{
    free(s->pold); free(s->pcur); free(s->pnew);
    free(s->zA); free(s->xA); free(s->yA); free(s->zPhi); free(s->xPhi); free(s->yPhi);
    free(s->zPsi); free(s->xPsi); free(s->yPsi);
    free(s->xb); free(s->yb); free(s->zb); free(s->vdtSq); free(s->src);
}

static void cpml3d_step(cpml3d_t *s)
{
    const int64_t nz = s->nz, nx = s->nx, ny = s->ny, npz = s->npz, npx = s->npx, npy = s->npy;
    const int N = s->N;
    const int64_t l = s->l;
    const double *c = s->c, dz = s->dz, dx = s->dx, dy = s->dy;
    const double *P = s->pcur;
#define PI3(i, j, k) (((int64_t)(k) * npx + (j)) * npz + (i))
#define CI(i, j, k) (((int64_t)(k) * nx + (j)) * nz + (i))
    /* z: zPhi(izi,:,:) and zA(izl,:,:) on array (npz, nx, ny) */
    #pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < ny; ++k)
        for (int64_t j = 0; j < nx; ++j) {
            const double *col = P + PI3(0, j + l, k + l);
            for (int64_t ip = N - 1; ip <= npz - N - 1; ++ip) {
                double D = 0.0;
                for (int q = 0; q < N; ++q) D = D + c[q] * (col[ip + 1 + q] - col[ip - q]) / dz;
                int64_t a = (k * nx + j) * npz + ip;
                if (ip >= l && ip < nz + l) {
                    double b = s->zb[CI(ip - l, j, k)];
                    s->zPhi[a] = b * s->zPhi[a] + (b - 1) * D;
                }
                s->zA[a] = D + s->zPhi[a];
            }
        }
    /* x: xPhi/xA on array (nz, npx, ny) */
    #pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < ny; ++k)
        for (int64_t jp = N - 1; jp <= npx - N - 1; ++jp)
            for (int64_t i = 0; i < nz; ++i) {
                double D = 0.0;
                for (int q = 0; q < N; ++q)
                    D = D + c[q] * (P[PI3(i + l, jp + 1 + q, k + l)] - P[PI3(i + l, jp - q, k + l)]) / dx;
                int64_t a = (k * npx + jp) * nz + i;
                if (jp >= l && jp < nx + l) {
                    double b = s->xb[CI(i, jp - l, k)];
                    s->xPhi[a] = b * s->xPhi[a] + (b - 1) * D;
                }
                s->xA[a] = D + s->xPhi[a];
            }
    /* y: yPhi/yA on array (nz, nx, npy) */
    #pragma omp parallel for collapse(2) schedule(static)
    for (int64_t kp = N - 1; kp <= npy - N - 1; ++kp)
        for (int64_t j = 0; j < nx; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                double D = 0.0;
                for (int q = 0; q < N; ++q)
                    D = D + c[q] * (P[PI3(i + l, j + l, kp + 1 + q)] - P[PI3(i + l, j + l, kp - q)]) / dy;
                int64_t a = (kp * nx + j) * nz + i;
                if (kp >= l && kp < ny + l) {
                    double b = s->yb[CI(i, j, kp - l)];
                    s->yPhi[a] = b * s->yPhi[a] + (b - 1) * D;
                }
                s->yA[a] = D + s->yPhi[a];
            }
    /* Psi, P and the update */
    #pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < ny; ++k)
        for (int64_t j = 0; j < nx; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                const int64_t ip = i + l, jp = j + l, kp = k + l, ci = CI(i, j, k);
                const double *zA = s->zA + (k * nx + j) * npz;
                double Dz = 0.0;
                for (int q = 0; q < N; ++q) Dz = Dz + c[q] * (zA[ip + q] - zA[ip - 1 - q]) / dz;
                double bz = s->zb[ci];
                double *zpsi = s->zPsi + (k * nx + j) * (nz + l) + ip;
                *zpsi = bz * (*zpsi) + (bz - 1) * Dz;
                double zP = Dz + *zpsi;

                double Dx = 0.0;
                for (int q = 0; q < N; ++q)
                    Dx = Dx + c[q] * (s->xA[(k * npx + jp + q) * nz + i] -
                                      s->xA[(k * npx + jp - 1 - q) * nz + i]) / dx;
                double bx = s->xb[ci];
                double *xpsi = s->xPsi + (k * (nx + l) + jp) * nz + i;
                *xpsi = bx * (*xpsi) + (bx - 1) * Dx;
                double xP = Dx + *xpsi;

                double Dy = 0.0;
                for (int q = 0; q < N; ++q)
                    Dy = Dy + c[q] * (s->yA[((kp + q) * nx + j) * nz + i] -
                                      s->yA[((kp - 1 - q) * nx + j) * nz + i]) / dy;
                double by = s->yb[ci];
                double *ypsi = s->yPsi + (kp * nx + j) * nz + i;
                *ypsi = by * (*ypsi) + (by - 1) * Dy;
                double yP = Dy + *ypsi;

                s->pnew[PI3(ip, jp, kp)] = s->vdtSq[ci] * (zP + xP + yP + s->src[ci]) +
                                          2 * P[PI3(ip, jp, kp)] - s->pold[PI3(ip, jp, kp)];
            }
#undef PI3
#undef CI
    double *t = s->pold;
    s->pold = s->pcur;
    s->pcur = s->pnew;
    s->pnew = t;
}

int orc_fwd_cpml_3d(const double *v, int64_t nz, int64_t nx, int64_t ny, int64_t nt,
                    const double *dense,
                    int nsrc, const int *src_iz, const int *src_ix, const int *src_iy,
                    const double *wavelet,
                    int N, int nB, double dz, double dx, double dy, double dt,
                    double *data, double *snap, int64_t snap_stride)
{
    if (!v || nz <= 0 || nx <= 0 || ny <= 0 || nt < 1 || N < 1 || N > 8 || nB < 0 ||
        2 * (int64_t)nB > nx || 2 * (int64_t)nB > ny || nB > nz)
        return ORC_EINVAL;
    if (!dense && check_points(nsrc, src_iz, src_ix, src_iy, nz, nx, ny, wavelet)) return ORC_EINVAL;
    if (snap && snap_stride < 1) return ORC_EINVAL;
    cpml3d_t s;
    memset(&s, 0, sizeof(s));
    s.nz = nz; s.nx = nx; s.ny = ny; s.N = N; s.l = 2 * N - 1;
    const int64_t l = s.l;
    s.npz = nz + 2 * l; s.npx = nx + 2 * l; s.npy = ny + 2 * l;
    s.dz = dz; s.dx = dx; s.dy = dy;
    orc_dcoef(N, s.c);
    const int64_t nn = nz * nx * ny, np = s.npz * s.npx * s.npy;
    s.pold = calloc(np, sizeof(double));
    s.pcur = calloc(np, sizeof(double));
    s.pnew = calloc(np, sizeof(double));
    s.zA = calloc(s.npz * nx * ny, sizeof(double));
    s.zPhi = calloc(s.npz * nx * ny, sizeof(double));
    s.xA = calloc(nz * s.npx * ny, sizeof(double));
    s.xPhi = calloc(nz * s.npx * ny, sizeof(double));
    s.yA = calloc(nz * nx * s.npy, sizeof(double));
    s.yPhi = calloc(nz * nx * s.npy, sizeof(double));
    s.zPsi = calloc((nz + l) * nx * ny, sizeof(double));
    s.xPsi = calloc(nz * (nx + l) * ny, sizeof(double));
    s.yPsi = calloc(nz * nx * (ny + l), sizeof(double));
    s.xb = malloc(nn * sizeof(double));
    s.yb = malloc(nn * sizeof(double));
    s.zb = malloc(nn * sizeof(double));
    s.vdtSq = malloc(nn * sizeof(double));
    s.src = calloc(nn, sizeof(double));
    if (!s.pold || !s.pcur || !s.pnew || !s.zA || !s.zPhi || !s.xA || !s.xPhi || !s.yA ||
        !s.yPhi || !s.zPsi || !s.xPsi || !s.yPsi || !s.xb || !s.yb || !s.zb || !s.vdtSq || !s.src) {
        cpml3d_free(&s);
        return ORC_ENOMEM;
    }
    orc_cpml_profiles_3d(v, nz, nx, ny, nB, dt, s.xb, s.yb, s.zb);
    for (int64_t q = 0; q < nn; ++q) s.vdtSq[q] = (v[q] * dt) * (v[q] * dt);
    for (int64_t t = 0; t < nt; ++t) {
        if (dense) {
            memcpy(s.src, dense + t * nn, nn * sizeof(double));
        } else {
            memset(s.src, 0, nn * sizeof(double));
            for (int q = 0; q < nsrc; ++q)
                s.src[((int64_t)src_iy[q] * nx + src_ix[q]) * nz + src_iz[q]] += wavelet[t * nsrc + q];
        }
        cpml3d_step(&s);
        /* data(:, :, it) = fdm(1+k, ixi, iyi, 2)  (.m:144) */
        if (data)
            for (int64_t k = 0; k < ny; ++k)
                for (int64_t j = 0; j < nx; ++j)
                    data[(t * ny + k) * nx + j] = s.pcur[((k + l) * s.npx + j + l) * s.npz + l];
        if (snap && t % snap_stride == 0) {
            double *o = snap + (t / snap_stride) * nn;
            for (int64_t k = 0; k < ny; ++k)
                for (int64_t j = 0; j < nx; ++j)
                    memcpy(o + (k * nx + j) * nz, s.pcur + ((k + l) * s.npx + j + l) * s.npz + l,
                           nz * sizeof(double));
        }
    }
    cpml3d_free(&s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Sponge ABC, src/fwdTimeSpongeFor2dAw.m:37-106 and src/rvsTimeSpongeFor2dAw.m.
 * Padded arrays (nz+2) x (nx+2), column-major.  Damping factor
 *   g(m) = exp(-(alpha*(nB-m))^2)^10, alpha = 0.015, m = 0..nB-1,
 * applied after the rotation to BOTH the previous and the current level:
 * columns j<nB (rows i<nz-nB) by g(j), columns j>=nx-nB (rows i<nz-nB) by
 * g(nx-1-j), rows i>=nz-nB (all columns) by g(nz-1-i). */
typedef struct {
    int64_t nz, nx, npz, npx;
    int nB;
    double *p1, *p2, *p3, *ax, *az, *vdtSq, *b, *g, *src;
} sponge_t;

static void sponge_free(sponge_t *s)
{
    free(s->p1); free(s->p2); free(s->p3); free(s->ax); free(s->az); free(s->vdtSq);
    free(s->b); free(s->g); free(s->src);
}

static int sponge_init(sponge_t *s, const double *v, int64_t nz, int64_t nx, int nB,
                       double dz, double dx, double dt)
{
    memset(s, 0, sizeof(*s));
    if (!v || nz <= 0 || nx <= 0 || nB < 0 || 2 * (int64_t)nB > nx || nB > nz) return ORC_EINVAL;
    s->nz = nz; s->nx = nx; s->npz = nz + 2; s->npx = nx + 2; s->nB = nB;
    const int64_t nn = nz * nx, np = s->npz * s->npx;
    s->p1 = calloc(np, sizeof(double));
    s->p2 = calloc(np, sizeof(double));
    s->p3 = calloc(np, sizeof(double));
    s->ax = malloc(nn * sizeof(double));
    s->az = malloc(nn * sizeof(double));
    s->vdtSq = malloc(nn * sizeof(double));
    s->b = malloc(nn * sizeof(double));
    s->g = malloc((nB > 0 ? nB : 1) * sizeof(double));
    s->src = calloc(nn, sizeof(double));
    if (!s->p1 || !s->p2 || !s->p3 || !s->ax || !s->az || !s->vdtSq || !s->b || !s->g || !s->src) {
        sponge_free(s);
        return ORC_ENOMEM;
    }
    const double alpha = 0.015;
    for (int m = 0; m < nB; ++m) {
        double e = exp(-((alpha * (nB - m)) * (alpha * (nB - m))));
        s->g[m] = pow(e, 10);
    }
    for (int64_t q = 0; q < nn; ++q) {
        double a = v[q] * dt / dx, c = v[q] * dt / dz;
        s->ax[q] = a * a;
        s->az[q] = c * c;
        s->vdtSq[q] = (v[q] * dt) * (v[q] * dt);
        s->b[q] = 2 - 2 * s->ax[q] - 2 * s->az[q];
    }
    return 0;
}

static void sponge_damp(sponge_t *s, double *p)
{
    const int64_t nz = s->nz, nx = s->nx, npz = s->npz;
    const int nB = s->nB;
    for (int m = 0; m < nB; ++m) {
        for (int64_t i = 0; i < nz - nB; ++i) {
            p[(int64_t)(m + 1) * npz + i + 1] *= s->g[m];
            int64_t j = nx - nB + m;
            p[(j + 1) * npz + i + 1] *= s->g[nx - 1 - j];
        }
        int64_t i = nz - nB + m;
        for (int64_t jp = 0; jp < s->npx; ++jp) p[jp * npz + i + 1] *= s->g[nz - 1 - i];
    }
}

static void sponge_step(sponge_t *s)
{
    const int64_t nz = s->nz, nx = s->nx, npz = s->npz;
    const double *P2 = s->p2, *P1 = s->p1;
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nx; ++j)
        for (int64_t i = 0; i < nz; ++i) {
            int64_t q = j * nz + i, a = (j + 1) * npz + (i + 1);
            s->p3[a] = s->b[q] * P2[a] - P1[a] + s->ax[q] * (P2[a + npz] + P2[a - npz]) +
                       s->az[q] * (P2[a + 1] + P2[a - 1]) - s->vdtSq[q] * s->src[q];
        }
    double *t = s->p1;
    s->p1 = s->p2;
    s->p2 = s->p3;
    s->p3 = t;
    sponge_damp(s, s->p1);
    sponge_damp(s, s->p2);
}

int orc_fwd_sponge_2d(const double *v, int64_t nz, int64_t nx, int64_t nt,
                      const double *dense,
                      int nsrc, const int *src_iz, const int *src_ix, const double *wavelet,
                      int nB, double dz, double dx, double dt,
                      double *data, double *snap, int64_t snap_stride)
{
    if (nt < 1 || (!dense && check_points(nsrc, src_iz, src_ix, NULL, nz, nx, 0, wavelet)))
        return ORC_EINVAL;
    if (snap && snap_stride < 1) return ORC_EINVAL;
    sponge_t s;
    int rc = sponge_init(&s, v, nz, nx, nB, dz, dx, dt);
    if (rc) return rc;
    const int64_t nn = nz * nx;
    for (int64_t t = 0; t < nt; ++t) {
        if (dense) memcpy(s.src, dense + t * nn, nn * sizeof(double));
        else fill_point_sources(s.src, nz, nsrc, src_iz, src_ix, wavelet + t * nsrc, nn);
        sponge_step(&s);
        if (data)
            for (int64_t j = 0; j < nx; ++j) data[t * nx + j] = s.p2[(j + 1) * s.npz + 1];
        if (snap && t % snap_stride == 0) {
            double *o = snap + (t / snap_stride) * nn;
            for (int64_t j = 0; j < nx; ++j)
                memcpy(o + j * nz, s.p2 + (j + 1) * s.npz + 1, nz * sizeof(double));
        }
    }
    sponge_free(&s);
    return 0;
}

int orc_rvs_sponge_2d(const double *v, int64_t nz, int64_t nx, int64_t nt,
                      const double *data, int nB, double dz, double dx, double dt,
                      double *model, double *snap, int64_t snap_stride)
{
    if (nt < 1 || !data) return ORC_EINVAL;
    if (snap && snap_stride < 1) return ORC_EINVAL;
    sponge_t s;
    int rc = sponge_init(&s, v, nz, nx, nB, dz, dx, dt);
    if (rc) return rc;
    const int64_t nn = nz * nx;
    for (int64_t t = nt - 1; t >= 0; --t) {
        memset(s.src, 0, nn * sizeof(double));
        for (int64_t j = 0; j < nx; ++j) s.src[j * nz] = data[t * nx + j];
        sponge_step(&s);
        if (snap && t % snap_stride == 0) {
            double *o = snap + (t / snap_stride) * nn;
            for (int64_t j = 0; j < nx; ++j)
                memcpy(o + j * nz, s.p2 + (j + 1) * s.npz + 1, nz * sizeof(double));
        }
    }
    if (model) memcpy(model, s.p1, s.npz * s.npx * sizeof(double));
    sponge_free(&s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Imaging condition of mainRtmTimeCpmlFor2dAw.m:268-273 on stored snapshots:
 *   M = f.*r + M;  s2 = f.^2 + s2 (s2 starts at eps);  image = M./s2
 * summed over t = 0, stride, 2*stride, ... */
int orc_xcorr_image(const double *fwd_snap, const double *rvs_snap, int64_t ncell, int64_t nt,
                    int64_t stride, double eps, double *num, double *den, double *image)
{
    if (!fwd_snap || !rvs_snap || ncell <= 0 || nt < 1 || stride < 1) return ORC_EINVAL;
    for (int64_t q = 0; q < ncell; ++q) {
        double M = 0.0, s2 = eps;
        for (int64_t t = 0; t < nt; t += stride) {
            double f = fwd_snap[t * ncell + q], r = rvs_snap[t * ncell + q];
            M = f * r + M;
            s2 = f * f + s2;
        }
        if (num) num[q] = M;
        if (den) den[q] = s2;
        if (image) image[q] = M / s2;
    }
    return 0;
}

int orc_rtm_shot_cpml_2d(const double *V, const double *VS, int64_t nz, int64_t nx, int64_t nt,
                         int nsrc, const int *src_iz, const int *src_ix, const double *wavelet,
                         int N, int nB, double dz, double dx, double dt,
                         int64_t img_stride, double eps, const double *data_true_in,
                         double *data_true_out, double *num, double *den, double *image)
{
    if (!V || !VS || nt < 1 || img_stride < 1) return ORC_EINVAL;
    if (check_points(nsrc, src_iz, src_ix, NULL, nz, nx, 0, wavelet)) return ORC_EINVAL;
    const int64_t nn = nz * nx, nsnap = (nt + img_stride - 1) / img_stride;
    double *dtrue = malloc(nx * nt * sizeof(double));
    double *dsm = malloc(nx * nt * sizeof(double));
    double *fsnap = malloc(nsnap * nn * sizeof(double));
    double *rsnap = malloc(nsnap * nn * sizeof(double));
    int rc = ORC_ENOMEM;
    if (!dtrue || !dsm || !fsnap || !rsnap) goto done;
    if (data_true_in) {
        memcpy(dtrue, data_true_in, nx * nt * sizeof(double));
    } else {
        rc = orc_fwd_cpml_2d(V, nz, nx, nt, NULL, nsrc, src_iz, src_ix, wavelet, N, nB, dz, dx, dt,
                             dtrue, NULL, 1);
        if (rc) goto done;
    }
    if (data_true_out) memcpy(data_true_out, dtrue, nx * nt * sizeof(double));
    rc = orc_fwd_cpml_2d(VS, nz, nx, nt, NULL, nsrc, src_iz, src_ix, wavelet, N, nB, dz, dx, dt,
                         dsm, fsnap, img_stride);
    if (rc) goto done;
    /* dataDelta = dataTrue - dataSmooth (mainRtmTimeCpmlFor2dAw.m:252) */
    for (int64_t q = 0; q < nx * nt; ++q) dsm[q] = dtrue[q] - dsm[q];
    /* rvsTimeCpmlFor2dAw(V, dataDelta, ...) (mainRtmTimeCpmlFor2dAw.m:256) */
    rc = orc_rvs_cpml_2d(V, nz, nx, nt, dsm, N, nB, dz, dx, dt, NULL, rsnap, img_stride);
    if (rc) goto done;
    rc = orc_xcorr_image(fsnap, rsnap, nn, nsnap, 1, eps, num, den, image);
done:
    free(dtrue); free(dsm); free(fsnap); free(rsnap);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Traveltimes: src/fd_cmex/eikonal2d_mex.cpp:27-267 (Zhao's fast sweeping as
 * the reference codes it).  Every "sweep" reads only TOld and writes TNew, so
 * each of the four sweeps of an iteration is one Jacobi pass over the grid
 * (their loop orders do not change the result), followed by TNew(src) = 0 and
 * TOld = TNew.  Per cell, with neighbours taken inside the grid only (a row or
 * column at the edge uses its single inner neighbour):
 *   a = min(T(i-1,j), T(i+1,j)),  b = min(T(i,j-1), T(i,j+1))
 *   TT = |a-b| < s*dx ? 0.5*(a + b + sqrt(2*s*s*dx*dx - (a-b)*(a-b)))
 *                     : min(a,b) + s*dx
 *   TNew = min(TOld, TT)
 * s = 1/v; dx is truncated to an integer (eikonal2d_mex.cpp:35, `(int)`);
 * T starts at 65535 (:31) except T(src) = 0; sz, sx are 1-based.  The
 * operation order above is the MEX's (built with -ffp-contract=off here). */
int orc_eikonal2d(const double *v, int64_t nz, int64_t nx, double dx_in, int sz, int sx,
                  int iMax, double *T)
{
    if (!v || !T || nz < 2 || nx < 2 || sz < 1 || sz > nz || sx < 1 || sx > nx || iMax < 0)
        return ORC_EINVAL;
    const int64_t nn = nz * nx;
    const double dx = (double)(int)dx_in;
    double *slow = malloc(nn * sizeof(double));
    double *told = malloc(nn * sizeof(double));
    if (!slow || !told) { free(slow); free(told); return ORC_ENOMEM; }
    for (int64_t q = 0; q < nn; ++q) {
        slow[q] = 1 / v[q];
        told[q] = 65535;
        T[q] = 65535;
    }
    const int64_t src = (int64_t)(sz - 1) + (int64_t)(sx - 1) * nz;
    told[src] = 0;
    for (int it = 0; it < 4 * iMax; ++it) {
        #pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < nx; ++j)
            for (int64_t i = 0; i < nz; ++i) {
                const double *c = told + j * nz;
                double a, b, TT;
                if (i == 0) a = c[1];
                else if (i == nz - 1) a = c[nz - 2];
                else a = fmin(c[i - 1], c[i + 1]);
                if (j == 0) b = told[i + nz];
                else if (j == nx - 1) b = told[i + (nx - 2) * nz];
                else b = fmin(told[i + (j - 1) * nz], told[i + (j + 1) * nz]);
                const double s = slow[i + j * nz];
                if (fabs(a - b) < s * dx)
                    TT = 0.5 * (a + b + sqrt((2 * s * s * dx * dx) - (a - b) * (a - b)));
                else
                    TT = fmin(a, b) + (s * dx);
                T[i + j * nz] = fmin(c[i], TT);
            }
        T[src] = 0;
        memcpy(told, T, nn * sizeof(double));
    }
    free(slow);
    free(told);
    return 0;
}

/* Kirchhoff migration of one shot: src/migrate.m:1-26 with
 * src/shot2RecTime.m:1-6 (its nx == size(travelTime,2) branch).
 *   m = 0;  for each receiver xr in xRecGrid (in order):
 *       r  = position of xr in xShotAndRecGrid          (ismember, 1-based)
 *       it = round(T(:,:,ixs)/dt)+1 + round(T(:,:,r)/dt)+1   (1-based sample)
 *       m  = m + shot(it, xr)
 * travelTime (nz, nx, ntab) and shot (nt, nx) column-major; ixs, xRecGrid
 * and xShotAndRecGrid 1-based as in MATLAB.  round() is MATLAB's (half away
 * from zero), which C's round() is.  A sample index outside 1..nt is an
 * error, as MATLAB's indexing is. */
int orc_kirchhoff_migrate(const double *tt, int64_t nz, int64_t nx, int64_t ntab,
                          const int *xrec, int nrec, const int *xsr, int nsr,
                          const double *shot, int64_t nt, double dt, int ixs, double *m)
{
    if (!tt || !xrec || !xsr || !shot || !m || nz < 1 || nx < 1 || nt < 1 || nrec < 0 ||
        nsr < 1 || ixs < 1 || ixs > ntab || !(dt > 0))
        return ORC_EINVAL;
    const int64_t nn = nz * nx;
    int *ridx = malloc((nrec > 0 ? nrec : 1) * sizeof(int));
    if (!ridx) return ORC_ENOMEM;
    for (int r = 0; r < nrec; ++r) {
        int k = 0;
        while (k < nsr && xsr[k] != xrec[r]) ++k;
        if (k == nsr || k + 1 > ntab || xrec[r] < 1 || xrec[r] > nx) { free(ridx); return ORC_EINVAL; }
        ridx[r] = k;  /* 0-based table of receiver r */
    }
    int bad = 0;
    const double *Ts = tt + (int64_t)(ixs - 1) * nn;
    #pragma omp parallel for schedule(static) reduction(|:bad)
    for (int64_t q = 0; q < nn; ++q) {
        double acc = 0;
        for (int r = 0; r < nrec; ++r) {
            const double *Tr = tt + (int64_t)ridx[r] * nn;
            const double itd = (round(Ts[q] / dt) + 1) + (round(Tr[q] / dt) + 1);
            if (!(itd >= 1 && itd <= (double)nt)) { bad = 1; continue; }
            acc = acc + shot[(int64_t)(xrec[r] - 1) * nt + (int64_t)itd - 1];
        }
        m[q] = acc;
    }
    free(ridx);
    return bad ? ORC_EINVAL : 0;
}

/* ------------------------------------------------------------------------ */
/* Frequency-domain CPML Helmholtz: src/freqCpmlFor2dAw.m:43-133.
 *
 * Matrix (:43-123).  model is the squared slowness (nz, nx); v = sqrt(1/m).
 * Damping with dampPml (src/dampPml.m:16-18, d = d0 .* (u/L).^2, the square
 * taken first): x on the first nB columns with u = (nB-j) dx, the last nB with
 * u = (j+1) dx, z on the last nB rows with u = (i+1) dz, L = nB dx / nB dz;
 * no top layer.  Summed coefficients c(-k..k) of dCoef (k = 2N-1) in the
 * reference's loop order (:80-89).  Row (i, j) holds, for o = -k..-1, 1..k
 * (neighbours outside the grid dropped, :120-122):
 *   (i, j+o)  c(o) * (1j w / (dx (1j w + dx_ij)))^2
 *   (i+o, j)  c(o) * (1j w / (dz (1j w + dz_ij)))^2
 *   (i, j)    (m w^2 + c(0) (..dz..)^2) + c(0) (..dx..)^2
 * always with the row cell's own damping.  The complex factor is computed as
 *   t = (s d, s w);  q = (0, w) / t = (w t.im / |t|^2, w t.re / |t|^2);
 *   q^2 = (q.re q.re - q.im q.im, q.re q.im + q.im q.re),
 * each operation rounded separately (MATLAB's own complex division and power
 * are not reproducible at the last bit: parity unpinned there, pinned by the
 * closed form to rounding).  CSR columns ascend within a row; the index of
 * cell (i, j) is j nz + i, as MATLAB's column-major reshape.
 *
 * Solve (:127-133, snapshot = A \ (-s)).  MATLAB's sparse mldivide is not
 * restated; the solution is that of the banded LU with partial pivoting of
 * LAPACK's zgbtf2/zgbtrs, in the natural ordering (bandwidth kl = ku = k nz):
 * pivot = first row of maximal |re| + |im| among j..j+kl, rows swapped over
 * columns j..ju (ju the running reach of U), column scaled by the reciprocal
 * of the pivot, rank-1 update of columns j+1..ju, each element updated as
 * a - l u in pivot order; then the forward sweep with the interchanges and
 * the column-oriented backward sweep (b_j = b_j / u_jj by the textbook
 * division, b_i -= u_ij b_j).  Pinned by the residual and by an independent
 * sparse direct solver (SuperLU) in the tests. */
typedef struct { double re, im; } orc_cx;

static inline orc_cx cx_mul(orc_cx a, orc_cx b)
{
    orc_cx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline orc_cx cx_sub_mul(orc_cx a, orc_cx l, orc_cx u)  /* a - l u */
{
    orc_cx r = {a.re - (l.re * u.re - l.im * u.im), a.im - (l.re * u.im + l.im * u.re)};
    return r;
}
static inline orc_cx cx_div(orc_cx a, orc_cx b)
{
    const double den = b.re * b.re + b.im * b.im;
    orc_cx r = {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
    return r;
}
static inline orc_cx cx_recip(orc_cx b)
{
    const double den = b.re * b.re + b.im * b.im;
    orc_cx r = {b.re / den, -b.im / den};
    return r;
}

/* (1j w / (s (1j w + d)))^2 as specified above */
static orc_cx freq_factor(double w, double s, double d)
{
    const double tr = s * d, ti = s * w;
    const double den = tr * tr + ti * ti;
    const double qr = w * ti / den, qi = w * tr / den;
    orc_cx r = {qr * qr - qi * qi, qr * qi + qi * qr};
    return r;
}

/* Damping grids xd, zd (nz x nx) and summed coefficients c[0..2k]. */
static int freq_setup(const double *m, int64_t nz, int64_t nx, int N, int nB, double dz,
                      double dx, double *xd, double *zd, double *c)
{
    double coeff[8];
    if (orc_dcoef(N, coeff)) return ORC_EINVAL;
    const int k = 2 * N - 1;
    for (int o = 0; o <= 2 * k; ++o) c[o] = 0;
    for (int ii = 1; ii <= N; ++ii)
        for (int jj = 1; jj <= N; ++jj) {
            const int o1 = ii + jj - 1, o2 = ii - jj;
            const double p = coeff[ii - 1] * coeff[jj - 1];
            c[o1 + k] = c[o1 + k] + p;
            c[o2 + k] = c[o2 + k] - p;
            c[-o1 + k] = c[-o1 + k] + p;
            c[-o2 + k] = c[-o2 + k] - p;
        }
    const double logR = log(1e-6);
    for (int64_t q = 0; q < nz * nx; ++q) { xd[q] = 0; zd[q] = 0; }
    if (nB == 0) return 0;
    const double Lx = nB * dx, Lz = nB * dz;
    for (int j = 0; j < nB; ++j)
        for (int64_t i = 0; i < nz; ++i) {
            const int64_t jl = j, jr = nx - nB + j;
            const double ul = (nB - j) * dx, ur = (j + 1) * dx;
            const double vl = sqrt(1. / m[jl * nz + i]), vr = sqrt(1. / m[jr * nz + i]);
            const double d0l = -(3 * vl) / (2 * Lx) * logR, d0r = -(3 * vr) / (2 * Lx) * logR;
            const double sl = (ul / Lx) * (ul / Lx), sr = (ur / Lx) * (ur / Lx);
            xd[jl * nz + i] = d0l * sl;
            xd[jr * nz + i] = d0r * sr;
        }
    for (int64_t j = 0; j < nx; ++j)
        for (int i = 0; i < nB; ++i) {
            const int64_t ii = nz - nB + i;
            const double u = (i + 1) * dz;
            const double v = sqrt(1. / m[j * nz + ii]);
            const double d0 = -(3 * v) / (2 * Lz) * logR;
            zd[j * nz + ii] = d0 * ((u / Lz) * (u / Lz));
        }
    return 0;
}

int64_t orc_freq_cpml_2d_nnz(int64_t nz, int64_t nx, int N)
{
    const int k = 2 * N - 1;
    int64_t nnz = 0;
    for (int64_t i = 0; i < nz; ++i) {
        int64_t zc = 0;
        for (int o = -k; o <= k; ++o) zc += (o != 0 && i + o >= 0 && i + o < nz);
        nnz += zc * nx;
    }
    for (int64_t j = 0; j < nx; ++j) {
        int64_t xc = 0;
        for (int o = -k; o <= k; ++o) xc += (o != 0 && j + o >= 0 && j + o < nx);
        nnz += xc * nz;
    }
    return nnz + nz * nx;
}

int orc_freq_cpml_2d_matrix(const double *m, int64_t nz, int64_t nx, double w, int N, int nB,
                            double dz, double dx, int64_t *rowptr, int64_t *col, double *val)
{
    if (!m || !rowptr || !col || !val || nz < 1 || nx < 1 || N < 1 || N > 8 || nB < 0 ||
        2 * (int64_t)nB > nx || nB > nz)
        return ORC_EINVAL;
    const int k = 2 * N - 1;
    const int64_t n = nz * nx;
    double c[31];
    double *xd = malloc(n * sizeof(double)), *zd = malloc(n * sizeof(double));
    if (!xd || !zd) { free(xd); free(zd); return ORC_ENOMEM; }
    if (freq_setup(m, nz, nx, N, nB, dz, dx, xd, zd, c)) { free(xd); free(zd); return ORC_EINVAL; }
    int64_t e = 0;
    for (int64_t r = 0; r < n; ++r) {
        const int64_t j = r / nz, i = r - j * nz;
        const orc_cx fz = freq_factor(w, dz, zd[r]), fx = freq_factor(w, dx, xd[r]);
        rowptr[r] = e;
        /* ascending column order: x neighbours left, z neighbours, centre, z, x right */
        for (int o = -k; o <= -1; ++o)
            if (j + o >= 0) { col[e] = r + o * nz; val[2 * e] = c[o + k] * fx.re; val[2 * e + 1] = c[o + k] * fx.im; ++e; }
        for (int o = -k; o <= -1; ++o)
            if (i + o >= 0) { col[e] = r + o; val[2 * e] = c[o + k] * fz.re; val[2 * e + 1] = c[o + k] * fz.im; ++e; }
        {
            const double mw = m[r] * (w * w);
            col[e] = r;
            val[2 * e] = (mw + c[k] * fz.re) + c[k] * fx.re;
            val[2 * e + 1] = c[k] * fz.im + c[k] * fx.im;  /* m w^2 is real */
            ++e;
        }
        for (int o = 1; o <= k; ++o)
            if (i + o < nz) { col[e] = r + o; val[2 * e] = c[o + k] * fz.re; val[2 * e + 1] = c[o + k] * fz.im; ++e; }
        for (int o = 1; o <= k; ++o)
            if (j + o < nx) { col[e] = r + o * nz; val[2 * e] = c[o + k] * fx.re; val[2 * e + 1] = c[o + k] * fx.im; ++e; }
    }
    rowptr[n] = e;
    free(xd);
    free(zd);
    return 0;
}

/* A U = -S for nsrc right-hand sides (S, U complex interleaved (n, nsrc)). */
int orc_freq_cpml_2d_solve(const double *m, int64_t nz, int64_t nx, double w, int N, int nB,
                           double dz, double dx, const double *S, int nsrc, double *U)
{
    if (!S || !U || nsrc < 1) return ORC_EINVAL;
    const int64_t n = nz * nx, nnz = orc_freq_cpml_2d_nnz(nz, nx, N);
    const int64_t kl = (int64_t)(2 * N - 1) * nz, ku = kl, ldab = 2 * kl + ku + 1;
    int64_t *rp = malloc((n + 1) * sizeof(int64_t)), *ci = malloc(nnz * sizeof(int64_t));
    double *va = malloc(2 * nnz * sizeof(double));
    orc_cx *ab = calloc((size_t)ldab * n, sizeof(orc_cx));
    int64_t *piv = malloc(n * sizeof(int64_t));
    int rc = 0;
    if (!rp || !ci || !va || !ab || !piv) { rc = ORC_ENOMEM; goto done; }
    if ((rc = orc_freq_cpml_2d_matrix(m, nz, nx, w, N, nB, dz, dx, rp, ci, va))) goto done;
    /* band storage: A(i, j) at ab[(kl + ku + i - j) + j ldab] */
    for (int64_t r = 0; r < n; ++r)
        for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
            orc_cx v = {va[2 * e], va[2 * e + 1]};
            ab[(kl + ku + r - ci[e]) + ci[e] * ldab] = v;
        }
#define AB(i, j) ab[(kl + ku + (i) - (j)) + (j) * ldab]
    int64_t ju = 0;
    for (int64_t j = 0; j < n; ++j) {
        const int64_t km = kl < n - 1 - j ? kl : n - 1 - j;
        int64_t jp = 0;
        double best = -1;
        for (int64_t t = 0; t <= km; ++t) {
            const orc_cx a = AB(j + t, j);
            const double s = fabs(a.re) + fabs(a.im);
            if (s > best) { best = s; jp = t; }
        }
        piv[j] = j + jp;
        const int64_t reach = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
        if (reach > ju) ju = reach;
        if (best == 0) { rc = ORC_EINVAL; goto done; }  /* singular */
        if (jp != 0)
            for (int64_t cc = j; cc <= ju; ++cc) {
                orc_cx t = AB(j, cc);
                AB(j, cc) = AB(j + jp, cc);
                AB(j + jp, cc) = t;
            }
        if (km > 0) {
            const orc_cx r = cx_recip(AB(j, j));
            for (int64_t t = 1; t <= km; ++t) AB(j + t, j) = cx_mul(AB(j + t, j), r);
            #pragma omp parallel for schedule(static) if (km * (ju - j) > 65536)
            for (int64_t cc = j + 1; cc <= ju; ++cc) {
                const orc_cx u = AB(j, cc);
                for (int64_t t = 1; t <= km; ++t) AB(j + t, cc) = cx_sub_mul(AB(j + t, cc), AB(j + t, j), u);
            }
        }
    }
    #pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nsrc; ++s) {
        orc_cx *b = (orc_cx *)(U + 2 * (int64_t)s * n);
        for (int64_t q = 0; q < n; ++q) { b[q].re = -S[2 * ((int64_t)s * n + q)]; b[q].im = -S[2 * ((int64_t)s * n + q) + 1]; }
        for (int64_t j = 0; j < n - 1; ++j) {
            const int64_t lm = kl < n - 1 - j ? kl : n - 1 - j;
            const int64_t l = piv[j];
            if (l != j) { orc_cx t = b[l]; b[l] = b[j]; b[j] = t; }
            const orc_cx bj = b[j];
            for (int64_t t = 1; t <= lm; ++t) b[j + t] = cx_sub_mul(b[j + t], AB(j + t, j), bj);
        }
        for (int64_t j = n - 1; j >= 0; --j) {
            b[j] = cx_div(b[j], AB(j, j));
            const orc_cx bj = b[j];
            const int64_t i0 = j - kl - ku > 0 ? j - kl - ku : 0;
            for (int64_t i = i0; i < j; ++i) b[i] = cx_sub_mul(b[i], AB(i, j), bj);
        }
    }
#undef AB
done:
    free(rp); free(ci); free(va); free(ab); free(piv);
    return rc;
}
