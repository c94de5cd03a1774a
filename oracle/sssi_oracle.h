/*
 * sssi_oracle.h -- FP64 CPU restatement of the SSSI acoustic propagators.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark may load this library, and only as the
 * checker.  The product path (libsssi, HIP) never links or calls it.
 *
 * Arrays follow the reference's MATLAB convention: column-major, z fastest.
 *   v(nz,nx)             v[j*nz + i]
 *   source(nz,nx,nt)     src[t*nz*nx + j*nz + i]
 *   data(nx,nt)          data[t*nx + j]
 *   snapshot(nz,nx,nt)   snap[t*nz*nx + j*nz + i]
 *   3-D: v(nz,nx,ny)     v[(k*nx + j)*nz + i], data(nx,ny,nt) data[(t*ny+k)*nx + j]
 *
 * Sources are given either dense (exactly the reference argument) or as a
 * point list (iz[], ix[] [, iy[]], wavelet[t*nsrc + s]); the point list is
 * the same arithmetic as a dense array that is zero everywhere else.
 *
 * Parity status of every function is stated next to it in the .c file.
 */
#ifndef SSSI_ORACLE_H
#define SSSI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int  orc_num_threads(void);
void orc_set_threads(int n);

/* Staggered-grid coefficients c[0..N-1] (dCoef(N,'s')), 1 <= N <= 8. */
int orc_dcoef(int N, double *c);

/* 2-D CPML b = exp(-d*dt) profiles, full nz*nx, column-major. */
int orc_cpml_profiles_2d(const double *v, int64_t nz, int64_t nx, int nB,
                         double dz, double dx, double dt, double *xb, double *zb);

/* 3-D CPML profiles with the reference's index-unit quirks, full grids. */
int orc_cpml_profiles_3d(const double *v, int64_t nz, int64_t nx, int64_t ny, int nB,
                         double dt, double *xb, double *yb, double *zb);

/* fwdTimeCpmlFor2dAw.  dense may be NULL (then the point list is used).
 * snap may be NULL; when not NULL, snapshot slice t is stored for every t with
 * t % snap_stride == 0 at slot t/snap_stride. */
int orc_fwd_cpml_2d(const double *v, int64_t nz, int64_t nx, int64_t nt,
                    const double *dense,
                    int nsrc, const int *src_iz, const int *src_ix, const double *wavelet,
                    int N, int nB, double dz, double dx, double dt,
                    double *data, double *snap, int64_t snap_stride);

/* rvsTimeCpmlFor2dAw.  model is (nz+2l)x(nx+2l) column-major (may be NULL). */
int orc_rvs_cpml_2d(const double *v, int64_t nz, int64_t nx, int64_t nt,
                    const double *data, int N, int nB, double dz, double dx, double dt,
                    double *model, double *snap, int64_t snap_stride);

/* fwdTimeCpmlFor3dAw. */
int orc_fwd_cpml_3d(const double *v, int64_t nz, int64_t nx, int64_t ny, int64_t nt,
                    const double *dense,
                    int nsrc, const int *src_iz, const int *src_ix, const int *src_iy,
                    const double *wavelet,
                    int N, int nB, double dz, double dx, double dy, double dt,
                    double *data, double *snap, int64_t snap_stride);

/* fwdTimeSpongeFor2dAw / rvsTimeSpongeFor2dAw.  model is (nz+2)x(nx+2). */
int orc_fwd_sponge_2d(const double *v, int64_t nz, int64_t nx, int64_t nt,
                      const double *dense,
                      int nsrc, const int *src_iz, const int *src_ix, const double *wavelet,
                      int nB, double dz, double dx, double dt,
                      double *data, double *snap, int64_t snap_stride);
int orc_rvs_sponge_2d(const double *v, int64_t nz, int64_t nx, int64_t nt,
                      const double *data, int nB, double dz, double dx, double dt,
                      double *model, double *snap, int64_t snap_stride);

/* One RTM shot with the CPML propagators (mainRtmTimeCpmlFor2dAw.m:147-315):
 *   dataTrue  = fwd(V,  src)   (used as given when data_true_in != NULL)
 *   dataSmooth= fwd(VS, src)   with the source wavefield kept at t % img_stride == 0
 *   delta     = dataTrue - dataSmooth
 *   rvs(V, delta) and, at the same t, num += fwd*rvs, den += fwd^2 (den starts at eps)
 * Outputs (nz*nx each, may be NULL): num, den, image = num./den.
 * data_true_out (nx*nt, may be NULL) receives dataTrue. */
int orc_rtm_shot_cpml_2d(const double *V, const double *VS, int64_t nz, int64_t nx, int64_t nt,
                         int nsrc, const int *src_iz, const int *src_ix, const double *wavelet,
                         int N, int nB, double dz, double dx, double dt,
                         int64_t img_stride, double eps, const double *data_true_in,
                         double *data_true_out, double *num, double *den, double *image);

/* Cross-correlation imaging on stored snapshots (mainRtmTimeCpmlFor2dAw.m:268-273). */
int orc_xcorr_image(const double *fwd_snap, const double *rvs_snap, int64_t ncell, int64_t nt,
                    int64_t stride, double eps, double *num, double *den, double *image);

/* Traveltimes by the reference's fast-sweeping eikonal solver
 * (src/fd_cmex/eikonal2d_mex.cpp:27-267): T (nz*nx, column-major) for a source
 * at the 1-based cell (sz, sx) after iMax iterations of four sweeps. */
int orc_eikonal2d(const double *v, int64_t nz, int64_t nx, double dx, int sz, int sx,
                  int iMax, double *T);

/* Kirchhoff migration of one shot (src/migrate.m:1-26, src/shot2RecTime.m):
 * m (nz*nx) from travelTime (nz, nx, ntab), 1-based grids and shot (nt, nx). */
int orc_kirchhoff_migrate(const double *tt, int64_t nz, int64_t nx, int64_t ntab,
                          const int *xrec, int nrec, const int *xsr, int nsr,
                          const double *shot, int64_t nt, double dt, int ixs, double *m);

/* freqCpmlFor2dAw (src/freqCpmlFor2dAw.m): the impedance matrix A in CSR
 * (rowptr[n+1], col[nnz], val[2 nnz] complex interleaved; nnz from
 * orc_freq_cpml_2d_nnz) for squared slowness m (nz, nx), angular frequency w;
 * and U = A \ (-S) for nsrc complex right-hand sides S (n, nsrc). */
int64_t orc_freq_cpml_2d_nnz(int64_t nz, int64_t nx, int N);
int orc_freq_cpml_2d_matrix(const double *m, int64_t nz, int64_t nx, double w, int N, int nB,
                            double dz, double dx, int64_t *rowptr, int64_t *col, double *val);
int orc_freq_cpml_2d_solve(const double *m, int64_t nz, int64_t nx, double w, int N, int nB,
                           double dz, double dx, const double *S, int nsrc, double *U);

#ifdef __cplusplus
}
#endif
#endif
