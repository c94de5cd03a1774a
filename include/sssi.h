/*
 * sssi.h -- C ABI of libsssi, the MI355X (gfx950) acoustic finite-difference
 * propagators (CPML and sponge boundaries, 2-D and 3-D) and the reverse-time
 * migration built on them.
 *
 * Conventions
 *   - Host arrays are FP64, column-major (z fastest), exactly what a MATLAB
 *     MEX gateway receives from mxGetPr:
 *       v(nz,nx)            v[j*nz + i]
 *       source(nz,nx,nt)    source[(t*nx + j)*nz + i]
 *       data(nx,nt)         data[t*nx + j]
 *       snapshot(nz,nx,ns)  snapshot[(s*nx + j)*nz + i]
 *       3-D v(nz,nx,ny)     v[(k*nx + j)*nz + i];  data(nx,ny,nt)
 *   - Outputs are caller-allocated.  No pointer is kept after a call returns.
 *   - Every call returns SSSI_OK (0) or a negative status; sssi_last_error()
 *     returns a thread-local message for the last failure on that thread.
 *   - Device arithmetic is FP32 (stencil, memory variables); damping profiles
 *     are formed in FP64 on the host and rounded once.
 *   - Calls are synchronous unless stated otherwise.  Handles are not shared
 *     between threads without external locking; distinct handles are
 *     independent.
 */
#ifndef SSSI_H
#define SSSI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSSI_OK 0
#define SSSI_EINVAL (-1)    /* bad argument (shape, range, NULL)          */
#define SSSI_ENOMEM (-2)    /* device or host allocation failed          */
#define SSSI_EDEVICE (-3)   /* HIP runtime error / no device              */
#define SSSI_ESTATE (-4)    /* call not valid in the handle's state       */
#define SSSI_EBUDGET (-5)   /* job does not fit the device memory budget  */
#define SSSI_EUNSTABLE (-6) /* NaN/Inf detected in the wavefield          */

#define SSSI_ABC_CPML 0
#define SSSI_ABC_SPONGE 1

/* Message for the last failing call on the calling thread ("" if none). */
const char *sssi_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int sssi_version(void);
/* Number of visible HIP devices (0 on a host without GPU). */
int sssi_device_count(int *count);
/* Select the HIP device used by subsequently created handles and calls. */
int sssi_set_device(int device);
/* Layout autotuner (no counterpart in the reference): with on = 1, every
 * propagation handle created afterwards times the instantiated kernel layouts
 * on its own grid (2-D: tile height and interior march groups; 3-D: block y
 * extent and z chunk) and keeps the fastest, cached per problem shape (and
 * appended to the file named by SSSI_TUNE_CACHE, if set).  Off by default
 * (SSSI_AUTOTUNE=1 turns it on); results of different layouts agree with the
 * oracle to the same tolerance, and may differ from each other in the last
 * bits near damping-strip edges. */
int sssi_set_autotune(int on);

/* ---------------------------------------------------------------------
 * Drop-in entry points: same arguments, order and outputs as the MEX
 * gateways they replace.  snapshot may be NULL (the gateway passes NULL
 * when nlhs < 2).  nDiffOrder in 1..8, 0 <= nBoundary, 2*nBoundary <= nx
 * (and ny), nBoundary <= nz, nt >= 1.
 * ------------------------------------------------------------------- */

/* Replaces fwdTimeCpmlFor2dAw_mex (src/fd_cmex/fwdTimeCpmlFor2dAw_mex.c:39;
 * wrapper src/fd_cmex/fwdTimeCpmlFor2dAw.m:37).  source is dense (nz,nx,nt).
 * data (nx,nt), snapshot (nz,nx,nt). */
int sssi_fwdTimeCpmlFor2dAw(const double *v, int64_t nz, int64_t nx, const double *source,
                            int64_t nt, int nDiffOrder, int nBoundary, double dz, double dx,
                            double dt, double *data, double *snapshot);

/* Replaces rvsTimeCpmlFor2dAw_mex (src/fd_cmex/rvsTimeCpmlFor2dAw_mex.c:38).
 * data (nx,nt) is injected at the top row for t = nt-1 ... 0.
 * model ((nz+2l),(nx+2l)), l = 2*nDiffOrder-1, is the field one step before the
 * last (rvs…_mex.c:285-287).  snapshot (nz,nx,nt).  model may be NULL. */
int sssi_rvsTimeCpmlFor2dAw(const double *v, int64_t nz, int64_t nx, const double *data,
                            int64_t nt, int nDiffOrder, int nBoundary, double dz, double dx,
                            double dt, double *model, double *snapshot);

/* Replaces fwdTimeCpmlFor3dAw (src/fwdTimeCpmlFor3dAw.m:1), including its
 * damping profile conventions (fwdTimeCpmlFor3dAw.m:46-70).  source dense
 * (nz,nx,ny,nt); data (nx,ny,nt); snapshot (nz,nx,ny,nt).  A dense source
 * becomes the list of its nonzero cells (any number: a plane or an initial
 * wavefield too), each cell's sources summed in the caller's order. */
int sssi_fwdTimeCpmlFor3dAw(const double *v, int64_t nz, int64_t nx, int64_t ny,
                            const double *source, int64_t nt, int nDiffOrder, int nBoundary,
                            double dz, double dx, double dy, double dt, double *data,
                            double *snapshot);

/* Replaces fwdTimeSpongeFor2dAw (src/fwdTimeSpongeFor2dAw.m:1). */
int sssi_fwdTimeSpongeFor2dAw(const double *v, int64_t nz, int64_t nx, const double *source,
                              int64_t nt, int nBoundary, double dz, double dx, double dt,
                              double *data, double *snapshot);

/* Replaces rvsTimeSpongeFor2dAw (src/rvsTimeSpongeFor2dAw.m:1).
 * model ((nz+2),(nx+2)) may be NULL. */
int sssi_rvsTimeSpongeFor2dAw(const double *v, int64_t nz, int64_t nx, const double *data,
                              int64_t nt, int nBoundary, double dz, double dx, double dt,
                              double *model, double *snapshot);

/* ---------------------------------------------------------------------
 * Signature-preserving extensions used by the gateways when the dense
 * source or the full snapshot stack would not fit (a 1000x3000 grid with
 * 4000 steps needs ~100 GB for each): a point-source list with one wavelet
 * column per point, and a snapshot stride (slice s holds step s*stride).
 * ------------------------------------------------------------------- */

/* Point sources: src_iz/src_ix (0-based, nsrc each), wavelet[t*nsrc + s]
 * (nt x nsrc).  Identical arithmetic to a dense source that is zero away
 * from the points (coincident points add).  snapshot holds ceil(nt/stride)
 * slices, or is NULL. */
int sssi_fwd_cpml_2d_points(const double *v, int64_t nz, int64_t nx, int64_t nt, int nsrc,
                            const int *src_iz, const int *src_ix, const double *wavelet,
                            int nDiffOrder, int nBoundary, double dz, double dx, double dt,
                            double *data, double *snapshot, int64_t snap_stride);

/* No counterpart in the reference: sssi_fwd_cpml_2d_points for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results, bit
 * for bit; every device buffer lives in one library-held array of 64 MiB per
 * GPU (allocated with the library's code, kept until the process ends), and
 * the call makes no device or host allocation.  Problems that need more than
 * the array (roughly nz*nx*12 B for the levels plus nt*nx*4 B of receivers
 * plus 2 MiB of staging), or more than 64 sources, are refused (SSSI_EINVAL).
 * Calls on one GPU share its array and are serialised by the library. */
int sssi_fwd_cpml_2d_points_noalloc(const double *v, int64_t nz, int64_t nx, int64_t nt,
                                    int nsrc, const int *src_iz, const int *src_ix,
                                    const double *wavelet, int nDiffOrder, int nBoundary,
                                    double dz, double dx, double dt, double *data,
                                    double *snapshot, int64_t snap_stride);

int sssi_rvs_cpml_2d_ex(const double *v, int64_t nz, int64_t nx, const double *data, int64_t nt,
                        int nDiffOrder, int nBoundary, double dz, double dx, double dt,
                        double *model, double *snapshot, int64_t snap_stride);

int sssi_fwd_cpml_3d_points(const double *v, int64_t nz, int64_t nx, int64_t ny, int64_t nt,
                            int nsrc, const int *src_iz, const int *src_ix, const int *src_iy,
                            const double *wavelet, int nDiffOrder, int nBoundary, double dz,
                            double dx, double dy, double dt, double *data, double *snapshot,
                            int64_t snap_stride);

int sssi_fwd_sponge_2d_points(const double *v, int64_t nz, int64_t nx, int64_t nt, int nsrc,
                              const int *src_iz, const int *src_ix, const double *wavelet,
                              int nBoundary, double dz, double dx, double dt, double *data,
                              double *snapshot, int64_t snap_stride);

int sssi_rvs_sponge_2d_ex(const double *v, int64_t nz, int64_t nx, const double *data,
                          int64_t nt, int nBoundary, double dz, double dx, double dt,
                          double *model, double *snapshot, int64_t snap_stride);

/* ---------------------------------------------------------------------
 * Device-resident propagation handle (2-D and 3-D).  Created once per
 * model; inputs stay in HBM between runs so repeated propagations (shots,
 * benchmark passes) pay no host transfer.  Runs are asynchronous on the
 * handle's own HIP stream until sssi_prop_sync / a download.
 * ------------------------------------------------------------------- */
typedef struct sssi_prop sssi_prop;

/* 2-D handle.  abc = SSSI_ABC_CPML or SSSI_ABC_SPONGE (nDiffOrder ignored
 * for the sponge).  v column-major (nz,nx). */
int sssi_prop2d_create(const double *v, int64_t nz, int64_t nx, int abc, int nDiffOrder,
                       int nBoundary, double dz, double dx, double dt, sssi_prop **out);
/* 3-D CPML handle, v column-major (nz,nx,ny). */
int sssi_prop3d_create(const double *v, int64_t nz, int64_t nx, int64_t ny, int nDiffOrder,
                       int nBoundary, double dz, double dx, double dy, double dt,
                       sssi_prop **out);
void sssi_prop_destroy(sssi_prop *p);

/* Point sources (2-D: src_iy ignored, may be NULL). wavelet[t*nsrc + s]. */
int sssi_prop_set_points(sssi_prop *p, int nsrc, const int *src_iz, const int *src_ix,
                         const int *src_iy, const double *wavelet, int64_t nt);
/* Surface data (nx,nt) injected at the top row on reverse runs (2-D). */
int sssi_prop_set_surface(sssi_prop *p, const double *data, int64_t nt);
/* Zero the wavefield and memory variables. */
int sssi_prop_reset(sssi_prop *p);
/* Run nt steps: forward (reverse = 0; point sources, t = 0..nt-1) or reverse
 * (reverse = 1; surface data, t = nt-1..0).  Receivers (top row) are
 * recorded on device when record != 0.  Asynchronous. */
int sssi_prop_run(sssi_prop *p, int64_t nt, int reverse, int record);
/* Continue a propagation: count steps from time index t_first (forward:
 * t_first .. t_first+count-1; reverse: t_first down to t_first-count+1).
 * A run of nt steps equals consecutive calls covering 0..nt-1 (forward) or
 * nt-1..0 (reverse), bit for bit.  Asynchronous. */
int sssi_prop_run_steps(sssi_prop *p, int64_t t_first, int64_t count, int reverse, int record);
/* Same as sssi_prop_run, timed with HIP events on the handle's stream;
 * returns elapsed milliseconds of the whole launch sequence and the
 * number of kernel launches it made.  Synchronous. */
int sssi_prop_run_timed(sssi_prop *p, int64_t nt, int reverse, int record, double *ms,
                        int64_t *launches);
int sssi_prop_sync(sssi_prop *p);
/* Receiver traces of the last recorded run, column-major (nx,nt) for 2-D,
 * (nx,ny,nt) for 3-D. */
int sssi_prop_get_data(sssi_prop *p, double *data, int64_t nt);
/* Current wavefield interior (nz,nx[,ny]) column-major, and the previous
 * one padded as the reference's model output (2-D). */
int sssi_prop_get_field(sssi_prop *p, double *field);
int sssi_prop_get_model(sssi_prop *p, double *model);
/* Whole-model statistics of the grid (cells, strip cells) for roofline
 * accounting: algorithmic bytes per step of the propagation kernel. */
int sssi_prop_alg_bytes_per_step(sssi_prop *p, double *bytes, double *cells);

/* Kernel layout in use: 2-D (a, b) = (tile height, interior march groups),
 * 3-D (a, b) = (block y extent, z chunk); 0, 0 for a sponge handle. */
int sssi_prop_layout(sssi_prop *p, int *a, int *b);

/* ---------------------------------------------------------------------
 * Reverse-time migration, 2-D CPML (the imaging loop of
 * applications/rtm_2d_acoustic/mainRtmTimeCpmlFor2dAw.m:147-315).
 * The model V and the migration (smooth) model VS are physical grids
 * (nz,nx) column-major.  Each shot works on the window [x_begin,
 * x_begin+x_count) extended by nBoundary cells (src/extBoundary.m): left,
 * right and below.  Per shot:
 *   dataTrue  = fwd(V, src)             (computed once and cached, :165)
 *   dataSmooth= fwd(VS, src), checkpoints of the source wavefield every
 *               img_stride steps (FP32, BF16, or FP16 with a
 *               per-checkpoint scale)
 *   rvs(V, dataTrue - dataSmooth)       (:252, :256) with the imaging
 *   condition M += f.*r, s2 += f.^2 at the checkpoint steps (:270-273)
 *   Stacked(window) += M ./ (eps + s2)  (:315), FP64, in shot order.
 * img_stride = 1 images every step as the reference does.
 * ------------------------------------------------------------------- */
#define SSSI_CK_F32 1
#define SSSI_CK_F16 2
#define SSSI_CK_BF16 3

typedef struct sssi_rtm sssi_rtm;

int sssi_rtm2d_create(const double *V, const double *VS, int64_t nz, int64_t nx, int nDiffOrder,
                      int nBoundary, double dz, double dx, double dt, int64_t nt, int img_stride,
                      int ck_format, double eps, sssi_rtm **out);
/* Source at depth 0, column src_ix of the window (physical, 0-based);
 * wavelet has nt samples.  Returns the shot index in *shot_id. */
int sssi_rtm2d_add_shot(sssi_rtm *r, int64_t x_begin, int64_t x_count, int src_ix,
                        const double *wavelet, int *shot_id);
/* Compute and cache dataTrue for the shots that lack it. */
int sssi_rtm2d_precompute_data(sssi_rtm *r, int batched);
/* Migrate every shot and stack.  batched = 0: one shot per launch;
 * batched = 1: one launch per time step covers all shots.  Returns the
 * elapsed milliseconds (HIP events, excluding the cached dataTrue) and the
 * number of kernel launches. */
int sssi_rtm2d_migrate(sssi_rtm *r, int batched, double *ms, int64_t *launches);
/* Stacked image (nz,nx) column-major. */
int sssi_rtm2d_get_stack(sssi_rtm *r, double *stack);
/* Per-shot results on the extended window (nz+nB, x_count+2nB) column-major
 * (image, M, s2 without eps) and dataTrue (nx_ext, nt); any may be NULL. */
int sssi_rtm2d_get_shot(sssi_rtm *r, int shot, double *image, double *num, double *den,
                        double *data_true);
/* Durable per-shot results with resume by existence, as the reference's RTM
 * apps keep them (applications/rtm_2d_acoustic/mainRtmTimeCpmlFor2dAw.m:170-184,
 * :259-264; mainRtmTimeSpongeFor2dAw.m:165-179, :255-259, :343-346): with a
 * cache directory set (created if missing; NULL or "" turns caching off),
 * dataTrue<xs>.mat (variable dataTrue, (nx_ext, nt)) and image<xs>.mat
 * (variable image, the shot's image on its extended window) are loaded instead
 * of recomputed when they exist, and written when they do not; after a
 * migration StackedRTM.mat (variable StackedRTM, (nz, nx)) is written if
 * absent.  xs = x_begin + src_ix + 1 is the shot's source column (1-based),
 * the reference's file key.  Files are Level 5 MAT-files (MATLAB `load`
 * reads them; the reference's -v7.3 needs HDF5, absent here), each with an
 * sssiShot variable of the parameters and input checksums it was made from; a
 * file made from other inputs fails the call (SSSI_ESTATE).  A resumed
 * migration equals an uninterrupted one bit for bit. */
int sssi_rtm2d_set_cache(sssi_rtm *r, const char *dir);
/* Shots whose dataTrue / image came from the cache in the last precompute /
 * migration, and files written by the handle so far. */
int sssi_rtm2d_cache_stats(sssi_rtm *r, int *loaded_data, int *loaded_images, int *saved);
/* Device memory held by the handle. */
int sssi_rtm2d_device_bytes(sssi_rtm *r, double *bytes);

/* ---------------------------------------------------------------------
 * Kirchhoff path: traveltimes by the fast-sweeping eikonal solver and
 * Kirchhoff migration (applications/kirchhoff_2d_acoustic/
 * mainKirchhoffCpmlFor2dAw.m:244-266, :278-293).  Arrays are column-major
 * FP64; grid positions and source cells are 1-based, as in MATLAB.
 * ------------------------------------------------------------------- */

/* Replaces eikonal2d_mex (src/fd_cmex/eikonal2d_mex.cpp:27; MATLAB form
 * src/eikonal2d.m): T = eikonal2d(V, dx, sz, sx, iMax), T (nz, nx).  Same
 * arithmetic bit for bit, including the reference's integer truncation of dx
 * (eikonal2d_mex.cpp:35) and its 65535 start value; nz, nx >= 2. */
int sssi_eikonal2d(const double *V, int64_t nz, int64_t nx, double dx, int sz, int sx, int iMax,
                   double *T);
/* The same for nsrc sources at once: T (nz, nx, nsrc). */
int sssi_eikonal2d_batch(const double *V, int64_t nz, int64_t nx, double dx, int nsrc,
                         const int *sz, const int *sx, int iMax, double *T);

/* Replaces migrate (src/migrate.m:1 with src/shot2RecTime.m:1):
 * m = migrate(travelTime, xRecGrid, xShotAndRecGrid, shot, dt, nz, ixs).
 * travelTime (nz, nx, ntab), shot (nt, nx), m (nz, nx); ixs indexes the
 * traveltime tables (as the reference's third dimension).  A sample index
 * outside 1..nt is SSSI_EINVAL (MATLAB's indexing error). */
int sssi_kirchhoff_migrate(const double *travelTime, int64_t nz, int64_t nx, int64_t ntab,
                           const int *xRecGrid, int nrec, const int *xShotAndRecGrid, int nsr,
                           const double *shot, int64_t nt, double dt, int ixs, double *m);

/* Resident form of the app's loops: traveltime tables of every surface
 * position in xShotAndRecGrid (sources at row 1, as the app's [1 xs]) by the
 * eikonal solver, kept on the device with their sample indices for dt; then
 * the migration of many shots, stacked in shot order as the app's
 * Stacked = Stacked + M (mainKirchhoffCpmlFor2dAw.m:289-291).  No
 * counterpart call in the reference (its app uses ray2d.m for the tables). */
typedef struct sssi_kirchhoff sssi_kirchhoff;
int sssi_kirchhoff_create(const double *V, int64_t nz, int64_t nx, double dx,
                          const int *xShotAndRecGrid, int nsr, int iMax, double dt,
                          sssi_kirchhoff **out);
/* travelTime (nz, nx, nsr). */
int sssi_kirchhoff_travel_time(sssi_kirchhoff *h, double *travelTime);
/* shots (nt, nx, nshot), ixs[nshot] table indices (1-based); stack (nz, nx),
 * images (nz, nx, nshot) or NULL. */
int sssi_kirchhoff_migrate_shots(sssi_kirchhoff *h, const int *xRecGrid, int nrec,
                                 const double *shots, int64_t nt, int nshot, const int *ixs,
                                 double *stack, double *images);
/* The same in two halves, for records that stay on the device: load the
 * records (nt, nx, nshot) into the handle (replacing any loaded before), then
 * migrate the loaded records as often as needed (ixs[nshot]). */
int sssi_kirchhoff_load_shots(sssi_kirchhoff *h, const double *shots, int64_t nt, int nshot);
int sssi_kirchhoff_migrate_loaded(sssi_kirchhoff *h, const int *xRecGrid, int nrec,
                                  const int *ixs, double *stack, double *images);
/* Kernel time (HIP events on the handle's stream) of the eikonal passes at
 * create, their count and the launches that ran them, and the kernel time of
 * the last migration. */
int sssi_kirchhoff_timing(sssi_kirchhoff *h, double *eikonal_ms, int64_t *passes,
                          int64_t *launches, double *migrate_ms);
void sssi_kirchhoff_destroy(sssi_kirchhoff *h);

/* ---------------------------------------------------------------------
 * Frequency-domain CPML Helmholtz (src/freqCpmlFor2dAw.m:1, called per
 * frequency by applications/fwi_2d_acoustic/mainFwiFreqCpml2dAw.m:221 and the
 * LSRTM/GN-FWI apps).  model is the squared slowness (nz, nx), column-major
 * FP64; complex arrays are interleaved (re, im) FP64 pairs, (nz*nx, nsrc)
 * column-major.  Cell (i, j) is unknown j*nz + i, as MATLAB's reshape.
 * ------------------------------------------------------------------- */

/* [~, snapshot] = freqCpmlFor2dAw(model, source, w, nDiffOrder, nBoundary,
 * dz, dx): snapshot = A \ (-source), nsrc right-hand sides, solved on the
 * device by banded LU with partial pivoting (bandwidth (2 nDiffOrder - 1) nz).
 * The band must fit the solve's LDS window: (2 nDiffOrder - 1) nz up to about
 * 5000 on gfx950 (SSSI_EINVAL beyond, with the limit in the message). */
int sssi_freqCpmlFor2dAw(const double *model, int64_t nz, int64_t nx, const double *source,
                         int nsrc, double w, int nDiffOrder, int nBoundary, double dz, double dx,
                         double *snapshot);
/* A = freqCpmlFor2dAw(...): the impedance matrix in CSR (rows ascending,
 * columns ascending within a row): *nnz always; with rowptr non-NULL also
 * rowptr[nz*nx+1], col[nnz] and val[2*nnz] (complex interleaved). */
int sssi_freq_cpml_2d_matrix(const double *model, int64_t nz, int64_t nx, double w,
                             int nDiffOrder, int nBoundary, double dz, double dx, int64_t *nnz,
                             int64_t *rowptr, int64_t *col, double *val);
/* Kernel time (HIP events) of the calling thread's last sssi_freqCpmlFor2dAw:
 * factorisation and solve. */
int sssi_freq_last_timing(double *factor_ms, double *solve_ms);

/* ---------------------------------------------------------------------
 * Calibration (no counterpart in the reference): the device's achievable
 * streaming rate for the access pattern of a stencil step at its minimum,
 * three FP32 arrays of ncells read and one written once per launch
 * (16 B per cell), averaged over reps launches with HIP events.
 * ------------------------------------------------------------------- */
int sssi_calib_stream(int64_t ncells, int reps, double *us_per_launch, double *gb_per_s);

/* ---------------------------------------------------------------------
 * MAT-file (Level 5) I/O of real double arrays: the durable format of the
 * results above, for callers without MATLAB.  sssi_mat_write writes nvar
 * variables (names[v], data[v] column-major, ndims[v] dimensions each, dims
 * concatenated) to a temporary file renamed over path.  sssi_mat_read finds
 * variable `name` and returns its dimensions (up to 8) and, with out != NULL,
 * its n values (any numeric storage type, zlib-compressed elements included).
 * No counterpart in the reference, whose apps call MATLAB's save/load.
 * ------------------------------------------------------------------- */
int sssi_mat_write(const char *path, int nvar, const char *const *names,
                   const double *const *data, const int *ndims, const int64_t *dims);
int sssi_mat_read(const char *path, const char *name, int *ndims, int64_t *dims, double *out,
                  int64_t n);
void sssi_rtm2d_destroy(sssi_rtm *r);

/* ---------------------------------------------------------------------
 * Serving: a request loop with continuous batching.  Jobs are submitted at
 * any time (inputs are copied, or borrowed where stated); a worker thread admits them when a slot is
 * free and their device footprint fits the HBM budget, advances every
 * resident job one time step per tick (the 2-D jobs of one nDiffOrder share
 * one launch), keeps wavefields and checkpoints resident until the job ends
 * and frees them when its result is copied out.  Results do not depend on
 * co-resident jobs.  Job kinds: 2-D CPML forward (result: data (nx,nt)),
 * 2-D CPML RTM shot on the given extended grid (source at depth 0, column
 * src_ix; result: image (nz,nx) = M ./ (eps + s2)), 3-D CPML forward
 * (result: data (nx,ny,nt)).  No counterpart in the reference (its shots run
 * one after another in MATLAB loops, mainRtmTimeCpmlFor2dAw.m:147).
 * ------------------------------------------------------------------- */
typedef struct sssi_server sssi_server;

int sssi_server_create(double hbm_budget_bytes, int max_active, sssi_server **out);
int sssi_server_submit_fwd2d(sssi_server *s, const double *v, int64_t nz, int64_t nx, int64_t nt,
                             int nDiffOrder, int nBoundary, double dz, double dx, double dt,
                             int src_iz, int src_ix, const double *wavelet, int64_t *job_id);
int sssi_server_submit_rtm2d(sssi_server *s, const double *V, const double *VS, int64_t nz,
                             int64_t nx, int64_t nt, int nDiffOrder, int nBoundary, double dz,
                             double dx, double dt, int src_ix, const double *wavelet,
                             int img_stride, int ck_format, double eps, int64_t *job_id);
/* An RTM shot whose observed data dataTrue (nx, nt) column-major is already
 * known -- the reference's apps cache it per shot and load it instead of
 * recomputing (applications/rtm_2d_acoustic/mainRtmTimeCpmlFor2dAw.m:170-184):
 * the job skips the fwd(V) propagation.  data_true = NULL computes it, as
 * sssi_server_submit_rtm2d does. */
int sssi_server_submit_rtm2d_data(sssi_server *s, const double *V, const double *VS, int64_t nz,
                                  int64_t nx, int64_t nt, int nDiffOrder, int nBoundary,
                                  double dz, double dx, double dt, int src_ix,
                                  const double *wavelet, const double *data_true, int img_stride,
                                  int ck_format, double eps, int64_t *job_id);
/* The same job with borrowed inputs: V, VS and data_true are read when the
 * job is set up on the device instead of copied at submission, so a burst of
 * submissions returns at once.  The caller keeps them valid and unchanged
 * until sssi_server_wait reports the job done.  No counterpart in the
 * reference (serving only). */
int sssi_server_submit_rtm2d_borrowed(sssi_server *s, const double *V, const double *VS,
                                      int64_t nz, int64_t nx, int64_t nt, int nDiffOrder,
                                      int nBoundary, double dz, double dx, double dt, int src_ix,
                                      const double *wavelet, const double *data_true,
                                      int img_stride, int ck_format, double eps,
                                      int64_t *job_id);
int sssi_server_submit_fwd3d(sssi_server *s, const double *v, int64_t nz, int64_t nx, int64_t ny,
                             int64_t nt, int nDiffOrder, int nBoundary, double dz, double dx,
                             double dy, double dt, int src_iz, int src_ix, int src_iy,
                             const double *wavelet, int64_t *job_id);
/* Wait for a job (timeout_s < 0: no limit); *done = 1 when finished. */
int sssi_server_wait(sssi_server *s, int64_t job_id, double timeout_s, int *done);
/* Wait until every submitted job has finished. */
int sssi_server_drain(sssi_server *s);
/* Result of a finished job (n values, layout as stated above). */
int sssi_server_result(sssi_server *s, int64_t job_id, double *out, int64_t n);
/* Host timestamps (seconds, steady clock) of arrival, admission and
 * completion, and the device bytes the job reserved. */
int sssi_server_job_info(sssi_server *s, int64_t job_id, double *t_arrive, double *t_admit,
                         double *t_done, double *bytes);
/* Totals: cell updates, seconds the worker spent issuing work, launches, ticks. */
int sssi_server_stats(sssi_server *s, double *cell_updates, double *busy_s, int64_t *launches,
                      int64_t *ticks);
void sssi_server_destroy(sssi_server *s);

#ifdef __cplusplus
}
#endif
#endif /* SSSI_H */
