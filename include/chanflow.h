/*
 * chanflow.h -- C ABI of the MI355X channel-flow DNS time step.
 *
 * One run = one incompressible channel flow (periodic x and y, walls in
 * z) advanced by the reference solver's low-storage RK3 step: staggered
 * 2nd/4th-order finite differences, implicit wall-normal diffusion and an
 * FFT + banded-solve pressure projection.  All state lives in device memory
 * of one GPU; host buffers are borrowed only for the duration of a call.
 *
 * Axis names follow the reference: x streamwise (periodic, nx cells),
 * y spanwise (periodic, ny cells), z wall-normal (walls, nz cells).  A
 * grid quoted "nx x nz x ny" (streamwise, wall-normal, spanwise, as in the
 * benchmark tables) is nx, ny, nz here in the order x, y, z.
 *
 * Host field layout ("io layout") is the reference's restart-file layout,
 * io.c:211-296: raw native-endian doubles, C-order [nx][ny][nz], interior
 * cells only (no ghosts); for w the nz values are the faces above each
 * cell, the last one being the top wall.
 *
 * Errors: every entry point returns a cf_status; cf_last_error() gives a
 * thread-local message.  Nothing aborts (the reference aborts the whole
 * MPI job in check(), check.h:8-12).
 *
 * Threading: calls on distinct run handles are independent; calls on one
 * handle must be serialised by the caller.
 */
#ifndef CHANFLOW_H
#define CHANFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cf_status {
    CF_OK = 0,
    CF_ERR_ARG = 1,          /* invalid argument / configuration */
    CF_ERR_UNSUPPORTED = 2,  /* valid but not built (e.g. FFT length with prime factor > 5) */
    CF_ERR_NOMEM = 3,        /* device or host allocation failed */
    CF_ERR_DEVICE = 4,       /* HIP runtime error or no GPU */
    CF_ERR_NUMERIC = 5,      /* zero pivot / non-finite value detected */
    CF_ERR_STATE = 6         /* call not valid in the handle's current state */
} cf_status;

enum { CF_FP64 = 0, CF_FP32 = 1 };
enum { CF_STRETCH_COSINE = 0, CF_STRETCH_TANH = 1, CF_STRETCH_UNIFORM = 2, CF_STRETCH_USER = 3 };
enum { CF_U = 0, CF_V = 1, CF_W = 2, CF_C = 3, CF_P = 4 };

/* Run configuration: the parameters of chanfast2/chanfast4
 * (chanfast.h:1-21, main.c:45-61) plus the grid stretching that the
 * reference hard-codes (chanfast.c:184-189) and the precision mode. */
typedef struct cf_config {
    int nx, ny, nz;          /* cells along x, y, z */
    int order;               /* 2 or 4 */
    int scalar;              /* 1: advance the scalar c (needed if any betg != 0) */
    int precision;           /* CF_FP64 or CF_FP32 */
    double lx, ly, lz;       /* box; walls at z = 0 and z = lz */
    int stretch;             /* CF_STRETCH_* */
    double stretch_beta;     /* tanh parameter for CF_STRETCH_TANH */
    const double *dz_user;   /* nz cell heights for CF_STRETCH_USER (copied at create) */
    double dt, want_cfl;
    int cflmode;             /* 1: rescale dt every step so that CFL = want_cfl */
    double nu, prandtl;
    double dpdx, want_uvolavg;
    int uvolavgmode;         /* 1: constant mass flux (bulk u = want_uvolavg) */
    double want_cvolavg;
    int cvolavgmode;         /* 1: constant bulk scalar */
    double uframe, vframe;   /* wall velocities are -uframe, -vframe */
    double betg_x, betg_y, betg_z;   /* Boussinesq buoyancy coefficients */
    int device;              /* HIP device ordinal */
    int driver;              /* CF_DRIVER_CHANFAST or CF_DRIVER_CHANNEL */
} cf_config;

/* The two time-step drivers of the reference: chanfast2/4 (chanfast.c,
 * x-y diffusion explicit, z diffusion implicit by 1-D banded solves) and
 * channel2/4 (channel.c:71-1968, every diffusion term implicit: the 3-D
 * Helmholtz operator on the right-hand side and a spectral solve
 * FFT3D_FWD -> per (kx, ky) banded z solve -> FFT3D_BWD per field,
 * diffuseu.c:460-614, diffusew.c:466-622).  CF_DRIVER_CHANNEL runs in
 * FP64 on single runs (not in batches; served runs step on their own). */
enum { CF_DRIVER_CHANFAST = 0, CF_DRIVER_CHANNEL = 1 };

/* Per-step log record, the values the reference appends to cfl.dat,
 * pstat.dat and cstat.dat (chanfast.c:392-393, :526-529, :613-616). */
#define CF_LOG_N 13
/* [0] completed step number, [1] dt, [2] cfl at step start,
 * [3] mdpdx, [4] nu dU/dz bottom wall, [5] nu dU/dz top wall, [6] bulk U, [7] centreline U,
 * [8] dcdt, [9] kappa dC/dz bottom, [10] kappa dC/dz top, [11] bulk C, [12] centreline C */

typedef struct cf_run cf_run;

/* ---- host-only helpers (no GPU needed) ---- */

/* Fill cfg with the reference main.c defaults scaled to a plain channel:
 * order 2, no scalar, FP64, lx = 2 pi, ly = pi, lz = 2, cosine stretching,
 * dt = 1e-3, nu = 1/2800, constant mass flux 1. */
void cf_config_default(cf_config *cfg);

/* Wall-normal cell heights (nz values) for cfg's stretching.
 * Cosine follows chanfast.c:184-189; tanh places the faces at
 * z_j = lz/2 (1 + tanh(beta (2j/nz - 1)) / tanh(beta)). */
cf_status cf_grid_dz(const cf_config *cfg, double *dz);

/* Seeded synthetic initial fields in io layout, not yet projected:
 * u = 1.5 Ub (1 - (2 zeta/lz - 1)^2) + amp n_u, v = amp n_v, w = amp n_w
 * (top-wall face 0), c = zeta/lz - 0.5 + camp n_c, with zeta the cell-centre
 * height and n_* uniform in [-1, 1) from a SplitMix64 counter keyed by
 * (seed, field, ix, iy, iz).  Any pointer may be NULL. */
cf_status cf_initial_fields(const cf_config *cfg, uint64_t seed, double amp, double camp,
                            double *u, double *v, double *w, double *c);

/* Raw restart-file IO in the reference byte format (io.c:211-296) and
 * its ascii logs (io.c:5-35, "%+e"). */
cf_status cf_io_write(const char *path, const double *buf, size_t n);
cf_status cf_io_read(const char *path, double *buf, size_t n);
cf_status cf_io_append_ascii(const char *path, const double *vals, int n);
/* io_write_ascii (io.c:22-34): n values, one "%+e" per line */
cf_status cf_io_write_ascii(const char *path, const double *vals, int n);

/* get_z_at_p / get_z_at_w (get.c:458-488) for nghost >= 1 ghost levels
 * per wall (heights mirrored, ghost.c:31-54): zp, zw of nz + 2 nghost
 * values each, cell centres and the faces above each level, measured from
 * the bottom wall (the zp.dat / zw.dat files of chanfast.c:855-856).
 * Either output may be NULL. */
cf_status cf_grid_z(const cf_config *cfg, int nghost, double *zp, double *zw);

const char *cf_last_error(void);
const char *cf_version(void);

/* ---- run API (replaces the chanfast2/chanfast4 driver, chanfast.c:71-994,
 *      :995-1916, and its MPI/FFTW pencil machinery) ---- */

cf_status cf_run_create(const cf_config *cfg, cf_run **out);
cf_status cf_run_destroy(cf_run *run);

/* Copy one field in/out in io layout.  CF_P is download-only (the
 * pressure of the last projection). */
cf_status cf_run_upload(cf_run *run, int field, const double *host_io);
cf_status cf_run_download(cf_run *run, int field, double *host_io);

/* Make u, v, w discretely divergence free: divergence -> FFT -> banded
 * Poisson with the given dt -> inverse FFT -> projection, as cases.c:417-470
 * does.  The w wall faces are set to 0 first. */
cf_status cf_run_project(cf_run *run, double dt);

/* Advance nsteps RK3 steps (asynchronous on the run's stream). */
cf_status cf_run_step(cf_run *run, int nsteps);
cf_status cf_run_sync(cf_run *run);

/* Per-step log records of all steps done so far (at most max_steps most
 * recent), CF_LOG_N doubles each. */
cf_status cf_run_log(cf_run *run, double *out, int max_steps, int *nsteps_out);

/* Max |divergence| over interior cells of the current velocity. */
cf_status cf_run_max_divergence(cf_run *run, double *out);

/* Current dt and completed step count. */
cf_status cf_run_info(cf_run *run, double *dt, int *steps_done);

/* Device bytes held by one run of this configuration (admission control). */
cf_status cf_run_footprint(const cf_config *cfg, size_t *bytes);

/* Average duration (ms) of the named kernel over the last cf_run_step call
 * when profiling is on (cf_run_set_profiling); for the roofline figures. */
cf_status cf_run_set_profiling(cf_run *run, int on);
cf_status cf_run_kernel_time(cf_run *run, const char *kernel, double *avg_ms, int *launches);

/* ---- plane statistics of the current state (the istat block of
 *      chanfast.c:658-861) ---- */

enum { CF_STAT_AVG = 0, CF_STAT_CORR = 1, CF_STAT_MOMENT = 2, CF_STAT_SPEC = 3 };

/* Statistics over the x-y plane at each of the nz interior levels of the
 * fields CF_U, CF_V, CF_W, CF_C (scalar runs) or CF_P:
 *   CF_STAT_AVG     get_xy_avg (get.c:281-307) of field_a: out[nz]
 *   CF_STAT_CORR    get_xy_corr (get.c:136-174), mean of the product of the
 *                   deviations of field_a and field_b: out[nz]
 *   CF_STAT_MOMENT  get_xy_moment (get.c:100-133), mean of
 *                   (field_a - mean)^en: out[nz]
 *   CF_STAT_SPEC    get_xy_spec (get.c:5-97) of the FFT3D spectra (1/nx,
 *                   1/ny scaled) of field_a and field_b: out = specx
 *                   [nx/2+1][nz][re, im], then specy [ny/2+1][nz][re, im]
 * The reference's profiles also list its ghost levels, which hold solver
 * work values rather than flow statistics; those are not computed. */
cf_status cf_run_xy_stats(cf_run *run, int kind, int field_a, int field_b, double en, double *out);

/* cf_run_xy_stats for kinds CF_STAT_AVG, CF_STAT_CORR and CF_STAT_MOMENT
 * without any device allocation: the work space (partial sums, means,
 * result) is a fixed array in device memory, one per GPU for the life of
 * the process, sized for nz <= 2048 levels (larger runs get CF_ERR_ARG).
 * No counterpart in the reference: for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Calls that overlap on one GPU share that array
 * and race; results equal cf_run_xy_stats bit for bit. */
cf_status cf_run_xy_stats_noalloc(cf_run *run, int kind, int field_a, int field_b, double en, double *out);

/* ---- operator-level entry points (parity tests of single stages) ---- */

/* Layout transpose io [nx][ny][nz] <-> device [nz][ny][nx] through the
 * device (upload + LDS-tiled transpose + back); bit-exact data movement.
 * Stands in for the pencil transposes of transpose.c:409-448. */
cf_status cf_op_transpose_roundtrip(cf_run *run, const double *in_io, double *dev_layout_out,
                                    double *io_again_out);

/* Forward FFT3D of a real interior field (io layout, nz levels):
 * r2c along x scaled 1/nx, c2c along y scaled 1/ny, output
 * [ix][iy][iz][re,im] with ix < nx/2+1 (chanfast.c:21-37 ordering,
 * fft.c:161-202 scaling).  Backward is the unscaled inverse. */
cf_status cf_op_fft3d_fwd(cf_run *run, const double *in_io, double *spec_out);
cf_status cf_op_fft3d_bwd(cf_run *run, const double *spec_in, double *out_io);

/* Banded Poisson of pressure_poisson2/4 (pressure.c:293-444) on a spectral
 * array [ix][iy][iz][re,im] (interior levels), in place, through the device. */
cf_status cf_op_poisson(cf_run *run, double *spec_inout, double dt);

/* ---- batches: independent runs stepped together ----
 *
 * The reference steps one simulation per process (chanfast2/chanfast4,
 * chanfast.c:71-994, :995-1916); a batch holds runs of one order, scalar
 * setting and precision whose grids, viscosities and time steps may all
 * differ, and steps them together: each stage of the substep is one
 * persistent launch over all runs' work, the step one captured graph.  At
 * creation an autotuner times the launch shapes on the batch's own grids
 * (persistent grid per launch; for order 2 the z-solve shape).  Each run
 * keeps the semantics of cf_run_* on the same configuration. */
typedef struct cf_batch cf_batch;

cf_status cf_batch_create(const cf_config *cfgs, int nrun, cf_batch **out);
cf_status cf_batch_destroy(cf_batch *batch);
cf_status cf_batch_upload(cf_batch *batch, int run, int field, const double *host_io);
cf_status cf_batch_download(cf_batch *batch, int run, int field, double *host_io);
/* project every run's (u, v, w) as cf_run_project does */
cf_status cf_batch_project(cf_batch *batch, double dt);
cf_status cf_batch_step(cf_batch *batch, int nsteps);
cf_status cf_batch_sync(cf_batch *batch);
cf_status cf_batch_log(cf_batch *batch, int run, double *out, int max_steps, int *nsteps_out);
cf_status cf_batch_max_divergence(cf_batch *batch, int run, double *out);
cf_status cf_batch_set_profiling(cf_batch *batch, int on);
cf_status cf_batch_kernel_time(cf_batch *batch, const char *kernel, double *avg_ms, int *launches);
/* the autotuner's choices, one line per launch, into buf */
cf_status cf_batch_tuning(cf_batch *batch, char *buf, size_t len);

/* ---- serving: continuous batching of independent runs ----
 *
 * No counterpart in the reference (one run per process).  Requests are
 * queued by cf_server_submit.  Device work goes out in quanta of a few
 * steps: one per group of active runs sharing order / scalar / precision
 * (at most CF_BATCH_MAX = 256 runs, batched launches on the group's own
 * stream) and one per run of at least the solo size
 * (cf_server_set_solo_points, default 4 Mi cells; its own single-run graph
 * on its own stream).  Each cf_server_tick takes in the quanta that have
 * completed (crediting steps, retiring runs that reached their count),
 * admits queued requests in arrival order while their footprint
 * (cf_run_footprint) fits the HBM budget and projects them once (dt = 1,
 * as cf_run_project), and enqueues the next quantum of every idle group
 * and of the solo runs with the least remaining cells x steps while fewer
 * than the solo limit (cf_server_set_solo_limit, default 1) are in flight;
 * if nothing moved it first waits for one quantum to complete.  Runs with
 * cflmode 1 or the fully implicit driver always step as solo runs.
 * Groups never wait for each other.  A request whose group is full waits
 * while later ones pass; a request that cannot be set up is marked failed
 * (cf_server_request_state) and dropped, and the server goes on.  Times
 * are seconds since cf_server_create.  A run's result does not depend on
 * which other runs share the device. */
typedef struct cf_server cf_server;

cf_status cf_server_create(size_t hbm_budget_bytes, int device, int quantum_steps, cf_server **out);
cf_status cf_server_destroy(cf_server *server);
/* runs with nx ny nz >= points step on their own (0: every run; SIZE_MAX:
 * none); applies to requests admitted afterwards */
cf_status cf_server_set_solo_points(cf_server *server, size_t points);
/* at most n solo runs advance at once (default 1): each quantum goes to
 * the solo runs with the least remaining cells x steps */
cf_status cf_server_set_solo_limit(cf_server *server, int n);
/* copies the io-layout initial fields (c may be NULL without the scalar);
 * retain != 0 keeps the finished run resident for cf_server_result */
cf_status cf_server_submit(cf_server *server, const cf_config *cfg, const double *u, const double *v,
                           const double *w, const double *c, int nsteps, int retain, int *request_id);
cf_status cf_server_tick(cf_server *server, int *n_active, int *n_queued, int *n_done);
cf_status cf_server_result(cf_server *server, int request_id, int field, double *host_io);
cf_status cf_server_log(cf_server *server, int request_id, double *out, int max_steps, int *nsteps_out);
cf_status cf_server_release(cf_server *server, int request_id);
cf_status cf_server_request_info(cf_server *server, int request_id, double *t_submit, double *t_admit,
                                 double *t_done, int *steps_done);
/* 0 queued, 1 active, 2 done (retained), 3 done (released), 4 failed;
 * msg (may be NULL) receives why a failed request failed */
cf_status cf_server_request_state(cf_server *server, int request_id, int *state, char *msg, size_t len);
cf_status cf_server_stats(cf_server *server, double *seconds, double *cell_steps, size_t *hbm_in_use);

#ifdef __cplusplus
}
#endif

#endif
