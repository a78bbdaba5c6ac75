/*
 * chanflow_ops.h -- the reference's operator layer, on the GPU.
 *
 * Host-pointer functions with the names and argument lists of the
 * reference's operator headers (pressure.h, diffuseu.h, diffusew.h,
 * decomp.h), so that code written against those headers -- the reference's
 * own cases.c invariants, for one -- can drive the HIP operators one stage
 * at a time.  Each call copies its arrays to the device, runs the stage's
 * HIP kernel on the reference's padded layout and copies the results back;
 * it is for parity checks and tools, not for the time step (the run API of
 * chanflow.h keeps everything resident).
 *
 * Layout and conventions are the reference's: fields are flat [n0][n1][n2]
 * doubles, n0 = local_nx + 2 nghost_x (likewise n1, n2), k fastest; dz[]
 * holds n2 cell heights (ghosts mirrored, ghost.c:31-54).  One GPU holds
 * the whole domain, so local_nx = nx, local_ny = ny and the start offsets
 * are 0; the three nghost values must be equal (3 for order 4).  Every
 * array entry the reference's loop leaves alone is left alone here.
 *
 * Errors: like check() (check.h:8-12), which prints file:line and aborts
 * the job, an invalid argument or a HIP failure prints the reason to
 * stderr and calls abort().
 *
 * The *_initialize / *_finalize calls of the reference set the module's
 * order and allocate its scratch; here the order is in the function name,
 * so they only release the cached spectral solver (pressure_finalize*).
 */
#ifndef CHANFLOW_OPS_H
#define CHANFLOW_OPS_H

#ifdef __cplusplus
extern "C" {
#endif

/* fftw_complex (fftw3.h): two doubles, real then imaginary */
typedef double cf_fftw_complex[2];

/* boundary kinds: diffuseu.h:1-2, diffusew.h:1 */
#ifndef DIRICHU
#define DIRICHU (0)
#define NEUMANU (1)
#define DIRICHW (20)
#endif

/* ---- pressure.h ---- */

/* pressure.c:40-79 / :80-118: d = discrete divergence at the cells
 * (order 4 first mirrors the w ghosts oddly about the wall faces, in w) */
void divergence2(double u[], double v[], double w[], double d[], int local_nx, int local_ny, int nz,
                 int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[]);
void divergence4(double u[], double v[], double w[], double d[], int local_nx, int local_ny, int nz,
                 int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[]);
/* pressure.c:120-157 / :158-195: u -= dt grad p (w on the wall faces too) */
void project2(double u[], double v[], double w[], double p[], int local_nx, int local_ny, int nz,
              int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[], double dt);
void project4(double u[], double v[], double w[], double p[], int local_nx, int local_ny, int nz,
              int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[], double dt);
/* pressure.c:197-230 / :231-291: l = dt lap(p) at the cells */
void pressure_laplacian2(double p[], double l[], int local_nx, int local_ny, int nz,
                         int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[], double dt);
void pressure_laplacian4(double p[], double l[], int local_nx, int local_ny, int nz,
                         int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[], double dt);
/* pressure.c:293-355 / :357-444: the banded Poisson of every (kx, ky) of
 * cout[(ix local_ny + iy) n2 + k], in place, ghost levels filled */
void pressure_poisson2(cf_fftw_complex cout[], int nx, int local_nx, int local_x_start, int ny, int local_ny,
                       int local_y_start, int nz, int nghost_z, double dx, double dy, double dz[], double dt);
void pressure_poisson4(cf_fftw_complex cout[], int nx, int local_nx, int local_x_start, int ny, int local_ny,
                       int local_y_start, int nz, int nghost_z, double dx, double dy, double dz[], double dt);
/* pressure.c:772-850 */
void pressure_initialize2(int nz, int nghost_z);
void pressure_initialize4(int nz, int nghost_z);
void pressure_finalize2(void);
void pressure_finalize4(void);

/* ---- diffuseu.h (cell-centred u, v, c) ---- */

/* diffuseu.c:37-104: wall ghosts of every column from bc(x[i], y[j]) */
void apply_bc_u2(double u[], int local_nx, int local_ny, int nz, int nghost_x, int nghost_y, int nghost_z,
                 double xu[], double yu[], double dz[], double (*bc_bot)(double, double),
                 double (*bc_top)(double, double), int bctype_bot, int bctype_top);
void apply_bc_u4(double u[], int local_nx, int local_ny, int nz, int nghost_x, int nghost_y, int nghost_z,
                 double xu[], double yu[], double dz[], double (*bc_bot)(double, double),
                 double (*bc_top)(double, double), int bctype_bot, int bctype_top);
/* diffuseu.c:107-170: h = u + nudt d2u/dz2 at the interior cells */
void helmholtz1d_operator_u2(double u[], double h[], int local_nx, int local_ny, int nz,
                             int nghost_x, int nghost_y, int nghost_z, double dz[], double nudt);
void helmholtz1d_operator_u4(double u[], double h[], int local_nx, int local_ny, int nz,
                             int nghost_x, int nghost_y, int nghost_z, double dz[], double nudt);
/* diffuseu.c:173-283: solve (1 + nudt d2/dz2) u = h per column with wall
 * values (or gradients) avg_bot / avg_top; u gets every level */
void solve_helmholtz1d_u2(double u[], double h[], int local_nx, int local_ny, int nz,
                          int nghost_x, int nghost_y, int nghost_z, double dz[], double nudt,
                          double avg_bot, double avg_top, int bctype_bot, int bctype_top);
void solve_helmholtz1d_u4(double u[], double h[], int local_nx, int local_ny, int nz,
                          int nghost_x, int nghost_y, int nghost_z, double dz[], double nudt,
                          double avg_bot, double avg_top, int bctype_bot, int bctype_top);
/* diffuseu.c:286-359: h += nu (d2/dx2 + d2/dy2) u */
void add_laplacexy_operator_u2(double u[], double h[], int local_nx, int local_ny, int nz,
                               int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double nu);
void add_laplacexy_operator_u4(double u[], double h[], int local_nx, int local_ny, int nz,
                               int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double nu);
/* diffuseu.c:362-457: h = u + nudt lap(u) */
void helmholtz_operator_u2(double u[], double h[], int local_nx, int local_ny, int nz,
                           int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[], double nudt);
void helmholtz_operator_u4(double u[], double h[], int local_nx, int local_ny, int nz,
                           int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[], double nudt);
void diffuseu_initialize2(int nz, int nghost_z);
void diffuseu_initialize4(int nz, int nghost_z);
void diffuseu_finalize2(void);
void diffuseu_finalize4(void);

/* ---- diffusew.h (face-centred w) ---- */

/* diffusew.c:37-104 */
void apply_bc_w2(double w[], int local_nx, int local_ny, int nz, int nghost_x, int nghost_y, int nghost_z,
                 double xw[], double yw[], double dz[], double (*bc_bot)(double, double),
                 double (*bc_top)(double, double), int bctype_bot, int bctype_top);
void apply_bc_w4(double w[], int local_nx, int local_ny, int nz, int nghost_x, int nghost_y, int nghost_z,
                 double xw[], double yw[], double dz[], double (*bc_bot)(double, double),
                 double (*bc_top)(double, double), int bctype_bot, int bctype_top);
/* diffusew.c:107-172 */
void helmholtz1d_operator_w2(double w[], double h[], int local_nx, int local_ny, int nz,
                             int nghost_x, int nghost_y, int nghost_z, double dz[], double nudt);
void helmholtz1d_operator_w4(double w[], double h[], int local_nx, int local_ny, int nz,
                             int nghost_x, int nghost_y, int nghost_z, double dz[], double nudt);
/* diffusew.c:175-287 */
void solve_helmholtz1d_w2(double w[], double h[], int local_nx, int local_ny, int nz,
                          int nghost_x, int nghost_y, int nghost_z, double dz[], double nudt,
                          double avg_bot, double avg_top, int bctype_bot, int bctype_top);
void solve_helmholtz1d_w4(double w[], double h[], int local_nx, int local_ny, int nz,
                          int nghost_x, int nghost_y, int nghost_z, double dz[], double nudt,
                          double avg_bot, double avg_top, int bctype_bot, int bctype_top);
/* diffusew.c:290-365 */
void add_laplacexy_operator_w2(double w[], double h[], int local_nx, int local_ny, int nz,
                               int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double nu);
void add_laplacexy_operator_w4(double w[], double h[], int local_nx, int local_ny, int nz,
                               int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double nu);
/* diffusew.c:366-465 */
void helmholtz_operator_w2(double w[], double h[], int local_nx, int local_ny, int nz,
                           int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[], double nudt);
void helmholtz_operator_w4(double w[], double h[], int local_nx, int local_ny, int nz,
                           int nghost_x, int nghost_y, int nghost_z, double dx, double dy, double dz[], double nudt);
void diffusew_initialize2(int nz, int nghost_z);
void diffusew_initialize4(int nz, int nghost_z);
void diffusew_finalize2(void);
void diffusew_finalize4(void);

/* ---- decomp.h: pencil index maps (1-based, larger blocks last) ---- */

/* decomp.c:18-66 */
void partition(int nx, int ny, int nz, int pdim[3], int dims[2], int coord[2], int lstart[3], int lend[3],
               int lsize[3]);
/* decomp.c:72-98 */
void distribute(int data1, int proc, int st[], int en[], int sz[]);
/* decomp.c:105-118 */
void get_dist(int nx, int ny, int nz, int dims[2], int x1st[], int x1en[], int x1dist[], int y1st[], int y1en[],
              int y1dist[], int y2st[], int y2en[], int y2dist[], int z2st[], int z2en[], int z2dist[]);

/* decomp.h:9-31 */
typedef struct decomp {
    int xst[3], xen[3], xsz[3];
    int yst[3], yen[3], ysz[3];
    int zst[3], zen[3], zsz[3];
    int *x1st, *x1en, *x1dist;
    int *y1st, *y1en, *y1dist;
    int *y2st, *y2en, *y2dist;
    int *z2st, *z2en, *z2dist;
} decomp_plan;

/* decomp.c:121-168, :171-186 */
decomp_plan *create_decomp_plan(int dims[2], int coord[2], int nx, int ny, int nz);
void destroy_decomp_plan(decomp_plan *p);

/* The pencil the rank at coord of a dims[0] x dims[1] process grid holds of
 * a global [nx][ny][nz] field (the reference's io layout): pencil 0, 1, 2 =
 * the x, y, z pencil of create_decomp_plan; out is [sz0][sz1][sz2], last
 * index fastest, as the reference's local arrays.  Gathered on the device.
 * Stands in for the redistribution of transpose.c:409-448 on one GPU.
 * Returns 0, or -1 on a bad argument (message in cf_last_error). */
int cf_pencil_gather(const double *global_io, int nx, int ny, int nz, int dims[2], int coord[2], int pencil,
                     double *out);

#ifdef __cplusplus
}
#endif

#endif
