/*
 * orc.h -- CPU oracle for the channel-flow DNS RK3 step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * C host code, the Python package) may link, load or call this code.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker / the timed CPU baseline, never as a product path.
 *
 * It is a from-scratch, single-process restatement of the reference
 * solver's arithmetic (reference files cited per function in orc_core.c /
 * orc_step.c).  Data layout is the reference's padded z-pencil block:
 * fields are [n0][n1][n2] doubles, k (wall-normal) fastest, with
 * n0 = nx + 2 ng, n1 = ny + 2 ng, n2 = nz + 2 ng and
 * ijk = (i * n1 + j) * n2 + k                        (chanfast.c:139-146).
 *
 * Axis names follow the reference: x streamwise (periodic), y spanwise
 * (periodic), z wall-normal (walls, stretched).
 */
#ifndef ORC_H
#define ORC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* grid of one run; dz has n2 entries (ghosts mirrored, ghost.c:31-54) */
typedef struct orc_grid {
    int nx, ny, nz, ng;
    int n0, n1, n2;
    double dx, dy;
    double *dz;
} orc_grid;

/* boundary-condition kinds (diffuseu.h, diffusew.h, pressure.c:23-24) */
#define ORC_DIRICH_CELL 0
#define ORC_NEUMAN_CELL 1
#define ORC_DIRICH_FACE 20
#define ORC_DIRICH_P    10
#define ORC_NEUMAN_P    11

/* field ids for orc_put / orc_get */
#define ORC_U 0
#define ORC_V 1
#define ORC_W 2
#define ORC_C 3
#define ORC_P 4

/* per-step log record: the values chanfast appends to cfl.dat, pstat.dat
 * and cstat.dat (chanfast.c:392-393, :526-529, :613-616) */
#define ORC_LOG_N 13
/* [0] it+1 (completed step count), [1] dt, [2] cfl at step start,
 * [3] mdpdx, [4] nu dU/dz bottom, [5] nu dU/dz top, [6] bulk U, [7] centreline U,
 * [8] dcdt, [9] kappa dC/dz bottom, [10] kappa dC/dz top, [11] bulk C, [12] centreline C */

typedef struct orc_cfg {
    int nx, ny, nz;
    int ng;
    int order;            /* 2 or 4 */
    int scalar;           /* 1: advance c (required if betg != 0) */
    double lx, ly;
    double dt, want_cfl;
    int cflmode;
    double nu, prandtl;
    double dpdx, want_uvolavg;
    int uvolavgmode;
    double want_cvolavg;
    int cvolavgmode;
    double uframe, vframe;
    double betg_x, betg_y, betg_z;
    int driver;           /* ORC_CHANFAST (x-y diffusion explicit) or ORC_CHANNEL (all implicit) */
} orc_cfg;

#define ORC_CHANFAST 0    /* chanfast2 / chanfast4, chanfast.c:71-1916 */
#define ORC_CHANNEL  1    /* channel2 / channel4, channel.c:71-1968 */

typedef struct orc_run orc_run;

/* ---- run-level driver (chanfast2 / chanfast4 step) ---- */
int  orc_create(const orc_cfg *cfg, const double *dz_interior, orc_run **out);
void orc_destroy(orc_run *r);
int  orc_put(orc_run *r, int field, const double *io);   /* io.c layout [nx][ny][nz] */
int  orc_get(orc_run *r, int field, double *io);
int  orc_steps(orc_run *r, int nsteps, double *log);     /* log: nsteps*ORC_LOG_N or NULL */
int  orc_project_ic(orc_run *r, double dt);              /* cases.c:426-470 projection */
double orc_max_div(orc_run *r);                          /* max |div u| over interior */
const orc_grid *orc_run_grid(const orc_run *r);
double orc_run_dt(const orc_run *r);

/* ---- grid metric helpers ---- */
void orc_grid_init(orc_grid *g, int nx, int ny, int nz, int ng,
                   double lx, double ly, double *dz_n2, const double *dz_interior);
void orc_ghost_even_z(double *u, const orc_grid *g);      /* ghost.c:31-54 */
void orc_ghost_odd_z(double *u, const orc_grid *g);       /* ghost.c:5-29 */
void orc_z_at_p(const double *dz, double *z, int nz, int ng);   /* get.c:458-472 */
void orc_z_at_w(const double *dz, double *z, int nz, int ng);   /* get.c:475-488 */
double orc_z_avg(const double *u, int nz, int ng, const double *dz); /* get.c:490-502 */

/* ---- banded solvers ---- */
int orc_tridag(const double *a, const double *b, const double *c,
               const double *r, double *x, int n);                       /* tridag.c:4-26 */
int orc_heptdag(const double *a, const double *b, const double *c, const double *d,
                const double *e, const double *f, const double *g,
                const double *h, double *x, int n);                      /* heptdag.c:5-61 */

/* ---- FFT (own mixed-radix Stockham; unnormalised both ways) ---- */
int orc_fft(double *z, int n, int sign);             /* complex, interleaved re/im, in place */
int orc_rfft(const double *x, double *X, int n);     /* r2c: X[0..n/2] interleaved */
int orc_irfft(const double *X, double *x, int n);    /* c2r: imag of X[0], X[n/2] ignored */

/* ---- pencil index maps (decomp.c) ---- */
void orc_distribute(int data1, int proc, int *st, int *en, int *sz);
void orc_partition(int nx, int ny, int nz, const int pdim[3], const int dims[2],
                   const int coord[2], int lstart[3], int lend[3], int lsize[3]);

/* ---- operators on padded arrays (reference signatures, own code) ---- */
void orc_periodic(double *u, const orc_grid *g);                         /* ghostz.c:107-149 */
void orc_apply_bc_cell(double *u, const orc_grid *g, int order,
                       double bcb, double bct, int typb, int typt);      /* diffuseu.c:37-104 */
void orc_apply_bc_face(double *w, const orc_grid *g, int order,
                       double bcb, double bct);                          /* diffusew.c:37-104 */
void orc_helmholtz1d_cell(const double *u, double *h, const orc_grid *g,
                          int order, double nudt);                       /* diffuseu.c:107-170 */
void orc_helmholtz1d_face(const double *w, double *h, const orc_grid *g,
                          int order, double nudt);                       /* diffusew.c:107-172 */
void orc_solve1d_cell(double *u, const double *h, const orc_grid *g, int order,
                      double nudt, double bcb, double bct, int typb, int typt); /* diffuseu.c:173-283 */
void orc_solve1d_face(double *w, const double *h, const orc_grid *g, int order,
                      double nudt, double bcb, double bct);              /* diffusew.c:175-287 */
void orc_laplacexy_cell(const double *u, double *h, const orc_grid *g, int order, double nu); /* diffuseu.c:286-359 */
void orc_laplacexy_face(const double *w, double *h, const orc_grid *g, int order, double nu); /* diffusew.c:290-365 */
void orc_massflux(int dir, double *u, double *v, double *w,
                  double *mu, double *mv, double *mw, const orc_grid *g, int order); /* advect.c:6-283 */
void orc_advect(int comp, const double *q, const double *mu, const double *mv,
                const double *mw, double *f, const orc_grid *g, int order);      /* advect.c:286-734 */
void orc_flux_quick(const double *u, const double *v, const double *w, const double *c,
                    double *mu, double *mv, double *mw, const orc_grid *g);       /* scalar.c:113-162 */
void orc_advect_scalar(const double *mu, const double *mv, const double *mw,
                       double *f, const orc_grid *g);                            /* scalar.c:79-110 */
void orc_add_buoyancy(const double *c, double *fu, double *fv, double *fw,
                      double bx, double by, double bz, const orc_grid *g, int order); /* scalar.c:6-76 */
void orc_divergence(const double *u, const double *v, double *w, double *d,
                    const orc_grid *g, int order);                       /* pressure.c:40-118 */
void orc_project(double *u, double *v, double *w, const double *p,
                 const orc_grid *g, int order, double dt);               /* pressure.c:120-195 */
void orc_pressure_laplacian(const double *p, double *l, const orc_grid *g,
                            int order, double dt);                       /* pressure.c:197-291 */
/* the 3-D Helmholtz operators h = q + nudt lap q (diffuseu.c:362-457,
 * diffusew.c:366-465) */
void orc_helmholtz3d_cell(const double *u, double *h, const orc_grid *g, int order, double nudt);
void orc_helmholtz3d_face(const double *w, double *h, const orc_grid *g, int order, double nudt);
/* spectral Helmholtz solves on cout[(ix*ny+iy)*n2+k][2] (diffuseu.c:460-614,
 * diffusew.c:466-622): wall values apply to the real part of mode (0,0) */
void orc_solve_helmholtz_cell(double *cout, const orc_grid *g, int order, double nudt,
                              double bcb, double bct, int typb, int typt);
void orc_solve_helmholtz_face(double *cout, const orc_grid *g, int order, double nudt,
                              double bcb, double bct);
/* spectral Poisson on cout[(ix*ny+iy)*n2+k][2] (pressure.c:293-444) */
void orc_poisson(double *cout, const orc_grid *g, int order, double dt);
/* FFT3D_FWD / FFT3D_BWD of a padded field over all n2 levels (chanfast.c:21-55, fft.c:161-208) */
void orc_fft3d_fwd(const double *a, double *cout, const orc_grid *g);
void orc_fft3d_bwd(const double *cout, double *a, const orc_grid *g);
void orc_xy_avg(const double *u, double *avg, const orc_grid *g);        /* get.c:281-307 */
void orc_xy_corr(const double *u, const double *v, double *corr, const orc_grid *g); /* get.c:136-174 */
void orc_xy_moment(const double *u, double *mom, double en, const orc_grid *g);     /* get.c:100-133 */
/* get.c:5-97; uhat, vhat [(nx/2+1)][ny][n2][2]; specx [(nx/2+1)][n2][2], specy [(ny/2+1)][n2][2] */
void orc_xy_spec(const double *uhat, const double *vhat, double *specx, double *specy,
                 int nx, int ny, int n2);
double orc_cfl(const double *u, const double *v, const double *w,
               const orc_grid *g, double dt);                            /* get.c:390-417 */

#ifdef __cplusplus
}
#endif

#endif
