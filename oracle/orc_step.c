/*
 * orc_step.c -- CPU oracle: the chanfast2 / chanfast4 RK3 time loop on
 * one process.  TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Follows chanfast.c:383-656 (order 2) and :1307-1580 (order 4) operation
 * by operation (and, with driver ORC_CHANNEL, the fully implicit
 * channel2 / channel4 of channel.c:383-668 and :1307-1593): f<->g swap, periodic halos + wall ghosts, explicit RHS,
 * per-field implicit z-Helmholtz solve with the boundary lift, the
 * mass-flux (and scalar-mean) constraint, w solve, divergence, FFT,
 * banded Poisson, inverse FFT, projection.  Per-step cfl/pstat/cstat
 * values go to a caller array instead of ascii files.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FILL_VALUE (-999.0)   /* check.h:13 */

/* RK3 tables, rk3.c:3-6 */
static const double RK_ALP[3] = { 29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0 };
static const double RK_BET[3] = { 37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0 };
static const double RK_GAM[3] = { 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0 };
static const double RK_ZET[3] = { 0.0, -17.0 / 60.0, -5.0 / 12.0 };
static const double A1 = 1.125 / 1.0;
static const double A2 = -0.125 / 3.0;

struct orc_run {
    orc_cfg cfg;
    orc_grid g;
    double *dz;
    size_t ntot;
    double *u, *v, *w, *c, *p, *h, *bc, *hbc, *mx, *my, *mz;
    double *fu, *fv, *fw, *fc, *gu, *gv, *gw, *gc;
    double *spec;
    double *avg;
    double dt;
    int it;
};

static double *fresh(size_t n)
{
    double *a = (double *) malloc(sizeof(double) * n);
    if (a) for (size_t i = 0; i < n; i++) a[i] = FILL_VALUE;
    return a;
}

int orc_create(const orc_cfg *cfg, const double *dz_interior, orc_run **out)
{
    *out = NULL;
    if (cfg->order != 2 && cfg->order != 4) return -1;
    if (cfg->ng < (cfg->order == 2 ? 1 : 3)) return -1;
    if (cfg->nx < 2 * cfg->ng || cfg->ny < 2 * cfg->ng || cfg->nz < (cfg->order == 2 ? 2 : 6)) return -1;
    if (!cfg->scalar && (cfg->betg_x != 0.0 || cfg->betg_y != 0.0 || cfg->betg_z != 0.0)) return -1;
    orc_run *r = (orc_run *) calloc(1, sizeof(orc_run));
    if (!r) return -1;
    r->cfg = *cfg;
    int n2 = cfg->nz + 2 * cfg->ng;
    r->dz = (double *) malloc(sizeof(double) * (size_t) n2);
    orc_grid_init(&r->g, cfg->nx, cfg->ny, cfg->nz, cfg->ng, cfg->lx, cfg->ly, r->dz, dz_interior);
    r->ntot = (size_t) r->g.n0 * r->g.n1 * r->g.n2;
    double **fields[] = { &r->u, &r->v, &r->w, &r->c, &r->p, &r->h, &r->bc, &r->hbc, &r->mx, &r->my,
                          &r->mz, &r->fu, &r->fv, &r->fw, &r->fc, &r->gu, &r->gv, &r->gw, &r->gc };
    for (size_t q = 0; q < sizeof(fields) / sizeof(fields[0]); q++) {
        *fields[q] = fresh(r->ntot);
        if (!*fields[q]) { orc_destroy(r); return -1; }
    }
    /* state arrays start at zero (chanfast.c:334-345 then io_read) */
    memset(r->u, 0, sizeof(double) * r->ntot);
    memset(r->v, 0, sizeof(double) * r->ntot);
    memset(r->w, 0, sizeof(double) * r->ntot);
    memset(r->c, 0, sizeof(double) * r->ntot);
    r->spec = (double *) calloc((size_t) (cfg->nx / 2 + 1) * cfg->ny * n2 * 2, sizeof(double));
    r->avg = (double *) calloc((size_t) n2, sizeof(double));
    if (!r->spec || !r->avg) { orc_destroy(r); return -1; }
    r->dt = cfg->dt;
    r->it = 0;
    *out = r;
    return 0;
}

void orc_destroy(orc_run *r)
{
    if (!r) return;
    double *a[] = { r->u, r->v, r->w, r->c, r->p, r->h, r->bc, r->hbc, r->mx, r->my, r->mz,
                    r->fu, r->fv, r->fw, r->fc, r->gu, r->gv, r->gw, r->gc, r->spec, r->avg, r->dz };
    for (size_t q = 0; q < sizeof(a) / sizeof(a[0]); q++) free(a[q]);
    free(r);
}

const orc_grid *orc_run_grid(const orc_run *r) { return &r->g; }
double orc_run_dt(const orc_run *r) { return r->dt; }

static double *field_ptr(orc_run *r, int f)
{
    switch (f) {
    case ORC_U: return r->u;
    case ORC_V: return r->v;
    case ORC_W: return r->w;
    case ORC_C: return r->c;
    case ORC_P: return r->p;
    default: return NULL;
    }
}

int orc_put(orc_run *r, int field, const double *io)
{
    /* io_read + ghostz_pad_double (io.c:259-296, ghostz.c:167-177):
     * [nx][ny][nz] interior values into k = ng .. ng + nz - 1 */
    double *a = field_ptr(r, field);
    if (!a) return -1;
    const orc_grid *g = &r->g;
    for (int ix = 0; ix < g->nx; ix++)
        for (int iy = 0; iy < g->ny; iy++)
            for (int iz = 0; iz < g->nz; iz++)
                a[(((size_t) (ix + g->ng)) * g->n1 + iy + g->ng) * g->n2 + iz + g->ng] =
                    io[((size_t) ix * g->ny + iy) * g->nz + iz];
    return 0;
}

int orc_get(orc_run *r, int field, double *io)
{
    /* ghostz_truncate_double + io_write (ghostz.c:152-164, io.c:211-256) */
    double *a = field_ptr(r, field);
    if (!a) return -1;
    const orc_grid *g = &r->g;
    for (int ix = 0; ix < g->nx; ix++)
        for (int iy = 0; iy < g->ny; iy++)
            for (int iz = 0; iz < g->nz; iz++)
                io[((size_t) ix * g->ny + iy) * g->nz + iz] =
                    a[(((size_t) (ix + g->ng)) * g->n1 + iy + g->ng) * g->n2 + iz + g->ng];
    return 0;
}

/* project u, v, w with the pressure that makes them divergence free
 * (cases.c:417-470 with dt as given); w wall faces are set to 0 first */
static void projection(orc_run *r, double dt)
{
    const orc_grid *g = &r->g;
    int order = r->cfg.order;
    orc_divergence(r->u, r->v, r->w, r->h, g, order);
    orc_fft3d_fwd(r->h, r->spec, g);
    orc_poisson(r->spec, g, order, dt);
    orc_fft3d_bwd(r->spec, r->p, g);
    orc_periodic(r->p, g);
    orc_project(r->u, r->v, r->w, r->p, g, order, dt);
}

int orc_project_ic(orc_run *r, double dt)
{
    const orc_grid *g = &r->g;
    orc_periodic(r->u, g);
    orc_periodic(r->v, g);
    orc_periodic(r->w, g);
    orc_apply_bc_face(r->w, g, r->cfg.order, 0.0, 0.0);
    projection(r, dt);
    return 0;
}

double orc_max_div(orc_run *r)
{
    const orc_grid *g = &r->g;
    orc_periodic(r->u, g);
    orc_periodic(r->v, g);
    orc_periodic(r->w, g);
    orc_apply_bc_face(r->w, g, r->cfg.order, 0.0, 0.0);
    orc_divergence(r->u, r->v, r->w, r->h, g, r->cfg.order);
    double m = 0.0;
    for (int i = g->ng; i < g->n0 - g->ng; i++)
        for (int j = g->ng; j < g->n1 - g->ng; j++)
            for (int k = g->ng; k < g->n2 - g->ng; k++) {
                double d = fabs(r->h[((size_t) i * g->n1 + j) * g->n2 + k]);
                if (d > m) m = d;
            }
    return m;
}

/* wall shear (nu dq/dz) from a plane-average profile, chanfast.c:489-494
 * (order 2) and :1413-1418 (order 4) */
static double wall_grad(const double *avg, const double *dz, int ng, int n2, int order, int top)
{
    if (!top) {
        double num = order == 2 ? (avg[ng] - avg[ng - 1])
                                : (A1 * (avg[ng] - avg[ng - 1]) + A2 * (avg[ng + 1] - avg[ng - 2]));
        return num / (0.5 * (dz[ng] + dz[ng - 1]));
    }
    double num = order == 2 ? (avg[n2 - ng] - avg[n2 - ng - 1])
                            : (A1 * (avg[n2 - ng] - avg[n2 - ng - 1]) + A2 * (avg[n2 - ng + 1] - avg[n2 - ng - 2]));
    return num / (0.5 * (dz[n2 - ng] + dz[n2 - ng - 1]));
}

static double centreline(const double *avg, int ng, int nz)
{
    return nz % 2 == 0 ? 0.5 * (avg[ng + nz / 2 - 1] + avg[ng + nz / 2]) : avg[ng + nz / 2];
}

/* one implicit cell-centred field update, chanfast.c:465-483:
 * h = H(q; nu a dt) + gam dt f + zet dt g - H(bc-only; -nu b dt); solve */
static void advance_cell(orc_run *r, double *q, double *f, double *gq, double kap, int s,
                         double bcb, double bct)
{
    const orc_grid *g = &r->g;
    int order = r->cfg.order;
    double dt = r->dt;
    double nuadt = kap * RK_ALP[s] * dt, nubdt = kap * RK_BET[s] * dt;
    double gamdt = RK_GAM[s] * dt, zetdt = RK_ZET[s] * dt;
    size_t n = r->ntot;
    double *h = r->h, *bc = r->bc, *hbc = r->hbc;
    orc_apply_bc_cell(q, g, order, bcb, bct, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
    orc_helmholtz1d_cell(q, h, g, order, nuadt);
    for (size_t i = 0; i < n; i++) h[i] += gamdt * f[i] + zetdt * gq[i];
    for (size_t i = 0; i < n; i++) bc[i] = 0.0;
    orc_apply_bc_cell(bc, g, order, bcb, bct, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
    orc_helmholtz1d_cell(bc, hbc, g, order, -nubdt);
    for (size_t i = 0; i < n; i++) h[i] -= hbc[i];
    orc_solve1d_cell(q, h, g, order, -nubdt, 0.0, 0.0, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
    orc_apply_bc_cell(q, g, order, bcb, bct, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
}

/* one fully implicit cell-centred field update of channel2 / channel4,
 * channel.c:453-475 (u; v and c likewise at :537-585):
 * h = H3(q; nuadt) + gam dt f + zet dt g - H3(bc-only; -nubdt), then the
 * spectral Helmholtz solve with -nubdt between FFT3D_FWD and FFT3D_BWD */
static void advance_cell_implicit(orc_run *r, double *q, double *f, double *gq, double nuadt, double nubdt,
                                  int s, double bcb, double bct)
{
    const orc_grid *g = &r->g;
    int order = r->cfg.order;
    double dt = r->dt, gamdt = RK_GAM[s] * dt, zetdt = RK_ZET[s] * dt;
    size_t n = r->ntot;
    double *h = r->h, *bc = r->bc, *hbc = r->hbc;
    orc_apply_bc_cell(q, g, order, bcb, bct, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
    orc_helmholtz3d_cell(q, h, g, order, nuadt);
    for (size_t i = 0; i < n; i++) h[i] += gamdt * f[i] + zetdt * gq[i];
    for (size_t i = 0; i < n; i++) bc[i] = 0.0;
    orc_apply_bc_cell(bc, g, order, bcb, bct, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
    orc_helmholtz3d_cell(bc, hbc, g, order, -nubdt);
    for (size_t i = 0; i < n; i++) h[i] -= hbc[i];
    orc_fft3d_fwd(h, r->spec, g);
    orc_solve_helmholtz_cell(r->spec, g, order, -nubdt, 0.0, 0.0, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
    orc_fft3d_bwd(r->spec, q, g);
    orc_periodic(q, g);
    orc_apply_bc_cell(q, g, order, bcb, bct, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
}

/* bulk constraint, chanfast.c:485-530 (u) and :572-617 (c): returns the
 * forcing and fills stats[0..4] = forcing, grad bot, grad top, bulk, centreline */
static double constrain(orc_run *r, double *q, double *f, double kap, int s, int mode,
                        double target, double fixed, double *stats)
{
    const orc_grid *g = &r->g;
    int order = r->cfg.order, ng = g->ng, n2 = g->n2, nz = g->nz;
    double dt = r->dt, nubdt = kap * RK_BET[s] * dt, gamdt = RK_GAM[s] * dt;
    size_t n = r->ntot;
    double *h = r->h, *avg = r->avg;
    const double *dz = g->dz;

    orc_xy_avg(q, avg, g);
    double vol0 = orc_z_avg(avg, nz, ng, dz);
    double bot0 = kap * wall_grad(avg, dz, ng, n2, order, 0);
    double top0 = kap * wall_grad(avg, dz, ng, n2, order, 1);
    double cl0 = centreline(avg, ng, nz);

    for (size_t i = 0; i < n; i++) h[i] = gamdt;
    orc_solve1d_cell(h, h, g, order, -nubdt, 0.0, 0.0, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
    orc_xy_avg(h, avg, g);
    double vol1 = orc_z_avg(avg, nz, ng, dz);
    double bot1 = kap * wall_grad(avg, dz, ng, n2, order, 0);
    double top1 = kap * wall_grad(avg, dz, ng, n2, order, 1);
    double cl1 = centreline(avg, ng, nz);

    double m = mode == 1 ? (target - vol0) / vol1 : fixed;
    vol0 += vol1 * m;
    bot0 += bot1 * m;
    top0 += top1 * m;
    cl0 += cl1 * m;
    for (size_t i = 0; i < n; i++) {
        q[i] += h[i] * m;
        f[i] += m;
    }
    stats[0] = m; stats[1] = bot0; stats[2] = top0; stats[3] = vol0; stats[4] = cl0;
    return m;
}

int orc_steps(orc_run *r, int nsteps, double *log)
{
    orc_cfg *C = &r->cfg;
    const orc_grid *g = &r->g;
    int order = C->order;
    size_t n = r->ntot;
    double ubc = 0.0 - C->uframe, vbc = 0.0 - C->vframe;     /* chanfast.c:59-64 */
    double cbot = -0.5, ctop = +0.5;                          /* chanfast.c:67-68 */

    for (int step = 0; step < nsteps; step++) {
        double cfl = orc_cfl(r->u, r->v, r->w, g, r->dt);
        if (C->cflmode == 1) {
            r->dt *= cfl > 0.0 ? C->want_cfl / cfl : 1.0;
            cfl = cfl > 0.0 ? C->want_cfl : cfl;
        }
        double *lg = log ? log + (size_t) step * ORC_LOG_N : NULL;
        if (lg) {
            for (int q = 0; q < ORC_LOG_N; q++) lg[q] = 0.0;
            lg[0] = (double) (r->it + 1); lg[1] = r->dt; lg[2] = cfl;
        }
        for (int s = 0; s < 3; s++) {
            double *t;
            t = r->fu; r->fu = r->gu; r->gu = t;
            t = r->fv; r->fv = r->gv; r->gv = t;
            t = r->fw; r->fw = r->gw; r->gw = t;
            t = r->fc; r->fc = r->gc; r->gc = t;

            orc_periodic(r->u, g);
            orc_apply_bc_cell(r->u, g, order, ubc, ubc, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
            orc_periodic(r->v, g);
            orc_apply_bc_cell(r->v, g, order, vbc, vbc, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
            if (C->scalar) {
                orc_periodic(r->c, g);
                orc_apply_bc_cell(r->c, g, order, cbot, ctop, ORC_DIRICH_CELL, ORC_DIRICH_CELL);
            }
            orc_periodic(r->w, g);
            orc_apply_bc_face(r->w, g, order, 0.0, 0.0);

            orc_massflux(0, r->u, r->v, r->w, r->mx, r->my, r->mz, g, order);
            orc_advect(0, r->u, r->mx, r->my, r->mz, r->fu, g, order);
            orc_massflux(1, r->u, r->v, r->w, r->mx, r->my, r->mz, g, order);
            orc_advect(1, r->v, r->mx, r->my, r->mz, r->fv, g, order);
            orc_massflux(2, r->u, r->v, r->w, r->mx, r->my, r->mz, g, order);
            orc_advect(2, r->w, r->mx, r->my, r->mz, r->fw, g, order);

            /* chanfast: x-y diffusion explicit in f (chanfast.c:431-442);
             * channel: all diffusion implicit, f is advection only
             * (channel.c:418-440) */
            const int implicit = C->driver == ORC_CHANNEL;
            if (!implicit) {
                orc_laplacexy_cell(r->u, r->fu, g, order, C->nu);
                orc_laplacexy_cell(r->v, r->fv, g, order, C->nu);
                orc_laplacexy_face(r->w, r->fw, g, order, C->nu);
            }

            if (C->scalar) {
                orc_flux_quick(r->u, r->v, r->w, r->c, r->mx, r->my, r->mz, g);
                orc_advect_scalar(r->mx, r->my, r->mz, r->fc, g);
                if (!implicit) orc_laplacexy_cell(r->c, r->fc, g, order, C->nu / C->prandtl);
                orc_add_buoyancy(r->c, r->fu, r->fv, r->fw, C->betg_x, C->betg_y, C->betg_z, g, order);
            }

            orc_periodic(r->fu, g);
            orc_periodic(r->fv, g);
            orc_periodic(r->fw, g);
            if (C->scalar) orc_periodic(r->fc, g);

            double dt = r->dt;
            double abdt = (RK_ALP[s] + RK_BET[s]) * dt;
            double st[5];

            const double nuadt = C->nu * RK_ALP[s] * dt, nubdt = C->nu * RK_BET[s] * dt;
            if (implicit) advance_cell_implicit(r, r->u, r->fu, r->gu, nuadt, nubdt, s, ubc, ubc);
            else advance_cell(r, r->u, r->fu, r->gu, C->nu, s, ubc, ubc);
            constrain(r, r->u, r->fu, C->nu, s, C->uvolavgmode, C->want_uvolavg - C->uframe, -C->dpdx, st);
            if (s == 2 && lg) for (int q = 0; q < 5; q++) lg[3 + q] = st[q];

            if (implicit) advance_cell_implicit(r, r->v, r->fv, r->gv, nuadt, nubdt, s, vbc, vbc);
            else advance_cell(r, r->v, r->fv, r->gv, C->nu, s, vbc, vbc);

            if (C->scalar) {
                double kap = C->nu / C->prandtl;
                if (implicit)   /* channel.c:562-581: nuadt / prandtl, -nubdt / prandtl */
                    advance_cell_implicit(r, r->c, r->fc, r->gc, nuadt / C->prandtl, nubdt / C->prandtl, s, cbot, ctop);
                else
                    advance_cell(r, r->c, r->fc, r->gc, kap, s, cbot, ctop);
                constrain(r, r->c, r->fc, kap, s, C->cvolavgmode, C->want_cvolavg, 0.0, st);
                if (s == 2 && lg) for (int q = 0; q < 5; q++) lg[8 + q] = st[q];
            }

            /* w, chanfast.c:619-637; channel.c:636-662 with the 3-D
             * operator and the spectral solve */
            {
                double gamdt = RK_GAM[s] * dt, zetdt = RK_ZET[s] * dt;
                double *h = r->h, *bc = r->bc, *hbc = r->hbc;
                orc_apply_bc_face(r->w, g, order, 0.0, 0.0);
                if (implicit) orc_helmholtz3d_face(r->w, h, g, order, nuadt);
                else orc_helmholtz1d_face(r->w, h, g, order, nuadt);
                for (size_t i = 0; i < n; i++) h[i] += gamdt * r->fw[i] + zetdt * r->gw[i];
                for (size_t i = 0; i < n; i++) bc[i] = 0.0;
                orc_apply_bc_face(bc, g, order, 0.0, 0.0);
                if (implicit) orc_helmholtz3d_face(bc, hbc, g, order, -nubdt);
                else orc_helmholtz1d_face(bc, hbc, g, order, -nubdt);
                for (size_t i = 0; i < n; i++) h[i] -= hbc[i];
                if (implicit) {
                    orc_fft3d_fwd(h, r->spec, g);
                    orc_solve_helmholtz_face(r->spec, g, order, -nubdt, 0.0, 0.0);
                    orc_fft3d_bwd(r->spec, r->w, g);
                    orc_periodic(r->w, g);
                } else {
                    orc_solve1d_face(r->w, h, g, order, -nubdt, 0.0, 0.0);
                }
                orc_apply_bc_face(r->w, g, order, 0.0, 0.0);
            }

            projection(r, abdt);
        }
        r->it++;
    }
    return 0;
}
