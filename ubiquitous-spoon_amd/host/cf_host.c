/*
 * cf_host.c -- host-only parts of the C ABI: configuration defaults, the
 * wall-normal grid, seeded initial fields and the reference's file
 * formats.  No GPU calls.
 */
#include "chanflow.h"
#include "cf_internal_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void cf_config_default(cf_config *cfg)
{
    memset(cfg, 0, sizeof(*cfg));
    cfg->nx = 64; cfg->ny = 64; cfg->nz = 65;
    cfg->order = 2;
    cfg->scalar = 0;
    cfg->precision = CF_FP64;
    cfg->lx = 2.0 * M_PI; cfg->ly = M_PI; cfg->lz = 2.0;
    cfg->stretch = CF_STRETCH_COSINE;
    cfg->stretch_beta = 2.0;
    cfg->dt = 1e-3; cfg->want_cfl = 0.5; cfg->cflmode = 0;
    cfg->nu = 1.0 / 2800.0; cfg->prandtl = 1.0;
    cfg->dpdx = 0.0; cfg->want_uvolavg = 1.0; cfg->uvolavgmode = 1;
    cfg->want_cvolavg = 0.0; cfg->cvolavgmode = 0;
    cfg->uframe = 0.0; cfg->vframe = 0.0;
    cfg->betg_x = cfg->betg_y = cfg->betg_z = 0.0;
    cfg->device = 0;
    cfg->driver = CF_DRIVER_CHANFAST;
}

cf_status cf_grid_dz(const cf_config *cfg, double *dz)
{
    if (!cfg || !dz || cfg->nz < 2 || !(cfg->lz > 0.0)) {
        cf_set_error("cf_grid_dz: need nz >= 2, lz > 0 and an output array");
        return CF_ERR_ARG;
    }
    int nz = cfg->nz;
    double lz = cfg->lz;
    switch (cfg->stretch) {
    case CF_STRETCH_COSINE:
        /* chanfast.c:184-189 */
        for (int iz = 0; iz < nz; iz++)
            dz[iz] = lz * 0.5 * (cos(M_PI * (double) iz / (double) nz)
                               - cos(M_PI * (double) (iz + 1) / (double) nz));
        return CF_OK;
    case CF_STRETCH_TANH: {
        /* faces z_j = lz/2 (1 + tanh(beta x_j) / tanh(beta)), x_j = (2j - nz) / nz:
         * x is exactly antisymmetric and tanh odd, so dz mirrors exactly */
        double b = cfg->stretch_beta;
        if (!(b > 0.0)) { cf_set_error("cf_grid_dz: tanh stretching needs beta > 0"); return CF_ERR_ARG; }
        double tb = tanh(b);
        double tlo = tanh(b * ((double) (0 - nz) / (double) nz));
        for (int iz = 0; iz < nz; iz++) {
            double thi = tanh(b * ((double) (2 * (iz + 1) - nz) / (double) nz));
            dz[iz] = 0.5 * lz * (thi - tlo) / tb;
            tlo = thi;
        }
        return CF_OK;
    }
    case CF_STRETCH_UNIFORM:
        for (int iz = 0; iz < nz; iz++) dz[iz] = lz / (double) nz;
        return CF_OK;
    case CF_STRETCH_USER:
        if (!cfg->dz_user) { cf_set_error("cf_grid_dz: CF_STRETCH_USER without dz_user"); return CF_ERR_ARG; }
        for (int iz = 0; iz < nz; iz++) {
            if (!(cfg->dz_user[iz] > 0.0)) { cf_set_error("cf_grid_dz: dz_user[%d] <= 0", iz); return CF_ERR_ARG; }
            dz[iz] = cfg->dz_user[iz];
        }
        return CF_OK;
    default:
        cf_set_error("cf_grid_dz: unknown stretch kind %d", cfg->stretch);
        return CF_ERR_ARG;
    }
}

/* SplitMix64 finaliser (Steele, Lea, Flood 2014) on a counter */
static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static double noise(uint64_t seed, int field, uint64_t cell)
{
    uint64_t ctr = seed * 0x9e3779b97f4a7c15ULL + ((uint64_t) field << 56) + cell + 1ULL;
    uint64_t z = mix64(ctr * 0x9e3779b97f4a7c15ULL);
    double u01 = (double) (z >> 11) * (1.0 / 9007199254740992.0);   /* [0, 1) */
    return 2.0 * u01 - 1.0;                                          /* [-1, 1) */
}

cf_status cf_initial_fields(const cf_config *cfg, uint64_t seed, double amp, double camp,
                            double *u, double *v, double *w, double *c)
{
    if (!cfg || cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 2) {
        cf_set_error("cf_initial_fields: bad grid");
        return CF_ERR_ARG;
    }
    int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    double dzs[4096];
    double *dz = dzs;
    if (nz > 4096) { cf_set_error("cf_initial_fields: nz > 4096"); return CF_ERR_ARG; }
    cf_status st = cf_grid_dz(cfg, dz);
    if (st != CF_OK) return st;
    double zeta[4096];
    double zf = 0.0;
    for (int iz = 0; iz < nz; iz++) { zeta[iz] = zf + 0.5 * dz[iz]; zf += dz[iz]; }
    double lz = zf;
    double ub = cfg->uvolavgmode == 1 ? cfg->want_uvolavg : 1.0;
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++)
            for (int iz = 0; iz < nz; iz++) {
                uint64_t cell = ((uint64_t) ix * (uint64_t) ny + (uint64_t) iy) * (uint64_t) nz + (uint64_t) iz;
                size_t o = (size_t) cell;
                double eta = 2.0 * zeta[iz] / lz - 1.0;
                if (u) u[o] = 1.5 * ub * (1.0 - eta * eta) + amp * noise(seed, 0, cell);
                if (v) v[o] = amp * noise(seed, 1, cell);
                if (w) w[o] = iz == nz - 1 ? 0.0 : amp * noise(seed, 2, cell);
                if (c) c[o] = zeta[iz] / lz - 0.5 + camp * noise(seed, 3, cell);
            }
    return CF_OK;
}

cf_status cf_io_write(const char *path, const double *buf, size_t n)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) { cf_set_error("cf_io_write: cannot open %s", path); return CF_ERR_ARG; }
    size_t wr = fwrite(buf, sizeof(double), n, fp);
    fclose(fp);
    if (wr != n) { cf_set_error("cf_io_write: short write on %s", path); return CF_ERR_ARG; }
    return CF_OK;
}

cf_status cf_io_read(const char *path, double *buf, size_t n)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) { cf_set_error("cf_io_read: cannot open %s", path); return CF_ERR_ARG; }
    size_t rd = fread(buf, sizeof(double), n, fp);
    fclose(fp);
    if (rd != n) { cf_set_error("cf_io_read: %s holds %zu of %zu values", path, rd, n); return CF_ERR_ARG; }
    return CF_OK;
}

cf_status cf_io_append_ascii(const char *path, const double *vals, int n)
{
    /* io.c:5-20: one line, "%+e" separated by single spaces */
    FILE *fp = fopen(path, "a");
    if (!fp) { cf_set_error("cf_io_append_ascii: cannot open %s", path); return CF_ERR_ARG; }
    for (int i = 0; i < n; i++) fprintf(fp, i ? " %+e" : "%+e", vals[i]);
    fprintf(fp, "\n");
    fclose(fp);
    return CF_OK;
}

cf_status cf_io_write_ascii(const char *path, const double *vals, int n)
{
    /* io.c:22-34: one value per line, "%+e" */
    FILE *fp = fopen(path, "w");
    if (!fp) { cf_set_error("cf_io_write_ascii: cannot open %s", path); return CF_ERR_ARG; }
    for (int i = 0; i < n; i++) fprintf(fp, "%+e\n", vals[i]);
    fclose(fp);
    return CF_OK;
}

cf_status cf_grid_z(const cf_config *cfg, int nghost, double *zp, double *zw)
{
    /* get_z_at_p / get_z_at_w (get.c:458-488) on the padded dz of
     * chanfast.c:184-190: ghost heights mirrored about each wall
     * (ghost_even_z, ghost.c:31-54); zp at cell centres, zw at the faces
     * above each level, both measured from the bottom wall */
    if (!cfg || nghost < 1 || cfg->nz < 1 || nghost > cfg->nz || cfg->nz > 8000) {
        cf_set_error("cf_grid_z: need 1 <= nghost <= nz <= 8000");
        return CF_ERR_ARG;
    }
    const int nz = cfg->nz, n2 = nz + 2 * nghost;
    double *dz = malloc(sizeof(double) * (size_t) n2);
    if (!dz) { cf_set_error("cf_grid_z: out of memory"); return CF_ERR_NOMEM; }
    cf_status st = cf_grid_dz(cfg, dz + nghost);
    if (st != CF_OK) { free(dz); return st; }
    for (int k = 0; k < nghost; k++) {
        dz[nghost - 1 - k] = dz[nghost + k];
        dz[nghost + nz + k] = dz[nghost + nz - 1 - k];
    }
    if (zp) {
        zp[0] = 0.5 * dz[0];
        for (int k = 1; k < n2; k++) zp[k] = zp[k - 1] + 0.5 * (dz[k - 1] + dz[k]);
        const double z0 = 0.5 * (zp[nghost - 1] + zp[nghost]);
        for (int k = 0; k < n2; k++) zp[k] -= z0;
    }
    if (zw) {
        zw[0] = dz[0];
        for (int k = 1; k < n2; k++) zw[k] = zw[k - 1] + dz[k];
        const double z0 = zw[nghost - 1];
        for (int k = 0; k < n2; k++) zw[k] -= z0;
    }
    free(dz);
    return CF_OK;
}
