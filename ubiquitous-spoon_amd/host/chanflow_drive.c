/*
 * chanflow_drive.c -- the reference's run driver on the chanflow C ABI.
 *
 * Reads outputdir/{u,v,w[,c]}_it<it>.dat (io.c:259-296 byte format), steps
 * it -> nt on one GPU, appends outputdir/cfl.dat, pstat.dat and (with the
 * scalar) cstat.dat per step in the records of chanfast.c:392-393,
 * :526-529 and :613-616, and whenever (it + 1) % istat == 0 (chanfast.c:658)
 * writes the statistics files of chanfast.c:659-856 -- plane averages
 * {u,v,w,c,p}_avg, correlations {uu,vv,ww,cc,pp,wu,uc,wc}_corr, the third
 * and fourth moments u_mom3/u_mom4, the spectra
 * {uu,vv,ww,cc,pp,wu,uc,wc}_spec{x,y}, zp.dat and zw.dat -- and the
 * fields {u,v,w[,c]}_it<it+1>.dat (chanfast.c:858-877).  Without the
 * scalar the files that involve c are not written.  Profiles and spectra
 * keep the reference's length, nz + 2 nghost levels (--nghost, default 3 as
 * in main.c:11-13), with the interior levels at nghost .. nghost + nz - 1;
 * the ghost entries, which hold the reference's solver work values rather
 * than flow statistics, are written as 0.  Parameters are those of
 * chanfast2 / chanfast4 (chanfast.h:1-21, main.c:45-61), given as options;
 * the MPI process grid has no counterpart.
 *
 *   chanflow_drive --dir outputdir --nx 64 --ny 64 --nz 65 --it 0 --nt 100
 *                  --istat 100 [--order 4] [--scalar] [--fp32] [--lx ..]
 *                  [--ly ..] [--lz ..] [--dt ..] [--nu ..] [--prandtl ..]
 *                  [--dpdx ..] [--ubulk ..] [--uvolavgmode 0|1]
 *                  [--cbulk ..] [--cvolavgmode 0|1] [--uframe ..]
 *                  [--vframe ..] [--betg bx by bz] [--stretch cosine|tanh]
 *                  [--beta ..] [--cflmode 0|1] [--want-cfl ..] [--driver chanfast|channel]
 *                  [--nghost ..] [--no-stats] [--project]
 */
#include "chanflow.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int fail(const char *what)
{
    fprintf(stderr, "chanflow_drive: %s: %s\n", what, cf_last_error());
    return 1;
}

/* one profile of nz interior values as the reference's n2-entry file */
static int write_profile(const char *dir, const char *name, int it, const double *v, int nz, int ng, double *pad)
{
    char path[4096];
    const int n2 = nz + 2 * ng;
    for (int k = 0; k < n2; k++) pad[k] = k >= ng && k < ng + nz ? v[k - ng] : 0.0;
    snprintf(path, sizeof path, "%s/%s_it%d.dat", dir, name, it);
    return cf_io_write_ascii(path, pad, n2) == CF_OK ? 0 : fail(path);
}

/* the istat statistics files of chanfast.c:659-856 */
static int write_stats(cf_run *run, const cf_config *cfg, const char *dir, int it, int ng)
{
    const int nz = cfg->nz, n2 = nz + 2 * ng, nkx = cfg->nx / 2 + 1, nky = cfg->ny / 2 + 1;
    const size_t nsx = (size_t) nkx * nz * 2, nsy = (size_t) nky * nz * 2;
    double *prof = malloc(sizeof(double) * (size_t) nz);
    double *spec = malloc(sizeof(double) * (nsx + nsy));
    double *pad = malloc(sizeof(double) * (size_t) (nkx > nky ? nkx : nky) * n2 * 2);
    if (!prof || !spec || !pad) { fprintf(stderr, "chanflow_drive: out of host memory\n"); return 1; }
    const int sc = cfg->scalar != 0;
    static const struct { const char *name; int f; } avg[] = {
        { "u_avg", CF_U }, { "v_avg", CF_V }, { "w_avg", CF_W }, { "c_avg", CF_C }, { "p_avg", CF_P } };
    static const struct { const char *name; int a, b; } pair[] = {
        { "uu", CF_U, CF_U }, { "vv", CF_V, CF_V }, { "ww", CF_W, CF_W }, { "cc", CF_C, CF_C },
        { "pp", CF_P, CF_P }, { "wu", CF_W, CF_U }, { "uc", CF_U, CF_C }, { "wc", CF_W, CF_C } };
    int rc = 0;
    char name[64], path[4096];
    for (int q = 0; q < 5 && !rc; q++) {
        if (avg[q].f == CF_C && !sc) continue;
        if (cf_run_xy_stats(run, CF_STAT_AVG, avg[q].f, avg[q].f, 0.0, prof) != CF_OK) { rc = fail("xy avg"); break; }
        rc = write_profile(dir, avg[q].name, it, prof, nz, ng, pad);
    }
    for (int q = 0; q < 8 && !rc; q++) {
        if ((pair[q].a == CF_C || pair[q].b == CF_C) && !sc) continue;
        if (cf_run_xy_stats(run, CF_STAT_CORR, pair[q].a, pair[q].b, 0.0, prof) != CF_OK) { rc = fail("xy corr"); break; }
        snprintf(name, sizeof name, "%s_corr", pair[q].name);
        rc = write_profile(dir, name, it, prof, nz, ng, pad);
    }
    for (int en = 3; en <= 4 && !rc; en++) {
        if (cf_run_xy_stats(run, CF_STAT_MOMENT, CF_U, CF_U, (double) en, prof) != CF_OK) { rc = fail("xy moment"); break; }
        snprintf(name, sizeof name, "u_mom%d", en);
        rc = write_profile(dir, name, it, prof, nz, ng, pad);
    }
    for (int q = 0; q < 8 && !rc; q++) {
        if ((pair[q].a == CF_C || pair[q].b == CF_C) && !sc) continue;
        if (cf_run_xy_stats(run, CF_STAT_SPEC, pair[q].a, pair[q].b, 0.0, spec) != CF_OK) { rc = fail("xy spec"); break; }
        /* [mode][level][re, im] with n2 levels per mode (get.c:5-97) */
        for (int d = 0; d < 2 && !rc; d++) {
            const int nm = d ? nky : nkx;
            const double *src = d ? spec + nsx : spec;
            for (int m = 0; m < nm; m++)
                for (int k = 0; k < n2; k++)
                    for (int t = 0; t < 2; t++)
                        pad[((size_t) m * n2 + k) * 2 + t] =
                            k >= ng && k < ng + nz ? src[((size_t) m * nz + (k - ng)) * 2 + t] : 0.0;
            snprintf(path, sizeof path, "%s/%s_spec%c_it%d.dat", dir, pair[q].name, d ? 'y' : 'x', it);
            if (cf_io_write_ascii(path, pad, nm * n2 * 2) != CF_OK) rc = fail(path);
        }
    }
    if (!rc) {
        double *zp = malloc(sizeof(double) * (size_t) n2), *zw = malloc(sizeof(double) * (size_t) n2);
        if (!zp || !zw || cf_grid_z(cfg, ng, zp, zw) != CF_OK) rc = fail("grid z");
        snprintf(path, sizeof path, "%s/zp.dat", dir);
        if (!rc && cf_io_write_ascii(path, zp, n2) != CF_OK) rc = fail(path);
        snprintf(path, sizeof path, "%s/zw.dat", dir);
        if (!rc && cf_io_write_ascii(path, zw, n2) != CF_OK) rc = fail(path);
        free(zp);
        free(zw);
    }
    free(prof);
    free(spec);
    free(pad);
    return rc;
}

int main(int argc, char **argv)
{
    cf_config cfg;
    cf_config_default(&cfg);
    const char *dir = "outputdir";
    int it = 0, nt = -1, istat = -1, project = 0, nghost = 3, stats = 1;
    for (int a = 1; a < argc; a++) {
        const char *k = argv[a];
        const char *v = a + 1 < argc ? argv[a + 1] : NULL;
#define OPT_I(name, dst) if (!strcmp(k, name) && v) { dst = atoi(v); a++; continue; }
#define OPT_D(name, dst) if (!strcmp(k, name) && v) { dst = atof(v); a++; continue; }
        if (!strcmp(k, "--dir") && v) { dir = v; a++; continue; }
        OPT_I("--nx", cfg.nx) OPT_I("--ny", cfg.ny) OPT_I("--nz", cfg.nz)
        OPT_I("--it", it) OPT_I("--nt", nt) OPT_I("--istat", istat) OPT_I("--order", cfg.order)
        OPT_D("--lx", cfg.lx) OPT_D("--ly", cfg.ly) OPT_D("--lz", cfg.lz) OPT_D("--dt", cfg.dt)
        OPT_D("--nu", cfg.nu) OPT_D("--prandtl", cfg.prandtl) OPT_D("--dpdx", cfg.dpdx)
        OPT_D("--ubulk", cfg.want_uvolavg) OPT_I("--uvolavgmode", cfg.uvolavgmode)
        OPT_D("--cbulk", cfg.want_cvolavg) OPT_I("--cvolavgmode", cfg.cvolavgmode)
        OPT_D("--uframe", cfg.uframe) OPT_D("--vframe", cfg.vframe) OPT_D("--beta", cfg.stretch_beta)
        OPT_I("--device", cfg.device) OPT_I("--cflmode", cfg.cflmode) OPT_D("--want-cfl", cfg.want_cfl)
        OPT_I("--nghost", nghost)
#undef OPT_I
#undef OPT_D
        if (!strcmp(k, "--scalar")) { cfg.scalar = 1; continue; }
        if (!strcmp(k, "--fp32")) { cfg.precision = CF_FP32; continue; }
        if (!strcmp(k, "--project")) { project = 1; continue; }
        if (!strcmp(k, "--no-stats")) { stats = 0; continue; }
        if (!strcmp(k, "--betg") && a + 3 < argc) {
            cfg.betg_x = atof(argv[a + 1]); cfg.betg_y = atof(argv[a + 2]); cfg.betg_z = atof(argv[a + 3]);
            a += 3;
            continue;
        }
        if (!strcmp(k, "--driver") && v) {     /* chanfast2/4 or channel2/4 (main.c:52-55) */
            if (!strcmp(v, "channel")) cfg.driver = CF_DRIVER_CHANNEL;
            else if (!strcmp(v, "chanfast")) cfg.driver = CF_DRIVER_CHANFAST;
            else { fprintf(stderr, "chanflow_drive: --driver chanfast|channel\n"); return 2; }
            a++;
            continue;
        }
        if (!strcmp(k, "--stretch") && v) {
            cfg.stretch = !strcmp(v, "tanh") ? CF_STRETCH_TANH : !strcmp(v, "uniform") ? CF_STRETCH_UNIFORM
                                                                                        : CF_STRETCH_COSINE;
            a++;
            continue;
        }
        fprintf(stderr, "chanflow_drive: unknown or incomplete option %s\n", k);
        return 2;
    }
    if (nt < it) { fprintf(stderr, "chanflow_drive: need --nt >= --it\n"); return 2; }
    if (istat <= 0) istat = nt - it > 0 ? nt - it : 1;
    if (nghost < 1 || nghost > cfg.nz) { fprintf(stderr, "chanflow_drive: need 1 <= --nghost <= nz\n"); return 2; }

    const size_t n = (size_t) cfg.nx * cfg.ny * cfg.nz;
    double *buf = malloc(n * sizeof(double));
    double *log = malloc((size_t) CF_LOG_N * 4096 * sizeof(double));
    if (!buf || !log) { fprintf(stderr, "chanflow_drive: out of host memory\n"); return 1; }
    cf_run *run;
    if (cf_run_create(&cfg, &run) != CF_OK) return fail("create");
    const char *names = cfg.scalar ? "uvwc" : "uvw";
    char path[4096];
    for (int f = 0; names[f]; f++) {
        snprintf(path, sizeof path, "%s/%c_it%d.dat", dir, names[f], it);
        if (cf_io_read(path, buf, n) != CF_OK) return fail("read");
        if (cf_run_upload(run, f, buf) != CF_OK) return fail("upload");
    }
    if (project && cf_run_project(run, 1.0) != CF_OK) return fail("project");

    while (it < nt) {
        /* advance to the next dump (or the end), at most the log capacity */
        int next = (it / istat + 1) * istat;
        if (next > nt) next = nt;
        if (next - it > 4096) next = it + 4096;
        const int k = next - it;
        if (cf_run_step(run, k) != CF_OK) return fail("step");
        int got = 0;
        if (cf_run_log(run, log, k, &got) != CF_OK) return fail("log");
        for (int s = 0; s < got; s++) {
            const double *r = log + (size_t) s * CF_LOG_N;
            double rec[6];
            rec[0] = (double) (it + s); rec[1] = r[1]; rec[2] = r[2];
            snprintf(path, sizeof path, "%s/cfl.dat", dir);
            if (cf_io_append_ascii(path, rec, 3) != CF_OK) return fail("cfl.dat");
            rec[0] = (double) (it + s + 1);
            for (int q = 0; q < 5; q++) rec[1 + q] = r[3 + q];
            snprintf(path, sizeof path, "%s/pstat.dat", dir);
            if (cf_io_append_ascii(path, rec, 6) != CF_OK) return fail("pstat.dat");
            if (cfg.scalar) {
                for (int q = 0; q < 5; q++) rec[1 + q] = r[8 + q];
                snprintf(path, sizeof path, "%s/cstat.dat", dir);
                if (cf_io_append_ascii(path, rec, 6) != CF_OK) return fail("cstat.dat");
            }
        }
        it = next;
        if (it % istat == 0) {
            if (stats && write_stats(run, &cfg, dir, it, nghost)) return 1;
            for (int f = 0; names[f]; f++) {
                if (cf_run_download(run, f, buf) != CF_OK) return fail("download");
                snprintf(path, sizeof path, "%s/%c_it%d.dat", dir, names[f], it);
                if (cf_io_write(path, buf, n) != CF_OK) return fail("write");
            }
        }
    }
    cf_run_destroy(run);
    free(buf);
    free(log);
    return 0;
}
