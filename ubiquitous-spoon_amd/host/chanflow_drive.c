/*
 * chanflow_drive.c -- the reference's run driver on the chanflow C ABI.
 *
 * Reads outputdir/{u,v,w[,c]}_it<it>.dat (io.c:259-296 byte format), steps
 * it -> nt on one GPU, appends outputdir/cfl.dat, pstat.dat and (with the
 * scalar) cstat.dat per step in the records of chanfast.c:392-393,
 * :526-529 and :613-616, and writes {u,v,w[,c]}_it<it+1>.dat whenever
 * (it + 1) % istat == 0 (chanfast.c:658, :862-877).  Parameters are those
 * of chanfast2 / chanfast4 (chanfast.h:1-21, main.c:45-61), given as
 * options; the MPI process grid and ghost widths have no counterpart.
 *
 *   chanflow_drive --dir outputdir --nx 64 --ny 64 --nz 65 --it 0 --nt 100
 *                  --istat 100 [--order 4] [--scalar] [--fp32] [--lx ..]
 *                  [--ly ..] [--lz ..] [--dt ..] [--nu ..] [--prandtl ..]
 *                  [--dpdx ..] [--ubulk ..] [--uvolavgmode 0|1]
 *                  [--cbulk ..] [--cvolavgmode 0|1] [--uframe ..]
 *                  [--vframe ..] [--betg bx by bz] [--stretch cosine|tanh]
 *                  [--beta ..] [--project]
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

int main(int argc, char **argv)
{
    cf_config cfg;
    cf_config_default(&cfg);
    const char *dir = "outputdir";
    int it = 0, nt = -1, istat = -1, project = 0;
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
        OPT_I("--device", cfg.device)
#undef OPT_I
#undef OPT_D
        if (!strcmp(k, "--scalar")) { cfg.scalar = 1; continue; }
        if (!strcmp(k, "--fp32")) { cfg.precision = CF_FP32; continue; }
        if (!strcmp(k, "--project")) { project = 1; continue; }
        if (!strcmp(k, "--betg") && a + 3 < argc) {
            cfg.betg_x = atof(argv[a + 1]); cfg.betg_y = atof(argv[a + 2]); cfg.betg_z = atof(argv[a + 3]);
            a += 3;
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
