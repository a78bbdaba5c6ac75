/*
 * orc_core.c -- CPU oracle: grid metrics, banded solvers, FFT, wall
 * closures and finite-difference operators.  TEST INFRASTRUCTURE ONLY
 * (see orc.h).  Every function restates the arithmetic of the reference
 * function cited beside it, on the reference's padded [n0][n1][n2] layout
 * with k fastest.  Expressions keep the reference's operand order so the
 * rounding matches a -O2, no-FMA x86 build of the reference.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* fourth-order coefficients (fourth.c:3-6) */
static const double A1 = 1.125 / 1.0;
static const double A2 = -0.125 / 3.0;
static const double B1 = 1.125 / 2.0;
static const double B2 = -0.125 / 2.0;

#define ID(g, i, j, k) ((((size_t)(i)) * (size_t)(g)->n1 + (size_t)(j)) * (size_t)(g)->n2 + (size_t)(k))

/* ------------------------------------------------------------------ grid */

void orc_ghost_even_z(double *u, const orc_grid *g)
{
    /* ghost.c:31-54 applied to a 1-D profile (n0 = n1 = 1 in the
     * reference's call at chanfast.c:190) */
    int ng = g->ng, n2 = g->n2;
    for (int k = 0; k < ng; k++) u[k] = u[2 * ng - k - 1];
    for (int k = n2 - ng; k < n2; k++) u[k] = u[2 * (n2 - ng) - k - 1];
}

void orc_ghost_odd_z(double *u, const orc_grid *g)
{
    /* ghost.c:5-29, 1-D */
    int ng = g->ng, n2 = g->n2;
    for (int k = 0; k < ng; k++) u[k] = -u[2 * ng - k - 1];
    for (int k = n2 - ng; k < n2; k++) u[k] = -u[2 * (n2 - ng) - k - 1];
}

void orc_grid_init(orc_grid *g, int nx, int ny, int nz, int ng,
                   double lx, double ly, double *dz_n2, const double *dz_interior)
{
    g->nx = nx; g->ny = ny; g->nz = nz; g->ng = ng;
    g->n0 = nx + 2 * ng; g->n1 = ny + 2 * ng; g->n2 = nz + 2 * ng;
    g->dx = lx / (double) nx;                       /* chanfast.c:148-149 */
    g->dy = ly / (double) ny;
    g->dz = dz_n2;
    for (int iz = 0; iz < nz; iz++) dz_n2[iz + ng] = dz_interior[iz];
    orc_ghost_even_z(dz_n2, g);                     /* chanfast.c:190 */
}

void orc_z_at_p(const double *dz, double *z, int nz, int ng)
{
    /* get.c:458-472 */
    int n2 = nz + 2 * ng;
    z[0] = 0.5 * dz[0];
    for (int k = 1; k < n2; k++) z[k] = z[k - 1] + 0.5 * (dz[k - 1] + dz[k]);
    double z0 = 0.5 * (z[ng - 1] + z[ng]);
    for (int k = 0; k < n2; k++) z[k] -= z0;
}

void orc_z_at_w(const double *dz, double *z, int nz, int ng)
{
    /* get.c:475-488 */
    int n2 = nz + 2 * ng;
    z[0] = dz[0];
    for (int k = 1; k < n2; k++) z[k] = z[k - 1] + dz[k];
    double z0 = z[ng - 1];
    for (int k = 0; k < n2; k++) z[k] -= z0;
}

double orc_z_avg(const double *u, int nz, int ng, const double *dz)
{
    /* get.c:490-502 */
    double avg = 0.0, lz = 0.0;
    for (int k = ng; k < nz + ng; k++) {
        avg += u[k] * dz[k];
        lz += dz[k];
    }
    return avg / lz;
}

/* --------------------------------------------------------- banded solvers */

int orc_tridag(const double *a, const double *b, const double *c,
               const double *r, double *x, int n)
{
    /* Thomas algorithm, tridag.c:4-26 */
    double *gam = (double *) malloc(sizeof(double) * (size_t) (n > 0 ? n : 1));
    if (!gam) return -1;
    if (b[0] == 0.0) { free(gam); return -2; }
    double bet = b[0];
    x[0] = r[0] / bet;
    for (int j = 1; j < n; j++) {
        gam[j] = c[j - 1] / bet;
        bet = b[j] - a[j] * gam[j];
        if (bet == 0.0) { free(gam); return -2; }
        x[j] = (r[j] - a[j] * x[j - 1]) / bet;
    }
    for (int j = n - 2; j >= 0; j--) x[j] -= gam[j + 1] * x[j + 1];
    free(gam);
    return 0;
}

int orc_heptdag(const double *a, const double *b, const double *c, const double *d,
                const double *e, const double *f, const double *g,
                const double *h, double *x, int n)
{
    /* seven-band LU without pivoting, three sub- and three super-diagonals;
     * heptdag.c:5-61.  s aliases x as in the reference. */
    double *p = (double *) malloc(sizeof(double) * (size_t) n);
    double *q = (double *) malloc(sizeof(double) * (size_t) n);
    double *r = (double *) malloc(sizeof(double) * (size_t) n);
    double *s = x;
    if (!p || !q || !r || n < 3) { free(p); free(q); free(r); return -1; }
    double den, alp, bet;

    den = d[0];
    if (den == 0.0) goto singular;
    p[0] = e[0] / den; q[0] = f[0] / den; r[0] = g[0] / den; s[0] = h[0] / den;

    den = d[1] - c[1] * p[0];
    if (den == 0.0) goto singular;
    p[1] = (e[1] - c[1] * q[0]) / den;
    q[1] = (f[1] - c[1] * r[0]) / den;
    r[1] = (g[1]) / den;
    s[1] = (h[1] - c[1] * s[0]) / den;

    bet = c[2] - b[2] * p[0];
    den = d[2] - b[2] * q[0] - bet * p[1];
    if (den == 0.0) goto singular;
    p[2] = (e[2] - b[2] * r[0] - bet * q[1]) / den;
    q[2] = (f[2] - bet * r[1]) / den;
    r[2] = (g[2]) / den;
    s[2] = (h[2] - b[2] * s[0] - bet * s[1]) / den;

    for (int i = 3; i < n; i++) {
        alp = b[i] - a[i] * p[i - 3];
        bet = c[i] - a[i] * q[i - 3] - alp * p[i - 2];
        den = d[i] - a[i] * r[i - 3] - alp * q[i - 2] - bet * p[i - 1];
        if (den == 0.0) goto singular;
        p[i] = (e[i] - alp * r[i - 2] - bet * q[i - 1]) / den;
        q[i] = (f[i] - bet * r[i - 1]) / den;
        r[i] = (g[i]) / den;
        s[i] = (h[i] - a[i] * s[i - 3] - alp * s[i - 2] - bet * s[i - 1]) / den;
    }
    x[n - 1] = s[n - 1];
    x[n - 2] = s[n - 2] - p[n - 2] * s[n - 1];
    x[n - 3] = s[n - 3] - p[n - 3] * s[n - 2] - q[n - 3] * s[n - 1];
    for (int i = n - 4; i >= 0; i--)
        x[i] = s[i] - p[i] * s[i + 1] - q[i] * s[i + 2] - r[i] * s[i + 3];
    free(p); free(q); free(r);
    return 0;
singular:
    free(p); free(q); free(r);
    return -2;
}

/* -------------------------------------------------------------------- FFT */
/* Own mixed-radix Stockham autosort transform.  Twiddles come from libm
 * cos/sin of 2 pi j / n in double.  The reference calls FFTW 3.3.4
 * (fft.c:38-119); FFTW is not part of this build, only its conventions
 * are kept: forward = exp(-2 pi i jk/n), both directions unnormalised. */

typedef struct {
    int n, nf, fac[40];
    double *w;      /* w[2j], w[2j+1] = cos, -sin of 2 pi j / n */
    double *buf;    /* 2 * 2n scratch */
} fplan;

#define NPLAN 16
static fplan plans[NPLAN];
static int nplans = 0;

static fplan *get_plan(int n)
{
    for (int i = 0; i < nplans; i++)
        if (plans[i].n == n) return &plans[i];
    fplan *p;
    if (nplans < NPLAN) p = &plans[nplans++];
    else { p = &plans[NPLAN - 1]; free(p->w); free(p->buf); }
    p->n = n;
    p->nf = 0;
    int m = n;
    static const int pref[] = {4, 2, 3, 5};
    for (int t = 0; t < 4; t++)
        while (m % pref[t] == 0 && (pref[t] != 4 || m != 2)) { p->fac[p->nf++] = pref[t]; m /= pref[t]; }
    for (int q = 7; m > 1; q += 2)
        while (m % q == 0) { p->fac[p->nf++] = q; m /= q; }
    p->w = (double *) malloc(sizeof(double) * 2 * (size_t) n);
    p->buf = (double *) malloc(sizeof(double) * 4 * (size_t) n);
    for (int j = 0; j < n; j++) {
        double th = 2.0 * M_PI * (double) j / (double) n;
        p->w[2 * j] = cos(th);
        p->w[2 * j + 1] = -sin(th);
    }
    return p;
}

int orc_fft(double *z, int n, int sign)
{
    if (n < 1) return -1;
    if (n == 1) return 0;
    fplan *P = get_plan(n);
    double *x = z, *y = P->buf;
    double s = sign < 0 ? 1.0 : -1.0;   /* conj factor on the -sin table */
    for (int f = 0; f < P->nf; f++)
        if (P->fac[f] > 64) {                   /* large prime: plain DFT */
            for (int q = 0; q < n; q++) {
                double sr = 0.0, si = 0.0;
                for (int m = 0; m < n; m++) {
                    long idx = ((long) q * m) % n;
                    double wr = P->w[2 * idx], wi = s * P->w[2 * idx + 1];
                    sr += z[2 * m] * wr - z[2 * m + 1] * wi;
                    si += z[2 * m] * wi + z[2 * m + 1] * wr;
                }
                y[2 * q] = sr; y[2 * q + 1] = si;
            }
            memcpy(z, y, sizeof(double) * 2 * (size_t) n);
            return 0;
        }
    int pcur = 1;
    double ar[64], ai[64];
    for (int f = 0; f < P->nf; f++) {
        int r = P->fac[f];
        int T = n / r;
        for (int i = 0; i < T; i++) {
            int k = i % pcur;
            for (int m = 0; m < r; m++) {
                double xr = x[2 * (i + m * T)], xi = x[2 * (i + m * T) + 1];
                long idx = ((long) k * m * (n / (pcur * r))) % n;
                double wr = P->w[2 * idx], wi = s * P->w[2 * idx + 1];
                ar[m] = xr * wr - xi * wi;
                ai[m] = xr * wi + xi * wr;
            }
            int j0 = (i - k) * r + k;
            for (int q = 0; q < r; q++) {
                double sr = 0.0, si = 0.0;
                for (int m = 0; m < r; m++) {
                    long idx = ((long) q * m * T) % n;
                    double wr = P->w[2 * idx], wi = s * P->w[2 * idx + 1];
                    sr += ar[m] * wr - ai[m] * wi;
                    si += ar[m] * wi + ai[m] * wr;
                }
                y[2 * (j0 + q * pcur)] = sr;
                y[2 * (j0 + q * pcur) + 1] = si;
            }
        }
        pcur *= r;
        double *t = x; x = y; y = t;
    }
    if (x != z) memcpy(z, x, sizeof(double) * 2 * (size_t) n);
    return 0;
}

int orc_rfft(const double *xin, double *X, int n)
{
    double *z = (double *) malloc(sizeof(double) * 2 * (size_t) n);
    if (!z) return -1;
    for (int j = 0; j < n; j++) { z[2 * j] = xin[j]; z[2 * j + 1] = 0.0; }
    orc_fft(z, n, -1);
    memcpy(X, z, sizeof(double) * 2 * (size_t) (n / 2 + 1));
    free(z);
    return 0;
}

int orc_irfft(const double *X, double *xout, int n)
{
    double *z = (double *) malloc(sizeof(double) * 2 * (size_t) n);
    if (!z) return -1;
    int h = n / 2;
    for (int k = 0; k <= h; k++) { z[2 * k] = X[2 * k]; z[2 * k + 1] = X[2 * k + 1]; }
    z[1] = 0.0;                                  /* Hermitian input: imag(X0) ignored */
    if (n % 2 == 0) z[2 * h + 1] = 0.0;          /* and imag(X[n/2]) for even n */
    for (int k = 1; k < n - h; k++) {            /* z[n-k] = conj(z[k]) */
        z[2 * (n - k)] = z[2 * k];
        z[2 * (n - k) + 1] = -z[2 * k + 1];
    }
    orc_fft(z, n, +1);
    for (int j = 0; j < n; j++) xout[j] = z[2 * j];
    free(z);
    return 0;
}

/* -------------------------------------------------------- pencil maps */

void orc_distribute(int data1, int proc, int *st, int *en, int *sz)
{
    /* decomp.c:72-98: even split, the remainder goes to the last blocks */
    int size1 = data1 / proc;
    int nu = data1 - size1 * proc;
    int nl = proc - nu;
    st[0] = 1; sz[0] = size1; en[0] = size1;
    for (int i = 1; i < nl; i++) { st[i] = en[i - 1] + 1; sz[i] = size1; en[i] = en[i - 1] + size1; }
    size1 += 1;
    for (int i = nl; i < proc; i++) { st[i] = en[i - 1] + 1; sz[i] = size1; en[i] = en[i - 1] + size1; }
}

void orc_partition(int nx, int ny, int nz, const int pdim[3], const int dims[2],
                   const int coord[2], int lstart[3], int lend[3], int lsize[3])
{
    /* decomp.c:18-65 */
    for (int i = 0; i < 3; i++) {
        int gsize = i == 0 ? nx : (i == 1 ? ny : nz);
        if (pdim[i] == 1) { lstart[i] = 1; lend[i] = gsize; lsize[i] = gsize; continue; }
        int np = pdim[i] == 2 ? dims[0] : dims[1];
        int c = pdim[i] == 2 ? coord[0] : coord[1];
        int *st = (int *) malloc(sizeof(int) * (size_t) np);
        int *en = (int *) malloc(sizeof(int) * (size_t) np);
        int *sz = (int *) malloc(sizeof(int) * (size_t) np);
        orc_distribute(gsize, np, st, en, sz);
        lstart[i] = st[c]; lend[i] = en[c]; lsize[i] = sz[c];
        free(st); free(en); free(sz);
    }
}

/* ------------------------------------------------ wall closure bands */
/* One column's band storage: row k holds a..g for columns k-3..k+3
 * (heptdag) or a, b, c for k-1..k+1 (tridag), r the right-hand side and
 * x the solution.  Ghost rows sit at the indices the reference uses
 * (diffuseu.c:617-940, diffusew.c:625-886, pressure.c:446-770). */

typedef struct { int n2; double *a, *b, *c, *d, *e, *f, *g, *r, *x; double *mem; } bands;

static int bands_init(bands *s, int n2)
{
    s->n2 = n2;
    s->mem = (double *) calloc((size_t) n2 * 9, sizeof(double));
    if (!s->mem) return -1;
    s->a = s->mem; s->b = s->a + n2; s->c = s->b + n2; s->d = s->c + n2;
    s->e = s->d + n2; s->f = s->e + n2; s->g = s->f + n2; s->r = s->g + n2; s->x = s->r + n2;
    return 0;
}
static void bands_free(bands *s) { free(s->mem); }

static int is_neumann(int t) { return t == ORC_NEUMAN_CELL || t == ORC_NEUMAN_P; }

/* cell-centred family (u, v, c and p), second order:
 * diffuseu.c:617-700 == pressure.c:446-531 */
static void cell2_rows_bot(bands *s, const double *dz, int k, double bc, int typ)
{
    double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
    if (is_neumann(typ)) { s->a[k] = 0.0; s->b[k] = -1.0 / dzp1h; s->c[k] = +1.0 / dzp1h; }
    else                 { s->a[k] = 0.0; s->b[k] = 0.5;          s->c[k] = 0.5; }
    s->r[k] = bc;
}
static void cell2_rows_top(bands *s, const double *dz, int k, double bc, int typ)
{
    double dzm1h = 0.5 * (dz[k - 1] + dz[k]);
    if (is_neumann(typ)) { s->a[k] = -1.0 / dzm1h; s->b[k] = +1.0 / dzm1h; s->c[k] = 0.0; }
    else                 { s->a[k] = 0.5;          s->b[k] = 0.5;          s->c[k] = 0.0; }
    s->r[k] = bc;
}
static void cell2_reduce_bot(bands *s, int k)
{
    int kp1 = k + 1;
    double fac = s->a[kp1] / s->b[k];
    s->a[kp1] -= s->b[k] * fac;
    s->b[kp1] -= s->c[k] * fac;
    s->r[kp1] -= s->r[k] * fac;
}
static void cell2_reduce_top(bands *s, int k)
{
    int km1 = k - 1;
    double fac = s->c[km1] / s->b[k];
    s->c[km1] -= s->b[k] * fac;
    s->b[km1] -= s->a[k] * fac;
    s->r[km1] -= s->r[k] * fac;
}
static void cell2_fill_bot(bands *s, int k) { s->x[k] = (s->c[k] * s->x[k + 1] - s->r[k]) / -s->b[k]; }
static void cell2_fill_top(bands *s, int k) { s->x[k] = (s->a[k] * s->x[k - 1] - s->r[k]) / -s->b[k]; }

/* face family (w), second order: diffusew.c:625-688 */
static void face2_rows(bands *s, int k, double bc)
{
    s->a[k] = 0.0; s->b[k] = 1.0; s->c[k] = 0.0; s->r[k] = bc;
}
static void face2_reduce_bot(bands *s, int k)
{
    int kp1 = k + 1;
    double fac = s->a[kp1] / s->b[k];
    s->a[kp1] -= s->b[k] * fac;
    s->r[kp1] -= s->r[k] * fac;
}
static void face2_reduce_top(bands *s, int k)
{
    int km1 = k - 1;
    double fac = s->c[km1] / s->b[k];
    s->c[km1] -= s->b[k] * fac;
    s->r[km1] -= s->r[k] * fac;
}
static void face2_fill(bands *s, int k) { s->x[k] = (-s->r[k]) / -s->b[k]; }

/* cell-centred family, fourth order: diffuseu.c:703-940 == pressure.c:534-770 */
static void cell4_rows_bot(bands *s, const double *dz, int k, double bc, int typ)
{
    int km2 = k - 2, km1 = k - 1;
    double dzp3h = 0.5 * (dz[k + 2] + dz[k + 1]);
    double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
    double dzm1h = 0.5 * (dz[k - 1] + dz[k]);
    s->a[k] = 0.0;
    s->b[k] = - A2 / dzm1h;
    s->c[k] = - A1 / dzm1h - 2.0 * - A2 / dzp1h;
    s->d[k] = + A1 / dzm1h - 2.0 * - A1 / dzp1h - A2 / dzp3h;
    s->e[k] = + A2 / dzm1h - 2.0 * + A1 / dzp1h - A1 / dzp3h;
    s->f[k] =              - 2.0 * + A2 / dzp1h + A1 / dzp3h;
    s->g[k] =                                   + A2 / dzp3h;
    s->r[k] = 0.0;
    s->a[km1] = 0.0; s->b[km1] = 0.0; s->c[km1] = 0.0;
    if (is_neumann(typ)) {
        s->d[km1] = - A2 / dzp1h; s->e[km1] = - A1 / dzp1h;
        s->f[km1] = + A1 / dzp1h; s->g[km1] = + A2 / dzp1h;
    } else {
        s->d[km1] = 0.5 * A2; s->e[km1] = 0.5 * (A1 + 2.0 * A2);
        s->f[km1] = 0.5 * (A1 + 2.0 * A2); s->g[km1] = 0.5 * A2;
    }
    s->r[km1] = bc;
    s->a[km2] = 0.0; s->b[km2] = 0.0; s->c[km2] = 0.0; s->d[km2] = 0.0; s->e[km2] = 0.0;
    s->f[km2] = -1.0; s->g[km2] = +1.0; s->r[km2] = 0.0;
}
static void cell4_rows_top(bands *s, const double *dz, int k, double bc, int typ)
{
    int kp2 = k + 2, kp1 = k + 1;
    double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
    double dzm1h = 0.5 * (dz[k - 1] + dz[k]);
    double dzm3h = 0.5 * (dz[k - 2] + dz[k - 1]);
    s->a[k] = - A2 / dzm3h;
    s->b[k] = - A1 / dzm3h - 2.0 * - A2 / dzm1h;
    s->c[k] = + A1 / dzm3h - 2.0 * - A1 / dzm1h - A2 / dzp1h;
    s->d[k] = + A2 / dzm3h - 2.0 * + A1 / dzm1h - A1 / dzp1h;
    s->e[k] =              - 2.0 * + A2 / dzm1h + A1 / dzp1h;
    s->f[k] =                                   + A2 / dzp1h;
    s->g[k] = 0.0;
    s->r[k] = 0.0;
    if (is_neumann(typ)) {
        s->a[kp1] = - A2 / dzm1h; s->b[kp1] = - A1 / dzm1h;
        s->c[kp1] = + A1 / dzm1h; s->d[kp1] = + A2 / dzm1h;
    } else {
        s->a[kp1] = 0.5 * A2; s->b[kp1] = 0.5 * (A1 + 2.0 * A2);
        s->c[kp1] = 0.5 * (A1 + 2.0 * A2); s->d[kp1] = 0.5 * A2;
    }
    s->e[kp1] = 0.0; s->f[kp1] = 0.0; s->g[kp1] = 0.0; s->r[kp1] = bc;
    s->a[kp2] = -1.0; s->b[kp2] = +1.0; s->c[kp2] = 0.0; s->d[kp2] = 0.0;
    s->e[kp2] = 0.0; s->f[kp2] = 0.0; s->g[kp2] = 0.0; s->r[kp2] = 0.0;
}
static void cell4_reduce_bot(bands *s, int k)
{
    double *a = s->a, *b = s->b, *c = s->c, *d = s->d, *e = s->e, *f = s->f, *g = s->g, *r = s->r;
    int km2 = k - 2, km1 = k - 1, kp1 = k + 1, kp2 = k + 2, kp3 = k + 3;
    double fac;
    fac = a[kp1] / b[k];
    a[kp1] -= b[k] * fac; b[kp1] -= c[k] * fac; c[kp1] -= d[k] * fac;
    d[kp1] -= e[k] * fac; e[kp1] -= f[k] * fac; f[kp1] -= g[k] * fac;
    fac = b[kp1] / d[km1];
    b[kp1] -= d[km1] * fac; c[kp1] -= e[km1] * fac; d[kp1] -= f[km1] * fac;
    e[kp1] -= g[km1] * fac; r[kp1] -= r[km1] * fac;
    fac = c[kp1] / f[km2];
    c[kp1] -= f[km2] * fac; d[kp1] -= g[km2] * fac;
    fac = a[kp2] / d[km1];
    a[kp2] -= d[km1] * fac; b[kp2] -= e[km1] * fac; c[kp2] -= f[km1] * fac;
    d[kp2] -= g[km1] * fac; r[kp2] -= r[km1] * fac;
    fac = b[kp2] / f[km2];
    b[kp2] -= f[km2] * fac; c[kp2] -= g[km2] * fac;
    fac = a[kp3] / f[km2];
    a[kp3] -= f[km2] * fac; b[kp3] -= g[km2] * fac;
}
static void cell4_reduce_top(bands *s, int k)
{
    double *a = s->a, *b = s->b, *c = s->c, *d = s->d, *e = s->e, *f = s->f, *g = s->g, *r = s->r;
    int km3 = k - 3, km2 = k - 2, km1 = k - 1, kp1 = k + 1, kp2 = k + 2;
    double fac;
    fac = g[km1] / f[k];
    g[km1] -= f[k] * fac; f[km1] -= e[k] * fac; e[km1] -= d[k] * fac;
    d[km1] -= c[k] * fac; c[km1] -= b[k] * fac; b[km1] -= a[k] * fac;
    fac = f[km1] / d[kp1];
    f[km1] -= d[kp1] * fac; e[km1] -= c[kp1] * fac; d[km1] -= b[kp1] * fac;
    c[km1] -= a[kp1] * fac; r[km1] -= r[kp1] * fac;
    fac = e[km1] / b[kp2];
    e[km1] -= b[kp2] * fac; d[km1] -= a[kp2] * fac;
    fac = g[km2] / d[kp1];
    g[km2] -= d[kp1] * fac; f[km2] -= c[kp1] * fac; e[km2] -= b[kp1] * fac;
    d[km2] -= a[kp1] * fac; r[km2] -= r[kp1] * fac;
    fac = f[km2] / b[kp2];
    f[km2] -= b[kp2] * fac; e[km2] -= a[kp2] * fac;
    fac = g[km3] / b[kp2];
    g[km3] -= b[kp2] * fac; f[km3] -= a[kp2] * fac;
}
static void cell4_fill_bot(bands *s, int k)
{
    double *b = s->b, *c = s->c, *d = s->d, *e = s->e, *f = s->f, *g = s->g, *r = s->r, *x = s->x;
    int km2 = k - 2, km1 = k - 1, kp1 = k + 1, kp2 = k + 2, kp3 = k + 3;
    x[k]   = (g[km2] * x[kp1]) / -f[km2];
    x[km1] = (e[km1] * x[k] + f[km1] * x[kp1] + g[km1] * x[kp2] - r[km1]) / -d[km1];
    x[km2] = (c[k] * x[km1] + d[k] * x[k] + e[k] * x[kp1] + f[k] * x[kp2] + g[k] * x[kp3]) / -b[k];
}
static void cell4_fill_top(bands *s, int k)
{
    double *a = s->a, *b = s->b, *c = s->c, *d = s->d, *e = s->e, *f = s->f, *r = s->r, *x = s->x;
    int km3 = k - 3, km2 = k - 2, km1 = k - 1, kp1 = k + 1, kp2 = k + 2;
    x[k]   = (a[kp2] * x[km1]) / -b[kp2];
    x[kp1] = (a[kp1] * x[km2] + b[kp1] * x[km1] + c[kp1] * x[k] - r[kp1]) / -d[kp1];
    x[kp2] = (a[k] * x[km3] + b[k] * x[km2] + c[k] * x[km1] + d[k] * x[k] + e[k] * x[kp1]) / -f[k];
}

/* face family, fourth order: diffusew.c:691-886 */
static void face4_row_wall(bands *s, const double *dz, int k)
{
    double dzp0 = dz[k], dzp1 = dz[k + 1];
    s->a[k] = 0.0;
    s->b[k] = - A2 / dzp0;
    s->c[k] = - A1 / dzp0 + A2 / dzp1;
    s->d[k] = + A1 / dzp0 + A1 / dzp1;
    s->e[k] = + A2 / dzp0 - A1 / dzp1;
    s->f[k] =             - A2 / dzp1;
    s->g[k] = 0.0;
    s->r[k] = 0.0;
}
static void face4_rows_bot(bands *s, const double *dz, int k, double bc)
{
    int km2 = k - 2, km1 = k - 1;
    face4_row_wall(s, dz, k);
    s->a[km1] = 0.0; s->b[km1] = 0.0; s->c[km1] = 0.0; s->d[km1] = -1.0;
    s->e[km1] = +2.0; s->f[km1] = -1.0; s->g[km1] = 0.0; s->r[km1] = 0.0;
    s->a[km2] = 0.0; s->b[km2] = 0.0; s->c[km2] = 0.0; s->d[km2] = 0.0;
    s->e[km2] = 0.0; s->f[km2] = 1.0; s->g[km2] = 0.0; s->r[km2] = bc;
}
static void face4_rows_top(bands *s, const double *dz, int k, double bc)
{
    int kp2 = k + 2, kp1 = k + 1;
    face4_row_wall(s, dz, k);
    s->a[kp1] = 0.0; s->b[kp1] = -1.0; s->c[kp1] = +2.0; s->d[kp1] = -1.0;
    s->e[kp1] = 0.0; s->f[kp1] = 0.0; s->g[kp1] = 0.0; s->r[kp1] = 0.0;
    s->a[kp2] = 0.0; s->b[kp2] = 1.0; s->c[kp2] = 0.0; s->d[kp2] = 0.0;
    s->e[kp2] = 0.0; s->f[kp2] = 0.0; s->g[kp2] = 0.0; s->r[kp2] = bc;
}
static void face4_reduce_bot(bands *s, int k)
{
    double *a = s->a, *b = s->b, *c = s->c, *d = s->d, *e = s->e, *f = s->f, *r = s->r;
    int km2 = k - 2, km1 = k - 1, kp1 = k + 1, kp2 = k + 2, kp3 = k + 3;
    double fac;
    fac = a[kp1] / b[k];
    a[kp1] -= b[k] * fac; b[kp1] -= c[k] * fac; c[kp1] -= d[k] * fac;
    d[kp1] -= e[k] * fac; e[kp1] -= f[k] * fac;
    fac = b[kp1] / d[km1];
    b[kp1] -= d[km1] * fac; c[kp1] -= e[km1] * fac; d[kp1] -= f[km1] * fac;
    fac = c[kp1] / f[km2];
    c[kp1] -= f[km2] * fac; r[kp1] -= r[km2] * fac;
    fac = a[kp2] / d[km1];
    a[kp2] -= d[km1] * fac; b[kp2] -= e[km1] * fac; c[kp2] -= f[km1] * fac;
    fac = b[kp2] / f[km2];
    b[kp2] -= f[km2] * fac; r[kp2] -= r[km2] * fac;
    fac = a[kp3] / f[km2];
    a[kp3] -= f[km2] * fac; r[kp3] -= r[km2] * fac;
}
static void face4_reduce_top(bands *s, int k)
{
    double *b = s->b, *c = s->c, *d = s->d, *e = s->e, *f = s->f, *g = s->g, *r = s->r;
    int km3 = k - 3, km2 = k - 2, km1 = k - 1, kp1 = k + 1, kp2 = k + 2;
    double fac;
    fac = g[km1] / f[k];
    g[km1] -= f[k] * fac; f[km1] -= e[k] * fac; e[km1] -= d[k] * fac;
    d[km1] -= c[k] * fac; c[km1] -= b[k] * fac;
    fac = f[km1] / d[kp1];
    f[km1] -= d[kp1] * fac; e[km1] -= c[kp1] * fac; d[km1] -= b[kp1] * fac;
    fac = e[km1] / b[kp2];
    e[km1] -= b[kp2] * fac; r[km1] -= r[kp2] * fac;
    fac = g[km2] / d[kp1];
    g[km2] -= d[kp1] * fac; f[km2] -= c[kp1] * fac; e[km2] -= b[kp1] * fac;
    fac = f[km2] / b[kp2];
    f[km2] -= b[kp2] * fac; r[km2] -= r[kp2] * fac;
    fac = g[km3] / b[kp2];
    g[km3] -= b[kp2] * fac; r[km3] -= r[kp2] * fac;
}
static void face4_fill_bot(bands *s, int k)
{
    double *b = s->b, *c = s->c, *d = s->d, *e = s->e, *f = s->f, *r = s->r, *x = s->x;
    int km2 = k - 2, km1 = k - 1, kp1 = k + 1, kp2 = k + 2;
    x[k]   = (- r[km2]) / -f[km2];
    x[km1] = (e[km1] * x[k] + f[km1] * x[kp1]) / -d[km1];
    x[km2] = (c[k] * x[km1] + d[k] * x[k] + e[k] * x[kp1] + f[k] * x[kp2]) / -b[k];
}
static void face4_fill_top(bands *s, int k)
{
    double *b = s->b, *c = s->c, *d = s->d, *e = s->e, *f = s->f, *r = s->r, *x = s->x;
    int km2 = k - 2, km1 = k - 1, kp1 = k + 1, kp2 = k + 2;
    x[k]   = (- r[kp2]) / -b[kp2];
    x[kp1] = (b[kp1] * x[km1] + c[kp1] * x[k]) / -d[kp1];
    x[kp2] = (b[k] * x[km2] + c[k] * x[km1] + d[k] * x[k] + e[k] * x[kp1]) / -f[k];
}

/* ----------------------------------------------------- periodic halos */

void orc_periodic(double *u, const orc_grid *g)
{
    /* ghostz.c:107-149 on one rank: x-planes first over the full padded
     * y-z extent, then y-planes over the full padded x-z extent, so the
     * corner ghosts are copies of interior points. */
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    size_t plane = (size_t) n1 * n2;
    for (int t = 0; t < ng; t++) {
        memcpy(u + (size_t) (n0 - ng + t) * plane, u + (size_t) (ng + t) * plane, plane * sizeof(double));
        memcpy(u + (size_t) t * plane, u + (size_t) (n0 - 2 * ng + t) * plane, plane * sizeof(double));
    }
    for (int i = 0; i < n0; i++)
        for (int t = 0; t < ng; t++) {
            memcpy(u + ID(g, i, n1 - ng + t, 0), u + ID(g, i, ng + t, 0), (size_t) n2 * sizeof(double));
            memcpy(u + ID(g, i, t, 0), u + ID(g, i, n1 - 2 * ng + t, 0), (size_t) n2 * sizeof(double));
        }
}

/* ------------------------------------------------------ wall ghosts */

void orc_apply_bc_cell(double *u, const orc_grid *g, int order,
                       double bcb, double bct, int typb, int typt)
{
    /* apply_bc_u2 / apply_bc_u4, diffuseu.c:37-104 */
    int ng = g->ng, n2 = g->n2;
    bands s;
    if (bands_init(&s, n2)) return;
    for (int i = 0; i < g->n0; i++)
        for (int j = 0; j < g->n1; j++) {
            double *col = u + ID(g, i, j, 0);
            memcpy(s.x, col, sizeof(double) * (size_t) n2);
            if (order == 2) {
                cell2_rows_bot(&s, g->dz, ng - 1, bcb, typb);
                cell2_rows_top(&s, g->dz, n2 - ng, bct, typt);
                cell2_fill_bot(&s, ng - 1);
                cell2_fill_top(&s, n2 - ng);
            } else {
                cell4_rows_bot(&s, g->dz, ng - 1, bcb, typb);
                cell4_rows_top(&s, g->dz, n2 - ng, bct, typt);
                cell4_fill_bot(&s, ng - 1);
                cell4_fill_top(&s, n2 - ng);
            }
            memcpy(col, s.x, sizeof(double) * (size_t) n2);
        }
    bands_free(&s);
}

void orc_apply_bc_face(double *w, const orc_grid *g, int order, double bcb, double bct)
{
    /* apply_bc_w2 / apply_bc_w4, diffusew.c:37-104 */
    int ng = g->ng, n2 = g->n2;
    bands s;
    if (bands_init(&s, n2)) return;
    for (int i = 0; i < g->n0; i++)
        for (int j = 0; j < g->n1; j++) {
            double *col = w + ID(g, i, j, 0);
            memcpy(s.x, col, sizeof(double) * (size_t) n2);
            if (order == 2) {
                face2_rows(&s, ng - 1, bcb);
                face2_rows(&s, n2 - ng - 1, bct);
                face2_fill(&s, ng - 1);
                face2_fill(&s, n2 - ng - 1);
            } else {
                face4_rows_bot(&s, g->dz, ng - 1, bcb);
                face4_rows_top(&s, g->dz, n2 - ng - 1, bct);
                face4_fill_bot(&s, ng - 1);
                face4_fill_top(&s, n2 - ng - 1);
            }
            memcpy(col, s.x, sizeof(double) * (size_t) n2);
        }
    bands_free(&s);
}

/* ------------------------------------------------ 1-D z Helmholtz */

void orc_helmholtz1d_cell(const double *u, double *h, const orc_grid *g, int order, double nudt)
{
    /* diffuseu.c:107-170 */
    int ng = g->ng, n2 = g->n2;
    const double *dz = g->dz;
    for (int i = 0; i < g->n0; i++)
        for (int j = 0; j < g->n1; j++) {
            const double *q = u + ID(g, i, j, 0);
            double *o = h + ID(g, i, j, 0);
            for (int k = ng; k < n2 - ng; k++) {
                double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
                double dzm1h = 0.5 * (dz[k - 1] + dz[k]);
                if (order == 2) {
                    o[k] = q[k] + nudt * (((q[k + 1] - q[k]) / dzp1h - (q[k] - q[k - 1]) / dzm1h) / dz[k]);
                } else {
                    double wp3h = A1 * (q[k + 2] - q[k + 1]) + A2 * (q[k + 3] - q[k]);
                    double wp1h = A1 * (q[k + 1] - q[k]) + A2 * (q[k + 2] - q[k - 1]);
                    double wm1h = A1 * (q[k] - q[k - 1]) + A2 * (q[k + 1] - q[k - 2]);
                    double wm3h = A1 * (q[k - 1] - q[k - 2]) + A2 * (q[k] - q[k - 3]);
                    double dzp3h = 0.5 * (dz[k + 2] + dz[k + 1]);
                    double dzm3h = 0.5 * (dz[k - 2] + dz[k - 1]);
                    o[k] = q[k] + nudt * ((A1 * (wp1h / dzp1h - wm1h / dzm1h)
                                         + A2 * (wp3h / dzp3h - wm3h / dzm3h)) / dz[k]);
                }
            }
        }
}

void orc_helmholtz1d_face(const double *w, double *h, const orc_grid *g, int order, double nudt)
{
    /* diffusew.c:107-172 */
    int ng = g->ng, n2 = g->n2;
    const double *dz = g->dz;
    for (int i = 0; i < g->n0; i++)
        for (int j = 0; j < g->n1; j++) {
            const double *q = w + ID(g, i, j, 0);
            double *o = h + ID(g, i, j, 0);
            for (int k = ng; k < n2 - ng - 1; k++) {
                double dzp1 = dz[k + 1], dzm0 = dz[k];
                double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
                if (order == 2) {
                    o[k] = q[k] + nudt * (((q[k + 1] - q[k]) / dzp1 - (q[k] - q[k - 1]) / dzm0) / dzp1h);
                } else {
                    double wp3h = A1 * (q[k + 2] - q[k + 1]) + A2 * (q[k + 3] - q[k]);
                    double wp1h = A1 * (q[k + 1] - q[k]) + A2 * (q[k + 2] - q[k - 1]);
                    double wm1h = A1 * (q[k] - q[k - 1]) + A2 * (q[k + 1] - q[k - 2]);
                    double wm3h = A1 * (q[k - 1] - q[k - 2]) + A2 * (q[k] - q[k - 3]);
                    double dzp2 = dz[k + 2], dzm1 = dz[k - 1];
                    o[k] = q[k] + nudt * ((A1 * (wp1h / dzp1 - wm1h / dzm0)
                                         + A2 * (wp3h / dzp2 - wm3h / dzm1)) / dzp1h);
                }
            }
        }
}

/* interior band rows of (I - nudt d2/dz2) * dvol for a cell-centred
 * column, diffuseu.c:192-203 (order 2) and :241-266 (order 4) */
static void cell_bands(bands *s, const double *dz, int ng, int n2, int order,
                       double nudt, double dxy, double lam)
{
    /* dxy = 1 (solve_helmholtz1d_*) or dx*dy (pressure / spectral);
     * lam = the horizontal modified-wavenumber term added to the diagonal
     * (0 for the 1-D solves) -- caller adds it, see orc_poisson. */
    (void) lam;
    for (int k = ng; k < n2 - ng; k++) {
        double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
        double dzm1h = 0.5 * (dz[k - 1] + dz[k]);
        double dvol = dxy == 1.0 ? dz[k] : dxy * dz[k];
        if (order == 2) {
            s->a[k] = + dvol * nudt / (dzm1h * dz[k]);
            s->c[k] = + dvol * nudt / (dzp1h * dz[k]);
            s->b[k] = - dvol * nudt / (dzm1h * dz[k])
                      - dvol * nudt / (dzp1h * dz[k]);
        } else {
            double dzp3h = 0.5 * (dz[k + 2] + dz[k + 1]);
            double dzm3h = 0.5 * (dz[k - 2] + dz[k - 1]);
            s->a[k] = + A2 * A2 * dvol * nudt / (dzm3h * dz[k]);
            s->g[k] = + A2 * A2 * dvol * nudt / (dzp3h * dz[k]);
            s->b[k] = + A1 * A2 * dvol * nudt / (dzm1h * dz[k])
                      + A1 * A2 * dvol * nudt / (dzm3h * dz[k]);
            s->f[k] = + A1 * A2 * dvol * nudt / (dzp1h * dz[k])
                      + A1 * A2 * dvol * nudt / (dzp3h * dz[k]);
            s->c[k] = - A1 * A2 * dvol * nudt / (dzp1h * dz[k])
                      - A1 * A2 * dvol * nudt / (dzm3h * dz[k])
                      + A1 * A1 * dvol * nudt / (dzm1h * dz[k]);
            s->e[k] = - A1 * A2 * dvol * nudt / (dzm1h * dz[k])
                      - A1 * A2 * dvol * nudt / (dzp3h * dz[k])
                      + A1 * A1 * dvol * nudt / (dzp1h * dz[k]);
            s->d[k] = - A2 * A2 * dvol * nudt / (dzm3h * dz[k])
                      - A2 * A2 * dvol * nudt / (dzp3h * dz[k])
                      - A1 * A1 * dvol * nudt / (dzm1h * dz[k])
                      - A1 * A1 * dvol * nudt / (dzp1h * dz[k]);
        }
    }
}

void orc_solve1d_cell(double *u, const double *h, const orc_grid *g, int order,
                      double nudt, double bcb, double bct, int typb, int typt)
{
    /* solve_helmholtz1d_u2 / _u4, diffuseu.c:173-283.  The diagonal is
     * "... + dvol", added after the band terms exactly as written there. */
    int ng = g->ng, n2 = g->n2, nz = g->nz;
    const double *dz = g->dz;
    bands s;
    if (bands_init(&s, n2)) return;
    for (int k = 0; k < n2; k++) s.x[k] = 0.0;
    for (int i = 0; i < g->n0; i++)
        for (int j = 0; j < g->n1; j++) {
            const double *hc = h + ID(g, i, j, 0);
            cell_bands(&s, dz, ng, n2, order, nudt, 1.0, 0.0);
            for (int k = ng; k < n2 - ng; k++) {
                double dvol = dz[k];
                if (order == 2) s.b[k] = s.b[k] + dvol; else s.d[k] = s.d[k] + dvol;
                s.r[k] = hc[k] * dvol;
            }
            int kb = ng - 1, kt = n2 - ng;
            if (order == 2) {
                cell2_rows_bot(&s, dz, kb, bcb, typb);
                cell2_rows_top(&s, dz, kt, bct, typt);
                cell2_reduce_bot(&s, kb);
                cell2_reduce_top(&s, kt);
                orc_tridag(s.a + ng, s.b + ng, s.c + ng, s.r + ng, s.x + ng, nz);
                cell2_fill_bot(&s, kb);
                cell2_fill_top(&s, kt);
            } else {
                cell4_rows_bot(&s, dz, kb, bcb, typb);
                cell4_rows_top(&s, dz, kt, bct, typt);
                cell4_reduce_bot(&s, kb);
                cell4_reduce_top(&s, kt);
                orc_heptdag(s.a + ng, s.b + ng, s.c + ng, s.d + ng, s.e + ng, s.f + ng,
                            s.g + ng, s.r + ng, s.x + ng, nz);
                cell4_fill_bot(&s, kb);
                cell4_fill_top(&s, kt);
            }
            memcpy(u + ID(g, i, j, 0), s.x, sizeof(double) * (size_t) n2);
        }
    bands_free(&s);
}

void orc_solve1d_face(double *w, const double *h, const orc_grid *g, int order,
                      double nudt, double bcb, double bct)
{
    /* solve_helmholtz1d_w2 / _w4, diffusew.c:175-287: nz-1 unknowns on the
     * interior faces, dvol = the face cell height 0.5 (dz_k + dz_k+1) */
    int ng = g->ng, n2 = g->n2, nz = g->nz;
    const double *dz = g->dz;
    bands s;
    if (bands_init(&s, n2)) return;
    for (int k = 0; k < n2; k++) s.x[k] = 0.0;
    for (int i = 0; i < g->n0; i++)
        for (int j = 0; j < g->n1; j++) {
            const double *hc = h + ID(g, i, j, 0);
            for (int k = ng; k < n2 - ng - 1; k++) {
                double dzp1 = dz[k + 1], dzm0 = dz[k];
                double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
                double dvol = dzp1h;
                if (order == 2) {
                    s.a[k] = + dvol * nudt / (dzm0 * dzp1h);
                    s.c[k] = + dvol * nudt / (dzp1 * dzp1h);
                    s.b[k] = - dvol * nudt / (dzm0 * dzp1h)
                             - dvol * nudt / (dzp1 * dzp1h)
                             + dvol;
                } else {
                    double dzp2 = dz[k + 2], dzm1 = dz[k - 1];
                    s.a[k] = + A2 * A2 * dvol * nudt / (dzm1 * dzp1h);
                    s.g[k] = + A2 * A2 * dvol * nudt / (dzp2 * dzp1h);
                    s.b[k] = + A1 * A2 * dvol * nudt / (dzm0 * dzp1h)
                             + A1 * A2 * dvol * nudt / (dzm1 * dzp1h);
                    s.f[k] = + A1 * A2 * dvol * nudt / (dzp1 * dzp1h)
                             + A1 * A2 * dvol * nudt / (dzp2 * dzp1h);
                    s.c[k] = - A1 * A2 * dvol * nudt / (dzp1 * dzp1h)
                             - A1 * A2 * dvol * nudt / (dzm1 * dzp1h)
                             + A1 * A1 * dvol * nudt / (dzm0 * dzp1h);
                    s.e[k] = - A1 * A2 * dvol * nudt / (dzm0 * dzp1h)
                             - A1 * A2 * dvol * nudt / (dzp2 * dzp1h)
                             + A1 * A1 * dvol * nudt / (dzp1 * dzp1h);
                    s.d[k] = - A2 * A2 * dvol * nudt / (dzm1 * dzp1h)
                             - A2 * A2 * dvol * nudt / (dzp2 * dzp1h)
                             - A1 * A1 * dvol * nudt / (dzm0 * dzp1h)
                             - A1 * A1 * dvol * nudt / (dzp1 * dzp1h)
                             + dvol;
                }
                s.r[k] = hc[k] * dvol;
            }
            int kb = ng - 1, kt = n2 - ng - 1;
            if (order == 2) {
                face2_rows(&s, kb, bcb);
                face2_rows(&s, kt, bct);
                face2_reduce_bot(&s, kb);
                face2_reduce_top(&s, kt);
                orc_tridag(s.a + ng, s.b + ng, s.c + ng, s.r + ng, s.x + ng, nz - 1);
                face2_fill(&s, kb);
                face2_fill(&s, kt);
            } else {
                face4_rows_bot(&s, dz, kb, bcb);
                face4_rows_top(&s, dz, kt, bct);
                face4_reduce_bot(&s, kb);
                face4_reduce_top(&s, kt);
                orc_heptdag(s.a + ng, s.b + ng, s.c + ng, s.d + ng, s.e + ng, s.f + ng,
                            s.g + ng, s.r + ng, s.x + ng, nz - 1);
                face4_fill_bot(&s, kb);
                face4_fill_top(&s, kt);
            }
            memcpy(w + ID(g, i, j, 0), s.x, sizeof(double) * (size_t) n2);
        }
    bands_free(&s);
}

/* ------------------------------------------- explicit x-y diffusion */

static void laplacexy(const double *u, double *h, const orc_grid *g, int order, double nu, int ktop)
{
    int ng = g->ng, n0 = g->n0, n1 = g->n1;
    double dx = g->dx, dy = g->dy;
    size_t sx = (size_t) g->n1 * g->n2, sy = (size_t) g->n2;
    int lo = order == 2 ? 1 : 3;
    for (int i = lo; i < n0 - lo; i++)
        for (int j = lo; j < n1 - lo; j++)
            for (int k = ng; k < ktop; k++) {
                size_t o = ID(g, i, j, k);
                const double *q = u + o;
                if (order == 2) {
                    h[o] += nu * (((q[sx] - q[0]) / dx - (q[0] - q[-(long) sx]) / dx) / dx
                                + ((q[sy] - q[0]) / dy - (q[0] - q[-(long) sy]) / dy) / dy);
                } else {
                    long X = (long) sx, Y = (long) sy;
                    double up3h = A1 * (q[2 * X] - q[X]) + A2 * (q[3 * X] - q[0]);
                    double up1h = A1 * (q[X] - q[0]) + A2 * (q[2 * X] - q[-X]);
                    double um1h = A1 * (q[0] - q[-X]) + A2 * (q[X] - q[-2 * X]);
                    double um3h = A1 * (q[-X] - q[-2 * X]) + A2 * (q[0] - q[-3 * X]);
                    double vp3h = A1 * (q[2 * Y] - q[Y]) + A2 * (q[3 * Y] - q[0]);
                    double vp1h = A1 * (q[Y] - q[0]) + A2 * (q[2 * Y] - q[-Y]);
                    double vm1h = A1 * (q[0] - q[-Y]) + A2 * (q[Y] - q[-2 * Y]);
                    double vm3h = A1 * (q[-Y] - q[-2 * Y]) + A2 * (q[0] - q[-3 * Y]);
                    h[o] += nu * ((A1 * (up1h / dx - um1h / dx) + A2 * (up3h / dx - um3h / dx)) / dx
                                + (A1 * (vp1h / dy - vm1h / dy) + A2 * (vp3h / dy - vm3h / dy)) / dy);
                }
            }
}

void orc_laplacexy_cell(const double *u, double *h, const orc_grid *g, int order, double nu)
{
    laplacexy(u, h, g, order, nu, g->n2 - g->ng);           /* diffuseu.c:286-359 */
}

void orc_laplacexy_face(const double *w, double *h, const orc_grid *g, int order, double nu)
{
    laplacexy(w, h, g, order, nu, g->n2 - g->ng - 1);       /* diffusew.c:290-365 */
}

/* -------------------------------------------------------- advection */

void orc_massflux(int dir, double *u, double *v, double *w,
                  double *mu, double *mv, double *mw, const orc_grid *g, int order)
{
    /* massflux{x,y,z}{2,4}, advect.c:6-283.  The fourth-order versions
     * first rewrite the wall ghosts of u, v and w in place
     * (advect.c:100-131): first ghost of u, v = first interior value,
     * w mirrored oddly about the wall face. */
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    long X = (long) n1 * n2, Y = n2;
    if (order == 2) {
        for (int i = 0; i < n0 - 1; i++)
            for (int j = 0; j < n1 - 1; j++)
                for (int k = 0; k < n2 - 1; k++) {
                    size_t o = ID(g, i, j, k);
                    double dzm0 = dz[k];
                    if (dir == 0) {
                        mu[o] = 0.5 * (dy * dzm0 * u[o] + dy * dzm0 * u[o + X]);
                        mv[o] = 0.5 * (dx * dzm0 * v[o] + dx * dzm0 * v[o + X]);
                        mw[o] = 0.5 * (dx * dy * w[o] + dx * dy * w[o + X]);
                    } else if (dir == 1) {
                        mu[o] = 0.5 * (dy * dzm0 * u[o] + dy * dzm0 * u[o + Y]);
                        mv[o] = 0.5 * (dx * dzm0 * v[o] + dx * dzm0 * v[o + Y]);
                        mw[o] = 0.5 * (dx * dy * w[o] + dx * dy * w[o + Y]);
                    } else {
                        double dzp1 = dz[k + 1];
                        mu[o] = 0.5 * (dy * dzm0 * u[o] + dy * dzp1 * u[o + 1]);
                        mv[o] = 0.5 * (dx * dzm0 * v[o] + dx * dzp1 * v[o + 1]);
                        mw[o] = 0.5 * (dx * dy * w[o] + dx * dy * w[o + 1]);
                    }
                }
        return;
    }
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++) {
            size_t o = ID(g, i, j, 0);
            int k = ng - 1;
            u[o + k] = u[o + k + 1];
            v[o + k] = v[o + k + 1];
            w[o + k - 1] = 2.0 * w[o + k] - w[o + k + 1];
            w[o + k - 2] = 2.0 * w[o + k] - w[o + k + 2];
            k = n2 - ng;
            u[o + k] = u[o + k - 1];
            v[o + k] = v[o + k - 1];
            k = n2 - ng - 1;
            w[o + k + 1] = 2.0 * w[o + k] - w[o + k - 1];
            w[o + k + 2] = 2.0 * w[o + k] - w[o + k - 2];
        }
    for (int i = 1; i < n0 - 2; i++)
        for (int j = 1; j < n1 - 2; j++)
            for (int k = 1; k < n2 - 2; k++) {
                size_t o = ID(g, i, j, k);
                double dzm0 = dz[k];
                if (dir == 0) {
                    mu[o] = B1 * (dy * dzm0 * u[o] + dy * dzm0 * u[o + X])
                          + B2 * (dy * dzm0 * u[o - X] + dy * dzm0 * u[o + 2 * X]);
                    mv[o] = B1 * (dx * dzm0 * v[o] + dx * dzm0 * v[o + X])
                          + B2 * (dx * dzm0 * v[o - X] + dx * dzm0 * v[o + 2 * X]);
                    mw[o] = B1 * (dx * dy * w[o] + dx * dy * w[o + X])
                          + B2 * (dx * dy * w[o - X] + dx * dy * w[o + 2 * X]);
                } else if (dir == 1) {
                    mu[o] = B1 * (dy * dzm0 * u[o] + dy * dzm0 * u[o + Y])
                          + B2 * (dy * dzm0 * u[o - Y] + dy * dzm0 * u[o + 2 * Y]);
                    mv[o] = B1 * (dx * dzm0 * v[o] + dx * dzm0 * v[o + Y])
                          + B2 * (dx * dzm0 * v[o - Y] + dx * dzm0 * v[o + 2 * Y]);
                    mw[o] = B1 * (dx * dy * w[o] + dx * dy * w[o + Y])
                          + B2 * (dx * dy * w[o - Y] + dx * dy * w[o + 2 * Y]);
                } else {
                    double dzm1 = dz[k - 1], dzp1 = dz[k + 1], dzp2 = dz[k + 2];
                    mu[o] = B1 * (dy * dzm0 * u[o] + dy * dzp1 * u[o + 1])
                          + B2 * (dy * dzm1 * u[o - 1] + dy * dzp2 * u[o + 2]);
                    mv[o] = B1 * (dx * dzm0 * v[o] + dx * dzp1 * v[o + 1])
                          + B2 * (dx * dzm1 * v[o - 1] + dx * dzp2 * v[o + 2]);
                    mw[o] = B1 * (dx * dy * w[o] + dx * dy * w[o + 1])
                          + B2 * (dx * dy * w[o - 1] + dx * dy * w[o + 2]);
                }
            }
}

void orc_advect(int comp, const double *q, const double *mu, const double *mv,
                const double *mw, double *f, const orc_grid *g, int order)
{
    /* advect{u,v,w}{2,4}, advect.c:286-734: divergence-form convective
     * tendency; comp 0 = u, 1 = v (cell rows k in [ng, n2-ng)), 2 = w
     * (face rows k in [ng, n2-ng-1)).  For u and v the wall-face flux uses
     * the interpolated wall value (advect.c:336-341, :539-554). */
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    long X = (long) n1 * n2, Y = n2;
    memset(f, 0, sizeof(double) * (size_t) n0 * n1 * n2);
    /* comp u reads the x-direction fluxes (muphx...), v the y ones, w the z
     * ones; the caller passes the matching set. */
    int lo = order == 2 ? 1 : 3;
    for (int i = lo; i < n0 - lo; i++)
        for (int j = lo; j < n1 - lo; j++) {
            size_t c0 = ID(g, i, j, 0);
            const double *s = q + c0;
            double bcb = 0.0, bct = 0.0;
            if (comp < 2) {
                int kb = ng - 1, kt = n2 - ng;
                if (order == 2) {
                    bcb = 0.5 * (s[kb] + s[kb + 1]);
                    bct = 0.5 * (s[kt - 1] + s[kt]);
                } else {
                    bcb = 0.5 * A2 * (s[kb - 1] + s[kb + 2]) + 0.5 * (A1 + 2.0 * A2) * (s[kb] + s[kb + 1]);
                    bct = 0.5 * A2 * (s[kt - 2] + s[kt + 1]) + 0.5 * (A1 + 2.0 * A2) * (s[kt - 1] + s[kt]);
                }
            }
            int khi = comp < 2 ? n2 - ng : n2 - ng - 1;
            for (int k = ng; k < khi; k++) {
                size_t o = c0 + (size_t) k;
                const double *Q = q + o;
                double xp = mu[o] * 0.5 * (Q[0] + Q[X]);
                double yp = mv[o] * 0.5 * (Q[0] + Q[Y]);
                double zp = mw[o] * 0.5 * (Q[0] + Q[1]);
                double xm = mu[o - X] * 0.5 * (Q[0] + Q[-X]);
                double ym = mv[o - Y] * 0.5 * (Q[0] + Q[-Y]);
                double zm = mw[o - 1] * 0.5 * (Q[0] + Q[-1]);
                double den = comp < 2 ? (dx * dy * dz[k]) : (dx * dy * (0.5 * (dz[k + 1] + dz[k])));
                if (order == 2) {
                    if (comp < 2) {
                        if (k == n2 - ng - 1) zp = mw[o] * bct;
                        if (k == ng) zm = mw[o - 1] * bcb;
                    }
                    f[o] = -((xp - xm) + (yp - ym) + (zp - zm)) / den;
                } else {
                    double xp3 = mu[o + X] * 0.5 * (Q[0] + Q[3 * X]);
                    double yp3 = mv[o + Y] * 0.5 * (Q[0] + Q[3 * Y]);
                    double zp3 = mw[o + 1] * 0.5 * (Q[0] + Q[3]);
                    double xm3 = mu[o - 2 * X] * 0.5 * (Q[0] + Q[-3 * X]);
                    double ym3 = mv[o - 2 * Y] * 0.5 * (Q[0] + Q[-3 * Y]);
                    double zm3 = mw[o - 2] * 0.5 * (Q[0] + Q[-3]);
                    if (comp < 2) {
                        if (k == n2 - ng - 1) {
                            zp = mw[o] * bct;
                            zp3 = -mw[o - 1] * 0.5 * (Q[-2] + Q[1]) + 2.0 * mw[o] * bct;
                        }
                        if (k == n2 - ng - 2) zp3 = mw[o + 1] * bct;
                        if (k == ng) {
                            zm = mw[o - 1] * bcb;
                            zm3 = -mw[o] * 0.5 * (Q[2] + Q[-1]) + 2.0 * mw[o - 1] * bcb;
                        }
                        if (k == ng + 1) zm3 = mw[o - 2] * bcb;
                    }
                    f[o] = -((A1 * (xp - xm) + A2 * (xp3 - xm3))
                           + (A1 * (yp - ym) + A2 * (yp3 - ym3))
                           + (A1 * (zp - zm) + A2 * (zp3 - zm3))) / den;
                }
            }
        }
}

void orc_flux_quick(const double *u, const double *v, const double *w, const double *c,
                    double *mu, double *mv, double *mw, const orc_grid *g)
{
    /* QUICK face fluxes, scalar.c:113-162 (q = -1/8, upwind by the sign of
     * the face volume flux) */
    int n0 = g->n0, n1 = g->n1, n2 = g->n2;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    long X = (long) n1 * n2, Y = n2;
    const double q = -0.125;
    for (int i = 1; i < n0 - 2; i++)
        for (int j = 1; j < n1 - 2; j++)
            for (int k = 1; k < n2 - 2; k++) {
                size_t o = ID(g, i, j, k);
                const double *C = c + o;
                double fu = dy * dz[k] * u[o];
                double fv = dx * dz[k] * v[o];
                double fw = dx * dy * w[o];
                double cx = 0.5 * (C[0] + C[X]);
                double cy = 0.5 * (C[0] + C[Y]);
                double cz = 0.5 * (C[0] + C[1]);
                cx += fu > 0.0 ? q * (C[X] - 2.0 * C[0] + C[-X]) : q * (C[2 * X] - 2.0 * C[X] + C[0]);
                cy += fv > 0.0 ? q * (C[Y] - 2.0 * C[0] + C[-Y]) : q * (C[2 * Y] - 2.0 * C[Y] + C[0]);
                cz += fw > 0.0 ? q * (C[1] - 2.0 * C[0] + C[-1]) : q * (C[2] - 2.0 * C[1] + C[0]);
                mu[o] = fu * cx;
                mv[o] = fv * cy;
                mw[o] = fw * cz;
            }
}

void orc_advect_scalar(const double *mu, const double *mv, const double *mw,
                       double *f, const orc_grid *g)
{
    /* scalar.c:79-110 */
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    double dx = g->dx, dy = g->dy;
    long X = (long) n1 * n2, Y = n2;
    memset(f, 0, sizeof(double) * (size_t) n0 * n1 * n2);
    for (int i = 1; i < n0; i++)
        for (int j = 1; j < n1; j++)
            for (int k = ng; k < n2 - ng; k++) {
                size_t o = ID(g, i, j, k);
                f[o] = -((mu[o] - mu[o - X]) + (mv[o] - mv[o - Y]) + (mw[o] - mw[o - 1]))
                     / (dx * dy * g->dz[k]);
            }
}

void orc_add_buoyancy(const double *c, double *fu, double *fv, double *fw,
                      double bx, double by, double bz, const orc_grid *g, int order)
{
    /* add_buoyancy2 / 4, scalar.c:6-76 */
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    long X = (long) n1 * n2, Y = n2;
    int lo = order == 2 ? 0 : 1, hi = order == 2 ? 1 : 2;
    for (int i = lo; i < n0 - hi; i++)
        for (int j = lo; j < n1 - hi; j++) {
            for (int k = ng; k < n2 - ng; k++) {
                size_t o = ID(g, i, j, k);
                const double *C = c + o;
                if (order == 2) {
                    fu[o] += 0.5 * (C[0] + C[X]) * bx;
                    fv[o] += 0.5 * (C[0] + C[Y]) * by;
                } else {
                    double cx = B1 * (C[0] + C[X]) + B2 * (C[-X] + C[2 * X]);
                    double cy = B1 * (C[0] + C[Y]) + B2 * (C[-Y] + C[2 * Y]);
                    fu[o] += cx * bx;
                    fv[o] += cy * by;
                }
            }
        }
    for (int i = lo; i < n0 - hi; i++)
        for (int j = lo; j < n1 - hi; j++)
            for (int k = ng; k < n2 - ng - 1; k++) {
                size_t o = ID(g, i, j, k);
                const double *C = c + o;
                if (order == 2) fw[o] += 0.5 * (C[0] + C[1]) * bz;
                else fw[o] += (B1 * (C[0] + C[1]) + B2 * (C[-1] + C[2])) * bz;
            }
}

/* --------------------------------------------------------- projection */

void orc_divergence(const double *u, const double *v, double *w, double *d,
                    const orc_grid *g, int order)
{
    /* divergence2 / 4, pressure.c:40-118 (order 4 mirrors the w ghosts
     * oddly about the wall faces first, in place) */
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    long X = (long) n1 * n2, Y = n2;
    if (order == 2) {
        for (int i = 1; i < n0; i++)
            for (int j = 1; j < n1; j++)
                for (int k = ng; k < n2 - ng; k++) {
                    size_t o = ID(g, i, j, k);
                    d[o] = (u[o] - u[o - X]) / dx + (v[o] - v[o - Y]) / dy + (w[o] - w[o - 1]) / dz[k];
                }
        return;
    }
    for (int i = 0; i < n0; i++)
        for (int j = 0; j < n1; j++) {
            size_t o = ID(g, i, j, 0);
            int k = ng - 1;
            w[o + k - 1] = 2.0 * w[o + k] - w[o + k + 1];
            k = n2 - ng - 1;
            w[o + k + 1] = 2.0 * w[o + k] - w[o + k - 1];
        }
    for (int i = 2; i < n0 - 1; i++)
        for (int j = 2; j < n1 - 1; j++)
            for (int k = ng; k < n2 - ng; k++) {
                size_t o = ID(g, i, j, k);
                d[o] = (A1 * (u[o] - u[o - X]) + A2 * (u[o + X] - u[o - 2 * X])) / dx
                     + (A1 * (v[o] - v[o - Y]) + A2 * (v[o + Y] - v[o - 2 * Y])) / dy
                     + (A1 * (w[o] - w[o - 1]) + A2 * (w[o + 1] - w[o - 2])) / dz[k];
            }
}

void orc_project(double *u, double *v, double *w, const double *p,
                 const orc_grid *g, int order, double dt)
{
    /* project2 / 4, pressure.c:120-195 */
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    long X = (long) n1 * n2, Y = n2;
    int lo = order == 2 ? 0 : 1, hi = order == 2 ? 1 : 2;
    for (int i = lo; i < n0 - hi; i++)
        for (int j = lo; j < n1 - hi; j++)
            for (int k = ng; k < n2 - ng; k++) {
                size_t o = ID(g, i, j, k);
                if (order == 2) {
                    u[o] -= (p[o + X] - p[o]) * dt / dx;
                    v[o] -= (p[o + Y] - p[o]) * dt / dy;
                } else {
                    double up1h = A1 * (p[o + X] - p[o]) + A2 * (p[o + 2 * X] - p[o - X]);
                    double vp1h = A1 * (p[o + Y] - p[o]) + A2 * (p[o + 2 * Y] - p[o - Y]);
                    u[o] -= up1h * dt / dx;
                    v[o] -= vp1h * dt / dy;
                }
            }
    for (int i = lo; i < n0 - hi; i++)
        for (int j = lo; j < n1 - hi; j++)
            for (int k = ng - 1; k < n2 - ng; k++) {
                size_t o = ID(g, i, j, k);
                double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
                if (order == 2) w[o] -= (p[o + 1] - p[o]) * dt / dzp1h;
                else {
                    double wp1h = A1 * (p[o + 1] - p[o]) + A2 * (p[o + 2] - p[o - 1]);
                    w[o] -= wp1h * dt / dzp1h;
                }
            }
}

void orc_pressure_laplacian(const double *p, double *l, const orc_grid *g, int order, double dt)
{
    /* pressure_laplacian2 / 4, pressure.c:197-291 */
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    long X = (long) n1 * n2, Y = n2;
    int lo = order == 2 ? 1 : 3;
    for (int i = lo; i < n0 - lo; i++)
        for (int j = lo; j < n1 - lo; j++)
            for (int k = ng; k < n2 - ng; k++) {
                size_t o = ID(g, i, j, k);
                const double *P = p + o;
                double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
                double dzm1h = 0.5 * (dz[k - 1] + dz[k]);
                if (order == 2) {
                    l[o] = dt * (((P[X] - P[0]) / dx - (P[0] - P[-X]) / dx) / dx
                               + ((P[Y] - P[0]) / dy - (P[0] - P[-Y]) / dy) / dy
                               + ((P[1] - P[0]) / dzp1h - (P[0] - P[-1]) / dzm1h) / dz[k]);
                } else {
                    double up3h = A1 * (P[2 * X] - P[X]) + A2 * (P[3 * X] - P[0]);
                    double up1h = A1 * (P[X] - P[0]) + A2 * (P[2 * X] - P[-X]);
                    double um1h = A1 * (P[0] - P[-X]) + A2 * (P[X] - P[-2 * X]);
                    double um3h = A1 * (P[-X] - P[-2 * X]) + A2 * (P[0] - P[-3 * X]);
                    double vp3h = A1 * (P[2 * Y] - P[Y]) + A2 * (P[3 * Y] - P[0]);
                    double vp1h = A1 * (P[Y] - P[0]) + A2 * (P[2 * Y] - P[-Y]);
                    double vm1h = A1 * (P[0] - P[-Y]) + A2 * (P[Y] - P[-2 * Y]);
                    double vm3h = A1 * (P[-Y] - P[-2 * Y]) + A2 * (P[0] - P[-3 * Y]);
                    double wp3h = A1 * (P[2] - P[1]) + A2 * (P[3] - P[0]);
                    double wp1h = A1 * (P[1] - P[0]) + A2 * (P[2] - P[-1]);
                    double wm1h = A1 * (P[0] - P[-1]) + A2 * (P[1] - P[-2]);
                    double wm3h = A1 * (P[-1] - P[-2]) + A2 * (P[0] - P[-3]);
                    double dzp3h = 0.5 * (dz[k + 2] + dz[k + 1]);
                    double dzm3h = 0.5 * (dz[k - 2] + dz[k - 1]);
                    l[o] = dt * ((A1 * (up1h / dx - um1h / dx) + A2 * (up3h / dx - um3h / dx)) / dx
                               + (A1 * (vp1h / dy - vm1h / dy) + A2 * (vp3h / dy - vm3h / dy)) / dy
                               + (A1 * (wp1h / dzp1h - wm1h / dzm1h) + A2 * (wp3h / dzp3h - wm3h / dzm3h)) / dz[k]);
                }
            }
}

static void helmholtz3d(const double *u, double *h, const orc_grid *g, int order, double nudt, int face)
{
    int ng = g->ng, n0 = g->n0, n1 = g->n1, n2 = g->n2;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    long X = (long) n1 * n2, Y = n2;
    int lo = order == 2 ? 1 : 3, ktop = n2 - ng - (face ? 1 : 0);
    for (int i = lo; i < n0 - lo; i++)
        for (int j = lo; j < n1 - lo; j++)
            for (int k = ng; k < ktop; k++) {
                size_t o = ID(g, i, j, k);
                const double *q = u + o;
                /* z metric: cell family divides by dz_k with centre spacings
                 * inside; the face family the reverse */
                double zin_p = face ? dz[k + 1] : 0.5 * (dz[k + 1] + dz[k]);
                double zin_m = face ? dz[k] : 0.5 * (dz[k - 1] + dz[k]);
                double zout = face ? 0.5 * (dz[k + 1] + dz[k]) : dz[k];
                if (order == 2) {
                    h[o] = q[0] + nudt * (((q[X] - q[0]) / dx - (q[0] - q[-X]) / dx) / dx
                                        + ((q[Y] - q[0]) / dy - (q[0] - q[-Y]) / dy) / dy
                                        + ((q[1] - q[0]) / zin_p - (q[0] - q[-1]) / zin_m) / zout);
                } else {
                    double zin_p3 = face ? dz[k + 2] : 0.5 * (dz[k + 2] + dz[k + 1]);
                    double zin_m3 = face ? dz[k - 1] : 0.5 * (dz[k - 2] + dz[k - 1]);
                    double up3h = A1 * (q[2 * X] - q[X]) + A2 * (q[3 * X] - q[0]);
                    double up1h = A1 * (q[X] - q[0]) + A2 * (q[2 * X] - q[-X]);
                    double um1h = A1 * (q[0] - q[-X]) + A2 * (q[X] - q[-2 * X]);
                    double um3h = A1 * (q[-X] - q[-2 * X]) + A2 * (q[0] - q[-3 * X]);
                    double vp3h = A1 * (q[2 * Y] - q[Y]) + A2 * (q[3 * Y] - q[0]);
                    double vp1h = A1 * (q[Y] - q[0]) + A2 * (q[2 * Y] - q[-Y]);
                    double vm1h = A1 * (q[0] - q[-Y]) + A2 * (q[Y] - q[-2 * Y]);
                    double vm3h = A1 * (q[-Y] - q[-2 * Y]) + A2 * (q[0] - q[-3 * Y]);
                    double wp3h = A1 * (q[2] - q[1]) + A2 * (q[3] - q[0]);
                    double wp1h = A1 * (q[1] - q[0]) + A2 * (q[2] - q[-1]);
                    double wm1h = A1 * (q[0] - q[-1]) + A2 * (q[1] - q[-2]);
                    double wm3h = A1 * (q[-1] - q[-2]) + A2 * (q[0] - q[-3]);
                    h[o] = q[0] + nudt * ((A1 * (up1h / dx - um1h / dx) + A2 * (up3h / dx - um3h / dx)) / dx
                                        + (A1 * (vp1h / dy - vm1h / dy) + A2 * (vp3h / dy - vm3h / dy)) / dy
                                        + (A1 * (wp1h / zin_p - wm1h / zin_m)
                                         + A2 * (wp3h / zin_p3 - wm3h / zin_m3)) / zout);
                }
            }
}

void orc_helmholtz3d_cell(const double *u, double *h, const orc_grid *g, int order, double nudt)
{
    helmholtz3d(u, h, g, order, nudt, 0);                   /* diffuseu.c:362-457 */
}

void orc_helmholtz3d_face(const double *w, double *h, const orc_grid *g, int order, double nudt)
{
    helmholtz3d(w, h, g, order, nudt, 1);                   /* diffusew.c:366-465 */
}

void orc_poisson(double *cout, const orc_grid *g, int order, double dt)
{
    /* pressure_poisson2 / 4, pressure.c:293-444: per (ix, iy) the z-banded
     * discrete Poisson with modified wavenumbers; real and imaginary parts
     * solved separately; mode (0,0) Dirichlet-mean bottom, Neumann top. */
    int ng = g->ng, n2 = g->n2, nz = g->nz, nx = g->nx, ny = g->ny;
    int nkx = nx / 2 + 1;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    bands s;
    if (bands_init(&s, n2)) return;
    for (int k = 0; k < n2; k++) s.x[k] = 0.0;
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            double lamx, lamy;
            if (order == 2) {
                double cosx1 = cos(2.0 * M_PI * 1.0 * (double) ix / (double) nx);
                double cosy1 = cos(2.0 * M_PI * 1.0 * (double) iy / (double) ny);
                lamx = 2.0 - 2.0 * cosx1;
                lamy = 2.0 - 2.0 * cosy1;
            } else {
                double cosx1 = cos(2.0 * M_PI * 1.0 * (double) ix / (double) nx);
                double cosx2 = cos(2.0 * M_PI * 2.0 * (double) ix / (double) nx);
                double cosx3 = cos(2.0 * M_PI * 3.0 * (double) ix / (double) nx);
                double cosy1 = cos(2.0 * M_PI * 1.0 * (double) iy / (double) ny);
                double cosy2 = cos(2.0 * M_PI * 2.0 * (double) iy / (double) ny);
                double cosy3 = cos(2.0 * M_PI * 3.0 * (double) iy / (double) ny);
                lamx = (2.0 * A2 * A2 + 2.0 * A1 * A1) + (4.0 * A1 * A2 - 2.0 * A1 * A1) * cosx1
                     + (-4.0 * A1 * A2) * cosx2 + (-2.0 * A2 * A2) * cosx3;
                lamy = (2.0 * A2 * A2 + 2.0 * A1 * A1) + (4.0 * A1 * A2 - 2.0 * A1 * A1) * cosy1
                     + (-4.0 * A1 * A2) * cosy2 + (-2.0 * A2 * A2) * cosy3;
            }
            double *col = cout + (size_t) (ix * ny + iy) * n2 * 2;
            for (int rc = 0; rc < 2; rc++) {
                cell_bands(&s, dz, ng, n2, order, dt, dx * dy, 0.0);
                for (int k = ng; k < n2 - ng; k++) {
                    double dvol = dx * dy * dz[k];
                    if (order == 2)
                        s.b[k] = s.b[k] - lamx * dvol * dt / (dx * dx) - lamy * dvol * dt / (dy * dy);
                    else
                        s.d[k] = s.d[k] - lamx * dvol * dt / (dx * dx) - lamy * dvol * dt / (dy * dy);
                    s.r[k] = col[2 * k + rc] * dvol;
                }
                int kb = ng - 1, kt = n2 - ng;
                int tb = (ix == 0 && iy == 0) ? ORC_DIRICH_P : ORC_NEUMAN_P;
                if (order == 2) {
                    cell2_rows_bot(&s, dz, kb, 0.0, tb);
                    cell2_rows_top(&s, dz, kt, 0.0, ORC_NEUMAN_P);
                    cell2_reduce_bot(&s, kb);
                    cell2_reduce_top(&s, kt);
                    orc_tridag(s.a + ng, s.b + ng, s.c + ng, s.r + ng, s.x + ng, nz);
                    cell2_fill_bot(&s, kb);
                    cell2_fill_top(&s, kt);
                } else {
                    cell4_rows_bot(&s, dz, kb, 0.0, tb);
                    cell4_rows_top(&s, dz, kt, 0.0, ORC_NEUMAN_P);
                    cell4_reduce_bot(&s, kb);
                    cell4_reduce_top(&s, kt);
                    orc_heptdag(s.a + ng, s.b + ng, s.c + ng, s.d + ng, s.e + ng, s.f + ng,
                                s.g + ng, s.r + ng, s.x + ng, nz);
                    cell4_fill_bot(&s, kb);
                    cell4_fill_top(&s, kt);
                }
                for (int k = 0; k < n2; k++) col[2 * k + rc] = s.x[k];
            }
        }
    bands_free(&s);
}

/* modified wavenumbers of the 2nd- or 4th-order staggered second
 * difference (pressure.c:318-320 and :382-389, diffuseu.c:481-484, :537-550) */
static double mod_wavenumber(int order, int i, int n)
{
    if (order == 2) return 2.0 - 2.0 * cos(2.0 * M_PI * 1.0 * (double) i / (double) n);
    double c1 = cos(2.0 * M_PI * 1.0 * (double) i / (double) n);
    double c2 = cos(2.0 * M_PI * 2.0 * (double) i / (double) n);
    double c3 = cos(2.0 * M_PI * 3.0 * (double) i / (double) n);
    return (2.0 * A2 * A2 + 2.0 * A1 * A1) + (4.0 * A1 * A2 - 2.0 * A1 * A1) * c1
         + (-4.0 * A1 * A2) * c2 + (-2.0 * A2 * A2) * c3;
}

void orc_solve_helmholtz_cell(double *cout, const orc_grid *g, int order, double nudt,
                              double bcb, double bct, int typb, int typt)
{
    /* solve_helmholtz_u2 / _u4, diffuseu.c:460-614: per (ix, iy) and part,
     * (I + nudt (d2/dz2 - lamx/dx2 - lamy/dy2)) x dvol = r dvol with
     * dvol = dx dy dz; the wall values bcb / bct apply to the real part of
     * the (0, 0) mode only, every other system is homogeneous */
    int ng = g->ng, n2 = g->n2, nz = g->nz, nx = g->nx, ny = g->ny, nkx = nx / 2 + 1;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    bands s;
    if (bands_init(&s, n2)) return;
    for (int k = 0; k < n2; k++) s.x[k] = 0.0;
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            double lamx = mod_wavenumber(order, ix, nx), lamy = mod_wavenumber(order, iy, ny);
            double *col = cout + (size_t) (ix * ny + iy) * n2 * 2;
            for (int rc = 0; rc < 2; rc++) {
                cell_bands(&s, dz, ng, n2, order, nudt, dx * dy, 0.0);
                for (int k = ng; k < n2 - ng; k++) {
                    double dvol = dx * dy * dz[k];
                    double *diag = order == 2 ? &s.b[k] : &s.d[k];
                    *diag = *diag - lamx * dvol * nudt / (dx * dx) - lamy * dvol * nudt / (dy * dy) + dvol;
                    s.r[k] = col[2 * k + rc] * dvol;
                }
                int kb = ng - 1, kt = n2 - ng;
                int inh = ix == 0 && iy == 0 && rc == 0;
                double vb = inh ? bcb : 0.0, vt = inh ? bct : 0.0;
                if (order == 2) {
                    cell2_rows_bot(&s, dz, kb, vb, typb);
                    cell2_rows_top(&s, dz, kt, vt, typt);
                    cell2_reduce_bot(&s, kb);
                    cell2_reduce_top(&s, kt);
                    orc_tridag(s.a + ng, s.b + ng, s.c + ng, s.r + ng, s.x + ng, nz);
                    cell2_fill_bot(&s, kb);
                    cell2_fill_top(&s, kt);
                } else {
                    cell4_rows_bot(&s, dz, kb, vb, typb);
                    cell4_rows_top(&s, dz, kt, vt, typt);
                    cell4_reduce_bot(&s, kb);
                    cell4_reduce_top(&s, kt);
                    orc_heptdag(s.a + ng, s.b + ng, s.c + ng, s.d + ng, s.e + ng, s.f + ng,
                                s.g + ng, s.r + ng, s.x + ng, nz);
                    cell4_fill_bot(&s, kb);
                    cell4_fill_top(&s, kt);
                }
                for (int k = 0; k < n2; k++) col[2 * k + rc] = s.x[k];
            }
        }
    bands_free(&s);
}

void orc_solve_helmholtz_face(double *cout, const orc_grid *g, int order, double nudt,
                              double bcb, double bct)
{
    /* solve_helmholtz_w2 / _w4, diffusew.c:466-622: nz - 1 interior faces,
     * dvol = dx dy (dz_k + dz_k+1) / 2 */
    int ng = g->ng, n2 = g->n2, nz = g->nz, nx = g->nx, ny = g->ny, nkx = nx / 2 + 1;
    double dx = g->dx, dy = g->dy;
    const double *dz = g->dz;
    bands s;
    if (bands_init(&s, n2)) return;
    for (int k = 0; k < n2; k++) s.x[k] = 0.0;
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            double lamx = mod_wavenumber(order, ix, nx), lamy = mod_wavenumber(order, iy, ny);
            double *col = cout + (size_t) (ix * ny + iy) * n2 * 2;
            for (int rc = 0; rc < 2; rc++) {
                for (int k = ng; k < n2 - ng - 1; k++) {
                    double dzp1 = dz[k + 1], dzm0 = dz[k];
                    double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
                    double dvol = dx * dy * dzp1h;
                    if (order == 2) {
                        s.a[k] = + dvol * nudt / (dzm0 * dzp1h);
                        s.c[k] = + dvol * nudt / (dzp1 * dzp1h);
                        s.b[k] = - dvol * nudt / (dzm0 * dzp1h)
                                 - dvol * nudt / (dzp1 * dzp1h)
                                 - lamx * dvol * nudt / (dx * dx)
                                 - lamy * dvol * nudt / (dy * dy)
                                 + dvol;
                    } else {
                        double dzp2 = dz[k + 2], dzm1 = dz[k - 1];
                        s.a[k] = + A2 * A2 * dvol * nudt / (dzm1 * dzp1h);
                        s.g[k] = + A2 * A2 * dvol * nudt / (dzp2 * dzp1h);
                        s.b[k] = + A1 * A2 * dvol * nudt / (dzm0 * dzp1h)
                                 + A1 * A2 * dvol * nudt / (dzm1 * dzp1h);
                        s.f[k] = + A1 * A2 * dvol * nudt / (dzp1 * dzp1h)
                                 + A1 * A2 * dvol * nudt / (dzp2 * dzp1h);
                        s.c[k] = - A1 * A2 * dvol * nudt / (dzp1 * dzp1h)
                                 - A1 * A2 * dvol * nudt / (dzm1 * dzp1h)
                                 + A1 * A1 * dvol * nudt / (dzm0 * dzp1h);
                        s.e[k] = - A1 * A2 * dvol * nudt / (dzm0 * dzp1h)
                                 - A1 * A2 * dvol * nudt / (dzp2 * dzp1h)
                                 + A1 * A1 * dvol * nudt / (dzp1 * dzp1h);
                        s.d[k] = - A2 * A2 * dvol * nudt / (dzm1 * dzp1h)
                                 - A2 * A2 * dvol * nudt / (dzp2 * dzp1h)
                                 - A1 * A1 * dvol * nudt / (dzm0 * dzp1h)
                                 - A1 * A1 * dvol * nudt / (dzp1 * dzp1h)
                                 - lamx * dvol * nudt / (dx * dx)
                                 - lamy * dvol * nudt / (dy * dy)
                                 + dvol;
                    }
                    s.r[k] = col[2 * k + rc] * dvol;
                }
                int kb = ng - 1, kt = n2 - ng - 1;
                int inh = ix == 0 && iy == 0 && rc == 0;
                double vb = inh ? bcb : 0.0, vt = inh ? bct : 0.0;
                if (order == 2) {
                    face2_rows(&s, kb, vb);
                    face2_rows(&s, kt, vt);
                    face2_reduce_bot(&s, kb);
                    face2_reduce_top(&s, kt);
                    orc_tridag(s.a + ng, s.b + ng, s.c + ng, s.r + ng, s.x + ng, nz - 1);
                    face2_fill(&s, kb);
                    face2_fill(&s, kt);
                } else {
                    face4_rows_bot(&s, dz, kb, vb);
                    face4_rows_top(&s, dz, kt, vt);
                    face4_reduce_bot(&s, kb);
                    face4_reduce_top(&s, kt);
                    orc_heptdag(s.a + ng, s.b + ng, s.c + ng, s.d + ng, s.e + ng, s.f + ng,
                                s.g + ng, s.r + ng, s.x + ng, nz - 1);
                    face4_fill_bot(&s, kb);
                    face4_fill_top(&s, kt);
                }
                for (int k = 0; k < n2; k++) col[2 * k + rc] = s.x[k];
            }
        }
    bands_free(&s);
}

void orc_fft3d_fwd(const double *a, double *cout, const orc_grid *g)
{
    /* FFT3D_FWD (chanfast.c:21-37): truncate to interior x, y over all n2
     * levels, r2c along x scaled by 1/nx (fft.c:161-177), c2c along y
     * scaled by 1/ny (fft.c:186-202).  Output [ix][iy][k][re,im]. */
    int nx = g->nx, ny = g->ny, n2 = g->n2, ng = g->ng, nkx = nx / 2 + 1;
    double *line = (double *) malloc(sizeof(double) * (size_t) nx);
    double *spec = (double *) malloc(sizeof(double) * 2 * (size_t) nkx);
    double *yl = (double *) malloc(sizeof(double) * 2 * (size_t) ny);
    double sx = 1.0 / (double) nx, sy = 1.0 / (double) ny;
    for (int iy = 0; iy < ny; iy++)
        for (int k = 0; k < n2; k++) {
            for (int ix = 0; ix < nx; ix++) line[ix] = a[ID(g, ix + ng, iy + ng, k)];
            orc_rfft(line, spec, nx);
            for (int ix = 0; ix < nkx; ix++) {
                size_t o = ((size_t) (ix * ny + iy) * n2 + k) * 2;
                cout[o] = spec[2 * ix] * sx;
                cout[o + 1] = spec[2 * ix + 1] * sx;
            }
        }
    for (int ix = 0; ix < nkx; ix++)
        for (int k = 0; k < n2; k++) {
            for (int iy = 0; iy < ny; iy++) {
                size_t o = ((size_t) (ix * ny + iy) * n2 + k) * 2;
                yl[2 * iy] = cout[o]; yl[2 * iy + 1] = cout[o + 1];
            }
            orc_fft(yl, ny, -1);
            for (int iy = 0; iy < ny; iy++) {
                size_t o = ((size_t) (ix * ny + iy) * n2 + k) * 2;
                cout[o] = yl[2 * iy] * sy; cout[o + 1] = yl[2 * iy + 1] * sy;
            }
        }
    free(line); free(spec); free(yl);
}

void orc_fft3d_bwd(const double *cout, double *a, const orc_grid *g)
{
    /* FFT3D_BWD (chanfast.c:39-55): unnormalised inverse y then x, padded
     * back into the interior x, y of all n2 levels. */
    int nx = g->nx, ny = g->ny, n2 = g->n2, ng = g->ng, nkx = nx / 2 + 1;
    double *work = (double *) malloc(sizeof(double) * 2 * (size_t) nkx * ny * n2);
    double *yl = (double *) malloc(sizeof(double) * 2 * (size_t) ny);
    double *spec = (double *) malloc(sizeof(double) * 2 * (size_t) nkx);
    double *line = (double *) malloc(sizeof(double) * (size_t) nx);
    memcpy(work, cout, sizeof(double) * 2 * (size_t) nkx * ny * n2);
    for (int ix = 0; ix < nkx; ix++)
        for (int k = 0; k < n2; k++) {
            for (int iy = 0; iy < ny; iy++) {
                size_t o = ((size_t) (ix * ny + iy) * n2 + k) * 2;
                yl[2 * iy] = work[o]; yl[2 * iy + 1] = work[o + 1];
            }
            orc_fft(yl, ny, +1);
            for (int iy = 0; iy < ny; iy++) {
                size_t o = ((size_t) (ix * ny + iy) * n2 + k) * 2;
                work[o] = yl[2 * iy]; work[o + 1] = yl[2 * iy + 1];
            }
        }
    for (int iy = 0; iy < ny; iy++)
        for (int k = 0; k < n2; k++) {
            for (int ix = 0; ix < nkx; ix++) {
                size_t o = ((size_t) (ix * ny + iy) * n2 + k) * 2;
                spec[2 * ix] = work[o]; spec[2 * ix + 1] = work[o + 1];
            }
            orc_irfft(spec, line, nx);
            for (int ix = 0; ix < nx; ix++) a[ID(g, ix + ng, iy + ng, k)] = line[ix];
        }
    free(work); free(yl); free(spec); free(line);
}

/* -------------------------------------------------------- reductions */

void orc_xy_avg(const double *u, double *avg, const orc_grid *g)
{
    /* get_xy_avg, get.c:281-307: interior x, y; all n2 levels */
    int ng = g->ng, n2 = g->n2;
    for (int k = 0; k < n2; k++) avg[k] = 0.0;
    for (int i = ng; i < g->n0 - ng; i++)
        for (int j = ng; j < g->n1 - ng; j++)
            for (int k = 0; k < n2; k++) avg[k] += u[ID(g, i, j, k)];
    double scale = 1.0 / ((double) g->nx * (double) g->ny);
    for (int k = 0; k < n2; k++) avg[k] *= scale;
}

void orc_xy_corr(const double *u, const double *v, double *corr, const orc_grid *g)
{
    /* get_xy_corr, get.c:136-174: per level (all n2), the plane mean of the
     * product of the deviations from the plane means */
    int ng = g->ng, n2 = g->n2;
    double *ua = (double *) malloc(sizeof(double) * (size_t) n2);
    double *va = (double *) malloc(sizeof(double) * (size_t) n2);
    if (!ua || !va) { free(ua); free(va); return; }
    orc_xy_avg(u, ua, g);
    orc_xy_avg(v, va, g);
    for (int k = 0; k < n2; k++) corr[k] = 0.0;
    for (int i = ng; i < g->n0 - ng; i++)
        for (int j = ng; j < g->n1 - ng; j++)
            for (int k = 0; k < n2; k++) {
                size_t o = ID(g, i, j, k);
                corr[k] += (u[o] - ua[k]) * (v[o] - va[k]);
            }
    double scale = 1.0 / ((double) g->nx * (double) g->ny);
    for (int k = 0; k < n2; k++) corr[k] *= scale;
    free(ua);
    free(va);
}

void orc_xy_moment(const double *u, double *mom, double en, const orc_grid *g)
{
    /* get_xy_moment, get.c:100-133: plane mean of (u - mean)^en per level */
    int ng = g->ng, n2 = g->n2;
    double *ua = (double *) malloc(sizeof(double) * (size_t) n2);
    if (!ua) return;
    orc_xy_avg(u, ua, g);
    for (int k = 0; k < n2; k++) mom[k] = 0.0;
    for (int i = ng; i < g->n0 - ng; i++)
        for (int j = ng; j < g->n1 - ng; j++)
            for (int k = 0; k < n2; k++) mom[k] += pow(u[ID(g, i, j, k)] - ua[k], en);
    double scale = 1.0 / ((double) g->nx * (double) g->ny);
    for (int k = 0; k < n2; k++) mom[k] *= scale;
    free(ua);
}

void orc_xy_spec(const double *uhat, const double *vhat, double *specx, double *specy,
                 int nx, int ny, int n2)
{
    /* get_xy_spec, get.c:5-97 on one rank: the cross-spectrum
     * conj(uhat) vhat ... as re = Re(u)Re(v) + Im(u)Im(v), im = Im(u)Re(v) -
     * Re(u)Im(v); specx[ix][k] sums it over ky (the (0,0) mode left out),
     * specy[iy][k] over kx for 0 <= iy <= ny/2, folding the modes above ny/2
     * onto ny - iy with the imaginary part negated (the kx = 0 column only
     * for 0 < iy <= ny/2, the kx = nx/2 column of an even nx once) */
    int nkx = nx / 2 + 1;
    for (int q = 0; q < nkx * n2 * 2; q++) specx[q] = 0.0;
    for (int q = 0; q < (ny / 2 + 1) * n2 * 2; q++) specy[q] = 0.0;
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < ny; iy++)
            for (int k = 0; k < n2; k++) {
                size_t o = ((size_t) (ix * ny + iy) * n2 + k) * 2;
                double ru = uhat[o], iu = uhat[o + 1], rv = vhat[o], iv = vhat[o + 1];
                double re = ru * rv + iu * iv;
                double im = iu * rv - ru * iv;
                int niy = (ny - iy) % ny;
                size_t xo = ((size_t) ix * n2 + k) * 2, yo = ((size_t) iy * n2 + k) * 2;
                size_t no = ((size_t) niy * n2 + k) * 2;
                if (!(ix == 0 && iy == 0)) { specx[xo] += re; specx[xo + 1] += im; }
                if (ix == 0) {
                    if (0 < iy && iy <= ny / 2) { specy[yo] += re; specy[yo + 1] += im; }
                } else if (ix == nx / 2 && nx % 2 == 0) {
                    if (iy <= ny / 2) { specy[yo] += re; specy[yo + 1] += im; }
                } else {
                    if (iy <= ny / 2) { specy[yo] += re; specy[yo + 1] += im; }
                    if (niy <= ny / 2) { specy[no] += re; specy[no + 1] += -im; }
                }
            }
}

double orc_cfl(const double *u, const double *v, const double *w, const orc_grid *g, double dt)
{
    /* get_cfl, get.c:390-417 */
    int ng = g->ng, n2 = g->n2;
    double cfl = 0.0, c;
    for (int i = ng; i < g->n0 - ng; i++)
        for (int j = ng; j < g->n1 - ng; j++)
            for (int k = ng; k < n2 - ng; k++) {
                size_t o = ID(g, i, j, k);
                double dzph = 0.5 * (g->dz[k] + g->dz[k + 1]);
                if (cfl < (c = fabs(u[o]) * dt / g->dx)) cfl = c;
                if (cfl < (c = fabs(v[o]) * dt / g->dy)) cfl = c;
                if (cfl < (c = fabs(w[o]) * dt / dzph)) cfl = c;
            }
    return cfl;
}
