/*
 * cf_tables.cpp -- host builders of the wall-normal tables of the general
 * path (see cf_tables.h).
 *
 * The reference assembles, per column, the banded z operator plus ghost
 * rows at each wall (diffuseu.c:173-283 with :617-940, diffusew.c:175-287
 * with :625-886, pressure.c:293-444 with :446-770), eliminates the ghost
 * rows (ghost_reduce_*), solves (tridag / heptdag) and recovers the ghosts
 * (ghost_fill_*).  Here the ghost rows of one wall are solved once for the
 * ghosts as affine functions of the nearest interior unknowns; the same
 * closure then substitutes the ghost columns of the interior rows (the
 * Schur complement the reference's row eliminations produce) and fills
 * ghosts wherever a stencil reaches past a wall.
 */
#include "cf_tables.h"

#include <math.h>
#include <string.h>

namespace cft {

static const double A1 = 1.125 / 1.0;      /* fourth.c:3-6 */
static const double A2 = -0.125 / 3.0;

#define DZ(k) dzl[(k)]

struct GRow {
    int lev;
    Row7 row;
    int carries_bc;
};

static void zero_row(Row7 &r) { memset(&r, 0, sizeof(r)); }

/* ghost rows of one wall, in the reference's row definitions */
static int ghost_rows(int fam, int order, int top, int bctype, const double *dzl, int nz, GRow *g)
{
    const int neu = bctype == NEUMANN;
    for (int q = 0; q < 3; q++) { zero_row(g[q].row); g[q].lev = 0; g[q].carries_bc = 0; }
    if (fam == CELL && order == 2) {
        if (!top) {                       /* diffuseu.c:617-636, pressure.c:446-466 */
            double dzp1h = 0.5 * (DZ(0) + DZ(-1));
            g[0].lev = -1;
            if (neu) { g[0].row.c[3] = -1.0 / dzp1h; g[0].row.c[4] = +1.0 / dzp1h; }
            else { g[0].row.c[3] = 0.5; g[0].row.c[4] = 0.5; }
        } else {                          /* diffuseu.c:638-656 */
            double dzm1h = 0.5 * (DZ(nz - 1) + DZ(nz));
            g[0].lev = nz;
            if (neu) { g[0].row.c[2] = -1.0 / dzm1h; g[0].row.c[3] = +1.0 / dzm1h; }
            else { g[0].row.c[2] = 0.5; g[0].row.c[3] = 0.5; }
        }
        g[0].carries_bc = 1;
        return 1;
    }
    if (fam == FACE && order == 2) {      /* diffusew.c:625-648 */
        g[0].lev = top ? nz - 1 : -1;
        g[0].row.c[3] = 1.0;
        g[0].carries_bc = 1;
        return 1;
    }
    if (fam == CELL) {
        if (!top) {                       /* diffuseu.c:703-745, pressure.c:534-576 */
            const int k = -1;
            double dzp3h = 0.5 * (DZ(k + 2) + DZ(k + 1));
            double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
            double dzm1h = 0.5 * (DZ(k - 1) + DZ(k));
            Row7 &r = g[0].row;
            g[0].lev = k;
            r.c[1] = - A2 / dzm1h;
            r.c[2] = - A1 / dzm1h - 2.0 * - A2 / dzp1h;
            r.c[3] = + A1 / dzm1h - 2.0 * - A1 / dzp1h - A2 / dzp3h;
            r.c[4] = + A2 / dzm1h - 2.0 * + A1 / dzp1h - A1 / dzp3h;
            r.c[5] =              - 2.0 * + A2 / dzp1h + A1 / dzp3h;
            r.c[6] =                                   + A2 / dzp3h;
            Row7 &d = g[1].row;
            g[1].lev = k - 1;
            g[1].carries_bc = 1;
            if (neu) { d.c[3] = - A2 / dzp1h; d.c[4] = - A1 / dzp1h; d.c[5] = + A1 / dzp1h; d.c[6] = + A2 / dzp1h; }
            else { d.c[3] = 0.5 * A2; d.c[4] = 0.5 * (A1 + 2.0 * A2); d.c[5] = 0.5 * (A1 + 2.0 * A2); d.c[6] = 0.5 * A2; }
            g[2].lev = k - 2;
            g[2].row.c[5] = -1.0;
            g[2].row.c[6] = +1.0;
        } else {                          /* diffuseu.c:747-802, pressure.c:578-632 */
            const int k = nz;
            double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
            double dzm1h = 0.5 * (DZ(k - 1) + DZ(k));
            double dzm3h = 0.5 * (DZ(k - 2) + DZ(k - 1));
            Row7 &r = g[0].row;
            g[0].lev = k;
            r.c[0] = - A2 / dzm3h;
            r.c[1] = - A1 / dzm3h - 2.0 * - A2 / dzm1h;
            r.c[2] = + A1 / dzm3h - 2.0 * - A1 / dzm1h - A2 / dzp1h;
            r.c[3] = + A2 / dzm3h - 2.0 * + A1 / dzm1h - A1 / dzp1h;
            r.c[4] =              - 2.0 * + A2 / dzm1h + A1 / dzp1h;
            r.c[5] =                                   + A2 / dzp1h;
            Row7 &d = g[1].row;
            g[1].lev = k + 1;
            g[1].carries_bc = 1;
            if (neu) { d.c[0] = - A2 / dzm1h; d.c[1] = - A1 / dzm1h; d.c[2] = + A1 / dzm1h; d.c[3] = + A2 / dzm1h; }
            else { d.c[0] = 0.5 * A2; d.c[1] = 0.5 * (A1 + 2.0 * A2); d.c[2] = 0.5 * (A1 + 2.0 * A2); d.c[3] = 0.5 * A2; }
            g[2].lev = k + 2;
            g[2].row.c[0] = -1.0;
            g[2].row.c[1] = +1.0;
        }
        return 3;
    }
    /* faces, fourth order: diffusew.c:691-766 */
    const int k = top ? nz - 1 : -1;
    double dzp0 = DZ(k), dzp1 = DZ(k + 1);
    Row7 &r = g[0].row;
    g[0].lev = k;
    r.c[1] = - A2 / dzp0;
    r.c[2] = - A1 / dzp0 + A2 / dzp1;
    r.c[3] = + A1 / dzp0 + A1 / dzp1;
    r.c[4] = + A2 / dzp0 - A1 / dzp1;
    r.c[5] =             - A2 / dzp1;
    if (!top) {
        g[1].lev = k - 1;
        g[1].row.c[3] = -1.0; g[1].row.c[4] = +2.0; g[1].row.c[5] = -1.0;
        g[2].lev = k - 2;
        g[2].row.c[5] = 1.0;
    } else {
        g[1].lev = k + 1;
        g[1].row.c[1] = -1.0; g[1].row.c[2] = +2.0; g[1].row.c[3] = -1.0;
        g[2].lev = k + 2;
        g[2].row.c[1] = 1.0;
    }
    g[2].carries_bc = 1;
    return 3;
}

static int invert3(double A[3][3], int m, double Ai[3][3])
{
    double M[3][6];
    for (int i = 0; i < m; i++)
        for (int j = 0; j < 2 * m; j++) M[i][j] = j < m ? A[i][j] : (j - m == i ? 1.0 : 0.0);
    for (int c = 0; c < m; c++) {
        int piv = c;
        for (int r = c + 1; r < m; r++) if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
        if (M[piv][c] == 0.0) return -1;
        if (piv != c) for (int j = 0; j < 2 * m; j++) { double t = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = t; }
        double inv = 1.0 / M[c][c];
        for (int j = 0; j < 2 * m; j++) M[c][j] *= inv;
        for (int r = 0; r < m; r++) {
            if (r == c || M[r][c] == 0.0) continue;
            double f = M[r][c];
            for (int j = 0; j < 2 * m; j++) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) Ai[i][j] = M[i][m + j];
    return 0;
}

int closure(int fam, int order, int top, int bctype, const double *dzl, int nz, Clos1 &out)
{
    memset(&out, 0, sizeof(out));
    const int n = fam == CELL ? nz : nz - 1;
    GRow g[3];
    const int ng = ghost_rows(fam, order, top, bctype, dzl, nz, g);
    double Agg[3][3] = {{0}}, Agx[3][3] = {{0}}, rb[3] = {0};
    for (int q = 0; q < ng; q++) {
        for (int t = 0; t < 7; t++) {
            double cv = g[q].row.c[t];
            if (cv == 0.0) continue;
            int L = g[q].lev - 3 + t;
            int m = !top ? -1 - L : L - n;            /* ghost index, < 0 for an interior unknown */
            if (m >= 0) {
                if (m >= ng) return -1;
                Agg[q][m] += cv;
            } else {
                int tt = !top ? L : n - 1 - L;        /* interior index counted from the wall */
                if (tt < 0 || tt > 2) return -1;
                Agx[q][tt] += cv;
            }
        }
        rb[q] = g[q].carries_bc ? 1.0 : 0.0;
    }
    double Ai[3][3];
    if (invert3(Agg, ng, Ai)) return -1;
    for (int m = 0; m < ng; m++) {
        for (int t = 0; t < 3; t++) {
            double acc = 0.0;
            for (int q = 0; q < ng; q++) acc += Ai[m][q] * Agx[q][t];
            out.G[m][t] = -acc;
        }
        double b = 0.0;
        for (int q = 0; q < ng; q++) b += Ai[m][q] * rb[q];
        out.B[m] = b;
    }
    return 0;
}

void op_rows(int fam, int order, const double *dzl, int nz, double nudt, double dxy, bool with_dvol,
             std::vector<Row7> &rows)
{
    const int n = fam == CELL ? nz : nz - 1;
    rows.assign(n, Row7());
    for (int k = 0; k < n; k++) {
        Row7 &r = rows[k];
        zero_row(r);
        if (fam == CELL) {
            double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
            double dzm1h = 0.5 * (DZ(k - 1) + DZ(k));
            double dvol = dxy == 1.0 ? DZ(k) : dxy * DZ(k);
            if (order == 2) {
                r.c[2] = + dvol * nudt / (dzm1h * DZ(k));
                r.c[4] = + dvol * nudt / (dzp1h * DZ(k));
                r.c[3] = - dvol * nudt / (dzm1h * DZ(k))
                         - dvol * nudt / (dzp1h * DZ(k));
                if (with_dvol) r.c[3] = r.c[3] + dvol;
            } else {
                double dzp3h = 0.5 * (DZ(k + 2) + DZ(k + 1));
                double dzm3h = 0.5 * (DZ(k - 2) + DZ(k - 1));
                r.c[0] = + A2 * A2 * dvol * nudt / (dzm3h * DZ(k));
                r.c[6] = + A2 * A2 * dvol * nudt / (dzp3h * DZ(k));
                r.c[1] = + A1 * A2 * dvol * nudt / (dzm1h * DZ(k))
                         + A1 * A2 * dvol * nudt / (dzm3h * DZ(k));
                r.c[5] = + A1 * A2 * dvol * nudt / (dzp1h * DZ(k))
                         + A1 * A2 * dvol * nudt / (dzp3h * DZ(k));
                r.c[2] = - A1 * A2 * dvol * nudt / (dzp1h * DZ(k))
                         - A1 * A2 * dvol * nudt / (dzm3h * DZ(k))
                         + A1 * A1 * dvol * nudt / (dzm1h * DZ(k));
                r.c[4] = - A1 * A2 * dvol * nudt / (dzm1h * DZ(k))
                         - A1 * A2 * dvol * nudt / (dzp3h * DZ(k))
                         + A1 * A1 * dvol * nudt / (dzp1h * DZ(k));
                r.c[3] = - A2 * A2 * dvol * nudt / (dzm3h * DZ(k))
                         - A2 * A2 * dvol * nudt / (dzp3h * DZ(k))
                         - A1 * A1 * dvol * nudt / (dzm1h * DZ(k))
                         - A1 * A1 * dvol * nudt / (dzp1h * DZ(k));
                if (with_dvol) r.c[3] = r.c[3] + dvol;
            }
        } else {
            double dzp1 = DZ(k + 1), dzm0 = DZ(k);
            double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
            double dvol = dzp1h;
            if (order == 2) {
                r.c[2] = + dvol * nudt / (dzm0 * dzp1h);
                r.c[4] = + dvol * nudt / (dzp1 * dzp1h);
                r.c[3] = - dvol * nudt / (dzm0 * dzp1h)
                         - dvol * nudt / (dzp1 * dzp1h)
                         + dvol;
            } else {
                double dzp2 = DZ(k + 2), dzm1 = DZ(k - 1);
                r.c[0] = + A2 * A2 * dvol * nudt / (dzm1 * dzp1h);
                r.c[6] = + A2 * A2 * dvol * nudt / (dzp2 * dzp1h);
                r.c[1] = + A1 * A2 * dvol * nudt / (dzm0 * dzp1h)
                         + A1 * A2 * dvol * nudt / (dzm1 * dzp1h);
                r.c[5] = + A1 * A2 * dvol * nudt / (dzp1 * dzp1h)
                         + A1 * A2 * dvol * nudt / (dzp2 * dzp1h);
                r.c[2] = - A1 * A2 * dvol * nudt / (dzp1 * dzp1h)
                         - A1 * A2 * dvol * nudt / (dzm1 * dzp1h)
                         + A1 * A1 * dvol * nudt / (dzm0 * dzp1h);
                r.c[4] = - A1 * A2 * dvol * nudt / (dzm0 * dzp1h)
                         - A1 * A2 * dvol * nudt / (dzp2 * dzp1h)
                         + A1 * A1 * dvol * nudt / (dzp1 * dzp1h);
                r.c[3] = - A2 * A2 * dvol * nudt / (dzm1 * dzp1h)
                         - A2 * A2 * dvol * nudt / (dzp2 * dzp1h)
                         - A1 * A1 * dvol * nudt / (dzm0 * dzp1h)
                         - A1 * A1 * dvol * nudt / (dzp1 * dzp1h)
                         + dvol;
            }
        }
    }
}

void eliminate(int fam, int nz, const Clos1 &bot, const Clos1 &top, std::vector<Row7> &rows)
{
    const int n = fam == CELL ? nz : nz - 1;
    (void) n;
    const int nr = (int) rows.size();
    for (int i = 0; i < nr; i++) {
        Row7 &r = rows[i];
        for (int t = 0; t < 7; t++) {
            const int L = i - 3 + t;
            if (L >= 0 && L < nr) continue;
            const double cv = r.c[t];
            r.c[t] = 0.0;
            if (cv == 0.0) continue;
            const bool lo = L < 0;
            const int m = lo ? -1 - L : L - nr;
            const Clos1 &C = lo ? bot : top;
            for (int tt = 0; tt < 3; tt++) {
                if (C.G[m][tt] == 0.0) continue;
                const int lev = lo ? tt : nr - 1 - tt;
                const int tp = lev - i + 3;
                if (tp >= 0 && tp < 7) r.c[tp] += cv * C.G[m][tt];
            }
        }
    }
}

int hept_factor(const std::vector<Row7> &rows, std::vector<double> &tab)
{
    const int n = (int) rows.size();
    tab.assign((size_t) 7 * n, 0.0);
    double *A = &tab[0], *AL = A + n, *BE = AL + n, *ID = BE + n, *P = ID + n, *Q = P + n, *R = Q + n;
    auto p_ = [&](int i) { return i >= 0 ? P[i] : 0.0; };
    auto q_ = [&](int i) { return i >= 0 ? Q[i] : 0.0; };
    auto r_ = [&](int i) { return i >= 0 ? R[i] : 0.0; };
    for (int i = 0; i < n; i++) {
        const double *c = rows[i].c;
        double a = i >= 3 ? c[0] : 0.0, b = i >= 2 ? c[1] : 0.0, cc = i >= 1 ? c[2] : 0.0;
        double d = c[3], e = c[4], f = c[5], g = c[6];
        double alp = b - a * p_(i - 3);
        double bet = cc - a * q_(i - 3) - alp * p_(i - 2);
        double den = d - a * r_(i - 3) - alp * q_(i - 2) - bet * p_(i - 1);
        if (den == 0.0) return -1;
        P[i] = (e - alp * r_(i - 2) - bet * q_(i - 1)) / den;
        Q[i] = (f - bet * r_(i - 1)) / den;
        R[i] = g / den;
        A[i] = a; AL[i] = alp; BE[i] = bet; ID[i] = 1.0 / den;
    }
    return 0;
}

void hept_solve(const std::vector<double> &tab, int n, const std::vector<double> &rhs, std::vector<double> &x)
{
    const double *A = &tab[0], *AL = A + n, *BE = AL + n, *ID = BE + n, *P = ID + n, *Q = P + n, *R = Q + n;
    x.assign(n, 0.0);
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int i = 0; i < n; i++) {
        double si = (rhs[i] - A[i] * s3 - AL[i] * s2 - BE[i] * s1) * ID[i];
        x[i] = si;
        s3 = s2; s2 = s1; s1 = si;
    }
    double x1 = 0.0, x2 = 0.0, x3 = 0.0;
    for (int i = n - 1; i >= 0; i--) {
        double xi = x[i] - P[i] * x1 - Q[i] * x2 - R[i] * x3;
        x[i] = xi;
        x3 = x2; x2 = x1; x1 = xi;
    }
}

void helm_apply(int fam, int order, const double *dzl, int nz, double nudt, const double *col, double *out)
{
    const int n = fam == CELL ? nz : nz - 1;
    const double *q = col + 3;
    for (int k = 0; k < n; k++) {
        if (fam == CELL) {
            double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
            double dzm1h = 0.5 * (DZ(k - 1) + DZ(k));
            if (order == 2) {
                out[k] = q[k] + nudt * (((q[k + 1] - q[k]) / dzp1h - (q[k] - q[k - 1]) / dzm1h) / DZ(k));
            } else {
                double wp3h = A1 * (q[k + 2] - q[k + 1]) + A2 * (q[k + 3] - q[k]);
                double wp1h = A1 * (q[k + 1] - q[k]) + A2 * (q[k + 2] - q[k - 1]);
                double wm1h = A1 * (q[k] - q[k - 1]) + A2 * (q[k + 1] - q[k - 2]);
                double wm3h = A1 * (q[k - 1] - q[k - 2]) + A2 * (q[k] - q[k - 3]);
                double dzp3h = 0.5 * (DZ(k + 2) + DZ(k + 1));
                double dzm3h = 0.5 * (DZ(k - 2) + DZ(k - 1));
                out[k] = q[k] + nudt * ((A1 * (wp1h / dzp1h - wm1h / dzm1h)
                                       + A2 * (wp3h / dzp3h - wm3h / dzm3h)) / DZ(k));
            }
        } else {
            double dzp1 = DZ(k + 1), dzm0 = DZ(k);
            double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
            if (order == 2) {
                out[k] = q[k] + nudt * (((q[k + 1] - q[k]) / dzp1 - (q[k] - q[k - 1]) / dzm0) / dzp1h);
            } else {
                double wp3h = A1 * (q[k + 2] - q[k + 1]) + A2 * (q[k + 3] - q[k]);
                double wp1h = A1 * (q[k + 1] - q[k]) + A2 * (q[k + 2] - q[k - 1]);
                double wm1h = A1 * (q[k] - q[k - 1]) + A2 * (q[k + 1] - q[k - 2]);
                double wm3h = A1 * (q[k - 1] - q[k - 2]) + A2 * (q[k] - q[k - 3]);
                double dzp2 = DZ(k + 2), dzm1 = DZ(k - 1);
                out[k] = q[k] + nudt * ((A1 * (wp1h / dzp1 - wm1h / dzm0)
                                       + A2 * (wp3h / dzp2 - wm3h / dzm1)) / dzp1h);
            }
        }
    }
}

}   /* namespace cft */
