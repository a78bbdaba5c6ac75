/*
 * cf_run.hip -- host runtime of the chanflow library: run handles,
 * device memory, the column-independent coefficient tables, the
 * hipGraph-captured RK3 step and the C ABI of include/chanflow.h.
 *
 * Replaces the chanfast2 driver's setup and time loop (chanfast.c:71-656)
 * and its MPI/FFTW pencil machinery with one GPU's worth of kernels.
 */
#include "chanflow.h"
#include "cf_internal_host.h"
#include "cf_kernels.h"

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <map>
#include <string>
#include <vector>

/* ------------------------------------------------------------ errors */

static thread_local char g_err[512] = "";

extern "C" void cf_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *cf_last_error(void) { return g_err; }
extern "C" const char *cf_version(void) { return "chanflow-mi355x 0.2 (gfx950, FP64, order 2 and 4, passive scalar)"; }

#define HIPCHK(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            cf_set_error("%s:%d: %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));   \
            return CF_ERR_DEVICE;                                                           \
        }                                                                                   \
    } while (0)

#define LAUNCHCHK() HIPCHK(hipGetLastError())

/* RK3 tables, rk3.c:3-6 */
static const double RK_ALP[3] = { 29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0 };
static const double RK_BET[3] = { 37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0 };
static const double RK_GAM[3] = { 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0 };
static const double RK_ZET[3] = { 0.0, -17.0 / 60.0, -5.0 / 12.0 };

static const int LOGCAP = 4096;
static const int MAXE_LINE = 2048;      /* elements per FFT block (8 per thread x 256) */
#ifndef CF_P4_THREADS
#define CF_P4_THREADS 64                /* order-4 Poisson: modes (threads) per workgroup of a single run */
#endif

/* ------------------------------------------------------------ run */

struct PoisHost {
    std::vector<double> a, c, bz, ex, ey, dvol;
    double db_neu = 0, db_dir = 0, db_top = 0;
};

struct cf_run {
    cf_config cfg;
    Grid g;
    int nz_tab;
    std::vector<double> dz;          /* nz + 6, level k at k + 3 */
    size_t N, P, M;                  /* cells, plane, spectral modes per level */
    hipStream_t stream = nullptr;
    /* device buffers */
    double *u = nullptr, *v = nullptr, *w = nullptr, *c = nullptr;
    double *ru = nullptr, *rv = nullptr, *rw = nullptr, *rc = nullptr;
    double *fa[4] = {nullptr, nullptr, nullptr, nullptr}, *fb[4] = {nullptr, nullptr, nullptr, nullptr};
    double *S4 = nullptr;            /* order-4 Poisson U factors (3 per spectral value) */
    double *p = nullptr, *G = nullptr, *stage = nullptr;
    double2 *X = nullptr;
    double *tab = nullptr;           /* all constant tables in one allocation */
    double *scal = nullptr, *logbuf = nullptr, *part = nullptr, *lsum = nullptr;
    size_t tab_n = 0;
    int nb_plane = 1, nb_red = 1;
    RunDev rd;
    PoisTab pt[4];                   /* 3 substeps + ad-hoc dt */
    PoisTab hs2[3][3];               /* fully implicit driver: spectral Helmholtz tables [s][type], order 2 */
    Pois4Tab hs4[3][3];              /* ... order 4 */
    Pois4Tab pt4[4];
    size_t off_pois4_adhoc = 0;
    /* general path: order 4 and/or the scalar (cf_general.hip) */
    bool general = false;
    int nfield = 3;
    int prec = CF_FP64;
    size_t esz = sizeof(double);     /* bytes per stored value of the state and work fields */
    GSel<double> g64;
    GSel<float> g32;
    int gblocks = 0;
    double pt_dt = -1.0;
    size_t off_pois_adhoc = 0;
    /* graphs: one RK3 step starting from state set 0 or 1 */
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int cur = 0;                     /* state set holding the current solve output */
    bool proj_pending = false;       /* p / forcings not yet folded into the state */
    double *cflpart = nullptr;
    int rhs_blocks = 0, rhs_kc = 1;
    int r4_nch = 0, r4_blocks = 0;   /* order 4: z-chunks and blocks of k_rhs4_uv / _wc (0: per-field kernels) */
    int steps_done = 0;
    bool use_spike = false;
    int spike_waves = 1, sp_L = 0;              /* chunks and rows per chunk of the z solves */
    bool use_spike4 = false;                    /* order 4: partitioned z solves (k_spike4) */
    int s4_L = 0, s4_waves = 1;
    KSel<double>::div_t kdiv = nullptr;
    k_fftx_plain_t kfxp = nullptr;
    KSel<double>::ffty_t kfy = nullptr;
    KSel<double>::fftxi_t kfxi = nullptr;
    KSel<float>::div_t kdiv32 = nullptr;
    KSel<float>::ffty_t kfy32 = nullptr;
    KSel<float>::fftxi_t kfxi32 = nullptr;
    /* cflmode 1: the host copy of the tables being uploaded, and the
     * projection left pending by the last step, with that step's dt */
    std::vector<double> tabhost;
    double *h1prev = nullptr;
    /* FP32 fused path: u stored relative to the profile ubar (device copy
     * with ghosts and its right-hand-side shares in uoffd) */
    std::vector<double> ubar;
    double *uoffd = nullptr;
    double dtp_prev = 0.0;
    bool prev_proj = false;
    /* profiling */
    bool profile = false;
    std::map<std::string, std::pair<double, int>> ktime;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

/* ------------------------------------------------------------ host tables */

static int factor_radix(int n, int *rad)
{
    int nr = 0, m = n;
    while (m % 4 == 0 && m > 4) { rad[nr++] = 4; m /= 4; }
    if (m == 4) { rad[nr++] = 4; m = 1; }
    while (m % 2 == 0) { rad[nr++] = 2; m /= 2; }
    while (m % 3 == 0) { rad[nr++] = 3; m /= 3; }
    while (m % 5 == 0) { rad[nr++] = 5; m /= 5; }
    if (m != 1 || nr > CF_MAX_RAD) return -1;
    return nr;
}

/* dz at level k (k in [-3, nz + 2]) */
#define DZ(k) dzl[(k)]

/* Thomas factors of the wall-reduced 1-D Helmholtz system of a
 * cell-centred field, 2nd order: diffuseu.c:192-210 with the Dirichlet
 * ghost rows of :617-682 eliminated as ghost_reduce_*2 does. */
static void cell2_system(const double *dzl, int nz, double nudt, std::vector<double> &a,
                         std::vector<double> &b, std::vector<double> &c)
{
    a.assign(nz, 0.0); b.assign(nz, 0.0); c.assign(nz, 0.0);
    for (int k = 0; k < nz; k++) {
        double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
        double dzm1h = 0.5 * (DZ(k - 1) + DZ(k));
        double dvol = DZ(k);
        a[k] = +dvol * nudt / (dzm1h * DZ(k));
        c[k] = +dvol * nudt / (dzp1h * DZ(k));
        b[k] = -dvol * nudt / (dzm1h * DZ(k)) - dvol * nudt / (dzp1h * DZ(k)) + dvol;
    }
    double fac = a[0] / 0.5;            /* bottom ghost row a=0, b=0.5, c=0.5 */
    a[0] -= 0.5 * fac;
    b[0] -= 0.5 * fac;
    fac = c[nz - 1] / 0.5;              /* top ghost row a=0.5, b=0.5, c=0 */
    c[nz - 1] -= 0.5 * fac;
    b[nz - 1] -= 0.5 * fac;
}

/* face-centred w, 2nd order: diffusew.c:198-214 (nz-1 rows), ghost rows b=1 */
static void face2_system(const double *dzl, int nz, double nudt, std::vector<double> &a,
                         std::vector<double> &b, std::vector<double> &c)
{
    int n = nz - 1;
    a.assign(n, 0.0); b.assign(n, 0.0); c.assign(n, 0.0);
    for (int k = 0; k < n; k++) {
        double dzp1 = DZ(k + 1), dzm0 = DZ(k);
        double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
        double dvol = dzp1h;
        a[k] = +dvol * nudt / (dzm0 * dzp1h);
        c[k] = +dvol * nudt / (dzp1 * dzp1h);
        b[k] = -dvol * nudt / (dzm0 * dzp1h) - dvol * nudt / (dzp1 * dzp1h) + dvol;
    }
    double fac = a[0] / 1.0;
    a[0] -= 1.0 * fac;
    fac = c[n - 1] / 1.0;
    c[n - 1] -= 1.0 * fac;
}

/* tridag.c:4-26 factors: ibet_j = 1 / bet_j, gam_j = c_{j-1} / bet_{j-1} */
static int thomas_factors(const std::vector<double> &a, const std::vector<double> &b,
                          const std::vector<double> &c, double *A, double *IB, double *GM)
{
    int n = (int) b.size();
    double bet = b[0];
    if (bet == 0.0) return -1;
    A[0] = a[0]; IB[0] = 1.0 / bet; GM[0] = 0.0;
    for (int j = 1; j < n; j++) {
        double gam = c[j - 1] / bet;
        bet = b[j] - a[j] * gam;
        if (bet == 0.0) return -1;
        A[j] = a[j]; IB[j] = 1.0 / bet; GM[j] = gam;
    }
    return 0;
}

/* Partitioned (spike) solve tables for one tridiagonal system (rows
 * a_j x_{j-1} + b_j x_j + c_j x_{j+1} = r_j, a_0 = c_{n-1} = 0): chunks of
 * L rows, chunk-local Thomas factors, spikes v = B^-1 a_lo e_lo and
 * w = B^-1 c_{hi-1} e_{hi-1}, and the inverse of the interface system
 * z + K z = y_int over z = (F_0, E_0, F_1, E_1, ...), F/E the first/last
 * unknown of each chunk. */
struct SpikeHost {
    int n = 0, S = 0, L = 0;
    std::vector<double> a, ib, gam, vl, wr, minv;
};

static int invert_dense(std::vector<double> &M, int m)
{
    std::vector<double> A(M), I((size_t) m * m, 0.0);
    for (int i = 0; i < m; i++) I[(size_t) i * m + i] = 1.0;
    for (int c = 0; c < m; c++) {
        int piv = c;
        for (int r = c + 1; r < m; r++)
            if (fabs(A[(size_t) r * m + c]) > fabs(A[(size_t) piv * m + c])) piv = r;
        if (A[(size_t) piv * m + c] == 0.0) return -1;
        if (piv != c)
            for (int q = 0; q < m; q++) {
                std::swap(A[(size_t) c * m + q], A[(size_t) piv * m + q]);
                std::swap(I[(size_t) c * m + q], I[(size_t) piv * m + q]);
            }
        double inv = 1.0 / A[(size_t) c * m + c];
        for (int q = 0; q < m; q++) { A[(size_t) c * m + q] *= inv; I[(size_t) c * m + q] *= inv; }
        for (int r = 0; r < m; r++) {
            if (r == c) continue;
            double f = A[(size_t) r * m + c];
            if (f == 0.0) continue;
            for (int q = 0; q < m; q++) { A[(size_t) r * m + q] -= f * A[(size_t) c * m + q]; I[(size_t) r * m + q] -= f * I[(size_t) c * m + q]; }
        }
    }
    M = I;
    return 0;
}

/* L rows per chunk (spike_L of the run's nz, shared by its u, v and w
 * systems so one kernel instantiation serves all three) */
static int spike_build(const std::vector<double> &a, const std::vector<double> &b,
                       const std::vector<double> &c, int Lrun, SpikeHost &T)
{
    const int n = (int) b.size();
    T.n = n;
    if (Lrun <= 0) return -1;
    T.L = Lrun;
    T.S = (n + T.L - 1) / T.L;
    if (T.S > CF_SPIKE_SMAX) return -1;
    T.a = a; T.ib.assign(n, 0.0); T.gam.assign(n, 0.0); T.vl.assign(n, 0.0); T.wr.assign(n, 0.0);
    const int S = T.S, L = T.L, m = 2 * S;
    std::vector<double> M((size_t) m * m, 0.0);
    for (int i = 0; i < m; i++) M[(size_t) i * m + i] = 1.0;
    for (int s = 0; s < S; s++) {
        int lo = s * L, hi = std::min(lo + L, n);
        double bet = b[lo];
        if (bet == 0.0) return -2;
        T.ib[lo] = 1.0 / bet;
        for (int j = lo + 1; j < hi; j++) {
            T.gam[j] = c[j - 1] / bet;
            bet = b[j] - a[j] * T.gam[j];
            if (bet == 0.0) return -2;
            T.ib[j] = 1.0 / bet;
        }
        auto solve = [&](std::vector<double> &rhs) {   /* in place on [lo, hi) */
            double d = 0.0;
            for (int j = lo; j < hi; j++) { d = (rhs[j] - (j > lo ? a[j] * d : 0.0)) * T.ib[j]; rhs[j] = d; }
            for (int j = hi - 2; j >= lo; j--) { d = rhs[j] - T.gam[j + 1] * d; rhs[j] = d; }
        };
        std::vector<double> v(n, 0.0), w(n, 0.0);
        if (s > 0) { v[lo] = a[lo]; solve(v); }
        if (s < S - 1) { w[hi - 1] = c[hi - 1]; solve(w); }
        for (int j = lo; j < hi; j++) { T.vl[j] = v[j]; T.wr[j] = w[j]; }
        int rf = 2 * s, re = 2 * s + 1;
        if (s > 0) { M[(size_t) rf * m + 2 * (s - 1) + 1] += v[lo]; M[(size_t) re * m + 2 * (s - 1) + 1] += v[hi - 1]; }
        if (s < S - 1) { M[(size_t) rf * m + 2 * (s + 1)] += w[lo]; M[(size_t) re * m + 2 * (s + 1)] += w[hi - 1]; }
    }
    if (invert_dense(M, m)) return -3;
    /* device layout: chunk-major blocks of L rows, identity rows (a = 0,
     * ib = 1, gam = vl = wr = 0) past the last chunk's end, so the kernel's
     * sweeps run the same straight-line code for every chunk; minv rows
     * padded to 2 CF_SPIKE_SMAX columns */
    auto padv = [&](const std::vector<double> &v, double fill) {
        std::vector<double> o((size_t) S * L, fill);
        for (int c = 0; c < S; c++)
            for (int t = 0; t < L && c * L + t < n; t++) o[(size_t) c * L + t] = v[(size_t) c * L + t];
        return o;
    };
    T.a = padv(T.a, 0.0);
    T.ib = padv(T.ib, 1.0);
    T.gam = padv(T.gam, 0.0);
    T.vl = padv(T.vl, 0.0);
    T.wr = padv(T.wr, 0.0);
    const int mp = 2 * CF_SPIKE_SMAX;
    T.minv.assign((size_t) m * mp, 0.0);
    for (int r = 0; r < m; r++)
        for (int q = 0; q < m; q++) T.minv[(size_t) r * mp + q] = M[(size_t) r * m + q];
    return 0;
}

/* Partitioned (spike) tables of one wall-reduced heptadiagonal system
 * (rows sum_d c[d + 3] x_{i+d} = r_i, d = -3..3): chunks of L rows with
 * their own heptdag factors (couplings that leave the chunk dropped), the
 * left spikes V_c = A^-1 (column c of the coupling of the chunk's first
 * three rows to the previous chunk's last three unknowns), the right
 * spikes W_c (last three rows to the next chunk's first three), and the
 * inverse of the interface system X + K X = Y over X = (first three, last
 * three unknowns) of every chunk.  Device layout: chunk-major blocks of L
 * rows, identity rows past the system's end; minv padded to
 * 6 CF_SPIKE4_SMAX columns. */
struct Spike4Host {
    int n = 0, S = 0, L = 0;
    bool ok = false;                 /* reproduces the sequential solve (checked at build) */
    std::vector<double> f, vw, minv;
};

static int spike4_build(const std::vector<cft::Row7> &rows, int L, Spike4Host &T)
{
    const int n = (int) rows.size();
    if (L <= 0) return -1;
    const int S = (n + L - 1) / L;
    if (S > CF_SPIKE4_SMAX || (S > 1 && n - (S - 1) * L < 3)) return -1;
    T.n = n; T.S = S; T.L = L;
    const size_t SL = (size_t) S * L;
    T.f.assign(7 * SL, 0.0);
    T.vw.assign(6 * SL, 0.0);
    for (size_t i = 0; i < SL; i++) T.f[3 * SL + i] = 1.0;        /* identity rows: iden = 1 */
    const int m6 = 6 * S;
    std::vector<double> M((size_t) m6 * m6, 0.0);
    for (int i = 0; i < m6; i++) M[(size_t) i * m6 + i] = 1.0;
    for (int c = 0; c < S; c++) {
        const int lo = c * L, hi = std::min(lo + L, n), m = hi - lo;
        double *A = &T.f[(size_t) c * L], *AL = A + SL, *BE = A + 2 * SL, *ID = A + 3 * SL;
        double *Pp = A + 4 * SL, *Qq = A + 5 * SL, *Rr = A + 6 * SL;
        auto p_ = [&](int i) { return i >= 0 ? Pp[i] : 0.0; };
        auto q_ = [&](int i) { return i >= 0 ? Qq[i] : 0.0; };
        auto r_ = [&](int i) { return i >= 0 ? Rr[i] : 0.0; };
        for (int i = 0; i < m; i++) {                 /* heptdag.c:5-61 on the chunk alone */
            const double *cf = rows[lo + i].c;
            const double a = i >= 3 ? cf[0] : 0.0, b = i >= 2 ? cf[1] : 0.0, cc = i >= 1 ? cf[2] : 0.0;
            const double d = cf[3], e = i + 1 < m ? cf[4] : 0.0, ff = i + 2 < m ? cf[5] : 0.0, g = i + 3 < m ? cf[6] : 0.0;
            const double alp = b - a * p_(i - 3);
            const double bet = cc - a * q_(i - 3) - alp * p_(i - 2);
            const double den = d - a * r_(i - 3) - alp * q_(i - 2) - bet * p_(i - 1);
            if (den == 0.0) return -2;
            Pp[i] = (e - alp * r_(i - 2) - bet * q_(i - 1)) / den;
            Qq[i] = (ff - bet * r_(i - 1)) / den;
            Rr[i] = g / den;
            A[i] = a; AL[i] = alp; BE[i] = bet; ID[i] = 1.0 / den;
        }
        auto solve = [&](std::vector<double> &v) {   /* in place, local rows 0..m-1 */
            double s1 = 0, s2 = 0, s3 = 0;
            for (int i = 0; i < m; i++) {
                const double t = (v[i] - A[i] * s3 - AL[i] * s2 - BE[i] * s1) * ID[i];
                v[i] = t; s3 = s2; s2 = s1; s1 = t;
            }
            double x1 = 0, x2 = 0, x3 = 0;
            for (int i = m - 1; i >= 0; i--) {
                const double t = v[i] - Pp[i] * x1 - Qq[i] * x2 - Rr[i] * x3;
                v[i] = t; x3 = x2; x2 = x1; x1 = t;
            }
        };
        for (int k = 0; k < 3; k++) {
            std::vector<double> v(m, 0.0), w(m, 0.0);
            if (c > 0) {                              /* row lo+r couples to x_{lo-3+k} through c[k - r] */
                for (int r = 0; r < 3 && r < m; r++) v[r] = k >= r ? rows[lo + r].c[k - r] : 0.0;
                solve(v);
            }
            if (c < S - 1) {                          /* row hi-3+r couples to x_{hi+k} through c[6 + k - r] */
                for (int r = 0; r < 3; r++) w[m - 3 + r] = k <= r ? rows[hi - 3 + r].c[6 + k - r] : 0.0;
                solve(w);
            }
            for (int i = 0; i < m; i++) {
                T.vw[(size_t) k * SL + (size_t) c * L + i] = v[i];
                T.vw[(size_t) (3 + k) * SL + (size_t) c * L + i] = w[i];
            }
            for (int r = 0; r < 3; r++) {
                const int rf = 6 * c + r, rl = 6 * c + 3 + r;
                if (c > 0) {
                    M[(size_t) rf * m6 + 6 * (c - 1) + 3 + k] += v[r];
                    M[(size_t) rl * m6 + 6 * (c - 1) + 3 + k] += v[m - 3 + r];
                }
                if (c < S - 1) {
                    M[(size_t) rf * m6 + 6 * (c + 1) + k] += w[r];
                    M[(size_t) rl * m6 + 6 * (c + 1) + k] += w[m - 3 + r];
                }
            }
        }
    }
    if (invert_dense(M, m6)) return -3;
    const int mp = 6 * CF_SPIKE4_SMAX;
    T.minv.assign((size_t) mp * mp, 0.0);
    for (int r = 0; r < m6; r++)
        for (int q = 0; q < m6; q++) T.minv[(size_t) r * mp + q] = M[(size_t) r * m6 + q];
    return 0;
}

/* the device algorithm of k_spike4 on the host, for the self-check of a
 * built table against the sequential solve */
static void spike4_apply_host(const Spike4Host &T, const std::vector<double> &h, std::vector<double> &x)
{
    const int n = T.n, S = T.S, L = T.L, mp = 6 * CF_SPIKE4_SMAX;
    const size_t SL = (size_t) S * L;
    std::vector<double> y((size_t) S * L, 0.0), Y((size_t) 6 * S, 0.0);
    for (int c = 0; c < S; c++) {
        const int lo = c * L, m = std::min(L, n - lo);
        const double *A = &T.f[(size_t) c * L], *AL = A + SL, *BE = A + 2 * SL, *ID = A + 3 * SL;
        const double *Pp = A + 4 * SL, *Qq = A + 5 * SL, *Rr = A + 6 * SL;
        double *yc = &y[(size_t) c * L];
        double s1 = 0, s2 = 0, s3 = 0;
        for (int i = 0; i < L; i++) {
            const double t = ((i < m ? h[lo + i] : 0.0) - A[i] * s3 - AL[i] * s2 - BE[i] * s1) * ID[i];
            yc[i] = t; s3 = s2; s2 = s1; s1 = t;
        }
        double x1 = 0, x2 = 0, x3 = 0;
        for (int i = L - 1; i >= 0; i--) {
            const double t = yc[i] - Pp[i] * x1 - Qq[i] * x2 - Rr[i] * x3;
            yc[i] = t; x3 = x2; x2 = x1; x1 = t;
        }
        for (int r = 0; r < 3; r++) {
            Y[6 * c + r] = yc[r];
            Y[6 * c + 3 + r] = m - 3 + r >= 0 ? yc[m - 3 + r] : 0.0;
        }
    }
    x.assign(n, 0.0);
    for (int c = 0; c < S; c++) {
        const int lo = c * L, m = std::min(L, n - lo);
        double E[3] = {0, 0, 0}, F[3] = {0, 0, 0};
        for (int k = 0; k < 3; k++)
            for (int q = 0; q < 6 * S; q++) {
                if (c > 0) E[k] += T.minv[(size_t) (6 * (c - 1) + 3 + k) * mp + q] * Y[q];
                if (c < S - 1) F[k] += T.minv[(size_t) (6 * (c + 1) + k) * mp + q] * Y[q];
            }
        for (int i = 0; i < m; i++) {
            double v = y[(size_t) c * L + i];
            for (int k = 0; k < 3; k++)
                v -= T.vw[(size_t) k * SL + (size_t) c * L + i] * E[k] + T.vw[(size_t) (3 + k) * SL + (size_t) c * L + i] * F[k];
            x[lo + i] = v;
        }
    }
}

/* literal Thomas solve (division form) for the host-side profiles */
static void tridag_host(const std::vector<double> &a, const std::vector<double> &b,
                        const std::vector<double> &c, const std::vector<double> &r, std::vector<double> &x)
{
    int n = (int) b.size();
    std::vector<double> gam(n);
    x.assign(n, 0.0);
    double bet = b[0];
    x[0] = r[0] / bet;
    for (int j = 1; j < n; j++) {
        gam[j] = c[j - 1] / bet;
        bet = b[j] - a[j] * gam[j];
        x[j] = (r[j] - a[j] * x[j - 1]) / bet;
    }
    for (int j = n - 2; j >= 0; j--) x[j] -= gam[j + 1] * x[j + 1];
}

/* Poisson z-tables for one dt (pressure.c:318-339 and closures :446-531) */
static void poisson_tables(const double *dzl, int nz, double dx, double dy, double dt, PoisHost &T)
{
    T.a.assign(nz, 0); T.c.assign(nz, 0); T.bz.assign(nz, 0);
    T.ex.assign(nz, 0); T.ey.assign(nz, 0); T.dvol.assign(nz, 0);
    for (int k = 0; k < nz; k++) {
        double dzp1h = 0.5 * (DZ(k + 1) + DZ(k));
        double dzm1h = 0.5 * (DZ(k - 1) + DZ(k));
        double dvol = dx * dy * DZ(k);
        T.a[k] = +dvol * dt / (dzm1h * DZ(k));
        T.c[k] = +dvol * dt / (dzp1h * DZ(k));
        T.bz[k] = -dvol * dt / (dzm1h * DZ(k)) - dvol * dt / (dzp1h * DZ(k));
        T.ex[k] = dvol * dt / (dx * dx);
        T.ey[k] = dvol * dt / (dy * dy);
        T.dvol[k] = dvol;
    }
    {   /* NEUMANP bottom: a=0, b=-1/h, c=+1/h; b[0] -= c_g * (a[0] / b_g) */
        double h = 0.5 * (DZ(0) + DZ(-1));
        double bg = -1.0 / h, cg = +1.0 / h;
        T.db_neu = cg * (T.a[0] / bg);
    }
    T.db_dir = 0.5 * (T.a[0] / 0.5);       /* DIRICHP bottom: b = c = 0.5 */
    {   /* NEUMANP top: a=-1/h, b=+1/h; b[nz-1] -= a_g * (c[nz-1] / b_g) */
        double h = 0.5 * (DZ(nz - 1) + DZ(nz));
        double ag = -1.0 / h, bg = +1.0 / h;
        T.db_top = ag * (T.c[nz - 1] / bg);
    }
}

#undef DZ

/* order-4 Poisson z-tables for one dt (pressure.c:357-444 with the
 * closures of :534-726): wall-reduced bands at zero horizontal
 * wavenumber, Neumann bottom (variant 0) or the Dirichlet-mean bottom of
 * the (0,0) mode (variant 1), Neumann top */
static int pois4_tables(const double *dzl, int nz, double dx, double dy, double dt, std::vector<double> &T)
{
    Clos1 nb, db, nt;
    if (cft::closure(cft::CELL, 4, 0, cft::NEUMANN, dzl, nz, nb) || cft::closure(cft::CELL, 4, 0, cft::DIRICHLET, dzl, nz, db)
        || cft::closure(cft::CELL, 4, 1, cft::NEUMANN, dzl, nz, nt))
        return -1;
    T.assign((size_t) 17 * nz, 0.0);
    for (int v = 0; v < 2; v++) {
        std::vector<cft::Row7> rows;
        cft::op_rows(cft::CELL, 4, dzl, nz, dt, dx * dy, false, rows);
        cft::eliminate(cft::CELL, nz, v ? db : nb, nt, rows);
        for (int t = 0; t < 7; t++)
            for (int k = 0; k < nz; k++) T[(size_t) (v * 7 + t) * nz + k] = rows[k].c[t];
    }
    for (int k = 0; k < nz; k++) {
        double dvol = dx * dy * dzl[k];
        T[(size_t) 14 * nz + k] = dvol * dt / (dx * dx);
        T[(size_t) 15 * nz + k] = dvol * dt / (dy * dy);
        T[(size_t) 16 * nz + k] = dvol;
    }
    return 0;
}

/* ------------------------------------------------------------ helpers */

static cf_status validate(const cf_config *c)
{
    if (!c) { cf_set_error("null config"); return CF_ERR_ARG; }
    if (c->nx < 2 || c->ny < 2 || c->nz < 2) { cf_set_error("need nx, ny, nz >= 2"); return CF_ERR_ARG; }
    if (c->order != 2 && c->order != 4) { cf_set_error("order must be 2 or 4"); return CF_ERR_ARG; }
    if (!(c->lx > 0) || !(c->ly > 0) || !(c->lz > 0) || !(c->dt > 0) || !(c->nu > 0)) {
        cf_set_error("lx, ly, lz, dt, nu must be positive");
        return CF_ERR_ARG;
    }
    if (c->scalar && !(c->prandtl > 0)) { cf_set_error("scalar needs prandtl > 0"); return CF_ERR_ARG; }
    if (!c->scalar && (c->betg_x != 0 || c->betg_y != 0 || c->betg_z != 0)) {
        cf_set_error("buoyancy (betg != 0) needs scalar = 1");
        return CF_ERR_ARG;
    }
    if (c->order == 4 && (c->nx < 6 || c->ny < 6 || c->nz < 6)) {
        cf_set_error("order 4 needs nx, ny, nz >= 6 (three ghost levels, pressure.c:798)");
        return CF_ERR_ARG;
    }
    if (c->scalar && (c->nx < 4 || c->ny < 4)) { cf_set_error("the scalar needs nx, ny >= 4"); return CF_ERR_ARG; }
    if (c->precision != CF_FP64 && c->precision != CF_FP32) { cf_set_error("precision must be CF_FP64 or CF_FP32"); return CF_ERR_ARG; }
    if (c->cflmode != 0 && c->cflmode != 1) { cf_set_error("cflmode must be 0 or 1"); return CF_ERR_ARG; }
    if (c->driver != CF_DRIVER_CHANFAST && c->driver != CF_DRIVER_CHANNEL) { cf_set_error("driver must be CF_DRIVER_CHANFAST or CF_DRIVER_CHANNEL"); return CF_ERR_ARG; }
    if (c->driver == CF_DRIVER_CHANNEL && c->precision != CF_FP64) {
        cf_set_error("the fully implicit driver (CF_DRIVER_CHANNEL) runs in FP64 only");
        return CF_ERR_UNSUPPORTED;
    }
    if (c->cflmode == 1 && !(c->want_cfl > 0)) { cf_set_error("cflmode 1 needs want_cfl > 0"); return CF_ERR_ARG; }
    int rad[CF_MAX_RAD];
    if (c->nx % 1 || factor_radix(c->nx, rad) < 0 || c->nx > MAXE_LINE) {
        cf_set_error("nx = %d: FFT lengths must be 2^a 3^b 5^c and <= %d", c->nx, MAXE_LINE);
        return CF_ERR_UNSUPPORTED;
    }
    if (factor_radix(c->ny, rad) < 0 || c->ny > MAXE_LINE) {
        cf_set_error("ny = %d: FFT lengths must be 2^a 3^b 5^c and <= %d", c->ny, MAXE_LINE);
        return CF_ERR_UNSUPPORTED;
    }
    if (c->nz > 8000) { cf_set_error("nz > 8000"); return CF_ERR_ARG; }
    return CF_OK;
}

extern "C" cf_status cf_run_footprint(const cf_config *cfg, size_t *bytes)
{
    cf_status st = validate(cfg);
    if (st != CF_OK) return st;
    size_t N = (size_t) cfg->nx * cfg->ny * cfg->nz;
    size_t M = (size_t) (cfg->nx / 2 + 1) * cfg->ny * cfg->nz;
    size_t nf = cfg->scalar ? 4 : 3;
    size_t es = cfg->precision == CF_FP32 ? sizeof(float) : sizeof(double);
    /* state, solve output and two forcing slots per field, p: (4 nf + 1) N values;
     * X: 2 M values; FP64 stage N, G M, order-4 Poisson factors 3 M */
    *bytes = es * ((4 * nf + 1) * N + 2 * M) + sizeof(double) * (N + (cfg->order == 4 ? 4 : 1) * M)
           + ((size_t) 16 << 20);
    return CF_OK;
}

static void free_run(cf_run *r)
{
    if (!r) return;
    for (int q = 0; q < 2; q++) {
        if (r->gexec[q]) (void) hipGraphExecDestroy(r->gexec[q]);
        if (r->graph[q]) (void) hipGraphDestroy(r->graph[q]);
    }
    double *bufs[] = { r->u, r->v, r->w, r->c, r->ru, r->rv, r->rw, r->rc, r->fa[0], r->fa[1], r->fa[2], r->fa[3],
                       r->fb[0], r->fb[1], r->fb[2], r->fb[3], r->p, r->G, r->S4, r->stage, r->tab, r->scal,
                       r->logbuf, r->part, r->lsum, r->cflpart, r->h1prev, r->uoffd };
    for (double *b : bufs) if (b) (void) hipFree(b);
    if (r->X) (void) hipFree(r->X);
    for (auto &pe : r->pending) { (void) hipEventDestroy(pe.second.first); (void) hipEventDestroy(pe.second.second); }
    if (r->stream) (void) hipStreamDestroy(r->stream);
    delete r;
}

/* ------------------------------------------------------------ create */

/* balanced z-chunks of the fused RHS for a target length: levels per
 * chunk and the launch's workgroup count (also its CFL partials) */
static void rhs_chunks(int nx, int ny, int nz, int target, int *kc, int *blocks)
{
    int nchunk = (nz + target - 1) / target;
    *kc = (nz + nchunk - 1) / nchunk;
    nchunk = (nz + *kc - 1) / *kc;
    *blocks = ((nx + CF_RT_X - 1) / CF_RT_X) * ((ny + CF_RT_Y - 1) / CF_RT_Y) * nchunk;
}

/* Every table of a run that depends on dt (the z-solve factors and spikes,
 * the unit-forcing profiles h1 and their statistics, the boundary lifts,
 * the Poisson z-tables) together with the dt-independent ones (metrics,
 * twiddles, modified wavenumbers), packed in the device layout of r->tab.
 * The layout depends on the grid alone, so a new dt (cflmode 1) repacks
 * into the same offsets. */
struct TabPack {
    std::vector<double> pack;
    SubCoef sc[3];
    double lift[3][4];
    SpikeHost sph[3][3];
    Spike4Host s4h[3][3];
    PoisHost ph[4];
    Clos clo;
    bool spike_ok = true;
    bool spike4_ok = false, spike4_packed = false;
    size_t o_dz = 0, o_th = 0, o_st = 0, o_h1 = 0, o_twx = 0, o_twy = 0, o_twx32 = 0, o_twy32 = 0, o_lx = 0, o_ly = 0;
    size_t o_rdz = 0, o_rdzh = 0, o_lift = 0;
    size_t o_p[4][6], o_sp[3][3][6], o_hp[3][3], o_p4[4], o_s4[3][3][3];
    size_t o_hs[3][3];               /* spectral Helmholtz z-tables (fully implicit driver) */
};

static cf_status build_tables(const cf_config *cfg, const std::vector<double> &dz, double dx, double dy, double dt,
                              TabPack &T)
{
    const double nu = cfg->nu;
    const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    const double *dzl = dz.data() + CF_NGT;

    /* ---- host tables ---- */
    const size_t nzt = (size_t) nz;
    std::vector<double> th(3 * 4 * 3 * nzt, 0.0), stat1(3 * 2 * 4, 0.0), h1(3 * 2 * (nzt + 2), 0.0);
    SpikeHost (&sph)[3][3] = T.sph;
    PoisHost (&ph)[4] = T.ph;
    for (int s = 0; s < 3; s++) {
        SubCoef &C = T.sc[s];
        C.nuadt = nu * RK_ALP[s] * dt;
        C.nubdt = nu * RK_BET[s] * dt;
        C.abdt = (RK_ALP[s] + RK_BET[s]) * dt;
        C.gamdt = RK_GAM[s] * dt;
        C.zetdt = RK_ZET[s] * dt;
        const double kap = cfg->scalar ? nu / cfg->prandtl : nu;   /* chanfast.c:540-547 */
        C.kadt = kap * RK_ALP[s] * dt;
        C.kbdt = kap * RK_BET[s] * dt;
        for (int type = 0; type < 3; type++) {
            std::vector<double> a, b, c;
            if (type == 0) cell2_system(dzl, nz, -C.nubdt, a, b, c);
            else if (type == 1) cell2_system(dzl, nz, -C.kbdt, a, b, c);
            else face2_system(dzl, nz, -C.nubdt, a, b, c);
            double *A = &th[((size_t) (s * 4 + type) * 3 + 0) * nzt];
            double *IB = &th[((size_t) (s * 4 + type) * 3 + 1) * nzt];
            double *GM = &th[((size_t) (s * 4 + type) * 3 + 2) * nzt];
            if (thomas_factors(a, b, c, A, IB, GM)) {
                cf_set_error("zero pivot in the implicit z system (substep %d)", s);
                return CF_ERR_NUMERIC;
            }
            if (spike_build(a, b, c, spike_L(nz), sph[s][type])) sph[s][type].n = 0;   /* falls back to Thomas */
            if (type < 2) {
                /* unit-forcing profile h1: h = gam dt, solve (chanfast.c:498-500) */
                std::vector<double> rr(nz), x;
                for (int k = 0; k < nz; k++) rr[k] = C.gamdt * dzl[k];
                tridag_host(a, b, c, rr, x);
                double *H = &h1[(size_t) (s * 2 + type) * (nzt + 2)];
                for (int k = 0; k < nz; k++) H[k + 1] = x[k];
                H[0] = (0.5 * x[0] - 0.0) / -0.5;              /* ghost_fill_bot2 */
                H[nz + 1] = (0.5 * x[nz - 1] - 0.0) / -0.5;    /* ghost_fill_top2 */
                double kap = type == 0 ? nu : nu / cfg->prandtl;
                double vol = 0.0, lz = 0.0;
                for (int k = 0; k < nz; k++) { vol += H[k + 1] * dzl[k]; lz += dzl[k]; }
                double *S = &stat1[(size_t) (s * 2 + type) * 4];
                S[0] = vol / lz;
                S[1] = kap * (H[1] - H[0]) / (0.5 * (dzl[0] + dzl[-1]));
                S[2] = kap * (H[nz + 1] - H[nz]) / (0.5 * (dzl[nz] + dzl[nz - 1]));
                S[3] = nz % 2 == 0 ? 0.5 * (H[nz / 2] + H[nz / 2 + 1]) : H[nz / 2 + 1];
            }
        }
        /* boundary lift: h -= H1D(bc-only field; -nu b dt) (chanfast.c:472-477) */
        double bcs[2] = { 0.0 - cfg->uframe, 0.0 - cfg->vframe };
        for (int f = 0; f < 2; f++) {
            double bc = bcs[f];
            double g = (0.5 * 0.0 - bc) / -0.5;                 /* apply_bc_u2 on a zero field */
            double nudt = -C.nubdt;
            double dzp1h = 0.5 * (dzl[1] + dzl[0]), dzm1h = 0.5 * (dzl[-1] + dzl[0]);
            double bot_next = nz > 1 ? 0.0 : g;
            T.lift[s][2 * f] = 0.0 + nudt * (((bot_next - 0.0) / dzp1h - (0.0 - g) / dzm1h) / dzl[0]);
            dzp1h = 0.5 * (dzl[nz] + dzl[nz - 1]);
            dzm1h = 0.5 * (dzl[nz - 2] + dzl[nz - 1]);
            T.lift[s][2 * f + 1] = 0.0 + nudt * (((g - 0.0) / dzp1h - (0.0 - 0.0) / dzm1h) / dzl[nz - 1]);
        }
        poisson_tables(dzl, nz, dx, dy, C.abdt, ph[s]);
    }
    poisson_tables(dzl, nz, dx, dy, 1.0, ph[3]);

    /* ---- general path: closures, boundary lifts, order-4 systems ---- */
    const int order = cfg->order;
    Clos &clo = T.clo;
    if (cft::closure(cft::CELL, order, 0, cft::DIRICHLET, dzl, nz, clo.cell[0])
        || cft::closure(cft::CELL, order, 1, cft::DIRICHLET, dzl, nz, clo.cell[1])
        || cft::closure(cft::FACE, order, 0, cft::DIRICHLET, dzl, nz, clo.face[0])
        || cft::closure(cft::FACE, order, 1, cft::DIRICHLET, dzl, nz, clo.face[1])) {
        cf_set_error("singular wall closure rows");
        return CF_ERR_NUMERIC;
    }
    std::vector<double> liftg((size_t) 3 * 4 * nz, 0.0);
    for (int s = 0; s < 3; s++)
        for (int f = 0; f < 4; f++) {
            /* h -= H1D(bc-only field; -kappa b dt) (chanfast.c:472-477, :556-563, :627-632) */
            const int fam = f == 2 ? cft::FACE : cft::CELL;
            const int n = fam == cft::CELL ? nz : nz - 1;
            const double bcb = f == 0 ? 0.0 - cfg->uframe : f == 1 ? 0.0 - cfg->vframe : f == 2 ? 0.0 : -0.5;
            const double bct = f == 3 ? 0.5 : bcb;
            const double nudt = f == 3 ? -T.sc[s].kbdt : -T.sc[s].nubdt;
            const Clos1 *cl = fam == cft::CELL ? clo.cell : clo.face;
            std::vector<double> col((size_t) n + 6, 0.0), outv((size_t) n, 0.0);
            for (int m = 0; m < 3; m++) {
                col[(size_t) (-1 - m + 3)] = cl[0].B[m] * bcb;
                col[(size_t) (n + m + 3)] = cl[1].B[m] * bct;
            }
            cft::helm_apply(fam, order, dzl, nz, nudt, col.data(), outv.data());
            for (int k = 0; k < n; k++) liftg[((size_t) s * 4 + f) * nz + k] = outv[k];
        }
    std::vector<double> hept[3][3], p4[4];
    if (order == 4) {
        for (int s = 0; s < 3; s++) {
            const SubCoef &C = T.sc[s];
            for (int type = 0; type < 3; type++) {
                const int fam = type == 2 ? cft::FACE : cft::CELL;
                const double nudt = type == 1 ? -C.kbdt : -C.nubdt;
                const Clos1 *cl = fam == cft::CELL ? clo.cell : clo.face;
                std::vector<cft::Row7> rows;
                cft::op_rows(fam, 4, dzl, nz, nudt, 1.0, true, rows);
                cft::eliminate(fam, nz, cl[0], cl[1], rows);
                if (cft::hept_factor(rows, hept[s][type])) {
                    cf_set_error("zero pivot in the order-4 implicit z system (substep %d)", s);
                        return CF_ERR_NUMERIC;
                }
                /* the partitioned form, kept only if it reproduces the
                 * sequential solve on a test vector */
                Spike4Host &S4 = T.s4h[s][type];
                if (spike4_build(rows, spike4_L(nz), S4)) S4.n = 0;
                else {
                    const int n4 = (int) rows.size();
                    std::vector<double> hv(n4), xs, xp;
                    for (int k = 0; k < n4; k++) hv[k] = 1.0 + 0.5 * sin(0.37 * k) + (k % 3) * 0.25;
                    cft::hept_solve(hept[s][type], n4, hv, xs);
                    spike4_apply_host(S4, hv, xp);
                    double num = 0.0, den = 0.0;
                    for (int k = 0; k < n4; k++) { num += (xp[k] - xs[k]) * (xp[k] - xs[k]); den += xs[k] * xs[k]; }
                    S4.ok = num <= 1e-26 * den;
                }
                if (type == 2) continue;
                /* unit-forcing profile (chanfast.c:1422-1424) and its statistics */
                std::vector<double> rr(nz), x;
                for (int k = 0; k < nz; k++) rr[k] = C.gamdt * dzl[k];
                cft::hept_solve(hept[s][type], nz, rr, x);
                std::vector<double> A((size_t) nz + 6, 0.0);
                for (int k = 0; k < nz; k++) A[k + 3] = x[k];
                for (int m = 0; m < 3; m++) {
                    A[2 - m] = cl[0].G[m][0] * x[0] + cl[0].G[m][1] * x[1] + cl[0].G[m][2] * x[2];
                    A[nz + 3 + m] = cl[1].G[m][0] * x[nz - 1] + cl[1].G[m][1] * x[nz - 2] + cl[1].G[m][2] * x[nz - 3];
                }
                const double *Ap = A.data() + 3;
                double *H = &h1[(size_t) (s * 2 + type) * (nzt + 2)];
                for (int k = -1; k <= nz; k++) H[k + 1] = Ap[k];
                const double kap = type == 0 ? nu : nu / cfg->prandtl;
                double vol = 0.0, lz = 0.0;
                for (int k = 0; k < nz; k++) { vol += Ap[k] * dzl[k]; lz += dzl[k]; }
                const double A1h = 1.125 / 1.0, A2h = -0.125 / 3.0;
                double *S = &stat1[(size_t) (s * 2 + type) * 4];
                S[0] = vol / lz;
                S[1] = kap * ((A1h * (Ap[0] - Ap[-1]) + A2h * (Ap[1] - Ap[-2])) / (0.5 * (dzl[0] + dzl[-1])));
                S[2] = kap * ((A1h * (Ap[nz] - Ap[nz - 1]) + A2h * (Ap[nz + 1] - Ap[nz - 2])) / (0.5 * (dzl[nz] + dzl[nz - 1])));
                S[3] = nz % 2 == 0 ? 0.5 * (Ap[nz / 2 - 1] + Ap[nz / 2]) : Ap[nz / 2];
            }
        }
        for (int q = 0; q < 4; q++) {
            if (pois4_tables(dzl, nz, dx, dy, q < 3 ? T.sc[q].abdt : 1.0, p4[q])) {
                cf_set_error("singular order-4 pressure closure rows");
                return CF_ERR_NUMERIC;
            }
        }
    }

    /* fully implicit driver (channel2/4): the spectral Helmholtz z-tables
     * of each substep and field type (diffuseu.c:460-614, diffusew.c:466-622):
     * the 1-D system's wall-reduced bands (homogeneous Dirichlet walls) at
     * zero horizontal wavenumber, dvol nudt / dx^2 and dvol nudt / dy^2 for
     * the modified wavenumbers, and dvol.  Order 2: a | c | b | ex | ey | dvol
     * (the Poisson kernel's layout, no extra wall terms); order 4: the seven
     * bands | ex | ey | dvol. */
    std::vector<double> hs[3][3];
    if (cfg->driver == CF_DRIVER_CHANNEL) {
        for (int s = 0; s < 3; s++)
            for (int type = 0; type < 3; type++) {
                const int fam = type == 2 ? cft::FACE : cft::CELL;
                const int n = fam == cft::CELL ? nz : nz - 1;
                const double nudt = type == 1 ? -T.sc[s].kbdt : -T.sc[s].nubdt;
                std::vector<double> &H = hs[s][type];
                H.assign((size_t) (order == 2 ? 6 : 10) * n, 0.0);
                const size_t bo = (size_t) (order == 2 ? 3 : 7) * n;
                for (int k = 0; k < n; k++) {
                    const double dvol = fam == cft::CELL ? dzl[k] : 0.5 * (dzl[k] + dzl[k + 1]);
                    H[bo + k] = dvol * nudt / (dx * dx);
                    H[bo + n + k] = dvol * nudt / (dy * dy);
                    H[bo + 2 * n + k] = dvol;
                }
                if (order == 2) {
                    std::vector<double> a, b, c;
                    if (fam == cft::CELL) cell2_system(dzl, nz, nudt, a, b, c);
                    else face2_system(dzl, nz, nudt, a, b, c);
                    for (int k = 0; k < n; k++) { H[k] = a[k]; H[n + k] = c[k]; H[2 * n + k] = b[k]; }
                } else {
                    const Clos1 *cl = fam == cft::CELL ? clo.cell : clo.face;
                    std::vector<cft::Row7> rows;
                    cft::op_rows(fam, 4, dzl, nz, nudt, 1.0, true, rows);
                    cft::eliminate(fam, nz, cl[0], cl[1], rows);
                    for (int t = 0; t < 7; t++)
                        for (int k = 0; k < n; k++) H[(size_t) t * n + k] = rows[k].c[t];
                }
            }
    }

    /* FFT plans: radices, twiddles (host libm), modified wavenumbers */
    const int nkx = nx / 2 + 1;
    std::vector<double> twx(2 * nx), twy(2 * ny), lamx(nkx), lamy(ny);
    for (int j = 0; j < nx; j++) {
        double th_ = 2.0 * M_PI * (double) j / (double) nx;
        twx[2 * j] = cos(th_); twx[2 * j + 1] = -sin(th_);
    }
    for (int j = 0; j < ny; j++) {
        double th_ = 2.0 * M_PI * (double) j / (double) ny;
        twy[2 * j] = cos(th_); twy[2 * j + 1] = -sin(th_);
    }
    /* modified wavenumbers: pressure.c:318-320 (order 2), :382-389 (order 4) */
    auto lam = [&](int i, int n) -> double {
        if (order == 2) return 2.0 - 2.0 * cos(2.0 * M_PI * 1.0 * (double) i / (double) n);
        const double A1h = 1.125 / 1.0, A2h = -0.125 / 3.0;
        double c1 = cos(2.0 * M_PI * 1.0 * (double) i / (double) n);
        double c2 = cos(2.0 * M_PI * 2.0 * (double) i / (double) n);
        double c3 = cos(2.0 * M_PI * 3.0 * (double) i / (double) n);
        return (2.0 * A2h * A2h + 2.0 * A1h * A1h) + (4.0 * A1h * A2h - 2.0 * A1h * A1h) * c1
             + (-4.0 * A1h * A2h) * c2 + (-2.0 * A2h * A2h) * c3;
    };
    for (int ix = 0; ix < nkx; ix++) lamx[ix] = lam(ix, nx);
    for (int iy = 0; iy < ny; iy++) lamy[iy] = lam(iy, ny);

    /* pack tables: dz | th | stat1 | h1 | twx | twy | lamx | lamy | pois x4 */
    std::vector<double> &pack = T.pack;
    pack.clear();
    auto put = [&](const double *src, size_t n) { size_t o = pack.size(); pack.insert(pack.end(), src, src + n); return o; };
    T.o_dz = put(dz.data(), dz.size());
    T.o_th = put(th.data(), th.size());
    T.o_st = put(stat1.data(), stat1.size());
    T.o_h1 = put(h1.data(), h1.size());
    if (pack.size() % 2) pack.push_back(0.0);       /* 16-byte align the complex twiddles */
    T.o_twx = put(twx.data(), twx.size());
    T.o_twy = put(twy.data(), twy.size());
    /* FP32 copies of the twiddles, two floats per double slot */
    auto put32 = [&](const std::vector<double> &src) {
        std::vector<float> f(src.begin(), src.end());
        if (f.size() % 2) f.push_back(0.f);
        std::vector<double> slots(f.size() / 2);
        memcpy(slots.data(), f.data(), sizeof(float) * f.size());
        return put(slots.data(), slots.size());
    };
    T.o_twx32 = put32(twx);
    T.o_twy32 = put32(twy);
    T.o_lx = put(lamx.data(), lamx.size());
    T.o_ly = put(lamy.data(), lamy.size());
    size_t (&o_p)[4][6] = T.o_p;
    for (int q = 0; q < 4; q++) {
        o_p[q][0] = put(ph[q].a.data(), nz); o_p[q][1] = put(ph[q].c.data(), nz);
        o_p[q][2] = put(ph[q].bz.data(), nz); o_p[q][3] = put(ph[q].ex.data(), nz);
        o_p[q][4] = put(ph[q].ey.data(), nz); o_p[q][5] = put(ph[q].dvol.data(), nz);
    }
    std::vector<double> rdz(nz + 6), rdzh(nz + 6, 0.0);
    for (int k = -3; k < nz + 3; k++) rdz[k + 3] = 1.0 / dzl[k];
    for (int k = -3; k < nz + 2; k++) rdzh[k + 3] = 1.0 / (0.5 * (dzl[k] + dzl[k + 1]));
    T.o_rdz = put(rdz.data(), rdz.size());
    T.o_rdzh = put(rdzh.data(), rdzh.size());
    size_t (&o_sp)[3][3][6] = T.o_sp;
    bool &spike_ok = T.spike_ok;
    spike_ok = true;
    for (int s = 0; s < 3; s++)
        for (int t = 0; t < 3; t++) {
            SpikeHost &S = sph[s][t];
            if (S.n == 0) { spike_ok = false; continue; }
            o_sp[s][t][0] = put(S.a.data(), S.a.size()); o_sp[s][t][1] = put(S.ib.data(), S.ib.size());
            o_sp[s][t][2] = put(S.gam.data(), S.gam.size()); o_sp[s][t][3] = put(S.vl.data(), S.vl.size());
            o_sp[s][t][4] = put(S.wr.data(), S.wr.size()); o_sp[s][t][5] = put(S.minv.data(), S.minv.size());
        }
    T.o_lift = put(liftg.data(), liftg.size());
    size_t (&o_hp)[3][3] = T.o_hp;
    size_t (&o_p4)[4] = T.o_p4;
    memset(o_hp, 0, sizeof(T.o_hp));
    memset(o_p4, 0, sizeof(T.o_p4));
    if (order == 4) {
        for (int s = 0; s < 3; s++)
            for (int t = 0; t < 3; t++) o_hp[s][t] = put(hept[s][t].data(), hept[s][t].size());
        /* the layout depends on the chunking alone (a function of nz); the
         * checks decide whether the run uses the tables */
        bool built = true;
        T.spike4_ok = true;
        for (int s = 0; s < 3; s++)
            for (int t = 0; t < 3; t++) {
                built = built && T.s4h[s][t].n > 0;
                T.spike4_ok = T.spike4_ok && T.s4h[s][t].ok;
            }
        T.spike4_ok = T.spike4_ok && built;
        T.spike4_packed = built;
        if (built)
            for (int s = 0; s < 3; s++)
                for (int t = 0; t < 3; t++) {
                    const Spike4Host &S4 = T.s4h[s][t];
                    T.o_s4[s][t][0] = put(S4.f.data(), S4.f.size());
                    T.o_s4[s][t][1] = put(S4.vw.data(), S4.vw.size());
                    T.o_s4[s][t][2] = put(S4.minv.data(), S4.minv.size());
                }
        for (int q = 0; q < 4; q++) o_p4[q] = put(p4[q].data(), p4[q].size());
    }
    for (int s = 0; s < 3; s++)
        for (int t = 0; t < 3; t++) T.o_hs[s][t] = hs[s][t].empty() ? 0 : put(hs[s][t].data(), hs[s][t].size());

    return CF_OK;
}


/* point the run's parameter block at the tables of TP (already uploaded to
 * r->tab) and take over its by-value coefficients */
static void apply_tables(cf_run *r, const TabPack &TP)
{
    RunDev &R = r->rd;
    const int nz = r->g.nz, order = r->g.order;
    const double *tab = r->tab;
    R.dzt = tab + TP.o_dz;
    R.th = tab + TP.o_th;
    R.rdz = tab + TP.o_rdz;
    R.rdzh = tab + TP.o_rdzh;
    R.cl = TP.clo;
    R.liftg = tab + TP.o_lift;
    for (int s = 0; s < 3; s++) {
        R.sc[s] = TP.sc[s];
        for (int f = 0; f < 4; f++) R.lift[s][f] = TP.lift[s][f];
    }
    for (int s = 0; s < 3; s++)
        for (int t = 0; t < 3; t++) {
            R.hp[s][t].n = t == 2 ? nz - 1 : nz;
            R.hp[s][t].t = order == 4 ? tab + TP.o_hp[s][t] : nullptr;
            Spike4Tab &S4 = R.s4[s][t];
            S4.n = 0; S4.S = 0; S4.L = 0; S4.f = S4.vw = S4.minv = nullptr;
            if (order == 4 && TP.spike4_packed) {
                const Spike4Host &H4 = TP.s4h[s][t];
                S4.n = H4.n; S4.S = H4.S; S4.L = H4.L;
                S4.f = tab + TP.o_s4[s][t][0]; S4.vw = tab + TP.o_s4[s][t][1]; S4.minv = tab + TP.o_s4[s][t][2];
            }
            SpikeTab &ST = R.sp[s][t];
            ST.n = TP.sph[s][t].n; ST.S = TP.sph[s][t].S; ST.L = TP.sph[s][t].L;
            if (!TP.spike_ok) continue;
            ST.a = tab + TP.o_sp[s][t][0]; ST.ib = tab + TP.o_sp[s][t][1]; ST.gam = tab + TP.o_sp[s][t][2];
            ST.vl = tab + TP.o_sp[s][t][3]; ST.wr = tab + TP.o_sp[s][t][4]; ST.minv = tab + TP.o_sp[s][t][5];
        }
    R.stat1 = tab + TP.o_st;
    R.h1 = tab + TP.o_h1;
    R.twx = (const double2 *) (tab + TP.o_twx);
    R.twy = (const double2 *) (tab + TP.o_twy);
    R.twx32 = (const float2 *) (tab + TP.o_twx32);
    R.twy32 = (const float2 *) (tab + TP.o_twy32);
    R.lamx = tab + TP.o_lx;
    R.lamy = tab + TP.o_ly;
    for (int q = 0; q < 4; q++) {
        PoisTab &T = r->pt[q];
        T.a = tab + TP.o_p[q][0]; T.c = tab + TP.o_p[q][1]; T.bz = tab + TP.o_p[q][2];
        T.ex = tab + TP.o_p[q][3]; T.ey = tab + TP.o_p[q][4]; T.dvol = tab + TP.o_p[q][5];
        T.db_neu = TP.ph[q].db_neu; T.db_dir = TP.ph[q].db_dir; T.db_top = TP.ph[q].db_top;
        Pois4Tab &T4 = r->pt4[q];
        if (order == 4) {
            const double *b = tab + TP.o_p4[q];
            T4.band[0] = b; T4.band[1] = b + (size_t) 7 * nz;
            T4.ex = b + (size_t) 14 * nz; T4.ey = b + (size_t) 15 * nz; T4.dvol = b + (size_t) 16 * nz;
        }
    }
    r->off_pois_adhoc = TP.o_p[3][0];
    r->off_pois4_adhoc = TP.o_p4[3];
    r->pt_dt = 1.0;
    if (r->cfg.driver == CF_DRIVER_CHANNEL)
        for (int s = 0; s < 3; s++)
            for (int t = 0; t < 3; t++) {
                const size_t n = (size_t) (t == 2 ? nz - 1 : nz);
                const double *b = tab + TP.o_hs[s][t];
                if (order == 2) {
                    PoisTab &H = r->hs2[s][t];
                    H.a = b; H.c = b + n; H.bz = b + 2 * n; H.ex = b + 3 * n; H.ey = b + 4 * n; H.dvol = b + 5 * n;
                    H.db_neu = H.db_dir = H.db_top = 0.0;
                } else {
                    Pois4Tab &H = r->hs4[s][t];
                    H.band[0] = H.band[1] = b;
                    H.ex = b + 7 * n; H.ey = b + 8 * n; H.dvol = b + 9 * n;
                }
            }
}

/* The FP32 fused path stores u - U(z) with U = r->ubar.  The device copy
 * holds U with its wall ghosts 2 ubc - U (apply_bc_u2 on a column: the
 * stored deviation's ghost is then its negative) and, per substep, U's
 * share of the implicit right-hand side, nu a dt Lz_bc U + nu b dt Lz_0 U:
 * Lz_bc with those ghosts (how the explicit operator sees U) and Lz_0 with
 * the homogeneous ghosts -U of the z solve, so that the solve of the
 * stored deviation's right-hand side gives the new deviation (the solve is
 * linear: S (u' + U) = h  <=>  S u' = h - U + nu b dt Lz_0 U). */
static cf_status set_uoff(cf_run *r)
{
    if (!r->uoffd) return CF_OK;
    const int nz = r->g.nz;
    const double *dzl = r->dz.data() + CF_NGT;
    const double ub = r->rd.ubc;
    const std::vector<double> &U = r->ubar;
    std::vector<double> H((size_t) nz + 2 + 3 * (size_t) nz, 0.0);
    H[0] = 2.0 * ub - U[0];
    for (int k = 0; k < nz; k++) H[(size_t) k + 1] = U[k];
    H[(size_t) nz + 1] = 2.0 * ub - U[nz - 1];
    for (int s = 0; s < 3; s++)
        for (int k = 0; k < nz; k++) {
            const double rzp = 1.0 / (0.5 * (dzl[k] + dzl[k + 1])), rzm = 1.0 / (0.5 * (dzl[k - 1] + dzl[k]));
            const double rzk = 1.0 / dzl[k];
            const double u0 = H[(size_t) k + 1], up = H[(size_t) k + 2], um = H[k];
            const double lbc = ((up - u0) * rzp - (u0 - um) * rzm) * rzk;
            const double up0 = k == nz - 1 ? -u0 : up, um0 = k == 0 ? -u0 : um;
            const double l0 = ((up0 - u0) * rzp - (u0 - um0) * rzm) * rzk;
            H[(size_t) nz + 2 + (size_t) s * nz + k] = r->rd.sc[s].nuadt * lbc + r->rd.sc[s].nubdt * l0;
        }
    HIPCHK(hipMemcpyAsync(r->uoffd, H.data(), sizeof(double) * H.size(), hipMemcpyHostToDevice, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return CF_OK;
}

static cf_status create_run(const cf_config *cfg, bool force_general, cf_run **out);

extern "C" cf_status cf_run_create(const cf_config *cfg, cf_run **out)
{
    return create_run(cfg, false, out);
}

static cf_status create_run(const cf_config *cfg, bool force_general, cf_run **out)
{
    if (!out) { cf_set_error("null output handle"); return CF_ERR_ARG; }
    *out = nullptr;
    cf_status st = validate(cfg);
    if (st != CF_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) {
        cf_set_error("no HIP device %d available", cfg->device);
        return CF_ERR_DEVICE;
    }
    HIPCHK(hipSetDevice(cfg->device));

    cf_run *r = new cf_run();
    r->cfg = *cfg;
    r->cfg.dz_user = nullptr;
    const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    r->g.nx = nx; r->g.ny = ny; r->g.nz = nz; r->g.nkx = nx / 2 + 1; r->g.order = cfg->order;
    r->g.dx = cfg->lx / (double) nx;            /* chanfast.c:148-149 */
    r->g.dy = cfg->ly / (double) ny;
    r->N = (size_t) nx * ny * nz;
    r->P = (size_t) nx * ny;
    r->M = (size_t) r->g.nkx * ny;

    /* wall-normal metric with mirrored ghosts (chanfast.c:184-190, ghost.c:31-54) */
    std::vector<double> dzi(nz);
    st = cf_grid_dz(cfg, dzi.data());
    if (st != CF_OK) { delete r; return st; }
    r->dz.assign(nz + 6, 0.0);
    double *dzl = r->dz.data() + CF_NGT;
    for (int k = 0; k < nz; k++) dzl[k] = dzi[k];
    for (int k = 1; k <= 3; k++) { dzl[-k] = dzl[k - 1]; dzl[nz - 1 + k] = dzl[nz - k]; }

    const double dx = r->g.dx, dy = r->g.dy;
    const int order = cfg->order;
    r->prec = cfg->precision;
    r->esz = cfg->precision == CF_FP32 ? sizeof(float) : sizeof(double);
    /* the fused path and the order-4 tiles address their arrays with 32-bit
     * byte offsets: a larger field takes the general path */
    const bool small = (size_t) nx * ny * nz * r->esz < ((size_t) 1 << 32);
    r->general = force_general || order == 4 || cfg->scalar || cfg->driver == CF_DRIVER_CHANNEL || !small;
    r->nfield = cfg->scalar ? 4 : 3;
    r->rd.m_prev_slot[0] = SC_M0 + 2;
    r->rd.m_prev_slot[1] = SC_M0 + 0;
    r->rd.m_prev_slot[2] = SC_M0 + 1;
    r->rd.nradx = factor_radix(nx, r->rd.radx);
    r->rd.nrady = factor_radix(ny, r->rd.rady);
    TabPack TP;
    st = build_tables(cfg, r->dz, dx, dy, cfg->dt, TP);
    if (st != CF_OK) { delete r; return st; }
    const std::vector<double> &pack = TP.pack;
    r->tab_n = pack.size();

    /* ---- device memory ---- */
    auto dalloc = [&](double **pp, size_t n) -> bool { return hipMalloc((void **) pp, sizeof(double) * (n ? n : 1)) == hipSuccess; };
    /* state and work fields in the run's storage precision */
    auto falloc = [&](double **pp, size_t n) -> bool { return hipMalloc((void **) pp, r->esz * (n ? n : 1)) == hipSuccess; };
    bool ok = true;
    ok = ok && falloc(&r->u, r->N) && falloc(&r->v, r->N) && falloc(&r->w, r->N);
    ok = ok && falloc(&r->ru, r->N) && falloc(&r->rv, r->N) && falloc(&r->rw, r->N);
    for (int f = 0; f < 3; f++) ok = ok && falloc(&r->fa[f], r->N) && falloc(&r->fb[f], r->N);
    if (cfg->scalar) ok = ok && falloc(&r->c, r->N) && falloc(&r->rc, r->N) && falloc(&r->fa[3], r->N) && falloc(&r->fb[3], r->N);
    if (order == 4) ok = ok && dalloc(&r->S4, 3 * r->M * nz);
    /* G: the order-2 Poisson's chunk checkpoints (3 per chunk of 16 levels) per mode */
    const size_t glev = std::max((size_t) nz, (size_t) 3 * ((nz + 15) / 16));
    ok = ok && falloc(&r->p, r->N) && dalloc(&r->stage, r->N) && dalloc(&r->G, r->M * glev);
    ok = ok && hipMalloc((void **) &r->X, 2 * r->esz * r->M * nz) == hipSuccess;
    ok = ok && dalloc(&r->tab, r->tab_n) && dalloc(&r->scal, SC_N) && dalloc(&r->logbuf, (size_t) LOGCAP * CF_LOG_N);
    r->nb_plane = (int) ((r->P + 2047) / 2048);
    if (r->nb_plane > 1024) r->nb_plane = 1024;
    r->nb_red = 1024;
    size_t npart = std::max((size_t) r->nb_plane * (2 * nz + 2), (size_t) ((r->P + 63) / 64) * nz);
    if (npart < (size_t) r->nb_red) npart = r->nb_red;
    ok = ok && dalloc(&r->part, npart) && dalloc(&r->lsum, (size_t) nz + 2);
    rhs_chunks(nx, ny, nz, CF_RHS_KC, &r->rhs_kc, &r->rhs_blocks);
    r->gblocks = ((nx + 63) / 64) * ((ny + 3) / 4) * nz;
    if (cfg->order == 4 && nx % CF_R4_X == 0 && ny % CF_R4_Y == 0
        && (size_t) nx * ny * nz * r->esz < ((size_t) 1 << 32)) {   /* 32-bit byte offsets */
        r->r4_nch = std::max(1, (nz + CF_R4_KC / 2) / CF_R4_KC);
        r->r4_blocks = (nx / CF_R4_X) * (ny / CF_R4_Y) * r->r4_nch;
    }
    ok = ok && dalloc(&r->cflpart, (size_t) std::max(std::max(r->rhs_blocks, r->gblocks), r->r4_blocks));
    if (cfg->cflmode == 1) ok = ok && dalloc(&r->h1prev, (size_t) nz + 2);
    if (r->prec == CF_FP32 && !r->general) ok = ok && dalloc(&r->uoffd, (size_t) nz + 2 + 3 * (size_t) nz);
    if (!ok) { free_run(r); cf_set_error("device allocation failed"); return CF_ERR_NOMEM; }
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) {
        free_run(r); cf_set_error("stream creation failed"); return CF_ERR_DEVICE;
    }
    /* on the run's own stream, so that setting up a run does not wait for
     * other runs' work in flight on the device */
    if (hipMemcpyAsync(r->tab, pack.data(), sizeof(double) * r->tab_n, hipMemcpyHostToDevice, r->stream) != hipSuccess) {
        free_run(r); cf_set_error("table upload failed"); return CF_ERR_DEVICE;
    }
    /* zero state, tendencies and scalars (finite values: g meets zeta_0 = 0) */
    std::vector<double> zeros(SC_N, 0.0);
    bool zok = hipMemcpyAsync(r->scal, zeros.data(), sizeof(double) * SC_N, hipMemcpyHostToDevice, r->stream) == hipSuccess;
    double *zb[] = { r->u, r->v, r->w, r->c, r->ru, r->rv, r->rw, r->rc, r->fa[0], r->fa[1], r->fa[2], r->fa[3],
                     r->fb[0], r->fb[1], r->fb[2], r->fb[3], r->p };
    for (double *b : zb) if (b) zok = zok && hipMemsetAsync(b, 0, r->esz * r->N, r->stream) == hipSuccess;
    zok = zok && hipMemsetAsync(r->logbuf, 0, sizeof(double) * (size_t) LOGCAP * CF_LOG_N, r->stream) == hipSuccess;
    if (!zok || hipStreamSynchronize(r->stream) != hipSuccess) { free_run(r); cf_set_error("device init failed"); return CF_ERR_DEVICE; }

    /* ---- kernel parameter block ---- */
    RunDev &R = r->rd;
    R.g = r->g;
    R.nu = cfg->nu;
    R.prandtl = cfg->prandtl;
    R.ubc = 0.0 - cfg->uframe;                  /* chanfast.c:59-62 */
    R.vbc = 0.0 - cfg->vframe;
    R.cbot = -0.5; R.ctop = 0.5;                /* chanfast.c:67-68 */
    r->kdiv = pick_div_fftx<double>(nx, order);
    r->kdiv32 = pick_div_fftx<float>(nx, order);
    r->g64 = pick_general<double>(order, cfg->scalar);
    r->g32 = pick_general<float>(order, cfg->scalar);
    R.scalar = cfg->scalar;
    R.betg[0] = cfg->betg_x; R.betg[1] = cfg->betg_y; R.betg[2] = cfg->betg_z;
    R.kap = cfg->scalar ? cfg->nu / cfg->prandtl : cfg->nu;
    R.implicit = cfg->driver == CF_DRIVER_CHANNEL;
    r->kfxp = pick_fftx_plain(nx);
    r->kfy = pick_ffty<double>(ny);
    r->kfxi = pick_fftx_inv<double>(nx);
    r->kfy32 = pick_ffty<float>(ny);
    r->kfxi32 = pick_fftx_inv<float>(nx);
    R.rdx = 1.0 / r->g.dx;
    R.rdy = 1.0 / r->g.dy;
    R.rdxdy = 1.0 / (r->g.dx * r->g.dy);
    R.cmode[0] = cfg->uvolavgmode; R.ctarget[0] = cfg->want_uvolavg - cfg->uframe; R.cfixed[0] = -cfg->dpdx;
    R.cmode[1] = cfg->cvolavgmode; R.ctarget[1] = cfg->want_cvolavg; R.cfixed[1] = 0.0;
    R.logcap = LOGCAP;
    R.fx_pairs = fx_pairs_of(nx, (int) r->esz);
    /* y lines per workgroup: 8 elements per thread (up to 32 short lines) */
    R.fy_cols = MAXE_LINE / ny >= 32 ? 32 : (MAXE_LINE / ny > 0 ? MAXE_LINE / ny : 1);
    /* the two-pass y FFT gives each of a workgroup's 256 threads at most
     * one (line, column) task per pass */
    if (fy2_p1(ny) && R.fy_cols * std::max(fy2_p1(ny), ny / fy2_p1(ny)) > 256) {
        free_run(r);
        cf_set_error("ny = %d: y-FFT workgroup shape", ny);
        return CF_ERR_UNSUPPORTED;
    }
    apply_tables(r, TP);
    r->use_spike = TP.spike_ok;
    r->sp_L = TP.spike_ok ? spike_L(nz) : 0;
    for (int s = 0; s < 3; s++)
        for (int t = 0; t < 3; t++) r->spike_waves = std::max(r->spike_waves, R.sp[s][t].S);
    r->use_spike4 = TP.spike4_ok;
    r->s4_L = TP.spike4_ok ? spike4_L(nz) : 0;
    for (int s = 0; s < 3; s++)
        for (int t = 0; t < 3; t++) r->s4_waves = std::max(r->s4_waves, R.s4[s][t].S);
    R.uoff = r->uoffd;
    r->ubar.assign(nz, 0.0);
    if ((st = set_uoff(r)) != CF_OK) { free_run(r); return st; }
    *out = r;
    return CF_OK;
}

extern "C" cf_status cf_run_destroy(cf_run *r)
{
    if (!r) return CF_OK;
    (void) hipSetDevice(r->cfg.device);
    if (r->stream) (void) hipStreamSynchronize(r->stream);
    free_run(r);
    return CF_OK;
}

/* ------------------------------------------------------------ launches */

static cf_status ev_begin(cf_run *r, const char *name)
{
    if (!r->profile) return CF_OK;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, r->stream));
    r->pending.push_back({ std::string(name), { a, b } });
    return CF_OK;
}

static cf_status ev_end(cf_run *r)
{
    if (!r->profile) return CF_OK;
    HIPCHK(hipEventRecord(r->pending.back().second.second, r->stream));
    return CF_OK;
}

#define KLAUNCH(name, kern, grid, block, lds, ...)                                 \
    do {                                                                            \
        cf_status s_ = ev_begin(r, name);                                           \
        if (s_ != CF_OK) return s_;                                                 \
        hipLaunchKernelGGL(kern, grid, block, lds, r->stream, __VA_ARGS__);         \
        LAUNCHCHK();                                                                \
        s_ = ev_end(r);                                                             \
        if (s_ != CF_OK) return s_;                                                 \
    } while (0)

static dim3 cell_grid(const cf_run *r) { return dim3((r->g.nx + 63) / 64, (r->g.ny + 3) / 4, r->g.nz); }

static size_t fx_lds(const cf_run *r)
{
    return 2 * r->esz * ((size_t) r->rd.fx_pairs * fx_ld(r->g.nx, (int) r->esz) + (size_t) r->g.nx);
}
static size_t fy_lds(const cf_run *r)
{
    return 2 * r->esz * ((size_t) r->rd.fy_cols * fy_ld(r->g.ny) + (size_t) r->g.ny);
}

/* the kernel sets and typed views of the run's buffers for precision T */
template <typename T> struct Prec;
template <> struct Prec<double> {
    static KSel<double>::div_t div(const cf_run *r) { return r->kdiv; }
    static KSel<double>::ffty_t ffty(const cf_run *r) { return r->kfy; }
    static KSel<double>::fftxi_t fftxi(const cf_run *r) { return r->kfxi; }
    static const GSel<double> &gen(const cf_run *r) { return r->g64; }
};
template <> struct Prec<float> {
    static KSel<float>::div_t div(const cf_run *r) { return r->kdiv32; }
    static KSel<float>::ffty_t ffty(const cf_run *r) { return r->kfy32; }
    static KSel<float>::fftxi_t fftxi(const cf_run *r) { return r->kfxi32; }
    static const GSel<float> &gen(const cf_run *r) { return r->g32; }
};
template <typename T> static T *tv(double *p) { return reinterpret_cast<T *>(p); }
template <typename T> static typename Cplx<T>::type *tx(double2 *p) { return reinterpret_cast<typename Cplx<T>::type *>(p); }
static unsigned fy_blocks(const cf_run *r)
{
    long lines = (long) r->g.nz * r->g.nkx;
    return (unsigned) ((lines + r->rd.fy_cols - 1) / r->rd.fy_cols);
}

static void state_set(cf_run *r, int q, double **u, double **v, double **w)
{
    if (q == 0) { *u = r->u; *v = r->v; *w = r->w; }
    else { *u = r->ru; *v = r->rv; *w = r->rw; }
}

/* divergence -> FFT -> Poisson -> inverse FFT into p (pressure.c:40-61,
 * chanfast.c:639-651) */
template <typename T>
static cf_status enqueue_pressure(cf_run *r, const T *us, const T *vs, const T *ws,
                                  const double *h1, const double *mptr, int tq)
{
    typedef typename Cplx<T>::type C;
    C *X = tx<T>(r->X);
    const RunDev &R = r->rd;
    long NL = (long) r->g.ny * r->g.nz;
    long npairs = (NL + 1) / 2;
    unsigned gx = (unsigned) ((npairs + R.fx_pairs - 1) / R.fx_pairs);
    KLAUNCH("div_fftx", Prec<T>::div(r), dim3(gx), dim3(256), fx_lds(r), R, us, vs, ws, h1, mptr, X);
    dim3 gy(fy_blocks(r));
    KLAUNCH("ffty_fwd", Prec<T>::ffty(r), gy, dim3(256), fy_lds(r), R, X, 1);
    if (r->g.order == 4)
        if (r->g.nz >= 6)
            KLAUNCH("poisson", k_poisson4t<T>, dim3((unsigned) ((r->M + 63) / 64)), dim3(128), 0, R, r->pt4[tq], X, r->S4);
        else
            KLAUNCH("poisson", k_poisson4<T>, dim3((unsigned) ((r->M + CF_P4_THREADS - 1) / CF_P4_THREADS)),
                    dim3(CF_P4_THREADS), 0, R, r->pt4[tq], X, r->S4);
    else
    {
        const int L = pois2_L(r->g.nz);
        KLAUNCH("poisson", pick_poisson2<T>(L), dim3(pois2_blocks(L, r->M)), dim3(CF_P2S_THREADS), pois2_lds(L, r->g.nz),
                R, r->pt[tq], X, r->G);
        if constexpr (std::is_same<T, float>::value)   /* the FP64 islands */
            KLAUNCH("poisson", pick_poisson2_islands(L), dim3(pois2_island_blocks(L, r->g.ny, r->M)), dim3(CF_P2S_THREADS),
                    pois2_lds(L, r->g.nz), R, r->pt[tq], (typename Cplx<float>::type *) X, r->G);
    }
    KLAUNCH("ffty_inv", Prec<T>::ffty(r), gy, dim3(256), fy_lds(r), R, X, 0);
    KLAUNCH("fftx_inv", Prec<T>::fftxi(r), dim3(gx), dim3(256), fx_lds(r), R, (const C *) X, tv<T>(r->p));
    return CF_OK;
}

/* the projection parameters in force for the current state: the last
 * substep's h1, forcing slot and (alpha + beta) dt */
template <typename T>
static ProjSrcT<T> current_projection(cf_run *r)
{
    ProjSrcT<T> S;
    double *su, *sv, *sw;
    state_set(r, r->cur, &su, &sv, &sw);
    S.us = tv<T>(su); S.vs = tv<T>(sv); S.ws = tv<T>(sw);
    S.p = tv<T>(r->p);
    S.h1 = r->rd.h1 + (size_t) (2 * 2 + 0) * (r->g.nz + 2);
    S.m = r->scal + SC_M0 + 2;
    S.dtp = r->rd.sc[2].abdt;
    return S;
}

/* materialise the projected velocity of the current state into (u, v, w) */
template <typename T>
static cf_status enqueue_materialise_t(cf_run *r, double *u, double *v, double *w)
{
    ProjSrcT<T> S = current_projection<T>(r);
    KLAUNCH("materialise", k_project2<T>, cell_grid(r), dim3(64, 4), 0, r->rd, S.us, S.vs, S.ws, S.h1, S.m,
            S.p, S.dtp, tv<T>(u), tv<T>(v), tv<T>(w));
    return CF_OK;
}

/* the projected velocity of the fast path's current state into (u, v, w),
 * buffers of the run's storage type */
static cf_status enqueue_materialise(cf_run *r, double *u, double *v, double *w)
{
    return r->prec == CF_FP32 ? enqueue_materialise_t<float>(r, u, v, w) : enqueue_materialise_t<double>(r, u, v, w);
}

/* one RK3 step from state set `from` to the other set */
template <typename T>
static cf_status enqueue_step_t(cf_run *r, int from)
{
    const RunDev &R = r->rd;
    const int nz = r->g.nz;
    const double dt = r->cfg.dt;
    for (int s = 0; s < 3; s++) {
        double **fnew = (s % 2 == 0) ? r->fa : r->fb;
        double **gold = (s % 2 == 0) ? r->fb : r->fa;
        double *iu_, *iv_, *iw_, *ou_, *ov_, *ow_;
        state_set(r, (from + s) % 2, &iu_, &iv_, &iw_);
        state_set(r, (from + s + 1) % 2, &ou_, &ov_, &ow_);
        T *iu = tv<T>(iu_), *iv = tv<T>(iv_), *iw = tv<T>(iw_), *ou = tv<T>(ou_), *ov = tv<T>(ov_), *ow = tv<T>(ow_);
        const int ps = (s + 2) % 3;
        ProjSrcT<T> S;
        S.us = iu; S.vs = iv; S.ws = iw; S.p = tv<T>(r->p);
        S.h1 = R.h1 + (size_t) (ps * 2 + 0) * (nz + 2);
        S.m = r->scal + SC_M0 + ps;
        S.dtp = R.sc[ps].abdt;
        if (s == 0 && r->prev_proj) {       /* the pending projection was made with the previous dt */
            S.h1 = r->h1prev;
            S.dtp = r->dtp_prev;
        }
        /* zeta = 0 (first substep): no previous tendency; the last substep's
         * tendency is never read (the next step starts with zeta = 0) */
        const bool useg = RK_ZET[s] != 0.0, keepf = s < 2;
        KLAUNCH("rhs", k_rhs2t<T>, dim3(r->rhs_blocks), dim3(256), 0, R, s, S,
                (const T *) (useg ? tv<T>(gold[0]) : nullptr), (const T *) tv<T>(gold[1]), (const T *) tv<T>(gold[2]),
                (const double *) r->scal, keepf ? tv<T>(fnew[0]) : nullptr, tv<T>(fnew[1]), tv<T>(fnew[2]), ou, ov, ow,
                s == 0 ? r->cflpart : (double *) nullptr, dt, r->rhs_kc);
        if (s == 0)
            KLAUNCH("cfl_finish", k_cfl_finish, dim3(1), dim3(256), 0, R, (const double *) r->cflpart,
                    r->rhs_blocks, dt, r->scal, r->logbuf);
        const int nbx = (int) ((r->P + 63) / 64);
        if (r->use_spike) {
            /* the solve with the first stage of the plane sums of u* fused in */
            KLAUNCH("spike", pick_spike_sums<T>(r->sp_L), dim3((unsigned) nbx, 3), dim3(64 * r->spike_waves), 0, R, s,
                    ou, ov, ow, r->part);
            /* the level sums and the constraint in one launch (its last
             * workgroup finishes); the last substep also counts the step */
            KLAUNCH("plane_sum", k_level_mass, dim3((unsigned) nz), dim3(256), sizeof(double) * (nz + 2), R, s,
                    (const double *) r->part, nbx, r->lsum, r->scal, r->logbuf, s == 2 ? 1 : 0);
        } else {
            KLAUNCH("thomas", k_thomas<T>, dim3((unsigned) ((r->P + 255) / 256), 3), dim3(256), 0, R, s, ou, ov, ow,
                    (T *) nullptr);
            KLAUNCH("plane_sum", k_gplane<T>, dim3(r->nb_plane, nz, 1), dim3(256), 0, R, (const T *) ou,
                    (const T *) nullptr, r->part, r->nb_plane);
            KLAUNCH("plane_sum", k_level_sum, dim3((unsigned) nz), dim3(256), 0, (const double *) r->part, r->nb_plane,
                    r->lsum);
        }
        if (!r->use_spike)
            KLAUNCH("mass", k_mass_finalize, dim3(1), dim3(256), sizeof(double) * (nz + 2), R, s,
                    (const double *) r->lsum, r->scal, r->logbuf, 0);
        const double *h1 = R.h1 + (size_t) (s * 2 + 0) * (nz + 2);
        cf_status st = enqueue_pressure<T>(r, ou, ov, ow, h1, r->scal + SC_M0 + s, s);
        if (st != CF_OK) return st;
    }
    if (!r->use_spike) KLAUNCH("step_end", k_step_end, dim3(1), dim3(64), 0, r->scal);
    return CF_OK;
}

static cf_status enqueue_step(cf_run *r, int from)
{
    return r->prec == CF_FP32 ? enqueue_step_t<float>(r, from) : enqueue_step_t<double>(r, from);
}

/* The fully implicit update of one field in place (channel2/4, e.g.
 * channel.c:458-467): FFT3D_FWD of its right-hand side h, per (kx, ky) the
 * wall-reduced banded z solve with the modified wavenumbers on the
 * diagonal (solve_helmholtz_u/w, diffuseu.c:460-614, diffusew.c:466-622),
 * FFT3D_BWD.  The x-y passes are the pressure path's; the z solve is the
 * Poisson kernel of the same order on the Helmholtz tables, over nz cells
 * or the nz - 1 interior faces. */
static cf_status enqueue_spectral_solve(cf_run *r, int s, int type, double *fld)
{
    const RunDev &R = r->rd;
    RunDev Rn = R;
    const int n = type == 2 ? r->g.nz - 1 : r->g.nz;
    Rn.g.nz = n;
    double2 *X = r->X;
    const long NL = (long) r->g.ny * r->g.nz, npairs = (NL + 1) / 2;
    const unsigned gx = (unsigned) ((npairs + R.fx_pairs - 1) / R.fx_pairs);
    const dim3 gy(fy_blocks(r));
    KLAUNCH("hsolve", r->kfxp, dim3(gx), dim3(256), fx_lds(r), R, (const double *) fld, X);
    KLAUNCH("hsolve", r->kfy, gy, dim3(256), fy_lds(r), R, X, 1);
    if (r->g.order == 4) {
        if (n >= 6)
            KLAUNCH("hsolve", k_poisson4t<double>, dim3((unsigned) ((r->M + 63) / 64)), dim3(128), 0, Rn, r->hs4[s][type],
                    X, r->S4);
        else
            KLAUNCH("hsolve", k_poisson4<double>, dim3((unsigned) ((r->M + CF_P4_THREADS - 1) / CF_P4_THREADS)),
                    dim3(CF_P4_THREADS), 0, Rn, r->hs4[s][type], X, r->S4);
    } else {
        const int L = pois2_L(n);
        KLAUNCH("hsolve", pick_poisson2<double>(L), dim3(pois2_blocks(L, r->M)), dim3(CF_P2S_THREADS), pois2_lds(L, n),
                Rn, r->hs2[s][type], X, r->G);
    }
    KLAUNCH("hsolve", r->kfy, gy, dim3(256), fy_lds(r), R, X, 0);
    KLAUNCH("hsolve", r->kfxi, dim3(gx), dim3(256), fx_lds(r), R, (const double2 *) X, fld);
    return CF_OK;
}

/* one RK3 step of the general path (order 4, the scalar, FP32 storage or
 * the fully implicit driver): the state (u, v, w, c) is materialised after
 * every substep */
template <typename T>
static cf_status enqueue_step_general_t(cf_run *r)
{
    const RunDev &R = r->rd;
    const int nz = r->g.nz;
    const double dt = r->cfg.dt;
    const bool sc = r->cfg.scalar != 0;
    const dim3 cg = cell_grid(r), cb(64, 4);
    const GSel<T> &K = Prec<T>::gen(r);
    T *u = tv<T>(r->u), *v = tv<T>(r->v), *w = tv<T>(r->w), *c = tv<T>(r->c);
    T *ru = tv<T>(r->ru), *rv = tv<T>(r->rv), *rw = tv<T>(r->rw), *rc = tv<T>(r->rc);
    for (int s = 0; s < 3; s++) {
        double **fnew = (s % 2 == 0) ? r->fa : r->fb;
        double **gold = (s % 2 == 0) ? r->fb : r->fa;
        if (r->r4_nch > 0) {
            /* order 4 on tiles that divide the grid: two z-marching launches */
            const dim3 g4((unsigned) r->r4_blocks);
            if (sc) {
                KLAUNCH("rhs", (k_rhs4_uv<true, T>), g4, dim3(256), 0, R, s, (const T *) u, (const T *) v, (const T *) w,
                        (const T *) c, (const T *) tv<T>(gold[0]), (const T *) tv<T>(gold[1]), (const double *) r->scal,
                        tv<T>(fnew[0]), tv<T>(fnew[1]), ru, rv, s == 0 ? r->cflpart : (double *) nullptr, dt, r->r4_nch);
                KLAUNCH("rhs", (k_rhs4_wc<true, T>), g4, dim3(256), 0, R, s, (const T *) u, (const T *) v, (const T *) w,
                        (const T *) c, (const T *) tv<T>(gold[2]), (const T *) tv<T>(gold[3]), (const double *) r->scal,
                        tv<T>(fnew[2]), tv<T>(fnew[3]), rw, rc, r->r4_nch);
            } else {
                KLAUNCH("rhs", (k_rhs4_uv<false, T>), g4, dim3(256), 0, R, s, (const T *) u, (const T *) v, (const T *) w,
                        (const T *) nullptr, (const T *) tv<T>(gold[0]), (const T *) tv<T>(gold[1]), (const double *) r->scal,
                        tv<T>(fnew[0]), tv<T>(fnew[1]), ru, rv, s == 0 ? r->cflpart : (double *) nullptr, dt, r->r4_nch);
                KLAUNCH("rhs", (k_rhs4_wc<false, T>), g4, dim3(256), 0, R, s, (const T *) u, (const T *) v, (const T *) w,
                        (const T *) nullptr, (const T *) tv<T>(gold[2]), (const T *) nullptr, (const double *) r->scal,
                        tv<T>(fnew[2]), (T *) nullptr, rw, (T *) nullptr, r->r4_nch);
            }
            if (s == 0)
                KLAUNCH("cfl_finish", k_cfl_finish, dim3(1), dim3(256), 0, R, (const double *) r->cflpart, r->r4_blocks,
                        dt, r->scal, r->logbuf);
        } else {
            KLAUNCH("rhs", K.u, cg, cb, 0, R, s, u, v, w, c, (const T *) tv<T>(gold[0]), (const double *) r->scal,
                    tv<T>(fnew[0]), ru, s == 0 ? r->cflpart : (double *) nullptr, dt);
            if (s == 0)
                KLAUNCH("cfl_finish", k_cfl_finish, dim3(1), dim3(256), 0, R, (const double *) r->cflpart, r->gblocks, dt,
                        r->scal, r->logbuf);
            KLAUNCH("rhs", K.v, cg, cb, 0, R, s, u, v, w, c, (const T *) tv<T>(gold[1]), (const double *) r->scal,
                    tv<T>(fnew[1]), rv, (double *) nullptr, dt);
            KLAUNCH("rhs", K.w, cg, cb, 0, R, s, u, v, w, c, (const T *) tv<T>(gold[2]), (const double *) r->scal,
                    tv<T>(fnew[2]), rw, (double *) nullptr, dt);
            if (sc)
                KLAUNCH("rhs", K.c, cg, cb, 0, R, s, u, v, w, c, (const T *) tv<T>(gold[3]), (const double *) r->scal,
                        tv<T>(fnew[3]), rc, (double *) nullptr, dt);
        }
        const unsigned nf = (unsigned) r->nfield;
        if constexpr (std::is_same<T, double>::value) {
            if (R.implicit) {
                cf_status st = enqueue_spectral_solve(r, s, 0, ru);
                if (st == CF_OK) st = enqueue_spectral_solve(r, s, 0, rv);
                if (st == CF_OK && sc) st = enqueue_spectral_solve(r, s, 1, rc);
                if (st == CF_OK) st = enqueue_spectral_solve(r, s, 2, rw);
                if (st != CF_OK) return st;
            }
        }
        if (R.implicit) {
            /* done above */
        } else if (r->g.order == 4 && r->use_spike4)
            KLAUNCH("solve", pick_spike4<T>(r->s4_L), dim3((unsigned) ((r->P + 63) / 64), nf), dim3(64 * r->s4_waves), 0,
                    R, s, ru, rv, rw, rc);
        else if (r->g.order == 4)
            KLAUNCH("solve", k_gsolve4<T>, dim3((unsigned) ((r->P + 255) / 256), nf), dim3(256), 0, R, s, ru, rv, rw, rc);
        else if (r->use_spike)
            KLAUNCH("solve", pick_spike<T>(r->sp_L), dim3((unsigned) ((r->P + 63) / 64), nf), dim3(64 * r->spike_waves), 0, R, s,
                    ru, rv, rw, rc);
        else
            KLAUNCH("solve", k_thomas<T>, dim3((unsigned) ((r->P + 255) / 256), nf), dim3(256), 0, R, s, ru, rv, rw, rc);
        KLAUNCH("plane_sum", k_gplane<T>, dim3(r->nb_plane, nz, sc ? 2 : 1), dim3(256), 0, R, (const T *) ru,
                (const T *) rc, r->part, r->nb_plane);
        KLAUNCH("mass", k_gmass, dim3(sc ? 2 : 1), dim3(256), sizeof(double) * (nz + 6), R, s,
                (const double *) r->part, r->nb_plane, r->scal, r->logbuf);
        const double *h1u = R.h1 + (size_t) (s * 2 + 0) * (nz + 2);
        const double *h1c = R.h1 + (size_t) (s * 2 + 1) * (nz + 2);
        cf_status st = enqueue_pressure<T>(r, ru, rv, rw, h1u, r->scal + SC_M0 + s, s);
        if (st != CF_OK) return st;
        KLAUNCH("project", K.proj, cg, cb, 0, R, (const T *) ru, (const T *) rv, (const T *) rw, (const T *) rc, h1u,
                (const double *) (r->scal + SC_M0 + s), sc ? h1c : (const double *) nullptr,
                (const double *) (r->scal + SC_MC0 + s), (const T *) tv<T>(r->p), R.sc[s].abdt, u, v, w,
                sc ? c : (T *) nullptr);
    }
    KLAUNCH("step_end", k_step_end, dim3(1), dim3(64), 0, r->scal);
    return CF_OK;
}

static cf_status enqueue_step_general(cf_run *r)
{
    return r->prec == CF_FP32 ? enqueue_step_general_t<float>(r) : enqueue_step_general_t<double>(r);
}

static cf_status collect_profile(cf_run *r)
{
    if (!r->profile) return CF_OK;
    HIPCHK(hipStreamSynchronize(r->stream));
    for (auto &pe : r->pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
        auto &acc = r->ktime[pe.first];
        acc.first += ms;
        acc.second += 1;
        (void) hipEventDestroy(pe.second.first);
        (void) hipEventDestroy(pe.second.second);
    }
    r->pending.clear();
    return CF_OK;
}

/* ------------------------------------------------------------ C ABI */

#define CHECK_RUN(r)                                                  \
    do {                                                              \
        if (!(r)) { cf_set_error("null run handle"); return CF_ERR_ARG; } \
        HIPCHK(hipSetDevice((r)->cfg.device));                        \
    } while (0)

/* After an upload or an explicit projection the stored solve output is
 * the velocity itself: zero the pending pressure and forcings.  The other
 * fields of the current state are materialised first so they keep their
 * projected values. */
static cf_status reset_projection(cf_run *r)
{
    if (r->general || !r->proj_pending) return CF_OK;
    double *su, *sv, *sw;
    state_set(r, r->cur, &su, &sv, &sw);
    cf_status st = enqueue_materialise(r, r->fa[0], r->fa[1], r->fa[2]);
    if (st != CF_OK) return st;
    HIPCHK(hipMemcpyAsync(su, r->fa[0], r->esz * r->N, hipMemcpyDeviceToDevice, r->stream));
    HIPCHK(hipMemcpyAsync(sv, r->fa[1], r->esz * r->N, hipMemcpyDeviceToDevice, r->stream));
    HIPCHK(hipMemcpyAsync(sw, r->fa[2], r->esz * r->N, hipMemcpyDeviceToDevice, r->stream));
    HIPCHK(hipMemsetAsync(r->p, 0, r->esz * r->N, r->stream));
    HIPCHK(hipMemsetAsync(r->scal + SC_M0, 0, sizeof(double) * 6, r->stream));
    r->proj_pending = false;
    return CF_OK;
}

extern "C" cf_status cf_run_upload(cf_run *r, int field, const double *host)
{
    CHECK_RUN(r);
    if (!host) { cf_set_error("null host buffer"); return CF_ERR_ARG; }
    double *su, *sv, *sw;
    state_set(r, r->cur, &su, &sv, &sw);
    double *dst = field == CF_U ? su : field == CF_V ? sv : field == CF_W ? sw : (field == CF_C ? r->c : nullptr);
    if (!dst) { cf_set_error("cf_run_upload: field %d is not uploadable in this configuration", field); return CF_ERR_ARG; }
    cf_status rs = reset_projection(r);
    if (rs != CF_OK) return rs;
    const double *off = nullptr;
    if (field == CF_U && r->uoffd) {
        /* the stored deviation's profile: the plane means of the new u */
        const int nx = r->g.nx, ny = r->g.ny, nz = r->g.nz;
        std::vector<double> m((size_t) nz, 0.0);
        for (size_t ij = 0; ij < (size_t) nx * ny; ij++)
            for (int k = 0; k < nz; k++) m[k] += host[ij * nz + k];
        for (int k = 0; k < nz; k++) m[k] /= (double) nx * (double) ny;
        r->ubar.swap(m);
        cf_status so = set_uoff(r);
        if (so != CF_OK) return so;
        off = r->uoffd + 1;
    }
    HIPCHK(hipMemcpyAsync(r->stage, host, sizeof(double) * r->N, hipMemcpyHostToDevice, r->stream));
    dim3 grid((r->g.nz + 31) / 32, (r->g.nx + 31) / 32, r->g.ny);
    if (r->prec == CF_FP32)
        hipLaunchKernelGGL(k_io_to_dev<float>, grid, dim3(256), 0, r->stream, (const double *) r->stage, tv<float>(dst),
                           r->g.nx, r->g.ny, r->g.nz, off);
    else
        hipLaunchKernelGGL(k_io_to_dev<double>, grid, dim3(256), 0, r->stream, (const double *) r->stage, dst,
                           r->g.nx, r->g.ny, r->g.nz, off);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(r->stream));
    return CF_OK;
}

static cf_status download_dev(cf_run *r, const double *src, double *host, const double *off = nullptr)
{
    dim3 grid((r->g.nx + 31) / 32, (r->g.nz + 31) / 32, r->g.ny);
    if (r->prec == CF_FP32)
        hipLaunchKernelGGL(k_dev_to_io<float>, grid, dim3(256), 0, r->stream, (const float *) src, r->stage,
                           r->g.nx, r->g.ny, r->g.nz, off);
    else
        hipLaunchKernelGGL(k_dev_to_io<double>, grid, dim3(256), 0, r->stream, src, r->stage, r->g.nx, r->g.ny, r->g.nz,
                           off);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(host, r->stage, sizeof(double) * r->N, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    return CF_OK;
}

extern "C" cf_status cf_run_download(cf_run *r, int field, double *host)
{
    CHECK_RUN(r);
    if (!host) { cf_set_error("null host buffer"); return CF_ERR_ARG; }
    if (field == CF_P) return download_dev(r, r->p, host);
    if (field == CF_C && r->c) return download_dev(r, r->c, host);
    if (field != CF_U && field != CF_V && field != CF_W) {
        cf_set_error("cf_run_download: field %d not held in this configuration", field);
        return CF_ERR_ARG;
    }
    if (r->general) return download_dev(r, field == CF_U ? r->u : field == CF_V ? r->v : r->w, host);
    cf_status st = enqueue_materialise(r, r->fa[0], r->fa[1], r->fa[2]);
    if (st != CF_OK) return st;
    return download_dev(r, r->fa[field], host, field == CF_U && r->uoffd ? r->uoffd + 1 : nullptr);
}

/* divergence -> Poisson(dt) -> projection of the general path's state, in place */
template <typename T>
static cf_status project_general(cf_run *r, double dt)
{
    T *u = tv<T>(r->u), *v = tv<T>(r->v), *w = tv<T>(r->w);
    cf_status st = enqueue_pressure<T>(r, u, v, w, nullptr, nullptr, 3);
    if (st != CF_OK) return st;
    KLAUNCH("project", Prec<T>::gen(r).proj, cell_grid(r), dim3(64, 4), 0, r->rd, (const T *) u, (const T *) v,
            (const T *) w, (const T *) nullptr, (const double *) nullptr, (const double *) nullptr,
            (const double *) nullptr, (const double *) nullptr, (const T *) tv<T>(r->p), dt, u, v, w, (T *) nullptr);
    return CF_OK;
}

static cf_status set_adhoc_poisson(cf_run *r, double dt)
{
    if (r->pt_dt == dt) return CF_OK;
    if (r->g.order == 4) {
        std::vector<double> T;
        if (pois4_tables(r->dz.data() + CF_NGT, r->g.nz, r->g.dx, r->g.dy, dt, T)) {
            cf_set_error("singular order-4 pressure closure rows");
            return CF_ERR_NUMERIC;
        }
        HIPCHK(hipMemcpy(r->tab + r->off_pois4_adhoc, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice));
    }
    PoisHost ph;
    poisson_tables(r->dz.data() + CF_NGT, r->g.nz, r->g.dx, r->g.dy, dt, ph);
    int nz = r->g.nz;
    std::vector<double> pk;
    for (auto *vv : { &ph.a, &ph.c, &ph.bz, &ph.ex, &ph.ey, &ph.dvol }) pk.insert(pk.end(), vv->begin(), vv->end());
    HIPCHK(hipMemcpy(r->tab + r->off_pois_adhoc, pk.data(), sizeof(double) * 6 * nz, hipMemcpyHostToDevice));
    r->pt[3].db_neu = ph.db_neu; r->pt[3].db_dir = ph.db_dir; r->pt[3].db_top = ph.db_top;
    r->pt_dt = dt;
    return CF_OK;
}

/* the initial projection on the fast path: materialise, divergence ->
 * Poisson(dt) -> projection into the current state set */
template <typename T>
static cf_status project_fast(cf_run *r, double dt)
{
    cf_status st = enqueue_materialise(r, r->fa[0], r->fa[1], r->fa[2]);
    if (st != CF_OK) return st;
    T *a = tv<T>(r->fa[0]), *b = tv<T>(r->fa[1]), *c = tv<T>(r->fa[2]);
    st = enqueue_pressure<T>(r, a, b, c, nullptr, nullptr, 3);
    if (st != CF_OK) return st;
    double *su, *sv, *sw;
    state_set(r, r->cur, &su, &sv, &sw);
    KLAUNCH("project", k_project2<T>, cell_grid(r), dim3(64, 4), 0, r->rd, (const T *) a, (const T *) b,
            (const T *) c, (const double *) nullptr, (const double *) nullptr, (const T *) tv<T>(r->p), dt,
            tv<T>(su), tv<T>(sv), tv<T>(sw));
    return CF_OK;
}

extern "C" cf_status cf_run_project(cf_run *r, double dt)
{
    CHECK_RUN(r);
    if (!(dt > 0)) { cf_set_error("cf_run_project: dt must be positive"); return CF_ERR_ARG; }
    cf_status st = set_adhoc_poisson(r, dt);
    if (st != CF_OK) return st;
    if (r->general) {
        st = r->prec == CF_FP32 ? project_general<float>(r, dt) : project_general<double>(r, dt);
        if (st != CF_OK) return st;
        HIPCHK(hipStreamSynchronize(r->stream));
        return collect_profile(r);
    }
    st = r->prec == CF_FP32 ? project_fast<float>(r, dt) : project_fast<double>(r, dt);
    if (st != CF_OK) return st;
    HIPCHK(hipMemsetAsync(r->p, 0, r->esz * r->N, r->stream));
    HIPCHK(hipMemsetAsync(r->scal + SC_M0, 0, sizeof(double) * 6, r->stream));
    r->proj_pending = false;
    HIPCHK(hipStreamSynchronize(r->stream));
    return collect_profile(r);
}

/* Rebuild every dt-dependent table for a new dt and upload it into the
 * run's table block (same layout).  The stream must be idle: the kernels
 * of earlier steps read the old tables. */
static cf_status set_dt(cf_run *r, double dt)
{
    TabPack TP;
    cf_config c = r->cfg;
    c.dt = dt;
    cf_status st = build_tables(&c, r->dz, r->g.dx, r->g.dy, dt, TP);
    if (st != CF_OK) return st;
    if (TP.pack.size() != r->tab_n || TP.spike_ok != r->use_spike) {
        cf_set_error("table layout changed with dt");
        return CF_ERR_STATE;
    }
    /* the partitioned order-4 solve serves the new dt only if its tables
     * pass their check there; the sequential sweep serves it otherwise */
    r->use_spike4 = TP.spike4_ok && r->s4_L > 0;
    r->tabhost.swap(TP.pack);
    HIPCHK(hipMemcpyAsync(r->tab, r->tabhost.data(), sizeof(double) * r->tab_n, hipMemcpyHostToDevice, r->stream));
    apply_tables(r, TP);
    r->cfg.dt = dt;
    return set_uoff(r);                           /* U's right-hand-side shares depend on dt */
}

/* One step with cflmode 1 (chanfast.c:383-393): the cfl of the current
 * state with the current dt, dt *= want_cfl / cfl, the tables for the new
 * dt, then the step launched directly (no graph: the tables change). */
static cf_status adaptive_step(cf_run *r)
{
    const double *mu = r->u, *mv = r->v, *mw = r->w;
    if (!r->general) {
        cf_status st = enqueue_materialise(r, r->fa[0], r->fa[1], r->fa[2]);
        if (st != CF_OK) return st;
        mu = r->fa[0]; mv = r->fa[1]; mw = r->fa[2];
    }
    const dim3 cg = cell_grid(r);
    if (r->prec == CF_FP32)
        KLAUNCH("cfl", k_cfl_plain<float>, cg, dim3(64, 4), 0, r->rd, (const float *) mu, (const float *) mv,
                (const float *) mw, r->cfg.dt, r->cflpart);
    else
        KLAUNCH("cfl", k_cfl_plain<double>, cg, dim3(64, 4), 0, r->rd, mu, mv, mw, r->cfg.dt, r->cflpart);
    KLAUNCH("cfl", k_cfl_reduce, dim3(1), dim3(256), 0, (const double *) r->cflpart, (int) (cg.x * cg.y * cg.z),
            r->scal);
    double cfl = 0.0;
    HIPCHK(hipMemcpyAsync(&cfl, r->scal + SC_CFL0, sizeof(double), hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    if (!std::isfinite(cfl)) { cf_set_error("non-finite cfl %g", cfl); return CF_ERR_NUMERIC; }
    const double dt = r->cfg.dt * (cfl > 0.0 ? r->cfg.want_cfl / cfl : 1.0);
    r->rd.cfl_log = cfl > 0.0 ? r->cfg.want_cfl : 0.0;
    r->prev_proj = false;
    if (dt != r->cfg.dt) {
        if (!r->general && r->proj_pending) {
            /* the next step projects the current state with the old dt's h1 and (alpha + beta) dt */
            HIPCHK(hipMemcpyAsync(r->h1prev, r->rd.h1 + (size_t) (2 * 2 + 0) * (r->g.nz + 2),
                                  sizeof(double) * (r->g.nz + 2), hipMemcpyDeviceToDevice, r->stream));
            r->dtp_prev = r->rd.sc[2].abdt;
            r->prev_proj = true;
        }
        cf_status st = set_dt(r, dt);
        if (st != CF_OK) return st;
    }
    cf_status st = r->general ? enqueue_step_general(r) : enqueue_step(r, r->cur);
    r->prev_proj = false;
    if (st != CF_OK) return st;
    if (!r->general) {
        r->cur ^= 1;
        r->proj_pending = true;
    }
    r->steps_done += 1;
    /* the graphs hold the old tables' coefficients */
    for (int q = 0; q < 2; q++) {
        if (r->gexec[q]) { (void) hipGraphExecDestroy(r->gexec[q]); r->gexec[q] = nullptr; }
        if (r->graph[q]) { (void) hipGraphDestroy(r->graph[q]); r->graph[q] = nullptr; }
    }
    return CF_OK;
}

extern "C" cf_status cf_run_step(cf_run *r, int nsteps)
{
    CHECK_RUN(r);
    if (nsteps < 0) { cf_set_error("negative step count"); return CF_ERR_ARG; }
    if (nsteps == 0) return CF_OK;
    if (r->cfg.cflmode == 1) {
        for (int n = 0; n < nsteps; n++) {
            cf_status st = adaptive_step(r);
            if (st != CF_OK) return st;
        }
        return collect_profile(r);
    }
    if (r->profile) {
        for (int n = 0; n < nsteps; n++) {
            cf_status st = r->general ? enqueue_step_general(r) : enqueue_step(r, r->cur);
            if (st != CF_OK) return st;
            if (!r->general) {
                r->cur ^= 1;
                r->proj_pending = true;
            }
        }
        r->steps_done += nsteps;
        return collect_profile(r);
    }
    for (int q = 0; q < (r->general ? 1 : 2); q++) {
        if (r->gexec[q]) continue;
        HIPCHK(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
        cf_status st = r->general ? enqueue_step_general(r) : enqueue_step(r, q);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(r->stream, &gr);
        if (st != CF_OK) { if (gr) (void) hipGraphDestroy(gr); return st; }
        HIPCHK(e);
        r->graph[q] = gr;
        HIPCHK(hipGraphInstantiate(&r->gexec[q], r->graph[q], nullptr, nullptr, 0));
    }
    if (r->general) {
        for (int n = 0; n < nsteps; n++) HIPCHK(hipGraphLaunch(r->gexec[0], r->stream));
        r->steps_done += nsteps;
        return CF_OK;
    }
    for (int n = 0; n < nsteps; n++) {
        HIPCHK(hipGraphLaunch(r->gexec[r->cur], r->stream));
        r->cur ^= 1;
    }
    r->proj_pending = true;
    r->steps_done += nsteps;
    return CF_OK;
}

extern "C" cf_status cf_run_sync(cf_run *r)
{
    CHECK_RUN(r);
    HIPCHK(hipStreamSynchronize(r->stream));
    return CF_OK;
}

extern "C" cf_status cf_run_log(cf_run *r, double *out, int max_steps, int *nsteps_out)
{
    CHECK_RUN(r);
    HIPCHK(hipStreamSynchronize(r->stream));
    int have = r->steps_done < LOGCAP ? r->steps_done : LOGCAP;
    int n = have < max_steps ? have : max_steps;
    if (nsteps_out) *nsteps_out = n;
    if (!out || n <= 0) return CF_OK;
    std::vector<double> all((size_t) LOGCAP * CF_LOG_N);
    HIPCHK(hipMemcpy(all.data(), r->logbuf, sizeof(double) * all.size(), hipMemcpyDeviceToHost));
    for (int q = 0; q < n; q++) {
        int step = r->steps_done - n + q;
        memcpy(out + (size_t) q * CF_LOG_N, all.data() + (size_t) (step % LOGCAP) * CF_LOG_N,
               sizeof(double) * CF_LOG_N);
    }
    return CF_OK;
}

extern "C" cf_status cf_run_max_divergence(cf_run *r, double *out)
{
    CHECK_RUN(r);
    if (!out) { cf_set_error("null output"); return CF_ERR_ARG; }
    const double *mu = r->u, *mv = r->v, *mw = r->w;
    if (!r->general) {
        cf_status st = enqueue_materialise(r, r->fa[0], r->fa[1], r->fa[2]);
        if (st != CF_OK) return st;
        mu = r->fa[0]; mv = r->fa[1]; mw = r->fa[2];
    }
    if (r->prec == CF_FP32)
        hipLaunchKernelGGL(k_maxdiv_part<float>, dim3(r->nb_red), dim3(256), 0, r->stream, r->rd, (const float *) mu,
                           (const float *) mv, (const float *) mw, r->part, r->nb_red);
    else
        hipLaunchKernelGGL(k_maxdiv_part<double>, dim3(r->nb_red), dim3(256), 0, r->stream, r->rd, mu, mv, mw, r->part,
                           r->nb_red);
    LAUNCHCHK();
    std::vector<double> part(r->nb_red);
    HIPCHK(hipMemcpyAsync(part.data(), r->part, sizeof(double) * r->nb_red, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    double m = 0.0;
    for (double x : part) m = fmax(m, x);
    *out = m;
    return CF_OK;
}

/* ------------------------------------------------------------ statistics */

/* the device buffer of a field of the current state (storage precision of
 * the run); the fast path's velocity is materialised into the forcing
 * slots first, which are free between steps */
static cf_status stat_source(cf_run *r, int field, const double **src, bool *mat)
{
    if (field == CF_P) { *src = r->p; return CF_OK; }
    if (field == CF_C) {
        if (!r->c) { cf_set_error("statistics: field c needs a scalar run"); return CF_ERR_ARG; }
        *src = r->c;
        return CF_OK;
    }
    if (field < CF_U || field > CF_W) { cf_set_error("statistics: bad field %d", field); return CF_ERR_ARG; }
    if (r->general) {
        *src = field == CF_U ? r->u : field == CF_V ? r->v : r->w;
        return CF_OK;
    }
    if (!*mat) {
        cf_status st = enqueue_materialise(r, r->fa[0], r->fa[1], r->fa[2]);
        if (st != CF_OK) return st;
        *mat = true;
    }
    *src = r->fa[field];
    return CF_OK;
}

/* a plane statistic of kind AVG, CORR or MOMENT into dev_out[nz], with
 * part (CF_STAT_NBL nz doubles) and means (2 nz doubles) as device work
 * space; enqueued on the run's stream */
template <typename T>
static cf_status stat_plane_t(cf_run *r, int kind, const T *a, const T *b, double en, double *part, double *means,
                              double *dev_out, const double *aoff)
{
    const int nz = r->g.nz, P = (int) r->P, nbl = CF_STAT_NBL;
    const double scale = 1.0 / ((double) r->g.nx * (double) r->g.ny);
    const dim3 grid(nbl, nz), fin((nz + 255) / 256);
    double *ma = means, *mb = means + nz;
    const double *nul = nullptr;
    hipLaunchKernelGGL((k_stat_plane<T, 0>), grid, dim3(256), 0, r->stream, a, (const T *) nullptr, nul, nul, 0.0, P,
                       nbl, part);
    /* a stored relative to a profile: the average adds it back; the
     * central statistics need only the deviations */
    hipLaunchKernelGGL(k_stat_finish, fin, dim3(256), 0, r->stream, (const double *) part, nbl, nz, scale,
                       kind == CF_STAT_AVG ? dev_out : ma, kind == CF_STAT_AVG ? aoff : nul);
    if (kind == CF_STAT_CORR) {
        hipLaunchKernelGGL((k_stat_plane<T, 0>), grid, dim3(256), 0, r->stream, b, (const T *) nullptr, nul, nul, 0.0,
                           P, nbl, part);
        hipLaunchKernelGGL(k_stat_finish, fin, dim3(256), 0, r->stream, (const double *) part, nbl, nz, scale, mb, nul);
        hipLaunchKernelGGL((k_stat_plane<T, 1>), grid, dim3(256), 0, r->stream, a, b, (const double *) ma,
                           (const double *) mb, 0.0, P, nbl, part);
        hipLaunchKernelGGL(k_stat_finish, fin, dim3(256), 0, r->stream, (const double *) part, nbl, nz, scale, dev_out, nul);
    } else if (kind == CF_STAT_MOMENT) {
        hipLaunchKernelGGL((k_stat_plane<T, 2>), grid, dim3(256), 0, r->stream, a, (const T *) nullptr,
                           (const double *) ma, nul, en, P, nbl, part);
        hipLaunchKernelGGL(k_stat_finish, fin, dim3(256), 0, r->stream, (const double *) part, nbl, nz, scale, dev_out, nul);
    }
    LAUNCHCHK();
    return CF_OK;
}

static cf_status stat_plane(cf_run *r, int kind, int fa, int fb, double en, double *part, double *means,
                            double *dev_out)
{
    bool mat = false;
    const double *a = nullptr, *b = nullptr;
    cf_status st = stat_source(r, fa, &a, &mat);
    if (st == CF_OK && kind == CF_STAT_CORR) st = stat_source(r, fb, &b, &mat);
    if (st != CF_OK) return st;
    const double *aoff = fa == CF_U && r->uoffd ? r->uoffd + 1 : nullptr;
    if (r->prec == CF_FP32)
        return stat_plane_t<float>(r, kind, (const float *) a, (const float *) b, en, part, means, dev_out, aoff);
    return stat_plane_t<double>(r, kind, a, b, en, part, means, dev_out, aoff);
}

/* FFT3D_FWD (1/nx, 1/ny scaled) of an FP64 field in the device layout into
 * dst [nz][nkx][ny]: the run path's x and y FFT kernels */
static cf_status stat_fft(cf_run *r, const double *src, double2 *dst)
{
    const RunDev &R = r->rd;
    const long NL = (long) r->g.ny * r->g.nz, npairs = (NL + 1) / 2;
    const unsigned gx = (unsigned) ((npairs + R.fx_pairs - 1) / R.fx_pairs);
    const size_t lx = 2 * sizeof(double) * ((size_t) R.fx_pairs * fx_ld(r->g.nx, (int) r->esz) + (size_t) r->g.nx);
    const size_t ly = 2 * sizeof(double) * ((size_t) R.fy_cols * fy_ld(r->g.ny) + (size_t) r->g.ny);
    hipLaunchKernelGGL(r->kfxp, dim3(gx), dim3(256), lx, r->stream, R, src, dst);
    hipLaunchKernelGGL(r->kfy, dim3(fy_blocks(r)), dim3(256), ly, r->stream, R, dst, 1);
    LAUNCHCHK();
    return CF_OK;
}

extern "C" cf_status cf_run_xy_stats(cf_run *r, int kind, int fa, int fb, double en, double *out)
{
    CHECK_RUN(r);
    if (!out || kind < CF_STAT_AVG || kind > CF_STAT_SPEC) { cf_set_error("cf_run_xy_stats: bad kind or output"); return CF_ERR_ARG; }
    const int nz = r->g.nz;
    cf_status st;
    if (kind != CF_STAT_SPEC) {
        double *work = nullptr;
        HIPCHK(hipMalloc((void **) &work, sizeof(double) * ((size_t) CF_STAT_NBL * nz + 3 * (size_t) nz)));
        st = stat_plane(r, kind, fa, fb, en, work, work + (size_t) CF_STAT_NBL * nz,
                        work + (size_t) CF_STAT_NBL * nz + 2 * nz);
        if (st == CF_OK && hipMemcpyAsync(out, work + (size_t) CF_STAT_NBL * nz + 2 * nz, sizeof(double) * nz,
                                          hipMemcpyDeviceToHost, r->stream) != hipSuccess) {
            cf_set_error("cf_run_xy_stats: copy failed");
            st = CF_ERR_DEVICE;
        }
        (void) hipStreamSynchronize(r->stream);
        (void) hipFree(work);
        return st;
    }
    /* spectra: both fields widened to FP64 (FP32 runs), transformed, the
     * level sums of the cross-spectrum */
    const int nkx = r->g.nkx, ny = r->g.ny;
    const size_t M = r->M * (size_t) nz, nsx = (size_t) nkx * nz * 2, nsy = (size_t) (ny / 2 + 1) * nz * 2;
    double2 *X2 = nullptr;
    double *spec = nullptr, *wide = nullptr;
    bool ok = hipMalloc((void **) &X2, sizeof(double2) * 2 * M) == hipSuccess
           && hipMalloc((void **) &spec, sizeof(double) * (nsx + nsy)) == hipSuccess
           && hipMalloc((void **) &wide, sizeof(double) * r->N) == hipSuccess;
    st = ok ? CF_OK : CF_ERR_NOMEM;
    if (!ok) cf_set_error("cf_run_xy_stats: device allocation failed");
    bool mat = false;
    for (int q = 0; q < 2 && st == CF_OK; q++) {
        const double *src = nullptr;
        st = stat_source(r, q ? fb : fa, &src, &mat);
        if (st != CF_OK) break;
        if (r->prec == CF_FP32) {
            hipLaunchKernelGGL(k_stat_widen<float>, dim3((unsigned) ((r->N + 255) / 256)), dim3(256), 0, r->stream,
                               (const float *) src, wide, r->N);
            src = wide;
        }
        st = stat_fft(r, src, X2 + q * M);
    }
    if (st == CF_OK) {
        hipLaunchKernelGGL(k_stat_spec, dim3(nz), dim3(256), 0, r->stream, (const double2 *) X2,
                           (const double2 *) (X2 + M), r->g.nx, ny, nkx, nz, spec, spec + nsx);
        if (hipGetLastError() != hipSuccess
            || hipMemcpyAsync(out, spec, sizeof(double) * (nsx + nsy), hipMemcpyDeviceToHost, r->stream) != hipSuccess) {
            cf_set_error("cf_run_xy_stats: spectra failed");
            st = CF_ERR_DEVICE;
        }
    }
    (void) hipStreamSynchronize(r->stream);
    if (X2) (void) hipFree(X2);
    if (spec) (void) hipFree(spec);
    if (wide) (void) hipFree(wide);
    return st;
}

/* for entry points in other files: a plane statistic of kind AVG, CORR or
 * MOMENT into res[nz] with the caller's work space (part: CF_STAT_NBL nz
 * doubles, means: 2 nz), enqueued on the run's stream; nz, the stream and
 * the device come back for the caller */
extern "C" cf_status cf_run_stat_plane_into(cf_run *r, int kind, int fa, int fb, double en, double *part,
                                           double *means, double *res, int *nz, void **stream, int *device)
{
    CHECK_RUN(r);
    if (nz) *nz = r->g.nz;
    if (stream) *stream = (void *) r->stream;
    if (device) *device = r->cfg.device;
    if (!part) return CF_OK;                      /* shape query only */
    /* profiling would create events: off for this call */
    const bool prof = r->profile;
    r->profile = false;
    const cf_status st = stat_plane(r, kind, fa, fb, en, part, means, res);
    r->profile = prof;
    return st;
}

extern "C" cf_status cf_run_info(cf_run *r, double *dt, int *steps_done)
{
    if (!r) { cf_set_error("null run handle"); return CF_ERR_ARG; }
    if (dt) *dt = r->cfg.dt;
    if (steps_done) *steps_done = r->steps_done;
    return CF_OK;
}

extern "C" cf_status cf_run_set_profiling(cf_run *r, int on)
{
    if (!r) { cf_set_error("null run handle"); return CF_ERR_ARG; }
    r->profile = on != 0;
    r->ktime.clear();
    return CF_OK;
}

extern "C" cf_status cf_run_kernel_time(cf_run *r, const char *kernel, double *avg_ms, int *launches)
{
    if (!r || !kernel) { cf_set_error("null argument"); return CF_ERR_ARG; }
    auto it = r->ktime.find(kernel);
    if (it == r->ktime.end()) { if (avg_ms) *avg_ms = 0; if (launches) *launches = 0; return CF_OK; }
    if (avg_ms) *avg_ms = it->second.first / (it->second.second ? it->second.second : 1);
    if (launches) *launches = it->second.second;
    return CF_OK;
}

/* ------------------------------------------------------------ operators */

extern "C" cf_status cf_op_transpose_roundtrip(cf_run *r, const double *in_io, double *dev_out, double *io_out)
{
    CHECK_RUN(r);
    if (r->prec != CF_FP64) { cf_set_error("operator entry points take FP64 runs"); return CF_ERR_UNSUPPORTED; }
    if (!in_io) { cf_set_error("null input"); return CF_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(r->stage, in_io, sizeof(double) * r->N, hipMemcpyHostToDevice, r->stream));
    dim3 grid((r->g.nz + 31) / 32, (r->g.nx + 31) / 32, r->g.ny);
    hipLaunchKernelGGL(k_io_to_dev<double>, grid, dim3(256), 0, r->stream, (const double *) r->stage, r->fb[0],
                       r->g.nx, r->g.ny, r->g.nz, (const double *) nullptr);
    LAUNCHCHK();
    if (dev_out) HIPCHK(hipMemcpyAsync(dev_out, r->fb[0], sizeof(double) * r->N, hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    if (io_out) return download_dev(r, r->fb[0], io_out);
    return CF_OK;
}

/* host [ix][iy][iz][2] <-> device [iz][ix][iy] complex */
static void spec_host_to_dev(const cf_run *r, const double *h, std::vector<double2> &d)
{
    int nkx = r->g.nkx, ny = r->g.ny, nz = r->g.nz;
    d.resize(r->M * nz);
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < ny; iy++)
            for (int iz = 0; iz < nz; iz++) {
                size_t o = (((size_t) ix * ny + iy) * nz + iz) * 2;
                d[((size_t) iz * nkx + ix) * ny + iy] = make_double2(h[o], h[o + 1]);
            }
}

static void spec_dev_to_host(const cf_run *r, const std::vector<double2> &d, double *h)
{
    int nkx = r->g.nkx, ny = r->g.ny, nz = r->g.nz;
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < ny; iy++)
            for (int iz = 0; iz < nz; iz++) {
                size_t o = (((size_t) ix * ny + iy) * nz + iz) * 2;
                double2 val = d[((size_t) iz * nkx + ix) * ny + iy];
                h[o] = val.x; h[o + 1] = val.y;
            }
}

extern "C" cf_status cf_op_fft3d_fwd(cf_run *r, const double *in_io, double *spec_out)
{
    CHECK_RUN(r);
    if (r->prec != CF_FP64) { cf_set_error("operator entry points take FP64 runs"); return CF_ERR_UNSUPPORTED; }
    if (!in_io || !spec_out) { cf_set_error("null argument"); return CF_ERR_ARG; }
    HIPCHK(hipMemcpyAsync(r->stage, in_io, sizeof(double) * r->N, hipMemcpyHostToDevice, r->stream));
    dim3 grid((r->g.nz + 31) / 32, (r->g.nx + 31) / 32, r->g.ny);
    hipLaunchKernelGGL(k_io_to_dev<double>, grid, dim3(256), 0, r->stream, (const double *) r->stage, r->fb[0],
                       r->g.nx, r->g.ny, r->g.nz, (const double *) nullptr);
    LAUNCHCHK();
    const RunDev &R = r->rd;
    long NL = (long) r->g.ny * r->g.nz, npairs = (NL + 1) / 2;
    unsigned gx = (unsigned) ((npairs + R.fx_pairs - 1) / R.fx_pairs);
    hipLaunchKernelGGL(r->kfxp, dim3(gx), dim3(256), fx_lds(r), r->stream, R, (const double *) r->fb[0], r->X);
    LAUNCHCHK();
    dim3 gy(fy_blocks(r));
    hipLaunchKernelGGL(r->kfy, gy, dim3(256), fy_lds(r), r->stream, R, r->X, 1);
    LAUNCHCHK();
    std::vector<double2> d(r->M * r->g.nz);
    HIPCHK(hipMemcpyAsync(d.data(), r->X, sizeof(double2) * d.size(), hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    spec_dev_to_host(r, d, spec_out);
    return CF_OK;
}

extern "C" cf_status cf_op_fft3d_bwd(cf_run *r, const double *spec_in, double *out_io)
{
    CHECK_RUN(r);
    if (r->prec != CF_FP64) { cf_set_error("operator entry points take FP64 runs"); return CF_ERR_UNSUPPORTED; }
    if (!spec_in || !out_io) { cf_set_error("null argument"); return CF_ERR_ARG; }
    std::vector<double2> d;
    spec_host_to_dev(r, spec_in, d);
    HIPCHK(hipMemcpyAsync(r->X, d.data(), sizeof(double2) * d.size(), hipMemcpyHostToDevice, r->stream));
    const RunDev &R = r->rd;
    dim3 gy(fy_blocks(r));
    hipLaunchKernelGGL(r->kfy, gy, dim3(256), fy_lds(r), r->stream, R, r->X, 0);
    LAUNCHCHK();
    long NL = (long) r->g.ny * r->g.nz, npairs = (NL + 1) / 2;
    unsigned gx = (unsigned) ((npairs + R.fx_pairs - 1) / R.fx_pairs);
    hipLaunchKernelGGL(r->kfxi, dim3(gx), dim3(256), fx_lds(r), r->stream, R, (const double2 *) r->X, r->fb[0]);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(r->stream));
    return download_dev(r, r->fb[0], out_io);
}

extern "C" cf_status cf_op_poisson(cf_run *r, double *spec, double dt)
{
    CHECK_RUN(r);
    if (r->prec != CF_FP64) { cf_set_error("operator entry points take FP64 runs"); return CF_ERR_UNSUPPORTED; }
    if (!spec || !(dt > 0)) { cf_set_error("bad argument"); return CF_ERR_ARG; }
    cf_status st = set_adhoc_poisson(r, dt);
    if (st != CF_OK) return st;
    std::vector<double2> d;
    spec_host_to_dev(r, spec, d);
    HIPCHK(hipMemcpyAsync(r->X, d.data(), sizeof(double2) * d.size(), hipMemcpyHostToDevice, r->stream));
    if (r->g.order == 4) {
        if (r->g.nz >= 6)
            hipLaunchKernelGGL(k_poisson4t<double>, dim3((unsigned) ((r->M + 63) / 64)), dim3(128), 0, r->stream,
                               r->rd, r->pt4[3], r->X, r->S4);
        else
            hipLaunchKernelGGL(k_poisson4<double>, dim3((unsigned) ((r->M + CF_P4_THREADS - 1) / CF_P4_THREADS)),
                               dim3(CF_P4_THREADS), 0, r->stream, r->rd, r->pt4[3], r->X, r->S4);
    } else
        hipLaunchKernelGGL(pick_poisson2<double>(pois2_L(r->g.nz)), dim3(pois2_blocks(pois2_L(r->g.nz), r->M)),
                           dim3(CF_P2S_THREADS), pois2_lds(pois2_L(r->g.nz), r->g.nz), r->stream,
                           r->rd, r->pt[3], r->X, r->G);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(d.data(), r->X, sizeof(double2) * d.size(), hipMemcpyDeviceToHost, r->stream));
    HIPCHK(hipStreamSynchronize(r->stream));
    spec_dev_to_host(r, d, spec);
    return CF_OK;
}

/* ------------------------------------------------------------ batches
 *
 * A batch steps runs of one (order, scalar, precision) with any grid
 * sizes, viscosities and time steps together: every stage of the substep
 * is one persistent launch over the virtual blocks of all runs (the
 * kb_* kernels), the whole RK3 step one captured graph.  The persistent
 * grid of each launch and, for order 2, the z solve (several wavefronts
 * per system, or one lane per system) are chosen by timing the
 * candidates on the batch's own shapes at creation, while every field is
 * still zero (each stage then maps zeros to zeros, so tuning leaves the
 * state untouched).
 */

enum { KB_RHS_U = 0, KB_RHS_V, KB_RHS_W, KB_RHS_C, KB_SOLVE, KB_PLANE, KB_MASS, KB_CFL, KB_DIVX, KB_FFTY,
       KB_POIS, KB_FFTXI, KB_PROJ, KB_STEPEND, KB_FRHS, KB_FSPIKE, KB_FLEVEL, KB_FMASS, KB_N };
static const int KB_STAGE[KB_N] = { ST_CELL, ST_CELL, ST_CELL, ST_CELL, ST_SOLVE, ST_PLANE, ST_MASS, ST_ONE,
                                    ST_FFTX, ST_FFTY, ST_MODE, ST_FFTX, ST_CELL, ST_ONE,
                                    ST_RHS2, ST_SOLVE, ST_LEVEL, ST_ONE };
static const char *KB_NAME[KB_N] = { "rhs", "rhs", "rhs", "rhs", "solve", "plane_sum", "mass", "cfl_finish",
                                     "div_fftx", "ffty", "poisson", "fftx_inv", "project", "step_end",
                                     "rhs", "spike", "plane_sum", "mass" };
enum { SOLVE_HEPT = 0, SOLVE_SPIKE = 1, SOLVE_THOMAS = 2 };

struct cf_batch {
    std::vector<cf_run *> runs;
    int n = 0, prec = CF_FP64, order = 2, scalar = 0, device = 0;
    hipStream_t stream = nullptr;
    void *args = nullptr;                 /* device GArgs<T>[n] */
    unsigned *vstart = nullptr;           /* device [CF_VS_ROWS][n + 1] */
    std::vector<unsigned> hv;
    int solve_kind = SOLVE_HEPT;
    std::vector<int> p2L;                 /* distinct order-2 Poisson chunk lengths, one launch each */
    std::vector<int> spL, spW;            /* distinct spike chunk lengths (one launch each) and their waves */
    std::vector<int> fxK, fyK;            /* distinct x / y FFT plan keys (bfft_key), one launch each */
    unsigned kgrid[KB_N] = {0};
    int spike_waves = 1, ncu = 256;
    size_t lds_fx = 0, lds_fy = 0, lds_mass = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    /* order-2 runs (FP64 or FP32) without the scalar step on the fast path's fused
     * kernels: two argument tables (state sets swapped), one graph per
     * starting set */
    bool fast = false;
    int cur = 0, tab = 0;
    hipGraph_t fgraph[2] = { nullptr, nullptr };
    hipGraphExec_t fexec[2] = { nullptr, nullptr };
    int steps_done = 0;
    double dt_adhoc = -1.0;
    bool profile = false;
    std::map<std::string, std::pair<double, int>> ktime;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<std::string> tuning;      /* human-readable autotuner choices */
};

/* The order-2 Poisson of a batch goes out as one persistent launch per
 * distinct chunk length L (pois2_L of each run's nz), so that every run is
 * solved with its own partition, bit for bit as when stepped alone; the
 * prefix table of group g is row CF_NST + g of vstart. */
#define CF_P2_NG 20
#define CF_SP_NG 11                        /* spike launches: rows CF_NST + CF_P2_NG + g */
#define CF_P2I_ROW0 (CF_NST + CF_P2_NG + CF_SP_NG)  /* FP32 Poisson islands: rows CF_P2I_ROW0 + g */
#define CF_FFT_NG 8                        /* FFT length groups: bfft_key values (7 sizes and 0) */
#define CF_FX_ROW0 (CF_P2I_ROW0 + CF_P2_NG)      /* x-FFT groups (by nx): rows CF_FX_ROW0 + g */
#define CF_FY_ROW0 (CF_FX_ROW0 + CF_FFT_NG)      /* y-FFT groups (by ny) */
#define CF_TH_ROW (CF_FY_ROW0 + CF_FFT_NG)       /* order-2 runs without spike tables: Thomas */
#define CF_VS_ROWS (CF_TH_ROW + 1)
/* an order-2 run solves its z systems as it does alone: the spike where
 * its tables exist, Thomas otherwise (the two round differently) */
static bool run_uses_spike(const cf_batch *b, const cf_run *r) { return b->order == 2 && r->use_spike; }

static void batch_spike_groups(cf_batch *b)
{
    b->spL.clear();
    b->spW.clear();
    for (const cf_run *r : b->runs) {
        if (!run_uses_spike(b, r)) continue;
        const auto it = std::find(b->spL.begin(), b->spL.end(), r->sp_L);
        if (it == b->spL.end()) { b->spL.push_back(r->sp_L); b->spW.push_back(r->spike_waves); }
        else { int &w = b->spW[it - b->spL.begin()]; w = std::max(w, r->spike_waves); }
    }
}
static std::vector<int> batch_p2L(const std::vector<cf_run *> &runs)
{
    std::vector<int> ls;
    for (const cf_run *r : runs) {
        const int l = pois2_L(r->g.nz);
        if (std::find(ls.begin(), ls.end(), l) == ls.end()) ls.push_back(l);
    }
    std::sort(ls.begin(), ls.end());
    return ls;
}

static void batch_fft_groups(cf_batch *b)
{
    b->fxK.clear();
    b->fyK.clear();
    for (const cf_run *r : b->runs) {
        const int kx = bfft_key((int) r->g.nx), ky = bfft_key((int) r->g.ny);
        if (std::find(b->fxK.begin(), b->fxK.end(), kx) == b->fxK.end()) b->fxK.push_back(kx);
        if (std::find(b->fyK.begin(), b->fyK.end(), ky) == b->fyK.end()) b->fyK.push_back(ky);
    }
}

template <typename T>
static void batch_fill_args(cf_batch *b, std::vector<GArgs<T>> &A)
{
    A.assign(b->n, GArgs<T>());
    for (int q = 0; q < b->n; q++) {
        cf_run *r = b->runs[q];
        GArgs<T> &a = A[q];
        a.R = r->rd;
        a.u = tv<T>(r->u); a.v = tv<T>(r->v); a.w = tv<T>(r->w); a.c = tv<T>(r->c);
        a.ru = tv<T>(r->ru); a.rv = tv<T>(r->rv); a.rw = tv<T>(r->rw); a.rc = tv<T>(r->rc);
        for (int f = 0; f < 4; f++) { a.fa[f] = tv<T>(r->fa[f]); a.fb[f] = tv<T>(r->fb[f]); }
        a.p = tv<T>(r->p);
        a.X = tx<T>(r->X);
        a.G = r->G; a.S4 = r->S4; a.scal = r->scal; a.logbuf = r->logbuf; a.part = r->part; a.cflpart = r->cflpart;
        a.lsum = r->lsum;
        a.dt = r->cfg.dt;
        a.dt_adhoc = r->pt_dt;
        a.nb_plane = r->nb_plane;
        /* a batch has workgroups to spare: longer RHS z-chunks (fewer halo
         * re-reads) than a run alone; the chunking does not change results */
        int bkc = 1, bblocks = 0;
        rhs_chunks((int) r->g.nx, (int) r->g.ny, (int) r->g.nz, CF_RHS_KC_BATCH, &bkc, &bblocks);
        a.gblocks = b->fast ? bblocks : r->gblocks;
        a.rhs_kc = bkc;
        for (int t = 0; t < 4; t++) { a.pt[t] = r->pt[t]; a.pt4[t] = r->pt4[t]; }
        const unsigned nx = r->g.nx, ny = r->g.ny, nz = r->g.nz;
        const unsigned P = (unsigned) r->P, nf = (unsigned) r->nfield, kinds = r->cfg.scalar ? 2u : 1u;
        auto set = [&](int st, unsigned x, unsigned y, unsigned z) { a.dims[st][0] = x; a.dims[st][1] = y; a.dims[st][2] = z; };
        set(ST_CELL, (nx + 63) / 64, (ny + 3) / 4, nz);
        if (run_uses_spike(b, r)) set(ST_SOLVE, (P + 63) / 64, nf, 1);
        else set(ST_SOLVE, (P + 255) / 256, nf, 1);
        set(ST_PLANE, (unsigned) r->nb_plane, nz, kinds);
        set(ST_MASS, kinds, 1, 1);
        set(ST_ONE, 1, 1, 1);
        long NL = (long) ny * nz, npairs = (NL + 1) / 2;
        set(ST_FFTX, (unsigned) ((npairs + r->rd.fx_pairs - 1) / r->rd.fx_pairs), 1, 1);
        set(ST_FFTY, fy_blocks(r), 1, 1);
        set(ST_MODE, b->order == 2 ? pois2_blocks(pois2_L(r->g.nz), r->M) : (unsigned) ((r->M + 255) / 256), 1, 1);
        set(ST_RHS2, (unsigned) std::max(bblocks, 1), 1, 1);
        set(ST_LEVEL, nz, 1, 1);
    }
}

template <typename T>
static cf_status batch_upload_tables(cf_batch *b)
{
    std::vector<GArgs<T>> A;
    batch_fill_args<T>(b, A);
    const int ng = (int) b->p2L.size();
    if (ng > CF_P2_NG) { cf_set_error("too many Poisson chunk lengths in one batch"); return CF_ERR_UNSUPPORTED; }
    batch_spike_groups(b);
    const int nsg = (int) b->spL.size();
    if (nsg > CF_SP_NG) { cf_set_error("too many spike chunk lengths in one batch"); return CF_ERR_UNSUPPORTED; }
    batch_fft_groups(b);                  /* at most CF_FFT_NG keys each, by construction */
    const int nrows = CF_VS_ROWS;
    b->hv.assign((size_t) nrows * (b->n + 1), 0u);
    for (int st = 0; st < nrows; st++) {
        unsigned acc = 0;
        for (int q = 0; q < b->n; q++) {
            b->hv[(size_t) st * (b->n + 1) + q] = acc;
            int sd = st;
            bool in = true;
            if (st == CF_TH_ROW) {                /* Thomas runs of an order-2 batch */
                if (b->order == 2 && !run_uses_spike(b, b->runs[q]))
                    acc += A[q].dims[ST_SOLVE][0] * A[q].dims[ST_SOLVE][1] * A[q].dims[ST_SOLVE][2];
                continue;
            }
            if (st >= CF_FX_ROW0) {               /* FFT length group */
                const bool y = st >= CF_FY_ROW0;
                const int g = st - (y ? CF_FY_ROW0 : CF_FX_ROW0);
                const std::vector<int> &K = y ? b->fyK : b->fxK;
                const cf_run *rq = b->runs[q];
                if (g < (int) K.size() && bfft_key((int) (y ? rq->g.ny : rq->g.nx)) == K[g]) {
                    const int sd2 = y ? ST_FFTY : ST_FFTX;
                    acc += A[q].dims[sd2][0] * A[q].dims[sd2][1] * A[q].dims[sd2][2];
                }
                continue;
            }
            if (st >= CF_P2I_ROW0) {              /* FP32 Poisson island group */
                const int g = st - CF_P2I_ROW0;
                const cf_run *rq = b->runs[q];
                if (g < ng && pois2_L(rq->g.nz) == b->p2L[g])
                    acc += pois2_island_blocks(b->p2L[g], rq->g.ny, rq->M);
                continue;
            }
            if (st >= CF_NST + CF_P2_NG) {        /* spike group */
                const int g = st - CF_NST - CF_P2_NG;
                sd = ST_SOLVE;
                in = g < nsg && b->runs[q]->sp_L == b->spL[g];
            } else if (st >= CF_NST) {            /* Poisson group */
                const int g = st - CF_NST;
                sd = ST_MODE;
                in = g < ng && pois2_L(b->runs[q]->g.nz) == b->p2L[g];
            }
            if (in) acc += A[q].dims[sd][0] * A[q].dims[sd][1] * A[q].dims[sd][2];
        }
        b->hv[(size_t) st * (b->n + 1) + b->n] = acc;
    }
    HIPCHK(hipMemcpy(b->args, A.data(), sizeof(GArgs<T>) * A.size(), hipMemcpyHostToDevice));
    if (b->fast) {
        /* table 1: the other state set is the substep's input */
        for (GArgs<T> &a : A) {
            std::swap(a.u, a.ru);
            std::swap(a.v, a.rv);
            std::swap(a.w, a.rw);
        }
        HIPCHK(hipMemcpy((GArgs<T> *) b->args + b->n, A.data(), sizeof(GArgs<T>) * A.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(b->vstart, b->hv.data(), sizeof(unsigned) * b->hv.size(), hipMemcpyHostToDevice));
    return CF_OK;
}

static unsigned batch_total(const cf_batch *b, int st) { return b->hv[(size_t) st * (b->n + 1) + b->n]; }

static cf_status bev_begin(cf_batch *b, const char *name)
{
    if (!b->profile) return CF_OK;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, b->stream));
    b->pending.push_back({ std::string(name), { e0, e1 } });
    return CF_OK;
}

static cf_status bev_end(cf_batch *b)
{
    if (!b->profile) return CF_OK;
    HIPCHK(hipEventRecord(b->pending.back().second.second, b->stream));
    return CF_OK;
}

/* launch kernel slot k of the batch for substep s (or FFT direction) */
/* the spike solve of a batch: one launch per chunk length, each over its
 * own prefix table and as many waves per workgroup as its runs need */
template <typename T>
static cf_status batch_launch_spike(cf_batch *b, int k, int s, bool sums)
{
    cf_status e = bev_begin(b, KB_NAME[k]);
    if (e != CF_OK) return e;
    for (size_t g = 0; g < b->spL.size(); g++) {
        typename BSel<T>::kb_t fn = nullptr;
        fn = sums ? pick_batched_spike_sums<T>(b->spL[g]) : pick_batched_spike<T>(b->spL[g]);
        if (!fn) { cf_set_error("no spike kernel for %d rows per chunk", b->spL[g]); return CF_ERR_UNSUPPORTED; }
        const int row = CF_NST + CF_P2_NG + (int) g;
        const unsigned tg = batch_total(b, row);
        const unsigned grid = b->kgrid[k] ? std::min(b->kgrid[k], tg) : tg;
        if (grid == 0) continue;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(64 * b->spW[g]), 0, b->stream,
                           (const GArgs<T> *) b->args + (size_t) b->tab * b->n,
                           (const unsigned *) (b->vstart + (size_t) row * (b->n + 1)), b->n, s);
        HIPCHK(hipGetLastError());
    }
    return bev_end(b);
}

/* the FFT passes of a batch: one launch per length group, each over its
 * own prefix table, with the group's plan and LDS */
template <typename T>
static cf_status batch_launch_fft(cf_batch *b, int k, int s)
{
    cf_status e = bev_begin(b, KB_NAME[k]);
    if (e != CF_OK) return e;
    const bool y = k == KB_FFTY;
    const std::vector<int> &K = y ? b->fyK : b->fxK;
    for (size_t g = 0; g < K.size(); g++) {
        typename BSel<T>::kb_t fn = y ? pick_batched_ffty<T>(K[g])
                                  : (k == KB_DIVX ? pick_batched_divx_n<T>(b->order, K[g]) : pick_batched_fftxi<T>(K[g]));
        size_t lds = 0;
        for (const cf_run *r : b->runs)
            if (bfft_key((int) (y ? r->g.ny : r->g.nx)) == K[g]) lds = std::max(lds, y ? fy_lds(r) : fx_lds(r));
        const int row = (y ? CF_FY_ROW0 : CF_FX_ROW0) + (int) g;
        const unsigned tg = batch_total(b, row);
        const unsigned grid = b->kgrid[k] ? std::min(b->kgrid[k], tg) : tg;
        if (grid == 0) continue;
        hipLaunchKernelGGL(fn, dim3(grid), dim3(256), lds, b->stream, (const GArgs<T> *) b->args + (size_t) b->tab * b->n,
                           (const unsigned *) (b->vstart + (size_t) row * (b->n + 1)), b->n, s);
        HIPCHK(hipGetLastError());
    }
    return bev_end(b);
}

template <typename T>
static cf_status batch_launch(cf_batch *b, int k, int s)
{
    typedef void (*kb_t)(const GArgs<T> *, const unsigned *, int, int);
    const BSel<T> S = pick_batched<T>(b->order, b->scalar);
    kb_t fn = nullptr;
    dim3 block(256);
    size_t lds = 0;
    switch (k) {
    case KB_RHS_U: fn = S.u; block = dim3(64, 4); break;
    case KB_RHS_V: fn = S.v; block = dim3(64, 4); break;
    case KB_RHS_W: fn = S.w; block = dim3(64, 4); break;
    case KB_RHS_C: fn = S.c; block = dim3(64, 4); break;
    case KB_PROJ: fn = S.proj; block = dim3(64, 4); break;
    case KB_SOLVE:
        if (b->solve_kind == SOLVE_HEPT) { fn = kb_solve4<T>; break; }
        {
            /* order 2: the spike launches (one per chunk length), then
             * Thomas over the runs without spike tables */
            cf_status e = batch_launch_spike<T>(b, k, s, false);
            if (e != CF_OK) return e;
            const unsigned tt = batch_total(b, CF_TH_ROW);
            const unsigned gt = b->kgrid[k] ? std::min(b->kgrid[k], tt) : tt;
            if (gt == 0) return CF_OK;
            if ((e = bev_begin(b, KB_NAME[k])) != CF_OK) return e;
            hipLaunchKernelGGL(kb_thomas<T>, dim3(gt), dim3(256), 0, b->stream,
                               (const GArgs<T> *) b->args + (size_t) b->tab * b->n,
                               (const unsigned *) (b->vstart + (size_t) CF_TH_ROW * (b->n + 1)), b->n, s);
            HIPCHK(hipGetLastError());
            return bev_end(b);
        }
    case KB_PLANE: fn = kb_plane<T>; break;
    case KB_MASS: fn = kb_mass<T>; lds = b->lds_mass; break;
    case KB_CFL: fn = kb_cfl<T>; break;
    case KB_DIVX:
        /* FP64: one runtime-plan launch over all runs (its compiled forms
         * spill beyond length 64, and split launches measured slower) */
        if constexpr (std::is_same<T, double>::value) { fn = S.divx; lds = b->lds_fx; break; }
        else return batch_launch_fft<T>(b, k, s);
    case KB_FFTY:
    case KB_FFTXI:
        return batch_launch_fft<T>(b, k, s);
    case KB_POIS:
        if (b->order == 4) { fn = kb_poisson4<T>; break; }
        {
            /* one launch per chunk-length group, each over its own prefix table */
            block = dim3(CF_P2S_THREADS);
            cf_status e = bev_begin(b, KB_NAME[k]);
            if (e != CF_OK) return e;
            for (size_t g = 0; g < b->p2L.size(); g++) {
                const int row = CF_NST + (int) g;
                const unsigned tg = batch_total(b, row);
                const unsigned grid = b->kgrid[k] ? std::min(b->kgrid[k], tg) : tg;
                if (grid == 0) continue;
                size_t lp = 0;
                for (cf_run *rr : b->runs)
                    if (pois2_L(rr->g.nz) == b->p2L[g]) lp = std::max(lp, pois2_lds(b->p2L[g], rr->g.nz));
                hipLaunchKernelGGL(pick_batched_poisson<T>(b->p2L[g]), dim3(grid), block, lp, b->stream,
                                   (const GArgs<T> *) b->args + (size_t) b->tab * b->n,
                                   (const unsigned *) (b->vstart + (size_t) row * (b->n + 1)), b->n, s);
                HIPCHK(hipGetLastError());
                if constexpr (std::is_same<T, float>::value) {    /* the FP64 islands */
                    const int irow = CF_P2I_ROW0 + (int) g;
                    const unsigned ti = batch_total(b, irow);
                    const unsigned gi = b->kgrid[k] ? std::min(b->kgrid[k], ti) : ti;
                    if (gi == 0) continue;
                    hipLaunchKernelGGL(pick_batched_poisson_islands(b->p2L[g]), dim3(gi), block, lp, b->stream,
                                       (const GArgs<T> *) b->args + (size_t) b->tab * b->n,
                                       (const unsigned *) (b->vstart + (size_t) irow * (b->n + 1)), b->n, s);
                    HIPCHK(hipGetLastError());
                }
            }
            return bev_end(b);
        }
    case KB_STEPEND: fn = kb_step_end<T>; break;
    case KB_FRHS: fn = kb_rhs2t<T>; break;
    case KB_FSPIKE:
        return batch_launch_spike<T>(b, k, s, true);
    case KB_FLEVEL: fn = kb_level_sum<T>; break;
    case KB_FMASS: fn = kb_mass_fast<T>; lds = b->lds_mass; break;
    default: cf_set_error("bad batch kernel slot %d", k); return CF_ERR_ARG;
    }
    if (!fn) return CF_OK;
    const int st = KB_STAGE[k];
    const unsigned total = batch_total(b, st);
    unsigned grid = b->kgrid[k] ? std::min(b->kgrid[k], total) : total;
    if (grid == 0) return CF_OK;
    cf_status e = bev_begin(b, KB_NAME[k]);
    if (e != CF_OK) return e;
    hipLaunchKernelGGL(fn, dim3(grid), block, lds, b->stream, (const GArgs<T> *) b->args + (size_t) b->tab * b->n,
                       (const unsigned *) (b->vstart + (size_t) st * (b->n + 1)), b->n, s);
    HIPCHK(hipGetLastError());
    return bev_end(b);
}

template <typename T>
static cf_status batch_enqueue_pressure(cf_batch *b, int s)
{
    cf_status st;
    if ((st = batch_launch<T>(b, KB_DIVX, s)) != CF_OK) return st;
    if ((st = batch_launch<T>(b, KB_FFTY, 1)) != CF_OK) return st;
    if ((st = batch_launch<T>(b, KB_POIS, s)) != CF_OK) return st;
    if ((st = batch_launch<T>(b, KB_FFTY, 0)) != CF_OK) return st;
    return batch_launch<T>(b, KB_FFTXI, s);
}

template <typename T>
static cf_status batch_enqueue_step(cf_batch *b)
{
    cf_status st;
    for (int s = 0; s < 3; s++) {
        if ((st = batch_launch<T>(b, KB_RHS_U, s)) != CF_OK) return st;
        if (s == 0 && (st = batch_launch<T>(b, KB_CFL, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_RHS_V, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_RHS_W, s)) != CF_OK) return st;
        if (b->scalar && (st = batch_launch<T>(b, KB_RHS_C, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_SOLVE, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_PLANE, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_MASS, s)) != CF_OK) return st;
        if ((st = batch_enqueue_pressure<T>(b, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_PROJ, s)) != CF_OK) return st;
    }
    return batch_launch<T>(b, KB_STEPEND, 0);
}

/* one RK3 step of a fast batch starting from state set `from`: the fast
 * path's sequence (enqueue_step) with every launch covering all runs */
template <typename T>
static cf_status fbatch_enqueue_step_t(cf_batch *b, int from)
{
    cf_status st;
    for (int s = 0; s < 3; s++) {
        b->tab = (from + s) & 1;
        if ((st = batch_launch<T>(b, KB_FRHS, s)) != CF_OK) return st;
        if (s == 0 && (st = batch_launch<T>(b, KB_CFL, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_FSPIKE, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_FLEVEL, s)) != CF_OK) return st;
        if ((st = batch_launch<T>(b, KB_FMASS, s)) != CF_OK) return st;
        if ((st = batch_enqueue_pressure<T>(b, s)) != CF_OK) return st;
    }
    b->tab = 0;
    return batch_launch<T>(b, KB_STEPEND, 0);
}

static cf_status fbatch_enqueue_step(cf_batch *b, int from)
{
    return b->prec == CF_FP32 ? fbatch_enqueue_step_t<float>(b, from) : fbatch_enqueue_step_t<double>(b, from);
}

static cf_status batch_collect(cf_batch *b)
{
    if (!b->profile) return CF_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto &pe : b->pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
        auto &acc = b->ktime[pe.first];
        acc.first += ms;
        acc.second += 1;
        (void) hipEventDestroy(pe.second.first);
        (void) hipEventDestroy(pe.second.second);
    }
    b->pending.clear();
    return CF_OK;
}

/* time one launch slot (median of 3) under the current tables and grid */
template <typename T>
static cf_status batch_time_slot(cf_batch *b, int k, float *ms)
{
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float t[3];
    cf_status st = batch_launch<T>(b, k, 0);              /* warm */
    for (int i = 0; i < 3 && st == CF_OK; i++) {
        HIPCHK(hipEventRecord(e0, b->stream));
        st = batch_launch<T>(b, k, 0);
        HIPCHK(hipEventRecord(e1, b->stream));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&t[i], e0, e1));
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    if (st != CF_OK) return st;
    std::sort(t, t + 3);
    *ms = t[1];
    return CF_OK;
}

/* the autotuner: for each launch slot, the persistent grid (one
 * workgroup per virtual block, or a multiple of the CU count) and for
 * order 2 the z-solve shape, by measured time on the batch's shapes */
template <typename T>
static cf_status batch_autotune(cf_batch *b)
{
    cf_status st;
    b->tuning.clear();
    if (b->fast) {
        b->solve_kind = SOLVE_SPIKE;
        b->tuning.push_back("fast path: fused RHS, spike with plane sums, FFT Poisson");
    } else if (b->order == 2) {
        /* the solve is fixed per run, never tuned: the spike where the run
         * has tables, Thomas otherwise, as the run does alone */
        b->solve_kind = SOLVE_SPIKE;
        b->tuning.push_back("solve: per run, spike where its tables exist (as alone), else Thomas");
    } else {
        b->solve_kind = SOLVE_HEPT;
    }
    if ((st = batch_upload_tables<T>(b)) != CF_OK) return st;
    /* slots whose repeated launch is harmless on a zero state and that do real work */
    const int gslots[] = { KB_RHS_U, KB_RHS_V, KB_RHS_W, KB_RHS_C, KB_SOLVE, KB_PLANE, KB_DIVX, KB_FFTY, KB_POIS,
                           KB_FFTXI, KB_PROJ };
    const int fslots[] = { KB_FRHS, KB_FSPIKE, KB_DIVX, KB_FFTY, KB_POIS, KB_FFTXI };
    std::vector<int> slots = b->fast ? std::vector<int>(fslots, fslots + 6) : std::vector<int>(gslots, gslots + 11);
    for (int k : slots) {
        if (k == KB_RHS_C && !b->scalar) continue;
        const unsigned total = batch_total(b, KB_STAGE[k]);
        std::vector<unsigned> cand = { 0 };
        for (int m : { 1, 2, 4, 8, 16 })
            if ((unsigned) (b->ncu * m) < total) cand.push_back((unsigned) (b->ncu * m));
        float best = 1e30f;
        unsigned bg = 0;
        for (unsigned g : cand) {
            b->kgrid[k] = g;
            float ms;
            if ((st = batch_time_slot<T>(b, k, &ms)) != CF_OK) return st;
            if (ms < best * 0.97f) { best = ms; bg = g; }
        }
        b->kgrid[k] = bg;
        char line[160];
        snprintf(line, sizeof line, "%s[%d]: grid %u of %u virtual blocks (%.1f us)", KB_NAME[k], k,
                 bg ? bg : total, total, 1e3 * best);
        b->tuning.push_back(line);
    }
    b->kgrid[KB_RHS_C] = b->kgrid[KB_RHS_C] ? b->kgrid[KB_RHS_C] : b->kgrid[KB_RHS_U];
    /* the tuning launches wrote forcings, log entries and (through the
     * scalar's boundary lift) state values: restore the zero state */
    for (cf_run *r : b->runs) {
        HIPCHK(hipMemset(r->scal, 0, sizeof(double) * SC_N));
        HIPCHK(hipMemset(r->logbuf, 0, sizeof(double) * (size_t) LOGCAP * CF_LOG_N));
        double *zb[] = { r->u, r->v, r->w, r->c, r->ru, r->rv, r->rw, r->fa[0], r->fa[1], r->fa[2], r->fa[3],
                         r->fb[0], r->fb[1], r->fb[2], r->fb[3], r->p };
        for (double *z : zb)
            if (z) HIPCHK(hipMemset(z, 0, r->esz * r->N));
    }
    HIPCHK(hipDeviceSynchronize());
    return CF_OK;
}

static void free_batch(cf_batch *b)
{
    if (!b) return;
    if (b->gexec) (void) hipGraphExecDestroy(b->gexec);
    if (b->graph) (void) hipGraphDestroy(b->graph);
    for (int q = 0; q < 2; q++) {
        if (b->fexec[q]) (void) hipGraphExecDestroy(b->fexec[q]);
        if (b->fgraph[q]) (void) hipGraphDestroy(b->fgraph[q]);
    }
    if (b->args) (void) hipFree(b->args);
    if (b->vstart) (void) hipFree(b->vstart);
    for (auto &pe : b->pending) { (void) hipEventDestroy(pe.second.first); (void) hipEventDestroy(pe.second.second); }
    if (b->stream) (void) hipStreamDestroy(b->stream);
    for (cf_run *r : b->runs) free_run(r);
    delete b;
}

extern "C" cf_status cf_batch_create(const cf_config *cfgs, int nrun, cf_batch **out)
{
    if (!out || !cfgs) { cf_set_error("null argument"); return CF_ERR_ARG; }
    *out = nullptr;
    if (nrun < 1 || nrun > CF_BATCH_MAX) { cf_set_error("a batch holds 1..%d runs", CF_BATCH_MAX); return CF_ERR_ARG; }
    for (int q = 0; q < nrun; q++)
        if (cfgs[q].driver != CF_DRIVER_CHANFAST) {
            cf_set_error("run %d: the fully implicit driver steps as a run of its own", q);
            return CF_ERR_UNSUPPORTED;
        }
    for (int q = 0; q < nrun; q++)
        if (cfgs[q].cflmode != 0) {
            cf_set_error("run %d: cflmode 1 changes dt every step; step it as a run of its own", q);
            return CF_ERR_UNSUPPORTED;
        }
    for (int q = 1; q < nrun; q++) {
        if (cfgs[q].order != cfgs[0].order || cfgs[q].scalar != cfgs[0].scalar
            || cfgs[q].precision != cfgs[0].precision || cfgs[q].device != cfgs[0].device) {
            cf_set_error("run %d: runs of a batch share order, scalar, precision and device", q);
            return CF_ERR_ARG;
        }
    }
    cf_batch *b = new cf_batch();
    b->n = nrun;
    b->prec = cfgs[0].precision;
    b->order = cfgs[0].order;
    b->scalar = cfgs[0].scalar;
    b->device = cfgs[0].device;
    /* order 2, no scalar, z-solves the spike covers: the fast path */
    b->fast = b->order == 2 && !b->scalar;
    for (int q = 0; q < nrun && b->fast; q++)
        b->fast = spike_L(cfgs[q].nz) > 0 && cfgs[q].cflmode == 0;
    for (int q = 0; q < nrun; q++) {
        cf_run *r = nullptr;
        cf_status st = create_run(&cfgs[q], !b->fast, &r);
        if (st == CF_OK && b->fast && (r->general || !r->use_spike)) {
            /* a system the spike cannot take: every run goes general */
            free_run(r);
            for (cf_run *o : b->runs) free_run(o);
            b->runs.clear();
            b->fast = false;
            b->spike_waves = 1;
            b->lds_fx = b->lds_fy = b->lds_mass = 0;
            q = -1;
            continue;
        }
        if (st != CF_OK) {
            char msg[512];
            snprintf(msg, sizeof msg, "run %d: %s", q, cf_last_error());
            free_batch(b);
            cf_set_error("%s", msg);
            return st;
        }
        b->runs.push_back(r);
        b->spike_waves = std::max(b->spike_waves, r->spike_waves);
        b->lds_fx = std::max(b->lds_fx, fx_lds(r));
        b->lds_fy = std::max(b->lds_fy, fy_lds(r));
        b->lds_mass = std::max(b->lds_mass, sizeof(double) * (size_t) (r->g.nz + 6));
    }
    b->p2L = batch_p2L(b->runs);
    HIPCHK(hipSetDevice(b->device));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->ncu = prop.multiProcessorCount;
    size_t asz = b->prec == CF_FP32 ? sizeof(GArgs<float>) : sizeof(GArgs<double>);
    if (hipMalloc(&b->args, asz * nrun * (b->fast ? 2 : 1)) != hipSuccess
        || hipMalloc((void **) &b->vstart, sizeof(unsigned) * CF_VS_ROWS * (nrun + 1)) != hipSuccess
        || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        free_batch(b);
        cf_set_error("batch allocation failed");
        return CF_ERR_NOMEM;
    }
    cf_status st = b->prec == CF_FP32 ? batch_autotune<float>(b) : batch_autotune<double>(b);
    if (st != CF_OK) { free_batch(b); return st; }
    *out = b;
    return CF_OK;
}

extern "C" cf_status cf_batch_destroy(cf_batch *b)
{
    if (!b) return CF_OK;
    (void) hipSetDevice(b->device);
    if (b->stream) (void) hipStreamSynchronize(b->stream);
    free_batch(b);
    return CF_OK;
}

#define CHECK_BATCH(b, run)                                                              \
    do {                                                                                 \
        if (!(b)) { cf_set_error("null batch handle"); return CF_ERR_ARG; }              \
        if ((run) < 0 || (run) >= (b)->n) { cf_set_error("run %d out of range", (int) (run)); return CF_ERR_ARG; } \
        HIPCHK(hipSetDevice((b)->device));                                               \
        HIPCHK(hipStreamSynchronize((b)->stream));                                       \
    } while (0)

extern "C" cf_status cf_batch_upload(cf_batch *b, int run, int field, const double *host_io)
{
    CHECK_BATCH(b, run);
    return cf_run_upload(b->runs[run], field, host_io);
}

extern "C" cf_status cf_batch_download(cf_batch *b, int run, int field, double *host_io)
{
    CHECK_BATCH(b, run);
    return cf_run_download(b->runs[run], field, host_io);
}

extern "C" cf_status cf_batch_project(cf_batch *b, double dt)
{
    CHECK_BATCH(b, 0);
    if (!(dt > 0)) { cf_set_error("cf_batch_project: dt must be positive"); return CF_ERR_ARG; }
    if (b->fast) {
        /* the fast path's own projection, run by run (once per batch) */
        for (cf_run *r : b->runs) {
            cf_status st = cf_run_project(r, dt);
            if (st != CF_OK) return st;
        }
        return CF_OK;
    }
    for (cf_run *r : b->runs) {
        cf_status st = set_adhoc_poisson(r, dt);
        if (st != CF_OK) return st;
    }
    b->dt_adhoc = dt;
    cf_status st = b->prec == CF_FP32 ? batch_upload_tables<float>(b) : batch_upload_tables<double>(b);
    if (st != CF_OK) return st;
    if (b->prec == CF_FP32) {
        if ((st = batch_enqueue_pressure<float>(b, 3)) != CF_OK) return st;
        if ((st = batch_launch<float>(b, KB_PROJ, 3)) != CF_OK) return st;
    } else {
        if ((st = batch_enqueue_pressure<double>(b, 3)) != CF_OK) return st;
        if ((st = batch_launch<double>(b, KB_PROJ, 3)) != CF_OK) return st;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return batch_collect(b);
}

extern "C" cf_status cf_batch_step(cf_batch *b, int nsteps)
{
    CHECK_BATCH(b, 0);
    if (nsteps < 0) { cf_set_error("negative step count"); return CF_ERR_ARG; }
    if (nsteps == 0) return CF_OK;
    cf_status st;
    if (b->fast) {
        for (int n = 0; n < nsteps; n++) {
            const int from = b->cur;
            if (b->profile) {
                if ((st = fbatch_enqueue_step(b, from)) != CF_OK) return st;
            } else {
                if (!b->fexec[from]) {
                    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                    st = fbatch_enqueue_step(b, from);
                    hipGraph_t gr = nullptr;
                    hipError_t e = hipStreamEndCapture(b->stream, &gr);
                    if (st != CF_OK) { if (gr) (void) hipGraphDestroy(gr); return st; }
                    HIPCHK(e);
                    b->fgraph[from] = gr;
                    HIPCHK(hipGraphInstantiate(&b->fexec[from], b->fgraph[from], nullptr, nullptr, 0));
                }
                HIPCHK(hipGraphLaunch(b->fexec[from], b->stream));
            }
            b->cur ^= 1;
        }
        b->steps_done += nsteps;
        for (cf_run *r : b->runs) {
            r->steps_done += nsteps;
            r->cur = b->cur;
            r->proj_pending = true;
        }
        return b->profile ? batch_collect(b) : CF_OK;
    }
    if (b->profile) {
        for (int n = 0; n < nsteps; n++) {
            st = b->prec == CF_FP32 ? batch_enqueue_step<float>(b) : batch_enqueue_step<double>(b);
            if (st != CF_OK) return st;
        }
    } else {
        if (!b->gexec) {
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            st = b->prec == CF_FP32 ? batch_enqueue_step<float>(b) : batch_enqueue_step<double>(b);
            hipGraph_t gr = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &gr);
            if (st != CF_OK) { if (gr) (void) hipGraphDestroy(gr); return st; }
            HIPCHK(e);
            b->graph = gr;
            HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        for (int n = 0; n < nsteps; n++) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    }
    b->steps_done += nsteps;
    for (cf_run *r : b->runs) r->steps_done += nsteps;
    return b->profile ? batch_collect(b) : CF_OK;
}

extern "C" cf_status cf_batch_sync(cf_batch *b)
{
    CHECK_BATCH(b, 0);
    return CF_OK;
}

extern "C" cf_status cf_batch_log(cf_batch *b, int run, double *out, int max_steps, int *nsteps_out)
{
    CHECK_BATCH(b, run);
    return cf_run_log(b->runs[run], out, max_steps, nsteps_out);
}

extern "C" cf_status cf_batch_max_divergence(cf_batch *b, int run, double *out)
{
    CHECK_BATCH(b, run);
    return cf_run_max_divergence(b->runs[run], out);
}

extern "C" cf_status cf_batch_set_profiling(cf_batch *b, int on)
{
    if (!b) { cf_set_error("null batch handle"); return CF_ERR_ARG; }
    b->profile = on != 0;
    b->ktime.clear();
    return CF_OK;
}

extern "C" cf_status cf_batch_kernel_time(cf_batch *b, const char *kernel, double *avg_ms, int *launches)
{
    if (!b || !kernel) { cf_set_error("null argument"); return CF_ERR_ARG; }
    auto it = b->ktime.find(kernel);
    if (it == b->ktime.end()) { if (avg_ms) *avg_ms = 0; if (launches) *launches = 0; return CF_OK; }
    if (avg_ms) *avg_ms = it->second.first / (it->second.second ? it->second.second : 1);
    if (launches) *launches = it->second.second;
    return CF_OK;
}

extern "C" cf_status cf_batch_tuning(cf_batch *b, char *buf, size_t len)
{
    if (!b || !buf || len == 0) { cf_set_error("null argument"); return CF_ERR_ARG; }
    std::string all;
    for (auto &l : b->tuning) { all += l; all += "\n"; }
    snprintf(buf, len, "%s", all.c_str());
    return CF_OK;
}

/* ------------------------------------------------------------ serving
 *
 * Continuous batching of independent channel runs.  Requests queue on
 * submit.  The unit of device work is a quantum: a few RK3 steps of one
 * group (the active runs sharing order, scalar setting and precision, in
 * one batch whose every stage is one launch) or of one solo run (a run of
 * at least solo_points cells, which fills the GPU alone and keeps its
 * single-run graph), each on its own stream and closed by an event.  A
 * tick
 *   1. takes in the quanta whose events have completed: credits their
 *      steps and retires the runs that reached their step count;
 *   2. admits queued requests in arrival order while their footprint fits
 *      the HBM budget (a request whose group already holds CF_BATCH_MAX
 *      runs waits and lets later ones pass; one that fails to set up is
 *      marked failed and dropped);
 *   3. rebuilds the batch of every idle group whose membership changed;
 *   4. enqueues the next quantum of every idle group, and of the solo runs
 *      with the least remaining work while fewer than solo_limit are in
 *      flight;
 * and, when none of that made progress, waits for the first quantum in
 * flight to complete and does 1, 3 and 4 once more.  Groups therefore
 * never wait for each other: a group of small runs re-enqueues as soon as
 * its own quantum is done.  A quantum is the smallest remaining step count
 * of its runs, at most the quantum size, so runs retire exactly at their
 * step count.  Per-run arithmetic does not depend on the other runs of a
 * group (per-run reductions, the z solve fixed per run, one Poisson launch
 * per chunk length, FFT instantiations fixed per length), so a run's
 * result is the same alone and in any mix.
 */

#include <chrono>
#include <thread>

enum { RQ_QUEUED = 0, RQ_ACTIVE = 1, RQ_DONE = 2, RQ_RELEASED = 3, RQ_FAILED = 4 };

struct cf_server {
    size_t budget = 0, used = 0;
    int device = 0, quantum = 20;
    std::chrono::steady_clock::time_point t0;
    struct Req {
        cf_config cfg;
        std::vector<double> init[4];
        int nsteps = 0, done = 0, state = RQ_QUEUED;
        bool retain = false;
        bool solo = false;                         /* stepped through its own graph, not a group batch */
        int gkey = 0;                              /* group of a batched run (group_key, + 8 on the fast path) */
        size_t bytes = 0;
        cf_run *run = nullptr;
        double t_submit = 0, t_admit = -1, t_done = -1;
        /* a solo run's quantum in flight */
        hipEvent_t ev = nullptr;
        int inflight = 0;
        std::string err;                           /* why a failed request failed */
    };
    struct Group {
        cf_batch *b = nullptr;                     /* non-owning batch of the members */
        hipEvent_t ev = nullptr;
        int inflight = 0;                          /* steps of the quantum in flight (0: idle) */
        bool dirty = false;                        /* membership changed: rebuild before the next quantum */
        std::vector<int> members;                  /* request ids in b->runs order */
    };
    std::vector<Req> reqs;
    std::vector<int> queue;
    std::map<int, Group> groups;                   /* key -> the group's batch and quantum state */
    std::vector<std::pair<cf_run *, size_t>> retired;   /* runs to free once the device is idle */
    size_t retired_bytes = 0;
    size_t solo_points = (size_t) 4 << 20;         /* runs of at least this many cells step on their own */
    int solo_limit = 1;                            /* solo runs advancing at once */
    double pt_steps = 0.0;                         /* completed cell-steps */
};

/* device memory of retired runs is returned while no quantum is in flight
 * (hipFree waits for the whole device), or at once past this many bytes */
static const size_t SRV_RETIRE_FLUSH = (size_t) 32 << 30;

static double srv_now(const cf_server *s)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

static int group_key(const cf_config &c) { return (c.order == 4 ? 4 : 0) + (c.scalar ? 2 : 0) + (c.precision == CF_FP32 ? 1 : 0); }
/* order-2 runs without the scalar whose z-solves the spike covers are
 * batched on the fast path (FP64 and FP32 in groups of their own) */
static bool fast_eligible(const cf_config &c)
{
    return c.order == 2 && !c.scalar && c.cflmode == 0 && spike_L(c.nz) > 0;
}
/* relabel a fast-path run's two state sets (host pointers only), so that
 * it joins a batch stepping from the other set; its single-run graphs,
 * captured with the old labels, are dropped */
static void swap_state_sets(cf_run *r)
{
    std::swap(r->u, r->ru);
    std::swap(r->v, r->rv);
    std::swap(r->w, r->rw);
    r->cur ^= 1;
    for (int q = 0; q < 2; q++) {
        if (r->gexec[q]) { (void) hipGraphExecDestroy(r->gexec[q]); r->gexec[q] = nullptr; }
        if (r->graph[q]) { (void) hipGraphDestroy(r->graph[q]); r->graph[q] = nullptr; }
    }
}

extern "C" cf_status cf_server_set_solo_limit(cf_server *s, int n)
{
    if (!s || n < 1) { cf_set_error("cf_server_set_solo_limit: need a server and n >= 1"); return CF_ERR_ARG; }
    s->solo_limit = n;
    return CF_OK;
}

extern "C" cf_status cf_server_set_solo_points(cf_server *s, size_t points)
{
    if (!s) { cf_set_error("null server"); return CF_ERR_ARG; }
    s->solo_points = points;
    return CF_OK;
}

extern "C" cf_status cf_server_create(size_t hbm_budget_bytes, int device, int quantum_steps, cf_server **out)
{
    if (!out || hbm_budget_bytes == 0 || quantum_steps < 1) { cf_set_error("bad server arguments"); return CF_ERR_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        cf_set_error("no HIP device %d available", device);
        return CF_ERR_DEVICE;
    }
    cf_server *s = new cf_server();
    s->budget = hbm_budget_bytes;
    s->device = device;
    s->quantum = quantum_steps;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return CF_OK;
}

static void srv_drop_batch(cf_batch *b)
{
    if (!b) return;
    b->runs.clear();                               /* runs belong to their requests */
    free_batch(b);
}

static void srv_flush_retired(cf_server *s)
{
    for (auto &pr : s->retired) free_run(pr.first);
    s->retired.clear();
    s->retired_bytes = 0;
}

extern "C" cf_status cf_server_destroy(cf_server *s)
{
    if (!s) return CF_OK;
    (void) hipSetDevice(s->device);
    (void) hipDeviceSynchronize();
    for (auto &g : s->groups) {
        srv_drop_batch(g.second.b);
        if (g.second.ev) (void) hipEventDestroy(g.second.ev);
    }
    for (auto &r : s->reqs) {
        if (r.run) free_run(r.run);
        if (r.ev) (void) hipEventDestroy(r.ev);
    }
    srv_flush_retired(s);
    delete s;
    return CF_OK;
}

extern "C" cf_status cf_server_submit(cf_server *s, const cf_config *cfg, const double *u, const double *v,
                                      const double *w, const double *c, int nsteps, int retain, int *request_id)
{
    if (!s || !cfg || !u || !v || !w || nsteps < 1) { cf_set_error("bad submit arguments"); return CF_ERR_ARG; }
    if (cfg->scalar && !c) { cf_set_error("a scalar run needs its c field"); return CF_ERR_ARG; }
    if (cfg->device != s->device) { cf_set_error("request device %d differs from the server's", cfg->device); return CF_ERR_ARG; }
    size_t bytes = 0;
    cf_status st = cf_run_footprint(cfg, &bytes);
    if (st != CF_OK) return st;
    if (bytes > s->budget) { cf_set_error("request needs %zu bytes, more than the whole budget", bytes); return CF_ERR_NOMEM; }
    cf_server::Req r;
    r.cfg = *cfg;
    r.cfg.dz_user = nullptr;
    if (cfg->stretch == CF_STRETCH_USER) { cf_set_error("the server takes built-in stretchings only"); return CF_ERR_UNSUPPORTED; }
    const size_t n = (size_t) cfg->nx * cfg->ny * cfg->nz;
    const double *src[4] = { u, v, w, cfg->scalar ? c : nullptr };
    for (int f = 0; f < 4; f++)
        if (src[f]) r.init[f].assign(src[f], src[f] + n);
    r.nsteps = nsteps;
    r.retain = retain != 0;
    r.bytes = bytes;
    r.t_submit = srv_now(s);
    s->reqs.push_back(std::move(r));
    s->queue.push_back((int) s->reqs.size() - 1);
    if (request_id) *request_id = (int) s->reqs.size() - 1;
    return CF_OK;
}

/* rebuild the batch of an idle group from its currently active runs */
static cf_status srv_regroup(cf_server *s, int k, cf_server::Group &G)
{
    std::vector<int> ids;
    for (size_t q = 0; q < s->reqs.size(); q++) {
        const cf_server::Req &r = s->reqs[q];
        if (r.state == RQ_ACTIVE && !r.solo && r.gkey == k) ids.push_back((int) q);
    }
    if (ids.size() > (size_t) CF_BATCH_MAX) {      /* admission keeps groups within the limit */
        cf_set_error("group %d: %zu active runs, more than %d", k, ids.size(), CF_BATCH_MAX);
        return CF_ERR_STATE;
    }
    G.dirty = false;
    G.members = ids;
    if (ids.empty()) {                             /* keep the idle batch: freeing would wait for the device */
        if (G.b) { G.b->runs.clear(); G.b->n = 0; }
        return CF_OK;
    }
    std::vector<cf_run *> runs;
    for (int q : ids) runs.push_back(s->reqs[q].run);
    cf_batch *b = G.b;
    if (!b) {
        b = new cf_batch();
        b->prec = runs[0]->cfg.precision;
        b->order = runs[0]->cfg.order;
        b->scalar = runs[0]->cfg.scalar;
        b->device = s->device;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, s->device) == hipSuccess && prop.multiProcessorCount > 0)
            b->ncu = prop.multiProcessorCount;
        /* argument and prefix tables sized for the largest group once */
        const size_t asz = b->prec == CF_FP32 ? sizeof(GArgs<float>) : sizeof(GArgs<double>);
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess
            || hipMalloc(&b->args, asz * CF_BATCH_MAX * 2) != hipSuccess
            || hipMalloc((void **) &b->vstart, sizeof(unsigned) * CF_VS_ROWS * (CF_BATCH_MAX + 1)) != hipSuccess) {
            free_batch(b);
            cf_set_error("group batch allocation failed");
            return CF_ERR_NOMEM;
        }
        G.b = b;
    }
    b->runs = runs;
    b->n = (int) runs.size();
    if (k >= 8) {
        /* fast group: every run steps from the batch's state set */
        if (!b->fast) { b->fast = true; b->cur = runs[0]->cur; }
        for (cf_run *r : runs)
            if (r->cur != b->cur) swap_state_sets(r);
    }
    b->spike_waves = 1;
    b->lds_fx = b->lds_fy = b->lds_mass = 0;
    for (cf_run *r : runs) {
        b->spike_waves = std::max(b->spike_waves, r->spike_waves);
        b->lds_fx = std::max(b->lds_fx, fx_lds(r));
        b->lds_fy = std::max(b->lds_fy, fy_lds(r));
        b->lds_mass = std::max(b->lds_mass, sizeof(double) * (size_t) (r->g.nz + 6));
    }
    b->solve_kind = b->order == 4 ? SOLVE_HEPT : SOLVE_SPIKE;   /* per run: spike or Thomas as alone */
    b->p2L = batch_p2L(runs);
    for (int q = 0; q < KB_N; q++) b->kgrid[q] = 0;
    return b->prec == CF_FP32 ? batch_upload_tables<float>(b) : batch_upload_tables<double>(b);
}

/* credit the steps of a completed quantum and retire finished runs */
static void srv_credit(cf_server *s, int q, int k, double now, std::vector<int> &changed)
{
    cf_server::Req &r = s->reqs[q];
    r.done += k;
    if (!r.solo) r.run->steps_done += k;
    s->pt_steps += (double) r.cfg.nx * r.cfg.ny * r.cfg.nz * k;
    if (r.done < r.nsteps) return;
    r.t_done = now;
    r.state = r.retain ? RQ_DONE : RQ_RELEASED;
    if (!r.retain) {
        s->retired.push_back({ r.run, r.bytes });
        s->retired_bytes += r.bytes;
        r.run = nullptr;
        s->used -= r.bytes;
    }
    if (!r.solo) changed.push_back(r.gkey);
}

/* take in every completed quantum; returns the number taken in */
static cf_status srv_poll(cf_server *s, int *ncompleted)
{
    *ncompleted = 0;
    const double now = srv_now(s);
    std::vector<int> changed;
    for (auto &g : s->groups) {
        cf_server::Group &G = g.second;
        if (G.inflight == 0) continue;
        hipError_t e = hipEventQuery(G.ev);
        if (e == hipErrorNotReady) continue;
        HIPCHK(e);
        for (int q : G.members) srv_credit(s, q, G.inflight, now, changed);
        G.inflight = 0;
        ++*ncompleted;
    }
    for (size_t q = 0; q < s->reqs.size(); q++) {
        cf_server::Req &r = s->reqs[q];
        if (r.inflight == 0) continue;
        hipError_t e = hipEventQuery(r.ev);
        if (e == hipErrorNotReady) continue;
        HIPCHK(e);
        const int k = r.inflight;
        r.inflight = 0;
        srv_credit(s, (int) q, k, now, changed);
        ++*ncompleted;
    }
    for (int k : changed) s->groups[k].dirty = true;
    return CF_OK;
}

static bool srv_busy(const cf_server *s)
{
    for (const auto &g : s->groups)
        if (g.second.inflight) return true;
    for (const auto &r : s->reqs)
        if (r.inflight) return true;
    return false;
}

/* active batched runs of a request's group family (fast and general
 * variants of one group_key together) */
static int srv_group_count(const cf_server *s, const cf_config &c)
{
    const int key = group_key(c);
    int n = 0;
    for (const auto &r : s->reqs)
        if (r.state == RQ_ACTIVE && !r.solo && (r.gkey & 7) == key) n++;
    return n;
}

/* set up one admitted request's run: create, upload, project */
static cf_status srv_setup(cf_server *s, cf_server::Req &r)
{
    cf_status st;
    cf_run *run = nullptr;
    /* large runs fill the GPU on their own: they keep the single-run graph
     * (the order-2 fast path where it applies) on their own stream; small
     * ones join their group's batched launches.  The choice depends on the
     * request alone, never on the mix. */
    r.solo = (size_t) r.cfg.nx * r.cfg.ny * r.cfg.nz >= s->solo_points || r.cfg.cflmode == 1
                            || r.cfg.driver != CF_DRIVER_CHANFAST;
    const bool fast = !r.solo && fast_eligible(r.cfg);
    if ((st = create_run(&r.cfg, !(r.solo || fast), &run)) != CF_OK) return st;
    r.gkey = group_key(r.cfg) + (fast && !run->general && run->use_spike ? 8 : 0);
    if (fast && r.gkey < 8) {                      /* the spike cannot take it: the general batch */
        free_run(run);
        run = nullptr;
        if ((st = create_run(&r.cfg, true, &run)) != CF_OK) return st;
    }
    for (int f = 0; f < 4; f++) {
        if (r.init[f].empty()) continue;
        if ((st = cf_run_upload(run, f, r.init[f].data())) != CF_OK) { free_run(run); return st; }
    }
    if ((st = cf_run_project(run, 1.0)) != CF_OK) { free_run(run); return st; }
    r.run = run;
    return CF_OK;
}

/* admit queued requests; returns the number admitted or failed */
static int srv_admit(cf_server *s)
{
    int n = 0;
    size_t qi = 0;
    while (qi < s->queue.size()) {
        cf_server::Req &r = s->reqs[s->queue[qi]];
        const bool solo = (size_t) r.cfg.nx * r.cfg.ny * r.cfg.nz >= s->solo_points || r.cfg.cflmode == 1
                            || r.cfg.driver != CF_DRIVER_CHANFAST;
        if (!solo && srv_group_count(s, r.cfg) >= CF_BATCH_MAX) { qi++; continue; }   /* its group is full */
        /* head-of-line for memory: a large request waits for memory rather
         * than being overtaken indefinitely */
        if (s->used + r.bytes > s->budget) break;
        if (s->used + r.bytes + s->retired_bytes > s->budget + SRV_RETIRE_FLUSH) srv_flush_retired(s);
        cf_status st = srv_setup(s, r);
        for (int f = 0; f < 4; f++) std::vector<double>().swap(r.init[f]);
        s->queue.erase(s->queue.begin() + (long) qi);
        n++;
        if (st != CF_OK) {
            r.state = RQ_FAILED;
            r.err = cf_last_error();
            r.t_done = srv_now(s);
            continue;
        }
        r.state = RQ_ACTIVE;
        r.t_admit = srv_now(s);
        s->used += r.bytes;
        if (!r.solo) s->groups[r.gkey].dirty = true;
    }
    return n;
}

/* enqueue the next quantum of every idle group and solo run */
static cf_status srv_launch(cf_server *s)
{
    cf_status st;
    for (auto &g : s->groups) {
        cf_server::Group &G = g.second;
        if (G.inflight) continue;
        if (G.dirty && (st = srv_regroup(s, g.first, G)) != CF_OK) return st;
        cf_batch *b = G.b;
        if (!b || b->n == 0) continue;
        int k = s->quantum;
        for (int q : G.members) k = std::min(k, s->reqs[q].nsteps - s->reqs[q].done);
        if (k <= 0) continue;
        for (int n = 0; n < k; n++) {
            if (b->fast) {
                st = fbatch_enqueue_step(b, b->cur);
                b->cur ^= 1;
            } else {
                st = b->prec == CF_FP32 ? batch_enqueue_step<float>(b) : batch_enqueue_step<double>(b);
            }
            if (st != CF_OK) return st;
        }
        if (b->fast)
            for (cf_run *r : b->runs) {
                r->cur = b->cur;
                r->proj_pending = true;
            }
        if (!G.ev) HIPCHK(hipEventCreateWithFlags(&G.ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(G.ev, b->stream));
        G.inflight = k;
    }
    /* Solo runs each fill the device, so running several at once only
     * shares it: at most solo_limit advance, the ones with the least
     * remaining work (cells x steps) first, ties in admission order.  Short
     * runs then finish early instead of every large run finishing late (the
     * mean latency of a shared server is least when it serves the shortest
     * remaining work first); the choice is made again every quantum. */
    int busy = 0;
    std::vector<std::pair<double, size_t>> ready;
    for (size_t q = 0; q < s->reqs.size(); q++) {
        const cf_server::Req &r = s->reqs[q];
        if (r.state != RQ_ACTIVE || !r.solo) continue;
        if (r.inflight) { busy++; continue; }
        if (r.nsteps - r.done <= 0) continue;
        ready.push_back({ (double) r.cfg.nx * r.cfg.ny * r.cfg.nz * (double) (r.nsteps - r.done), q });
    }
    std::sort(ready.begin(), ready.end());
    for (const auto &rq : ready) {
        if (busy >= s->solo_limit) break;
        cf_server::Req &r = s->reqs[rq.second];
        const int k = std::min(s->quantum, r.nsteps - r.done);
        busy++;
        if ((st = cf_run_step(r.run, k)) != CF_OK) return st;   /* counts its own steps */
        if (!r.ev) HIPCHK(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
        HIPCHK(hipEventRecord(r.ev, r.run->stream));
        r.inflight = k;
    }
    return CF_OK;
}

extern "C" cf_status cf_server_tick(cf_server *s, int *n_active, int *n_queued, int *n_done)
{
    if (!s) { cf_set_error("null server"); return CF_ERR_ARG; }
    HIPCHK(hipSetDevice(s->device));
    cf_status st;
    int ncomp = 0;
    if ((st = srv_poll(s, &ncomp)) != CF_OK) return st;
    const int nadm = srv_admit(s);
    if ((st = srv_launch(s)) != CF_OK) return st;
    if (ncomp == 0 && nadm == 0 && srv_busy(s)) {
        /* nothing moved: wait for the first quantum in flight to finish */
        while ((st = srv_poll(s, &ncomp)) == CF_OK && ncomp == 0)
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        if (st != CF_OK) return st;
        if ((st = srv_launch(s)) != CF_OK) return st;
    }
    if (!srv_busy(s) && !s->retired.empty()) srv_flush_retired(s);
    int na = 0, nd = 0;
    for (auto &r : s->reqs) { na += r.state == RQ_ACTIVE; nd += r.state >= RQ_DONE; }
    if (n_active) *n_active = na;
    if (n_queued) *n_queued = (int) s->queue.size();
    if (n_done) *n_done = nd;
    return CF_OK;
}

static cf_server::Req *srv_req(cf_server *s, int id)
{
    if (!s || id < 0 || id >= (int) s->reqs.size()) { cf_set_error("bad request id %d", id); return nullptr; }
    return &s->reqs[id];
}

extern "C" cf_status cf_server_result(cf_server *s, int request_id, int field, double *host_io)
{
    cf_server::Req *r = srv_req(s, request_id);
    if (!r) return CF_ERR_ARG;
    if (r->state != RQ_DONE || !r->run) { cf_set_error("request %d holds no retained result", request_id); return CF_ERR_STATE; }
    return cf_run_download(r->run, field, host_io);
}

extern "C" cf_status cf_server_log(cf_server *s, int request_id, double *out, int max_steps, int *nsteps_out)
{
    cf_server::Req *r = srv_req(s, request_id);
    if (!r) return CF_ERR_ARG;
    if (!r->run) { cf_set_error("request %d is not resident", request_id); return CF_ERR_STATE; }
    if (r->inflight) { cf_set_error("request %d has steps in flight", request_id); return CF_ERR_STATE; }
    return cf_run_log(r->run, out, max_steps, nsteps_out);
}

extern "C" cf_status cf_server_release(cf_server *s, int request_id)
{
    cf_server::Req *r = srv_req(s, request_id);
    if (!r) return CF_ERR_ARG;
    if (r->state == RQ_DONE && r->run) {
        s->retired.push_back({ r->run, r->bytes });
        s->retired_bytes += r->bytes;
        r->run = nullptr;
        s->used -= r->bytes;
        r->state = RQ_RELEASED;
        if (!srv_busy(s)) srv_flush_retired(s);
    }
    return CF_OK;
}

extern "C" cf_status cf_server_request_info(cf_server *s, int request_id, double *t_submit, double *t_admit,
                                            double *t_done, int *steps_done)
{
    cf_server::Req *r = srv_req(s, request_id);
    if (!r) return CF_ERR_ARG;
    if (t_submit) *t_submit = r->t_submit;
    if (t_admit) *t_admit = r->t_admit;
    if (t_done) *t_done = r->t_done;
    if (steps_done) *steps_done = r->done;
    return CF_OK;
}

extern "C" cf_status cf_server_request_state(cf_server *s, int request_id, int *state, char *msg, size_t len)
{
    cf_server::Req *r = srv_req(s, request_id);
    if (!r) return CF_ERR_ARG;
    if (state) *state = r->state;
    if (msg && len) snprintf(msg, len, "%s", r->err.c_str());
    return CF_OK;
}

extern "C" cf_status cf_server_stats(cf_server *s, double *seconds, double *cell_steps, size_t *hbm_in_use)
{
    if (!s) { cf_set_error("null server"); return CF_ERR_ARG; }
    if (seconds) *seconds = srv_now(s);
    if (cell_steps) *cell_steps = s->pt_steps;
    if (hbm_in_use) *hbm_in_use = s->used;
    return CF_OK;
}
