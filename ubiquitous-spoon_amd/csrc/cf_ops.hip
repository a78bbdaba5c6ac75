/*
 * cf_ops.hip -- the reference's operator layer on the GPU (chanflow_ops.h):
 * padded-layout HIP kernels behind the reference's function names and
 * argument lists, the pencil index maps and an on-device pencil gather.
 *
 * Layout: [n0][n1][n2] doubles, k fastest (chanfast.c:139-146).  The
 * stencil kernels run one thread per point, 64 consecutive k per wave so
 * every access is a coalesced row; the column solves run one thread per
 * (i, j) column on column-independent factors built on the host.
 */
#include "chanflow.h"
#include "chanflow_ops.h"
#include "cf_internal_host.h"
#include "cf_kernels.h"

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

static const double A1 = 1.125 / 1.0;       /* fourth.c:3-6 */
static const double A2 = -0.125 / 3.0;

/* check() of check.h:8-12: report where, then end the process */
#define SHIM_CHECK(cond, ...)                                                             \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            fprintf(stderr, "%s:%d: ", __FILE__, __LINE__);                               \
            fprintf(stderr, __VA_ARGS__);                                                 \
            fprintf(stderr, "\n");                                                        \
            abort();                                                                      \
        }                                                                                 \
    } while (0)
#define SHIM_HIP(x)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        SHIM_CHECK(e_ == hipSuccess, "%s: %s", #x, hipGetErrorString(e_));                \
    } while (0)

/* ------------------------------------------------------------ device */

struct PG {
    int n0, n1, n2, ng;
    double dx, dy;
    const double *dz;       /* n2 cell heights (device) */
};

static __device__ __forceinline__ size_t pid(const PG &g, int i, int j, int k)
{
    return ((size_t) i * g.n1 + j) * g.n2 + k;
}

/* thread -> (i, j, k); grid (ceil(n2 / 64), ceil(n1 / 4), n0) of 64 x 4 */
#define PG_POINT(g, i, j, k)                                              \
    const int k = blockIdx.x * 64 + threadIdx.x;                          \
    const int j = blockIdx.y * 4 + threadIdx.y;                           \
    const int i = blockIdx.z;                                             \
    if (k >= (g).n2 || j >= (g).n1) return;

/* divergence2 / 4, pressure.c:40-118 */
template <int ORD>
__global__ void k_sh_div(PG g, const double *u, const double *v, const double *w, double *d)
{
    PG_POINT(g, i, j, k);
    const int lo = ORD == 2 ? 1 : 2, hi = ORD == 2 ? 0 : 1;
    if (i < lo || i >= g.n0 - hi || j < lo || j >= g.n1 - hi || k < g.ng || k >= g.n2 - g.ng) return;
    const size_t o = pid(g, i, j, k), X = (size_t) g.n1 * g.n2, Y = g.n2;
    if (ORD == 2) {
        d[o] = (u[o] - u[o - X]) / g.dx + (v[o] - v[o - Y]) / g.dy + (w[o] - w[o - 1]) / g.dz[k];
    } else {
        d[o] = (A1 * (u[o] - u[o - X]) + A2 * (u[o + X] - u[o - 2 * X])) / g.dx
             + (A1 * (v[o] - v[o - Y]) + A2 * (v[o + Y] - v[o - 2 * Y])) / g.dy
             + (A1 * (w[o] - w[o - 1]) + A2 * (w[o + 1] - w[o - 2])) / g.dz[k];
    }
}

/* the odd mirror of w about the wall faces that divergence4 writes first
 * (pressure.c:80-97); one thread per column */
__global__ void k_sh_mirror_w(PG g, double *w)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1) return;
    double *col = w + (size_t) c * g.n2;
    int k = g.ng - 1;
    col[k - 1] = 2.0 * col[k] - col[k + 1];
    k = g.n2 - g.ng - 1;
    col[k + 1] = 2.0 * col[k] - col[k - 1];
}

/* project2 / 4, pressure.c:120-195 */
template <int ORD>
__global__ void k_sh_project(PG g, double *u, double *v, double *w, const double *p, double dt)
{
    PG_POINT(g, i, j, k);
    const int lo = ORD == 2 ? 0 : 1, hi = ORD == 2 ? 1 : 2;
    if (i < lo || i >= g.n0 - hi || j < lo || j >= g.n1 - hi) return;
    const size_t o = pid(g, i, j, k), X = (size_t) g.n1 * g.n2, Y = g.n2;
    if (k >= g.ng && k < g.n2 - g.ng) {
        if (ORD == 2) {
            u[o] -= (p[o + X] - p[o]) * dt / g.dx;
            v[o] -= (p[o + Y] - p[o]) * dt / g.dy;
        } else {
            const double up1h = A1 * (p[o + X] - p[o]) + A2 * (p[o + 2 * X] - p[o - X]);
            const double vp1h = A1 * (p[o + Y] - p[o]) + A2 * (p[o + 2 * Y] - p[o - Y]);
            u[o] -= up1h * dt / g.dx;
            v[o] -= vp1h * dt / g.dy;
        }
    }
    if (k >= g.ng - 1 && k < g.n2 - g.ng) {
        const double dzp1h = 0.5 * (g.dz[k + 1] + g.dz[k]);
        if (ORD == 2) {
            w[o] -= (p[o + 1] - p[o]) * dt / dzp1h;
        } else {
            const double wp1h = A1 * (p[o + 1] - p[o]) + A2 * (p[o + 2] - p[o - 1]);
            w[o] -= wp1h * dt / dzp1h;
        }
    }
}

/* d2/dx2 + d2/dy2 of q at o in the 2nd- or 4th-order compact-staggered form */
template <int ORD>
static __device__ __forceinline__ double lap_xy(const PG &g, const double *q, size_t o)
{
    const size_t X = (size_t) g.n1 * g.n2, Y = g.n2;
    const double *P = q + o;
    if (ORD == 2)
        return ((P[X] - P[0]) / g.dx - (P[0] - P[-(long) X]) / g.dx) / g.dx
             + ((P[Y] - P[0]) / g.dy - (P[0] - P[-(long) Y]) / g.dy) / g.dy;
    const long x = (long) X, y = (long) Y;
    const double up3h = A1 * (P[2 * x] - P[x]) + A2 * (P[3 * x] - P[0]);
    const double up1h = A1 * (P[x] - P[0]) + A2 * (P[2 * x] - P[-x]);
    const double um1h = A1 * (P[0] - P[-x]) + A2 * (P[x] - P[-2 * x]);
    const double um3h = A1 * (P[-x] - P[-2 * x]) + A2 * (P[0] - P[-3 * x]);
    const double vp3h = A1 * (P[2 * y] - P[y]) + A2 * (P[3 * y] - P[0]);
    const double vp1h = A1 * (P[y] - P[0]) + A2 * (P[2 * y] - P[-y]);
    const double vm1h = A1 * (P[0] - P[-y]) + A2 * (P[y] - P[-2 * y]);
    const double vm3h = A1 * (P[-y] - P[-2 * y]) + A2 * (P[0] - P[-3 * y]);
    return (A1 * (up1h / g.dx - um1h / g.dx) + A2 * (up3h / g.dx - um3h / g.dx)) / g.dx
         + (A1 * (vp1h / g.dy - vm1h / g.dy) + A2 * (vp3h / g.dy - vm3h / g.dy)) / g.dy;
}

/* d2/dz2 of q at level k: cell family (divided by the cell height, the
 * differences by the centre spacings) or face family (the reverse) */
template <int ORD, bool FACE>
static __device__ __forceinline__ double lap_z(const PG &g, const double *q, int k)
{
    const double *dz = g.dz;
    if (!FACE) {
        const double dzp1h = 0.5 * (dz[k + 1] + dz[k]), dzm1h = 0.5 * (dz[k - 1] + dz[k]);
        if (ORD == 2) return ((q[k + 1] - q[k]) / dzp1h - (q[k] - q[k - 1]) / dzm1h) / dz[k];
        const double wp3h = A1 * (q[k + 2] - q[k + 1]) + A2 * (q[k + 3] - q[k]);
        const double wp1h = A1 * (q[k + 1] - q[k]) + A2 * (q[k + 2] - q[k - 1]);
        const double wm1h = A1 * (q[k] - q[k - 1]) + A2 * (q[k + 1] - q[k - 2]);
        const double wm3h = A1 * (q[k - 1] - q[k - 2]) + A2 * (q[k] - q[k - 3]);
        const double dzp3h = 0.5 * (dz[k + 2] + dz[k + 1]), dzm3h = 0.5 * (dz[k - 2] + dz[k - 1]);
        return (A1 * (wp1h / dzp1h - wm1h / dzm1h) + A2 * (wp3h / dzp3h - wm3h / dzm3h)) / dz[k];
    }
    const double dzp1 = dz[k + 1], dzm0 = dz[k], dzp1h = 0.5 * (dz[k + 1] + dz[k]);
    if (ORD == 2) return ((q[k + 1] - q[k]) / dzp1 - (q[k] - q[k - 1]) / dzm0) / dzp1h;
    const double wp3h = A1 * (q[k + 2] - q[k + 1]) + A2 * (q[k + 3] - q[k]);
    const double wp1h = A1 * (q[k + 1] - q[k]) + A2 * (q[k + 2] - q[k - 1]);
    const double wm1h = A1 * (q[k] - q[k - 1]) + A2 * (q[k + 1] - q[k - 2]);
    const double wm3h = A1 * (q[k - 1] - q[k - 2]) + A2 * (q[k] - q[k - 3]);
    const double dzp2 = dz[k + 2], dzm1 = dz[k - 1];
    return (A1 * (wp1h / dzp1 - wm1h / dzm0) + A2 * (wp3h / dzp2 - wm3h / dzm1)) / dzp1h;
}

/* the 3-D operators: MODE 0 pressure_laplacian (pressure.c:197-291,
 * out = coef lap q), MODE 1 helmholtz_operator (diffuseu.c:362-457,
 * diffusew.c:366-465, out = q + coef lap q) */
template <int ORD, bool FACE, int MODE>
__global__ void k_sh_lap3(PG g, const double *q, double *out, double coef)
{
    PG_POINT(g, i, j, k);
    const int lo = ORD == 2 ? 1 : 3;
    const int ktop = g.n2 - g.ng - (FACE ? 1 : 0);
    if (i < lo || i >= g.n0 - lo || j < lo || j >= g.n1 - lo || k < g.ng || k >= ktop) return;
    const size_t o = pid(g, i, j, k);
    const double *col = q + pid(g, i, j, 0);
    const double L = lap_xy<ORD>(g, q, o) + lap_z<ORD, FACE>(g, col, k);
    out[o] = MODE == 0 ? coef * L : q[o] + coef * L;
}

/* helmholtz1d_operator (diffuseu.c:107-170, diffusew.c:107-172):
 * h = q + nudt d2q/dz2 at every column's interior levels */
template <int ORD, bool FACE>
__global__ void k_sh_helm1d(PG g, const double *q, double *h, double nudt)
{
    PG_POINT(g, i, j, k);
    const int ktop = g.n2 - g.ng - (FACE ? 1 : 0);
    if (k < g.ng || k >= ktop) return;
    const double *col = q + pid(g, i, j, 0);
    h[pid(g, i, j, k)] = col[k] + nudt * lap_z<ORD, FACE>(g, col, k);
}

/* add_laplacexy_operator (diffuseu.c:286-359, diffusew.c:290-365) */
template <int ORD, bool FACE>
__global__ void k_sh_lapxy(PG g, const double *q, double *h, double nu)
{
    PG_POINT(g, i, j, k);
    const int lo = ORD == 2 ? 1 : 3;
    const int ktop = g.n2 - g.ng - (FACE ? 1 : 0);
    if (i < lo || i >= g.n0 - lo || j < lo || j >= g.n1 - lo || k < g.ng || k >= ktop) return;
    const size_t o = pid(g, i, j, k);
    h[o] += nu * lap_xy<ORD>(g, q, o);
}

/* where one wall's ghosts and interior unknowns sit in a column */
struct WallMap {
    int ghost[2][3];        /* level of ghost m at the bottom / top wall */
    int inner[2][3];        /* level of interior unknown t counted from the wall */
    int nfill;              /* ghosts written per wall */
    double G[2][3][3], B[2][3];
};

static __device__ __forceinline__ void fill_ghosts(const WallMap &W, double *col, double bcb, double bct)
{
    for (int s = 0; s < 2; s++) {
        const double bc = s ? bct : bcb;
        const double x0 = col[W.inner[s][0]], x1 = col[W.inner[s][1]], x2 = col[W.inner[s][2]];
        for (int m = 0; m < W.nfill; m++)
            col[W.ghost[s][m]] = W.G[s][m][0] * x0 + W.G[s][m][1] * x1 + W.G[s][m][2] * x2 + W.B[s][m] * bc;
    }
}

/* apply_bc (diffuseu.c:37-104, diffusew.c:37-104): one thread per column,
 * the column's own wall values */
__global__ void k_sh_apply_bc(PG g, double *q, WallMap W, const double *bcb, const double *bct)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1) return;
    fill_ghosts(W, q + (size_t) c * g.n2, bcb[c], bct[c]);
}

/* solve_helmholtz1d (diffuseu.c:173-283, diffusew.c:175-287): one thread
 * per column.  rhs = dvol h minus the wall values' share (the closures'
 * B terms carried through the ghost columns), the 7-band recurrence of
 * the wall-reduced system on host-built factors, ghosts from the closures,
 * every other level 0 (the reference's zeroed work column). */
__global__ void k_sh_solve1d(PG g, const double *h, double *u, WallMap W, const double *tab, const double *dvol,
                             const double *eb, const double *et, int kfirst, int n, double bcb, double bct)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.n0 * g.n1) return;
    const double *hc = h + (size_t) c * g.n2;
    double *x = u + (size_t) c * g.n2;
    const double *TA = tab, *AL = TA + n, *BE = AL + n, *ID = BE + n, *P = ID + n, *Q = P + n, *R = Q + n;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int t = 0; t < n; t++) {
        const double r = hc[kfirst + t] * dvol[t] - eb[t] * bcb - et[t] * bct;
        const double si = (r - TA[t] * s3 - AL[t] * s2 - BE[t] * s1) * ID[t];
        x[kfirst + t] = si;
        s3 = s2; s2 = s1; s1 = si;
    }
    double x1 = 0.0, x2 = 0.0, x3 = 0.0;
    for (int t = n - 1; t >= 0; t--) {
        const double xi = x[kfirst + t] - P[t] * x1 - Q[t] * x2 - R[t] * x3;
        x[kfirst + t] = xi;
        x3 = x2; x2 = x1; x1 = xi;
    }
    for (int k = 0; k < kfirst; k++) x[k] = 0.0;
    for (int k = kfirst + n; k < g.n2; k++) x[k] = 0.0;
    fill_ghosts(W, x, bcb, bct);
}

/* one pencil of a global io-layout field: out[a][b][c] =
 * in[st0 + a][st1 + b][st2 + c] (0-based starts) */
__global__ void k_pencil_gather(const double *in, int ny, int nz, int st0, int st1, int st2, int sz0, int sz1, int sz2,
                                double *out)
{
    const int cc = blockIdx.x * 64 + threadIdx.x;
    const int b = blockIdx.y * 4 + threadIdx.y;
    const int a = blockIdx.z;
    if (cc >= sz2 || b >= sz1 || a >= sz0) return;
    out[((size_t) a * sz1 + b) * sz2 + cc] = in[((size_t) (st0 + a) * ny + st1 + b) * nz + st2 + cc];
}

/* ------------------------------------------------------------ host */

namespace {

/* device copies of the caller's arrays for one call */
struct Dev {
    std::vector<double *> bufs;
    double *put(const double *host, size_t n)
    {
        double *d = nullptr;
        SHIM_HIP(hipMalloc((void **) &d, sizeof(double) * (n ? n : 1)));
        if (host) SHIM_HIP(hipMemcpy(d, host, sizeof(double) * n, hipMemcpyHostToDevice));
        bufs.push_back(d);
        return d;
    }
    void get(double *host, const double *d, size_t n)
    {
        SHIM_HIP(hipMemcpy(host, d, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    ~Dev()
    {
        for (double *b : bufs) (void) hipFree(b);
    }
};

struct Shape {
    int n0, n1, n2, ng, nz;
    size_t N;
};

Shape shape_of(int local_nx, int local_ny, int nz, int gx, int gy, int gz, int order)
{
    SHIM_CHECK(gx == gy && gy == gz, "nghost_x = nghost_y = nghost_z required (got %d %d %d)", gx, gy, gz);
    SHIM_CHECK(gz >= (order == 2 ? 1 : 3), "order %d needs nghost >= %d (pressure.c:775, :798)", order,
               order == 2 ? 1 : 3);
    SHIM_CHECK(local_nx >= 1 && local_ny >= 1 && nz >= 4, "bad grid %d x %d x %d", local_nx, local_ny, nz);
    Shape s;
    s.ng = gz;
    s.n0 = local_nx + 2 * gx;
    s.n1 = local_ny + 2 * gy;
    s.n2 = nz + 2 * gz;
    s.nz = nz;
    s.N = (size_t) s.n0 * s.n1 * s.n2;
    return s;
}

PG pg_of(const Shape &s, double dx, double dy, const double *dz_dev)
{
    PG g;
    g.n0 = s.n0; g.n1 = s.n1; g.n2 = s.n2; g.ng = s.ng;
    g.dx = dx; g.dy = dy; g.dz = dz_dev;
    return g;
}

dim3 point_grid(const Shape &s) { return dim3((s.n2 + 63) / 64, (s.n1 + 3) / 4, s.n0); }
dim3 column_grid(const Shape &s) { return dim3((unsigned) ((s.n0 * s.n1 + 255) / 256)); }

/* the caller's n2 dz values as a level table valid on -3 .. nz + 2
 * (mirrored beyond the caller's ghosts, ghost.c:31-54); dzl = level 0 */
std::vector<double> level_dz(const Shape &s, const double *dz)
{
    std::vector<double> t((size_t) s.nz + 6, 0.0);
    double *dzl = t.data() + 3;
    for (int k = 0; k < s.nz; k++) dzl[k] = dz[s.ng + k];
    for (int k = 1; k <= 3; k++) {
        dzl[-k] = k <= s.ng ? dz[s.ng - k] : dzl[k - 1];
        dzl[s.nz - 1 + k] = k <= s.ng ? dz[s.ng + s.nz - 1 + k] : dzl[s.nz - k];
    }
    return t;
}

/* ghost / interior levels and closures of one column family */
WallMap wall_map(const Shape &s, int fam, int order, int typb, int typt, const double *dzl)
{
    WallMap W;
    memset(&W, 0, sizeof(W));
    W.nfill = order == 2 ? 1 : 3;
    Clos1 cb, ct;
    SHIM_CHECK(cft::closure(fam, order, 0, typb, dzl, s.nz, cb) == 0, "singular bottom ghost rows");
    SHIM_CHECK(cft::closure(fam, order, 1, typt, dzl, s.nz, ct) == 0, "singular top ghost rows");
    const int ng = s.ng, n2 = s.n2;
    for (int m = 0; m < 3; m++) {
        /* cells: ghosts at ng-1-m and n2-ng+m; faces: the wall face is
         * ghost 0 (ng-1 and n2-ng-1) */
        W.ghost[0][m] = ng - 1 - m;
        W.ghost[1][m] = fam == cft::CELL ? n2 - ng + m : n2 - ng - 1 + m;
        W.inner[0][m] = ng + m;
        W.inner[1][m] = fam == cft::CELL ? n2 - ng - 1 - m : n2 - ng - 2 - m;
        for (int t = 0; t < 3; t++) { W.G[0][m][t] = cb.G[m][t]; W.G[1][m][t] = ct.G[m][t]; }
        W.B[0][m] = cb.B[m];
        W.B[1][m] = ct.B[m];
    }
    return W;
}

int cell_bctype(int t)
{
    SHIM_CHECK(t == DIRICHU || t == NEUMANU, "cell bctype %d (DIRICHU or NEUMANU)", t);
    return t == DIRICHU ? cft::DIRICHLET : cft::NEUMANN;
}

void sync_check()
{
    SHIM_HIP(hipGetLastError());
    SHIM_HIP(hipDeviceSynchronize());
}

/* ---- stencil operators ---- */

template <int ORD>
void divergence(double u[], double v[], double w[], double d[], int lnx, int lny, int nz, int gx, int gy, int gz,
                double dx, double dy, double dz[])
{
    const Shape s = shape_of(lnx, lny, nz, gx, gy, gz, ORD);
    Dev D;
    double *du = D.put(u, s.N), *dv = D.put(v, s.N), *dw = D.put(w, s.N), *dd = D.put(d, s.N), *ddz = D.put(dz, s.n2);
    const PG g = pg_of(s, dx, dy, ddz);
    if (ORD == 4) hipLaunchKernelGGL(k_sh_mirror_w, column_grid(s), dim3(256), 0, 0, g, dw);
    hipLaunchKernelGGL(k_sh_div<ORD>, point_grid(s), dim3(64, 4), 0, 0, g, du, dv, dw, dd);
    sync_check();
    D.get(d, dd, s.N);
    if (ORD == 4) D.get(w, dw, s.N);
}

template <int ORD>
void project(double u[], double v[], double w[], double p[], int lnx, int lny, int nz, int gx, int gy, int gz,
             double dx, double dy, double dz[], double dt)
{
    const Shape s = shape_of(lnx, lny, nz, gx, gy, gz, ORD);
    Dev D;
    double *du = D.put(u, s.N), *dv = D.put(v, s.N), *dw = D.put(w, s.N), *dp = D.put(p, s.N), *ddz = D.put(dz, s.n2);
    hipLaunchKernelGGL(k_sh_project<ORD>, point_grid(s), dim3(64, 4), 0, 0, pg_of(s, dx, dy, ddz), du, dv, dw,
                       (const double *) dp, dt);
    sync_check();
    D.get(u, du, s.N);
    D.get(v, dv, s.N);
    D.get(w, dw, s.N);
}

template <int ORD, bool FACE, int MODE>
void lap3(double q[], double out[], int lnx, int lny, int nz, int gx, int gy, int gz, double dx, double dy,
          double dz[], double coef)
{
    const Shape s = shape_of(lnx, lny, nz, gx, gy, gz, ORD);
    Dev D;
    double *dq = D.put(q, s.N), *dout = D.put(out, s.N), *ddz = D.put(dz, s.n2);
    hipLaunchKernelGGL((k_sh_lap3<ORD, FACE, MODE>), point_grid(s), dim3(64, 4), 0, 0, pg_of(s, dx, dy, ddz),
                       (const double *) dq, dout, coef);
    sync_check();
    D.get(out, dout, s.N);
}

template <int ORD, bool FACE>
void helm1d(double q[], double h[], int lnx, int lny, int nz, int gx, int gy, int gz, double dz[], double nudt)
{
    const Shape s = shape_of(lnx, lny, nz, gx, gy, gz, ORD);
    Dev D;
    double *dq = D.put(q, s.N), *dh = D.put(h, s.N), *ddz = D.put(dz, s.n2);
    hipLaunchKernelGGL((k_sh_helm1d<ORD, FACE>), point_grid(s), dim3(64, 4), 0, 0, pg_of(s, 1.0, 1.0, ddz),
                       (const double *) dq, dh, nudt);
    sync_check();
    D.get(h, dh, s.N);
}

template <int ORD, bool FACE>
void lapxy(double q[], double h[], int lnx, int lny, int nz, int gx, int gy, int gz, double dx, double dy, double nu)
{
    const Shape s = shape_of(lnx, lny, nz, gx, gy, gz, ORD);
    Dev D;
    double *dq = D.put(q, s.N), *dh = D.put(h, s.N);
    hipLaunchKernelGGL((k_sh_lapxy<ORD, FACE>), point_grid(s), dim3(64, 4), 0, 0, pg_of(s, dx, dy, nullptr),
                       (const double *) dq, dh, nu);
    sync_check();
    D.get(h, dh, s.N);
}

/* ---- wall ghosts and column solves ---- */

void apply_bc(int fam, int order, double q[], int lnx, int lny, int nz, int gx, int gy, int gz, const double xq[],
              const double yq[], double dz[], double (*bc_bot)(double, double), double (*bc_top)(double, double),
              int typb, int typt)
{
    const Shape s = shape_of(lnx, lny, nz, gx, gy, gz, order);
    SHIM_CHECK(bc_bot && bc_top && xq && yq, "null boundary function or coordinates");
    if (fam == cft::FACE) {
        SHIM_CHECK(typb == DIRICHW && typt == DIRICHW, "w bctype must be DIRICHW");
    }
    const std::vector<double> dzt = level_dz(s, dz);
    const WallMap W = wall_map(s, fam, order, fam == cft::CELL ? cell_bctype(typb) : cft::DIRICHLET,
                               fam == cft::CELL ? cell_bctype(typt) : cft::DIRICHLET, dzt.data() + 3);
    std::vector<double> bb((size_t) s.n0 * s.n1), bt(bb.size());
    for (int i = 0; i < s.n0; i++)
        for (int j = 0; j < s.n1; j++) {
            bb[(size_t) i * s.n1 + j] = bc_bot(xq[i], yq[j]);
            bt[(size_t) i * s.n1 + j] = bc_top(xq[i], yq[j]);
        }
    Dev D;
    double *dq = D.put(q, s.N), *dbb = D.put(bb.data(), bb.size()), *dbt = D.put(bt.data(), bt.size());
    hipLaunchKernelGGL(k_sh_apply_bc, column_grid(s), dim3(256), 0, 0, pg_of(s, 1.0, 1.0, nullptr), dq, W,
                       (const double *) dbb, (const double *) dbt);
    sync_check();
    D.get(q, dq, s.N);
}

void solve1d(int fam, int order, double u[], double h[], int lnx, int lny, int nz, int gx, int gy, int gz,
             double dz[], double nudt, double bcb, double bct, int typb, int typt)
{
    const Shape s = shape_of(lnx, lny, nz, gx, gy, gz, order);
    if (fam == cft::FACE) {
        SHIM_CHECK(typb == DIRICHW && typt == DIRICHW, "w bctype must be DIRICHW");
    }
    const std::vector<double> dzt = level_dz(s, dz);
    const double *dzl = dzt.data() + 3;
    const int cb = fam == cft::CELL ? cell_bctype(typb) : cft::DIRICHLET;
    const int ct = fam == cft::CELL ? cell_bctype(typt) : cft::DIRICHLET;
    const WallMap W = wall_map(s, fam, order, cb, ct, dzl);
    Clos1 clb, clt;
    SHIM_CHECK(cft::closure(fam, order, 0, cb, dzl, nz, clb) == 0 && cft::closure(fam, order, 1, ct, dzl, nz, clt) == 0,
               "singular ghost rows");
    /* interior rows of (I + nudt d2/dz2) dvol; the wall values' share of
     * each row is what its ghost columns contribute through the closures'
     * B terms (moved to the right-hand side) */
    std::vector<cft::Row7> rows;
    cft::op_rows(fam, order, dzl, nz, nudt, 1.0, true, rows);
    const int n = (int) rows.size();
    std::vector<double> eb(n, 0.0), et(n, 0.0), dvol(n, 0.0);
    for (int i = 0; i < n; i++) {
        for (int t = 0; t < 7; t++) {
            const int L = i - 3 + t;
            if (L >= 0 && L < n) continue;
            const bool lo = L < 0;
            const int m = lo ? -1 - L : L - n;
            if (m < 0 || m > 2) continue;
            (lo ? eb : et)[i] += rows[i].c[t] * (lo ? clb : clt).B[m];
        }
        dvol[i] = fam == cft::CELL ? dzl[i] : 0.5 * (dzl[i] + dzl[i + 1]);
    }
    cft::eliminate(fam, nz, clb, clt, rows);
    std::vector<double> tab;
    SHIM_CHECK(cft::hept_factor(rows, tab) == 0, "zero pivot in the z system");
    Dev D;
    double *du = D.put(nullptr, s.N), *dh = D.put(h, s.N), *dtab = D.put(tab.data(), tab.size());
    double *ddv = D.put(dvol.data(), n), *deb = D.put(eb.data(), n), *det = D.put(et.data(), n);
    hipLaunchKernelGGL(k_sh_solve1d, column_grid(s), dim3(256), 0, 0, pg_of(s, 1.0, 1.0, nullptr), (const double *) dh,
                       du, W, (const double *) dtab, (const double *) ddv, (const double *) deb, (const double *) det,
                       s.ng, n, bcb, bct);
    sync_check();
    D.get(u, du, s.N);
}

/* ---- spectral Poisson: the run path's own kernels on a cached run ---- */

struct PoisCache {
    cf_run *run = nullptr;
    int nx = 0, ny = 0, nz = 0, order = 0;
    double dx = 0, dy = 0;
    std::vector<double> dz;
} g_pois;

void pois_release()
{
    if (g_pois.run) (void) cf_run_destroy(g_pois.run);
    g_pois = PoisCache();
}

void poisson(int order, cf_fftw_complex cout[], int nx, int lnx, int lxs, int ny, int lny, int lys, int nz, int gz,
             double dx, double dy, double dz[], double dt)
{
    SHIM_CHECK(lnx == nx / 2 + 1 && lxs == 0 && lny == ny && lys == 0,
               "one GPU holds every mode: local_nx = nx/2+1, local_ny = ny, starts 0");
    const Shape s = shape_of(nx, ny, nz, gz, gz, gz, order);
    std::vector<double> dzi(dz + gz, dz + gz + nz);
    if (!(g_pois.run && g_pois.nx == nx && g_pois.ny == ny && g_pois.nz == nz && g_pois.order == order
          && g_pois.dx == dx && g_pois.dy == dy && g_pois.dz == dzi)) {
        pois_release();
        cf_config c;
        cf_config_default(&c);
        c.nx = nx; c.ny = ny; c.nz = nz; c.order = order;
        c.lx = dx * nx; c.ly = dy * ny;
        c.lz = 0.0;
        for (double v : dzi) c.lz += v;
        c.stretch = CF_STRETCH_USER;
        c.dz_user = dzi.data();
        cf_status st = cf_run_create(&c, &g_pois.run);
        SHIM_CHECK(st == CF_OK, "pressure_poisson%d: %s", order, cf_last_error());
        g_pois.nx = nx; g_pois.ny = ny; g_pois.nz = nz; g_pois.order = order;
        g_pois.dx = dx; g_pois.dy = dy; g_pois.dz = dzi;
    }
    const int nkx = nx / 2 + 1, n2 = s.n2;
    std::vector<double> spec((size_t) nkx * ny * nz * 2);
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < ny; iy++)
            for (int k = 0; k < nz; k++)
                for (int rc = 0; rc < 2; rc++)
                    spec[(((size_t) ix * ny + iy) * nz + k) * 2 + rc] = cout[((size_t) ix * ny + iy) * n2 + gz + k][rc];
    cf_status st = cf_op_poisson(g_pois.run, spec.data(), dt);
    SHIM_CHECK(st == CF_OK, "pressure_poisson%d: %s", order, cf_last_error());
    /* ghost levels from the pressure closures (pressure.c:446-726): Neumann
     * at both walls, the (0, 0) mode Dirichlet (mean 0) at the bottom; the
     * levels the closures do not reach are 0, as in the reference */
    std::vector<double> dzt = level_dz(s, dz);
    const WallMap Wn = wall_map(s, cft::CELL, order, cft::NEUMANN, cft::NEUMANN, dzt.data() + 3);
    const WallMap Wd = wall_map(s, cft::CELL, order, cft::DIRICHLET, cft::NEUMANN, dzt.data() + 3);
    std::vector<double> col(n2);
    for (int ix = 0; ix < nkx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            const WallMap &W = (ix == 0 && iy == 0) ? Wd : Wn;
            for (int rc = 0; rc < 2; rc++) {
                std::fill(col.begin(), col.end(), 0.0);
                for (int k = 0; k < nz; k++) col[gz + k] = spec[(((size_t) ix * ny + iy) * nz + k) * 2 + rc];
                for (int sw = 0; sw < 2; sw++) {
                    const double x0 = col[W.inner[sw][0]], x1 = col[W.inner[sw][1]], x2 = col[W.inner[sw][2]];
                    for (int m = 0; m < W.nfill; m++)
                        col[W.ghost[sw][m]] = W.G[sw][m][0] * x0 + W.G[sw][m][1] * x1 + W.G[sw][m][2] * x2;
                }
                for (int k = 0; k < n2; k++) cout[((size_t) ix * ny + iy) * n2 + k][rc] = col[k];
            }
        }
}

}   /* namespace */

/* ------------------------------------------------------------ C ABI */

extern "C" {

#define SHIM_ARGS_G int local_nx, int local_ny, int nz, int nghost_x, int nghost_y, int nghost_z
#define SHIM_PASS_G local_nx, local_ny, nz, nghost_x, nghost_y, nghost_z

void divergence2(double u[], double v[], double w[], double d[], SHIM_ARGS_G, double dx, double dy, double dz[])
{ divergence<2>(u, v, w, d, SHIM_PASS_G, dx, dy, dz); }
void divergence4(double u[], double v[], double w[], double d[], SHIM_ARGS_G, double dx, double dy, double dz[])
{ divergence<4>(u, v, w, d, SHIM_PASS_G, dx, dy, dz); }
void project2(double u[], double v[], double w[], double p[], SHIM_ARGS_G, double dx, double dy, double dz[], double dt)
{ project<2>(u, v, w, p, SHIM_PASS_G, dx, dy, dz, dt); }
void project4(double u[], double v[], double w[], double p[], SHIM_ARGS_G, double dx, double dy, double dz[], double dt)
{ project<4>(u, v, w, p, SHIM_PASS_G, dx, dy, dz, dt); }
void pressure_laplacian2(double p[], double l[], SHIM_ARGS_G, double dx, double dy, double dz[], double dt)
{ lap3<2, false, 0>(p, l, SHIM_PASS_G, dx, dy, dz, dt); }
void pressure_laplacian4(double p[], double l[], SHIM_ARGS_G, double dx, double dy, double dz[], double dt)
{ lap3<4, false, 0>(p, l, SHIM_PASS_G, dx, dy, dz, dt); }
void pressure_poisson2(cf_fftw_complex cout[], int nx, int local_nx, int local_x_start, int ny, int local_ny,
                       int local_y_start, int nz, int nghost_z, double dx, double dy, double dz[], double dt)
{ poisson(2, cout, nx, local_nx, local_x_start, ny, local_ny, local_y_start, nz, nghost_z, dx, dy, dz, dt); }
void pressure_poisson4(cf_fftw_complex cout[], int nx, int local_nx, int local_x_start, int ny, int local_ny,
                       int local_y_start, int nz, int nghost_z, double dx, double dy, double dz[], double dt)
{ poisson(4, cout, nx, local_nx, local_x_start, ny, local_ny, local_y_start, nz, nghost_z, dx, dy, dz, dt); }
void pressure_initialize2(int nz, int nghost_z) { (void) nz; (void) nghost_z; }
void pressure_initialize4(int nz, int nghost_z) { (void) nz; (void) nghost_z; }
void pressure_finalize2(void) { pois_release(); }
void pressure_finalize4(void) { pois_release(); }

void apply_bc_u2(double u[], SHIM_ARGS_G, double xu[], double yu[], double dz[], double (*bc_bot)(double, double),
                 double (*bc_top)(double, double), int bctype_bot, int bctype_top)
{ apply_bc(cft::CELL, 2, u, SHIM_PASS_G, xu, yu, dz, bc_bot, bc_top, bctype_bot, bctype_top); }
void apply_bc_u4(double u[], SHIM_ARGS_G, double xu[], double yu[], double dz[], double (*bc_bot)(double, double),
                 double (*bc_top)(double, double), int bctype_bot, int bctype_top)
{ apply_bc(cft::CELL, 4, u, SHIM_PASS_G, xu, yu, dz, bc_bot, bc_top, bctype_bot, bctype_top); }
void helmholtz1d_operator_u2(double u[], double h[], SHIM_ARGS_G, double dz[], double nudt)
{ helm1d<2, false>(u, h, SHIM_PASS_G, dz, nudt); }
void helmholtz1d_operator_u4(double u[], double h[], SHIM_ARGS_G, double dz[], double nudt)
{ helm1d<4, false>(u, h, SHIM_PASS_G, dz, nudt); }
void solve_helmholtz1d_u2(double u[], double h[], SHIM_ARGS_G, double dz[], double nudt, double avg_bot,
                          double avg_top, int bctype_bot, int bctype_top)
{ solve1d(cft::CELL, 2, u, h, SHIM_PASS_G, dz, nudt, avg_bot, avg_top, bctype_bot, bctype_top); }
void solve_helmholtz1d_u4(double u[], double h[], SHIM_ARGS_G, double dz[], double nudt, double avg_bot,
                          double avg_top, int bctype_bot, int bctype_top)
{ solve1d(cft::CELL, 4, u, h, SHIM_PASS_G, dz, nudt, avg_bot, avg_top, bctype_bot, bctype_top); }
void add_laplacexy_operator_u2(double u[], double h[], SHIM_ARGS_G, double dx, double dy, double nu)
{ lapxy<2, false>(u, h, SHIM_PASS_G, dx, dy, nu); }
void add_laplacexy_operator_u4(double u[], double h[], SHIM_ARGS_G, double dx, double dy, double nu)
{ lapxy<4, false>(u, h, SHIM_PASS_G, dx, dy, nu); }
void helmholtz_operator_u2(double u[], double h[], SHIM_ARGS_G, double dx, double dy, double dz[], double nudt)
{ lap3<2, false, 1>(u, h, SHIM_PASS_G, dx, dy, dz, nudt); }
void helmholtz_operator_u4(double u[], double h[], SHIM_ARGS_G, double dx, double dy, double dz[], double nudt)
{ lap3<4, false, 1>(u, h, SHIM_PASS_G, dx, dy, dz, nudt); }
void diffuseu_initialize2(int nz, int nghost_z) { (void) nz; (void) nghost_z; }
void diffuseu_initialize4(int nz, int nghost_z) { (void) nz; (void) nghost_z; }
void diffuseu_finalize2(void) {}
void diffuseu_finalize4(void) {}

void apply_bc_w2(double w[], SHIM_ARGS_G, double xw[], double yw[], double dz[], double (*bc_bot)(double, double),
                 double (*bc_top)(double, double), int bctype_bot, int bctype_top)
{ apply_bc(cft::FACE, 2, w, SHIM_PASS_G, xw, yw, dz, bc_bot, bc_top, bctype_bot, bctype_top); }
void apply_bc_w4(double w[], SHIM_ARGS_G, double xw[], double yw[], double dz[], double (*bc_bot)(double, double),
                 double (*bc_top)(double, double), int bctype_bot, int bctype_top)
{ apply_bc(cft::FACE, 4, w, SHIM_PASS_G, xw, yw, dz, bc_bot, bc_top, bctype_bot, bctype_top); }
void helmholtz1d_operator_w2(double w[], double h[], SHIM_ARGS_G, double dz[], double nudt)
{ helm1d<2, true>(w, h, SHIM_PASS_G, dz, nudt); }
void helmholtz1d_operator_w4(double w[], double h[], SHIM_ARGS_G, double dz[], double nudt)
{ helm1d<4, true>(w, h, SHIM_PASS_G, dz, nudt); }
void solve_helmholtz1d_w2(double w[], double h[], SHIM_ARGS_G, double dz[], double nudt, double avg_bot,
                          double avg_top, int bctype_bot, int bctype_top)
{ solve1d(cft::FACE, 2, w, h, SHIM_PASS_G, dz, nudt, avg_bot, avg_top, bctype_bot, bctype_top); }
void solve_helmholtz1d_w4(double w[], double h[], SHIM_ARGS_G, double dz[], double nudt, double avg_bot,
                          double avg_top, int bctype_bot, int bctype_top)
{ solve1d(cft::FACE, 4, w, h, SHIM_PASS_G, dz, nudt, avg_bot, avg_top, bctype_bot, bctype_top); }
void add_laplacexy_operator_w2(double w[], double h[], SHIM_ARGS_G, double dx, double dy, double nu)
{ lapxy<2, true>(w, h, SHIM_PASS_G, dx, dy, nu); }
void add_laplacexy_operator_w4(double w[], double h[], SHIM_ARGS_G, double dx, double dy, double nu)
{ lapxy<4, true>(w, h, SHIM_PASS_G, dx, dy, nu); }
void helmholtz_operator_w2(double w[], double h[], SHIM_ARGS_G, double dx, double dy, double dz[], double nudt)
{ lap3<2, true, 1>(w, h, SHIM_PASS_G, dx, dy, dz, nudt); }
void helmholtz_operator_w4(double w[], double h[], SHIM_ARGS_G, double dx, double dy, double dz[], double nudt)
{ lap3<4, true, 1>(w, h, SHIM_PASS_G, dx, dy, dz, nudt); }
void diffusew_initialize2(int nz, int nghost_z) { (void) nz; (void) nghost_z; }
void diffusew_initialize4(int nz, int nghost_z) { (void) nz; (void) nghost_z; }
void diffusew_finalize2(void) {}
void diffusew_finalize4(void) {}

/* ---- decomp.h ---- */

void distribute(int data1, int proc, int st[], int en[], int sz[])
{
    /* proc blocks of data1 / proc points, the last data1 % proc one larger */
    const int base = data1 / proc, nlarge = data1 - base * proc, nsmall = proc - nlarge;
    int next = 1;
    for (int r = 0; r < proc; r++) {
        const int n = base + (r >= nsmall ? 1 : 0);
        st[r] = next;
        sz[r] = n;
        en[r] = next + n - 1;
        next += n;
    }
}

void partition(int nx, int ny, int nz, int pdim[3], int dims[2], int coord[2], int lstart[3], int lend[3],
               int lsize[3])
{
    const int gsize[3] = { nx, ny, nz };
    for (int a = 0; a < 3; a++) {
        if (pdim[a] == 1) {
            lstart[a] = 1;
            lend[a] = gsize[a];
            lsize[a] = gsize[a];
            continue;
        }
        const int axis = pdim[a] == 2 ? 0 : 1;      /* distributed over dims[0] or dims[1] */
        std::vector<int> st(dims[axis]), en(dims[axis]), sz(dims[axis]);
        distribute(gsize[a], dims[axis], st.data(), en.data(), sz.data());
        lstart[a] = st[coord[axis]];
        lend[a] = en[coord[axis]];
        lsize[a] = sz[coord[axis]];
    }
}

void get_dist(int nx, int ny, int nz, int dims[2], int x1st[], int x1en[], int x1dist[], int y1st[], int y1en[],
              int y1dist[], int y2st[], int y2en[], int y2dist[], int z2st[], int z2en[], int z2dist[])
{
    distribute(nx, dims[0], x1st, x1en, x1dist);
    distribute(ny, dims[0], y1st, y1en, y1dist);
    distribute(ny, dims[1], y2st, y2en, y2dist);
    distribute(nz, dims[1], z2st, z2en, z2dist);
}

decomp_plan *create_decomp_plan(int dims[2], int coord[2], int nx, int ny, int nz)
{
    SHIM_CHECK(dims && coord && dims[0] >= 1 && dims[1] >= 1, "bad process grid");
    decomp_plan *p = (decomp_plan *) calloc(1, sizeof(decomp_plan));
    SHIM_CHECK(p != nullptr, "out of memory");
    int **arr[] = { &p->x1st, &p->x1en, &p->x1dist, &p->y1st, &p->y1en, &p->y1dist,
                    &p->y2st, &p->y2en, &p->y2dist, &p->z2st, &p->z2en, &p->z2dist };
    for (int q = 0; q < 12; q++) {
        *arr[q] = (int *) malloc(sizeof(int) * (size_t) dims[q < 6 ? 0 : 1]);
        SHIM_CHECK(*arr[q] != nullptr, "out of memory");
    }
    get_dist(nx, ny, nz, dims, p->x1st, p->x1en, p->x1dist, p->y1st, p->y1en, p->y1dist, p->y2st, p->y2en,
             p->y2dist, p->z2st, p->z2en, p->z2dist);
    /* x pencils: x local, y over dims[0], z over dims[1]; y pencils: x over
     * dims[0], y local; z pencils: x over dims[0], y over dims[1] */
    int px[3] = { 1, 2, 3 }, py[3] = { 2, 1, 3 }, pz[3] = { 2, 3, 1 };
    partition(nx, ny, nz, px, dims, coord, p->xst, p->xen, p->xsz);
    partition(nx, ny, nz, py, dims, coord, p->yst, p->yen, p->ysz);
    partition(nx, ny, nz, pz, dims, coord, p->zst, p->zen, p->zsz);
    return p;
}

void destroy_decomp_plan(decomp_plan *p)
{
    if (!p) return;
    int *arr[] = { p->x1st, p->x1en, p->x1dist, p->y1st, p->y1en, p->y1dist,
                   p->y2st, p->y2en, p->y2dist, p->z2st, p->z2en, p->z2dist };
    for (int *a : arr) free(a);
    free(p);
}

int cf_pencil_gather(const double *global_io, int nx, int ny, int nz, int dims[2], int coord[2], int pencil,
                     double *out)
{
    if (!global_io || !out || !dims || !coord || pencil < 0 || pencil > 2 || nx < 1 || ny < 1 || nz < 1
        || dims[0] < 1 || dims[1] < 1 || coord[0] < 0 || coord[0] >= dims[0] || coord[1] < 0 || coord[1] >= dims[1]) {
        cf_set_error("cf_pencil_gather: bad argument");
        return -1;
    }
    decomp_plan *p = create_decomp_plan(dims, coord, nx, ny, nz);
    const int *st = pencil == 0 ? p->xst : pencil == 1 ? p->yst : p->zst;
    const int *sz = pencil == 0 ? p->xsz : pencil == 1 ? p->ysz : p->zsz;
    const int s0 = st[0] - 1, s1 = st[1] - 1, s2 = st[2] - 1, z0 = sz[0], z1 = sz[1], z2 = sz[2];
    destroy_decomp_plan(p);
    const size_t N = (size_t) nx * ny * nz, M = (size_t) z0 * z1 * z2;
    double *din = nullptr, *dout = nullptr;
    if (hipMalloc((void **) &din, sizeof(double) * N) != hipSuccess
        || hipMalloc((void **) &dout, sizeof(double) * (M ? M : 1)) != hipSuccess) {
        if (din) (void) hipFree(din);
        cf_set_error("cf_pencil_gather: device allocation failed");
        return -1;
    }
    int rc = 0;
    if (hipMemcpy(din, global_io, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess) rc = -1;
    if (rc == 0 && M) {
        hipLaunchKernelGGL(k_pencil_gather, dim3((z2 + 63) / 64, (z1 + 3) / 4, z0), dim3(64, 4), 0, 0,
                           (const double *) din, ny, nz, s0, s1, s2, z0, z1, z2, dout);
        if (hipGetLastError() != hipSuccess || hipMemcpy(out, dout, sizeof(double) * M, hipMemcpyDeviceToHost) != hipSuccess)
            rc = -1;
    }
    (void) hipFree(din);
    (void) hipFree(dout);
    if (rc) cf_set_error("cf_pencil_gather: HIP failure");
    return rc;
}

}   /* extern "C" */
