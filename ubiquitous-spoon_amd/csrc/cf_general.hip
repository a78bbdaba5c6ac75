/*
 * cf_general.hip -- the general substep path: order 2 or 4 (fourth.c
 * stencils), with or without the passive scalar.  Same device layout as
 * the order-2 fast path (z-slowest, no ghost cells): periodic neighbours
 * by index wrap, wall ghosts from the host-built closures (cf_tables.h).
 *
 * One substep (chanfast.c:395-656 / :1319-1580):
 *   k_grhs_{u,v,w,c}  explicit tendency f (advection, x-y diffusion,
 *                     buoyancy) and the implicit right-hand side
 *                     h = H(q) + gam dt f + zet dt (g + forcing) - lift
 *   k_spike / k_gsolve4  the wall-reduced z solves (order 2 / 4)
 *   k_gplane, k_gmass    plane sums and the mass-flux / scalar-mean forcing
 *   k_div_fftx<.., ORD>, k_ffty, k_poisson2|4, k_ffty, k_fftx_inv  pressure
 *   k_gproject           projection + forcing correction into the state
 * Each stencil kernel reads its neighbours straight from global memory
 * (the caches catch the reuse); one thread per output point.
 */
#include "cf_kernels.h"

static constexpr double A1d = 1.125 / 1.0;      /* fourth.c:3-6 */
static constexpr double A2d = -0.125 / 3.0;
static constexpr double B1d = 1.125 / 2.0;
static constexpr double B2d = -0.125 / 2.0;

/* ------------------------------------------------------------ columns */

/* a cell field's values at levels k-3..k+3 of column cc, ghosts from the
 * closure (ghost m of a wall = G[m] . (x0, x1, x2) + B[m] bc) */
template <int ORD, typename T>
static __device__ __forceinline__ void cell_col7(const T *__restrict__ q, size_t P, size_t cc, int k, int nz,
                                                 const Clos1 *C, double bcb, double bct, T o[7])
{
#pragma unroll
    for (int t = 0; t < 7; t++) {
        const int kk = k - 3 + t;
        o[t] = (kk >= 0 && kk < nz) ? q[(size_t) kk * P + cc] : (T) 0;
    }
    if (k < 3) {
        const Clos1 &B = C[0];
        const T x0 = q[cc];
        const T x1 = ORD == 4 ? q[P + cc] : (T) 0, x2 = ORD == 4 ? q[2 * P + cc] : (T) 0;
#pragma unroll
        for (int m = 0; m < (ORD == 4 ? 3 : 1); m++) {
            const int t = -1 - m - k + 3;
            if (t >= 0) o[t] = (T) B.G[m][0] * x0 + (T) B.G[m][1] * x1 + (T) B.G[m][2] * x2 + (T) (B.B[m] * bcb);
        }
    }
    if (k > nz - 4) {
        const Clos1 &Tp = C[1];
        const T x0 = q[(size_t) (nz - 1) * P + cc];
        const T x1 = ORD == 4 ? q[(size_t) (nz - 2) * P + cc] : (T) 0;
        const T x2 = ORD == 4 ? q[(size_t) (nz - 3) * P + cc] : (T) 0;
#pragma unroll
        for (int m = 0; m < (ORD == 4 ? 3 : 1); m++) {
            const int t = nz + m - k + 3;
            if (t <= 6) o[t] = (T) Tp.G[m][0] * x0 + (T) Tp.G[m][1] * x1 + (T) Tp.G[m][2] * x2 + (T) (Tp.B[m] * bct);
        }
    }
}

/* w at face level kk as the advection terms see it: the wall faces -1 and
 * nz-1 are 0 and the ghost faces mirror oddly about them (apply_bc_w and
 * the in-place ghost rewrite of massflux*4, advect.c:100-131) */
template <typename T>
static __device__ __forceinline__ T wadv(const T *__restrict__ w, size_t P, size_t cc, int kk, int nz)
{
    if (kk >= 0 && kk <= nz - 2) return w[(size_t) kk * P + cc];
    if (kk == -1 || kk == nz - 1) return (T) 0;
    if (kk < -1) return -w[(size_t) (-2 - kk) * P + cc];
    return -w[(size_t) (2 * (nz - 1) - kk) * P + cc];
}

/* w at levels k-3..k+3 for the implicit operator: closure ghosts
 * (apply_bc_w2/4 is re-applied before helmholtz1d_operator_w) */
template <int ORD, typename T>
static __device__ __forceinline__ void face_col7(const T *__restrict__ w, size_t P, size_t cc, int k, int nz,
                                                 const Clos1 *C, T o[7])
{
#pragma unroll
    for (int t = 0; t < 7; t++) {
        const int kk = k - 3 + t;
        o[t] = (kk >= 0 && kk <= nz - 2) ? w[(size_t) kk * P + cc] : (T) 0;
    }
    if (k < 3) {
        const Clos1 &B = C[0];
        const T x0 = w[cc];
        const T x1 = ORD == 4 ? w[P + cc] : (T) 0, x2 = ORD == 4 ? w[2 * P + cc] : (T) 0;
#pragma unroll
        for (int m = 0; m < (ORD == 4 ? 3 : 1); m++) {
            const int t = -1 - m - k + 3;
            if (t >= 0) o[t] = (T) B.G[m][0] * x0 + (T) B.G[m][1] * x1 + (T) B.G[m][2] * x2;
        }
    }
    if (k > nz - 5) {
        const Clos1 &Tp = C[1];
        const T x0 = w[(size_t) (nz - 2) * P + cc];
        const T x1 = ORD == 4 ? w[(size_t) (nz - 3) * P + cc] : (T) 0;
        const T x2 = ORD == 4 ? w[(size_t) (nz - 4) * P + cc] : (T) 0;
#pragma unroll
        for (int m = 0; m < (ORD == 4 ? 3 : 1); m++) {
            const int t = nz - 1 + m - k + 3;
            if (t <= 6) o[t] = (T) Tp.G[m][0] * x0 + (T) Tp.G[m][1] * x1 + (T) Tp.G[m][2] * x2;
        }
    }
}

static __device__ __forceinline__ size_t rowof(int k, int j, int nx, size_t P) { return (size_t) k * P + (size_t) j * nx; }

/* cell/face 1-D Helmholtz operators (diffuseu.c:107-170, diffusew.c:107-172) */
template <int ORD, typename T>
static __device__ __forceinline__ T helm_cell(const T *Q, const double *dzd, int k, double nudtd)
{
    const T A1 = (T) A1d, A2 = (T) A2d;
    const T nudt = (T) nudtd;
    auto dz = [&](int kk) -> T { return (T) dzd[kk]; };
    const T dzp1h = (T) 0.5 * (dz(k + 1) + dz(k));
    const T dzm1h = (T) 0.5 * (dz(k - 1) + dz(k));
    if (ORD == 2) return Q[3] + nudt * (((Q[4] - Q[3]) / dzp1h - (Q[3] - Q[2]) / dzm1h) / dz(k));
    const T wp3h = A1 * (Q[5] - Q[4]) + A2 * (Q[6] - Q[3]);
    const T wp1h = A1 * (Q[4] - Q[3]) + A2 * (Q[5] - Q[2]);
    const T wm1h = A1 * (Q[3] - Q[2]) + A2 * (Q[4] - Q[1]);
    const T wm3h = A1 * (Q[2] - Q[1]) + A2 * (Q[3] - Q[0]);
    const T dzp3h = (T) 0.5 * (dz(k + 2) + dz(k + 1));
    const T dzm3h = (T) 0.5 * (dz(k - 2) + dz(k - 1));
    return Q[3] + nudt * ((A1 * (wp1h / dzp1h - wm1h / dzm1h) + A2 * (wp3h / dzp3h - wm3h / dzm3h)) / dz(k));
}

template <int ORD, typename T>
static __device__ __forceinline__ T helm_face(const T *Q, const double *dzd, int k, double nudtd)
{
    const T A1 = (T) A1d, A2 = (T) A2d;
    const T nudt = (T) nudtd;
    auto dz = [&](int kk) -> T { return (T) dzd[kk]; };
    const T dzp1 = dz(k + 1), dzm0 = dz(k);
    const T dzp1h = (T) 0.5 * (dz(k + 1) + dz(k));
    if (ORD == 2) return Q[3] + nudt * (((Q[4] - Q[3]) / dzp1 - (Q[3] - Q[2]) / dzm0) / dzp1h);
    const T wp3h = A1 * (Q[5] - Q[4]) + A2 * (Q[6] - Q[3]);
    const T wp1h = A1 * (Q[4] - Q[3]) + A2 * (Q[5] - Q[2]);
    const T wm1h = A1 * (Q[3] - Q[2]) + A2 * (Q[4] - Q[1]);
    const T wm3h = A1 * (Q[2] - Q[1]) + A2 * (Q[3] - Q[0]);
    const T dzp2 = dz(k + 2), dzm1 = dz(k - 1);
    return Q[3] + nudt * ((A1 * (wp1h / dzp1 - wm1h / dzm0) + A2 * (wp3h / dzp2 - wm3h / dzm1)) / dzp1h);
}

/* x-y diffusion nu (dxx + dyy) q of one point from its x and y lines
 * (add_laplacexy_operator_*, diffuseu.c:286-359) */
template <int ORD, typename T>
static __device__ __forceinline__ T lapxy(const T *X, const T *Y, T dx, T dy, T nu)
{
    const T A1 = (T) A1d, A2 = (T) A2d;
    if (ORD == 2)
        return nu * (((X[4] - X[3]) / dx - (X[3] - X[2]) / dx) / dx + ((Y[4] - Y[3]) / dy - (Y[3] - Y[2]) / dy) / dy);
    const T up3h = A1 * (X[5] - X[4]) + A2 * (X[6] - X[3]);
    const T up1h = A1 * (X[4] - X[3]) + A2 * (X[5] - X[2]);
    const T um1h = A1 * (X[3] - X[2]) + A2 * (X[4] - X[1]);
    const T um3h = A1 * (X[2] - X[1]) + A2 * (X[3] - X[0]);
    const T vp3h = A1 * (Y[5] - Y[4]) + A2 * (Y[6] - Y[3]);
    const T vp1h = A1 * (Y[4] - Y[3]) + A2 * (Y[5] - Y[2]);
    const T vm1h = A1 * (Y[3] - Y[2]) + A2 * (Y[4] - Y[1]);
    const T vm3h = A1 * (Y[2] - Y[1]) + A2 * (Y[3] - Y[0]);
    return nu * ((A1 * (up1h / dx - um1h / dx) + A2 * (up3h / dx - um3h / dx)) / dx
               + (A1 * (vp1h / dy - vm1h / dy) + A2 * (vp3h / dy - vm3h / dy)) / dy);
}

/* wall values of u/v the advection of u/v uses on the wall faces
 * (advect.c:336-341, :539-554); Qb = levels -2..1, Qt = nz-2..nz+1 */
template <int ORD, typename T>
static __device__ __forceinline__ T wall_interp(T gm2, T gm1, T x0, T x1)
{
    if (ORD == 2) return (T) 0.5 * (gm1 + x0);
    return (T) (0.5 * A2d) * (gm2 + x1) + (T) (0.5 * (A1d + 2.0 * A2d)) * (gm1 + x0);
}

/* block max of the CFL partials (all threads of the block call it) */
static __device__ __forceinline__ void block_max_store(double v, double *__restrict__ out, unsigned bidx)
{
    __shared__ double red[256];
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] = fmax(red[t], red[t + w]);
        __syncthreads();
    }
    if (t == 0) out[bidx] = red[0];
}

/* ------------------------------------------------------------ u */

/* advectu2/4 + laplacexy_u + buoyancy, then the u right-hand side of
 * solve_helmholtz1d_u (chanfast.c:465-477); CFL partials at s == 0 */
template <int ORD, bool SC, typename T>
static __device__ __forceinline__ void rhsu_body(const RunDev &R, const int s,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ c, const T *__restrict__ g, const double *__restrict__ scal,
    T *__restrict__ fo, T *__restrict__ ro, double *__restrict__ cflpart, double cfldt, const Blk b)
{
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = b.x * 64 + threadIdx.x, j = b.y * 4 + threadIdx.y, k = b.z;
    double cflv = 0.0;
    if (i < nx && j < ny) {
        const size_t P = (size_t) nx * ny;
        const double *dz = R.dzt + CF_NGT;
    const T A1 = (T) A1d, A2 = (T) A2d, B1 = (T) B1d, B2 = (T) B2d;
    (void) A1; (void) A2; (void) B1; (void) B2;
    auto DZ = [&](int kk) -> T { return (T) dz[kk]; };
        const T dx = (T) R.g.dx, dy = (T) R.g.dy;
        const size_t cc = (size_t) j * nx + i, q = (size_t) k * P + cc;
        const size_t row = rowof(k, j, nx, P);
        T U[7], UX[7], UY[7];
        cell_col7<ORD>(u, P, cc, k, nz, R.cl.cell, R.ubc, R.ubc, U);
#pragma unroll
        for (int t = 0; t < 7; t++) {
            UX[t] = (ORD == 2 && (t < 2 || t > 4)) ? (T) 0 : u[row + wrapm(i - 3 + t, nx)];
            UY[t] = (ORD == 2 && (t < 2 || t > 4)) ? (T) 0 : u[rowof(k, wrapm(j - 3 + t, ny), nx, P) + i];
        }
        const T Q0 = U[3];
        const T dzk = DZ(k);
        /* massfluxx (advect.c:6-99): fluxes interpolated to the u-point in x */
        auto MU = [&](int d) -> T {     /* at x-position i + d */
            return ORD == 2 ? (T) 0.5 * (dy * dzk * UX[3 + d] + dy * dzk * UX[4 + d])
                            : B1 * (dy * dzk * UX[3 + d] + dy * dzk * UX[4 + d])
                            + B2 * (dy * dzk * UX[2 + d] + dy * dzk * UX[5 + d]);
        };
        const int im = wrapm(i - 1, nx), ip = wrapm(i + 1, nx), ip2 = wrapm(i + 2, nx);
        auto MV = [&](int e) -> T {     /* at y-position j + e */
            const size_t r2 = rowof(k, wrapm(j + e, ny), nx, P);
            if (ORD == 2) return (T) 0.5 * (dx * dzk * v[r2 + i] + dx * dzk * v[r2 + ip]);
            return B1 * (dx * dzk * v[r2 + i] + dx * dzk * v[r2 + ip])
                 + B2 * (dx * dzk * v[r2 + im] + dx * dzk * v[r2 + ip2]);
        };
        const size_t cj = (size_t) j * nx;
        auto MW = [&](int e) -> T {     /* at face level k + e */
            const int kk = k + e;
            if (ORD == 2) return (T) 0.5 * (dx * dy * wadv(w, P, cj + i, kk, nz) + dx * dy * wadv(w, P, cj + ip, kk, nz));
            return B1 * (dx * dy * wadv(w, P, cj + i, kk, nz) + dx * dy * wadv(w, P, cj + ip, kk, nz))
                 + B2 * (dx * dy * wadv(w, P, cj + im, kk, nz) + dx * dy * wadv(w, P, cj + ip2, kk, nz));
        };
        T f;
        const T den = dx * dy * dzk;
        if (ORD == 2) {
            T xp = MU(0) * (T) 0.5 * (Q0 + UX[4]), xm = MU(-1) * (T) 0.5 * (Q0 + UX[2]);
            T yp = MV(0) * (T) 0.5 * (Q0 + UY[4]), ym = MV(-1) * (T) 0.5 * (Q0 + UY[2]);
            const T mw0 = MW(0), mwm = MW(-1);
            T zp = mw0 * (T) 0.5 * (Q0 + U[4]), zm = mwm * (T) 0.5 * (Q0 + U[2]);
            if (k == nz - 1) zp = mw0 * wall_interp<2, T>((T) 0, U[4], U[3], (T) 0);
            if (k == 0) zm = mwm * wall_interp<2, T>((T) 0, U[2], U[3], (T) 0);
            f = -((xp - xm) + (yp - ym) + (zp - zm)) / den;
        } else {
            T xp = MU(0) * (T) 0.5 * (Q0 + UX[4]), xm = MU(-1) * (T) 0.5 * (Q0 + UX[2]);
            T xp3 = MU(1) * (T) 0.5 * (Q0 + UX[6]), xm3 = MU(-2) * (T) 0.5 * (Q0 + UX[0]);
            T yp = MV(0) * (T) 0.5 * (Q0 + UY[4]), ym = MV(-1) * (T) 0.5 * (Q0 + UY[2]);
            T yp3 = MV(1) * (T) 0.5 * (Q0 + UY[6]), ym3 = MV(-2) * (T) 0.5 * (Q0 + UY[0]);
            const T mw0 = MW(0), mwm = MW(-1), mwp = MW(1), mwm2 = MW(-2);
            T zp = mw0 * (T) 0.5 * (Q0 + U[4]), zm = mwm * (T) 0.5 * (Q0 + U[2]);
            T zp3 = mwp * (T) 0.5 * (Q0 + U[6]), zm3 = mwm2 * (T) 0.5 * (Q0 + U[0]);
            /* U[t] is level k - 3 + t; wall values from levels -2..1 / nz-2..nz+1 */
            if (k == nz - 1) {
                const T bct = wall_interp<4>(U[5], U[4], U[3], U[2]);
                zp = mw0 * bct;
                zp3 = -mwm * (T) 0.5 * (U[1] + U[4]) + (T) 2.0 * mw0 * bct;
            }
            if (k == nz - 2) zp3 = mwp * wall_interp<4>(U[6], U[5], U[4], U[3]);
            if (k == 0) {
                const T bcb = wall_interp<4>(U[1], U[2], U[3], U[4]);
                zm = mwm * bcb;
                zm3 = -mw0 * (T) 0.5 * (U[5] + U[2]) + (T) 2.0 * mwm * bcb;
            }
            if (k == 1) zm3 = mwm2 * wall_interp<4>(U[0], U[1], U[2], U[3]);
            f = -((A1 * (xp - xm) + A2 * (xp3 - xm3)) + (A1 * (yp - ym) + A2 * (yp3 - ym3))
                  + (A1 * (zp - zm) + A2 * (zp3 - zm3))) / den;
        }
        /* the x-y diffusion: explicit in f (chanfast, diffuseu.c:286-359), or
         * in the 3-D Helmholtz operator of h when every diffusion term is
         * implicit (channel, helmholtz_operator_u, diffuseu.c:362-457) */
        const T lxy = lapxy<ORD>(UX, UY, dx, dy, (T) 1);
        if (!R.implicit) f += (T) R.nu * lxy;
        if (SC && R.betg[0] != 0.0) {       /* add_buoyancy2/4, scalar.c:6-76 */
            const size_t r0 = row;
            const T cx = ORD == 2 ? (T) 0.5 * (c[r0 + i] + c[r0 + ip])
                                       : B1 * (c[r0 + i] + c[r0 + ip]) + B2 * (c[r0 + im] + c[r0 + ip2]);
            f += cx * (T) R.betg[0];
        }
        const SubCoef &C = R.sc[s];
        T h = helm_cell<ORD>(U, dz, k, C.nuadt);
        if (R.implicit) h += (T) C.nuadt * lxy;
        if (s > 0) h += (T) C.gamdt * f + (T) C.zetdt * (g[q] + (T) scal[R.m_prev_slot[s]]);
        else h += (T) C.gamdt * f;
        h -= (T) R.liftg[((size_t) s * 4 + 0) * nz + k];
        if (s < 2) fo[q] = f;                 /* the last substep's tendency is never read */
        /* the z solve's right-hand side is h dvol (diffuseu.c:202); the
         * spectral solve takes h and applies dvol itself */
        ro[q] = R.implicit ? h : h * dzk;
        if (cflpart) {                        /* get_cfl, get.c:390-417 */
            const double dzph = 0.5 * (dz[k] + dz[k + 1]);
            cflv = fmax(fmax(fabs((double) Q0) * cfldt / R.g.dx, fabs((double) v[q]) * cfldt / R.g.dy),
                        fabs((double) w[q]) * cfldt / dzph);
        }
    }
    if (cflpart) block_max_store(cflv, cflpart, (b.z * b.ny + b.y) * b.nx + b.x);
}

template <int ORD, bool SC, typename T>
__global__ __launch_bounds__(256) void k_grhs_u(const RunDev R, const int s,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ c, const T *__restrict__ g, const double *__restrict__ scal,
    T *__restrict__ fo, T *__restrict__ ro, double *__restrict__ cflpart, double cfldt)
{
    rhsu_body<ORD, SC, T>(R, s, u, v, w, c, g, scal, fo, ro, cflpart, cfldt, blk_native());
}

/* ------------------------------------------------------------ v */

template <int ORD, bool SC, typename T>
static __device__ __forceinline__ void rhsv_body(const RunDev &R, const int s,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ c, const T *__restrict__ g, const double *__restrict__ scal,
    T *__restrict__ fo, T *__restrict__ ro, double *__restrict__ cflpart, double cfldt, const Blk b)
{
    (void) cflpart; (void) cfldt; (void) scal;
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = b.x * 64 + threadIdx.x, j = b.y * 4 + threadIdx.y, k = b.z;
    if (i >= nx || j >= ny) return;
    const size_t P = (size_t) nx * ny;
    const double *dz = R.dzt + CF_NGT;
    const T A1 = (T) A1d, A2 = (T) A2d, B1 = (T) B1d, B2 = (T) B2d;
    (void) A1; (void) A2; (void) B1; (void) B2;
    auto DZ = [&](int kk) -> T { return (T) dz[kk]; };
    const T dx = (T) R.g.dx, dy = (T) R.g.dy;
    const size_t cc = (size_t) j * nx + i, q = (size_t) k * P + cc;
    const size_t row = rowof(k, j, nx, P);
    T V[7], VX[7], VY[7];
    cell_col7<ORD>(v, P, cc, k, nz, R.cl.cell, R.vbc, R.vbc, V);
#pragma unroll
    for (int t = 0; t < 7; t++) {
        VX[t] = (ORD == 2 && (t < 2 || t > 4)) ? (T) 0 : v[row + wrapm(i - 3 + t, nx)];
        VY[t] = (ORD == 2 && (t < 2 || t > 4)) ? (T) 0 : v[rowof(k, wrapm(j - 3 + t, ny), nx, P) + i];
    }
    const T Q0 = V[3];
    const T dzk = DZ(k);
    const int jm = wrapm(j - 1, ny), jp = wrapm(j + 1, ny), jp2 = wrapm(j + 2, ny);
    /* massfluxy (advect.c:132-200): fluxes interpolated to the v-point in y */
    auto MU = [&](int d) -> T {     /* at x-position i + d */
        const int ii = wrapm(i + d, nx);
        const T a0 = u[rowof(k, j, nx, P) + ii], a1 = u[rowof(k, jp, nx, P) + ii];
        if (ORD == 2) return (T) 0.5 * (dy * dzk * a0 + dy * dzk * a1);
        const T am = u[rowof(k, jm, nx, P) + ii], a2 = u[rowof(k, jp2, nx, P) + ii];
        return B1 * (dy * dzk * a0 + dy * dzk * a1) + B2 * (dy * dzk * am + dy * dzk * a2);
    };
    auto MV = [&](int e) -> T {
        return ORD == 2 ? (T) 0.5 * (dx * dzk * VY[3 + e] + dx * dzk * VY[4 + e])
                        : B1 * (dx * dzk * VY[3 + e] + dx * dzk * VY[4 + e])
                        + B2 * (dx * dzk * VY[2 + e] + dx * dzk * VY[5 + e]);
    };
    auto MW = [&](int e) -> T {
        const int kk = k + e;
        const T a0 = wadv(w, P, (size_t) j * nx + i, kk, nz), a1 = wadv(w, P, (size_t) jp * nx + i, kk, nz);
        if (ORD == 2) return (T) 0.5 * (dx * dy * a0 + dx * dy * a1);
        const T am = wadv(w, P, (size_t) jm * nx + i, kk, nz), a2 = wadv(w, P, (size_t) jp2 * nx + i, kk, nz);
        return B1 * (dx * dy * a0 + dx * dy * a1) + B2 * (dx * dy * am + dx * dy * a2);
    };
    T f;
    const T den = dx * dy * dzk;
    if (ORD == 2) {
        T xp = MU(0) * (T) 0.5 * (Q0 + VX[4]), xm = MU(-1) * (T) 0.5 * (Q0 + VX[2]);
        T yp = MV(0) * (T) 0.5 * (Q0 + VY[4]), ym = MV(-1) * (T) 0.5 * (Q0 + VY[2]);
        const T mw0 = MW(0), mwm = MW(-1);
        T zp = mw0 * (T) 0.5 * (Q0 + V[4]), zm = mwm * (T) 0.5 * (Q0 + V[2]);
        if (k == nz - 1) zp = mw0 * wall_interp<2, T>((T) 0, V[4], V[3], (T) 0);
        if (k == 0) zm = mwm * wall_interp<2, T>((T) 0, V[2], V[3], (T) 0);
        f = -((xp - xm) + (yp - ym) + (zp - zm)) / den;
    } else {
        T xp = MU(0) * (T) 0.5 * (Q0 + VX[4]), xm = MU(-1) * (T) 0.5 * (Q0 + VX[2]);
        T xp3 = MU(1) * (T) 0.5 * (Q0 + VX[6]), xm3 = MU(-2) * (T) 0.5 * (Q0 + VX[0]);
        T yp = MV(0) * (T) 0.5 * (Q0 + VY[4]), ym = MV(-1) * (T) 0.5 * (Q0 + VY[2]);
        T yp3 = MV(1) * (T) 0.5 * (Q0 + VY[6]), ym3 = MV(-2) * (T) 0.5 * (Q0 + VY[0]);
        const T mw0 = MW(0), mwm = MW(-1), mwp = MW(1), mwm2 = MW(-2);
        T zp = mw0 * (T) 0.5 * (Q0 + V[4]), zm = mwm * (T) 0.5 * (Q0 + V[2]);
        T zp3 = mwp * (T) 0.5 * (Q0 + V[6]), zm3 = mwm2 * (T) 0.5 * (Q0 + V[0]);
        if (k == nz - 1) {
            const T bct = wall_interp<4>(V[5], V[4], V[3], V[2]);
            zp = mw0 * bct;
            zp3 = -mwm * (T) 0.5 * (V[1] + V[4]) + (T) 2.0 * mw0 * bct;
        }
        if (k == nz - 2) zp3 = mwp * wall_interp<4>(V[6], V[5], V[4], V[3]);
        if (k == 0) {
            const T bcb = wall_interp<4>(V[1], V[2], V[3], V[4]);
            zm = mwm * bcb;
            zm3 = -mw0 * (T) 0.5 * (V[5] + V[2]) + (T) 2.0 * mwm * bcb;
        }
        if (k == 1) zm3 = mwm2 * wall_interp<4>(V[0], V[1], V[2], V[3]);
        f = -((A1 * (xp - xm) + A2 * (xp3 - xm3)) + (A1 * (yp - ym) + A2 * (yp3 - ym3))
              + (A1 * (zp - zm) + A2 * (zp3 - zm3))) / den;
    }
    const T lxy = lapxy<ORD>(VX, VY, dx, dy, (T) 1);
    if (!R.implicit) f += (T) R.nu * lxy;
    if (SC && R.betg[1] != 0.0) {
        const T c0 = c[q], c1 = c[rowof(k, jp, nx, P) + i];
        const T cy = ORD == 2 ? (T) 0.5 * (c0 + c1)
                                   : B1 * (c0 + c1) + B2 * (c[rowof(k, jm, nx, P) + i] + c[rowof(k, jp2, nx, P) + i]);
        f += cy * (T) R.betg[1];
    }
    const SubCoef &C = R.sc[s];
    T h = helm_cell<ORD>(V, dz, k, C.nuadt);
    if (R.implicit) h += (T) C.nuadt * lxy;
    if (s > 0) h += (T) C.gamdt * f + (T) C.zetdt * g[q];
    else h += (T) C.gamdt * f;
    h -= (T) R.liftg[((size_t) s * 4 + 1) * nz + k];
    if (s < 2) fo[q] = f;                     /* the last substep's tendency is never read */
    ro[q] = R.implicit ? h : h * dzk;
}

template <int ORD, bool SC, typename T>
__global__ __launch_bounds__(256) void k_grhs_v(const RunDev R, const int s,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ c, const T *__restrict__ g, const double *__restrict__ scal,
    T *__restrict__ fo, T *__restrict__ ro, double *__restrict__ cflpart, double cfldt)
{
    rhsv_body<ORD, SC, T>(R, s, u, v, w, c, g, scal, fo, ro, cflpart, cfldt, blk_native());
}

/* ------------------------------------------------------------ w */

template <int ORD, bool SC, typename T>
static __device__ __forceinline__ void rhsw_body(const RunDev &R, const int s,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ c, const T *__restrict__ g, const double *__restrict__ scal,
    T *__restrict__ fo, T *__restrict__ ro, double *__restrict__ cflpart, double cfldt, const Blk b)
{
    (void) cflpart; (void) cfldt; (void) scal;
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = b.x * 64 + threadIdx.x, j = b.y * 4 + threadIdx.y, k = b.z;
    if (i >= nx || j >= ny) return;
    const size_t P = (size_t) nx * ny;
    const size_t cc = (size_t) j * nx + i, q = (size_t) k * P + cc;
    if (k == nz - 1) {                    /* top wall face: f = 0 (advect.c:301), no unknown */
        if (s < 2) fo[q] = 0.0;
        ro[q] = 0.0;
        return;
    }
    const double *dz = R.dzt + CF_NGT;
    const T A1 = (T) A1d, A2 = (T) A2d, B1 = (T) B1d, B2 = (T) B2d;
    (void) A1; (void) A2; (void) B1; (void) B2;
    auto DZ = [&](int kk) -> T { return (T) dz[kk]; };
    const T dx = (T) R.g.dx, dy = (T) R.g.dy;
    const size_t row = rowof(k, j, nx, P);
    T WA[7], WH[7], WX[7], WY[7];
#pragma unroll
    for (int t = 0; t < 7; t++) WA[t] = (ORD == 2 && (t < 2 || t > 4)) ? (T) 0 : wadv(w, P, cc, k - 3 + t, nz);
    face_col7<ORD>(w, P, cc, k, nz, R.cl.face, WH);
#pragma unroll
    for (int t = 0; t < 7; t++) {
        WX[t] = (ORD == 2 && (t < 2 || t > 4)) ? (T) 0 : w[row + wrapm(i - 3 + t, nx)];
        WY[t] = (ORD == 2 && (t < 2 || t > 4)) ? (T) 0 : w[rowof(k, wrapm(j - 3 + t, ny), nx, P) + i];
    }
    const T Q0 = WA[3];
    /* u and v at their own levels with closure ghosts, for the z-interpolated fluxes */
    auto ucell = [&](int ii, int jj, int kk, const T *fld, double bcb, double bct) -> T {
        const size_t c2 = (size_t) jj * nx + ii;
        if (kk >= 0 && kk < nz) return fld[(size_t) kk * P + c2];
        const bool lo = kk < 0;
        const Clos1 &Cl = R.cl.cell[lo ? 0 : 1];
        const int m = lo ? -1 - kk : kk - nz;
        const T x0 = fld[(size_t) (lo ? 0 : nz - 1) * P + c2];
        const T x1 = ORD == 4 ? fld[(size_t) (lo ? 1 : nz - 2) * P + c2] : (T) 0;
        const T x2 = ORD == 4 ? fld[(size_t) (lo ? 2 : nz - 3) * P + c2] : (T) 0;
        return (T) Cl.G[m][0] * x0 + (T) Cl.G[m][1] * x1 + (T) Cl.G[m][2] * x2 + (T) (Cl.B[m] * (lo ? bcb : bct));
    };
    /* massfluxz (advect.c:201-283): fluxes interpolated to the w-point in z */
    auto MUZ = [&](int d) -> T {
        const int ii = wrapm(i + d, nx);
        if (ORD == 2) return (T) 0.5 * (dy * DZ(k) * ucell(ii, j, k, u, R.ubc, R.ubc) + dy * DZ(k + 1) * ucell(ii, j, k + 1, u, R.ubc, R.ubc));
        return B1 * (dy * DZ(k) * ucell(ii, j, k, u, R.ubc, R.ubc) + dy * DZ(k + 1) * ucell(ii, j, k + 1, u, R.ubc, R.ubc))
             + B2 * (dy * DZ(k - 1) * ucell(ii, j, k - 1, u, R.ubc, R.ubc) + dy * DZ(k + 2) * ucell(ii, j, k + 2, u, R.ubc, R.ubc));
    };
    auto MVZ = [&](int e) -> T {
        const int jj = wrapm(j + e, ny);
        if (ORD == 2) return (T) 0.5 * (dx * DZ(k) * ucell(i, jj, k, v, R.vbc, R.vbc) + dx * DZ(k + 1) * ucell(i, jj, k + 1, v, R.vbc, R.vbc));
        return B1 * (dx * DZ(k) * ucell(i, jj, k, v, R.vbc, R.vbc) + dx * DZ(k + 1) * ucell(i, jj, k + 1, v, R.vbc, R.vbc))
             + B2 * (dx * DZ(k - 1) * ucell(i, jj, k - 1, v, R.vbc, R.vbc) + dx * DZ(k + 2) * ucell(i, jj, k + 2, v, R.vbc, R.vbc));
    };
    auto MWZ = [&](int e) -> T {     /* at face k + e from WA[2+e..5+e] */
        return ORD == 2 ? (T) 0.5 * (dx * dy * WA[3 + e] + dx * dy * WA[4 + e])
                        : B1 * (dx * dy * WA[3 + e] + dx * dy * WA[4 + e])
                        + B2 * (dx * dy * WA[2 + e] + dx * dy * WA[5 + e]);
    };
    T f;
    const T den = dx * dy * ((T) 0.5 * (DZ(k + 1) + DZ(k)));
    if (ORD == 2) {
        T xp = MUZ(0) * (T) 0.5 * (Q0 + WX[4]), xm = MUZ(-1) * (T) 0.5 * (Q0 + WX[2]);
        T yp = MVZ(0) * (T) 0.5 * (Q0 + WY[4]), ym = MVZ(-1) * (T) 0.5 * (Q0 + WY[2]);
        T zp = MWZ(0) * (T) 0.5 * (Q0 + WA[4]), zm = MWZ(-1) * (T) 0.5 * (Q0 + WA[2]);
        f = -((xp - xm) + (yp - ym) + (zp - zm)) / den;
    } else {
        T xp = MUZ(0) * (T) 0.5 * (Q0 + WX[4]), xm = MUZ(-1) * (T) 0.5 * (Q0 + WX[2]);
        T xp3 = MUZ(1) * (T) 0.5 * (Q0 + WX[6]), xm3 = MUZ(-2) * (T) 0.5 * (Q0 + WX[0]);
        T yp = MVZ(0) * (T) 0.5 * (Q0 + WY[4]), ym = MVZ(-1) * (T) 0.5 * (Q0 + WY[2]);
        T yp3 = MVZ(1) * (T) 0.5 * (Q0 + WY[6]), ym3 = MVZ(-2) * (T) 0.5 * (Q0 + WY[0]);
        T zp = MWZ(0) * (T) 0.5 * (Q0 + WA[4]), zm = MWZ(-1) * (T) 0.5 * (Q0 + WA[2]);
        T zp3 = MWZ(1) * (T) 0.5 * (Q0 + WA[6]), zm3 = MWZ(-2) * (T) 0.5 * (Q0 + WA[0]);
        f = -((A1 * (xp - xm) + A2 * (xp3 - xm3)) + (A1 * (yp - ym) + A2 * (yp3 - ym3))
              + (A1 * (zp - zm) + A2 * (zp3 - zm3))) / den;
    }
    const T lxy = lapxy<ORD>(WX, WY, dx, dy, (T) 1);
    if (!R.implicit) f += (T) R.nu * lxy;
    if (SC && R.betg[2] != 0.0) {
        const T c0 = ucell(i, j, k, c, R.cbot, R.ctop), c1 = ucell(i, j, k + 1, c, R.cbot, R.ctop);
        T cz;
        if (ORD == 2) cz = (T) 0.5 * (c0 + c1);
        else {
            const T cm = ucell(i, j, k - 1, c, R.cbot, R.ctop);
            const T c2 = ucell(i, j, k + 2, c, R.cbot, R.ctop);
            cz = B1 * (c0 + c1) + B2 * (cm + c2);
        }
        f += cz * (T) R.betg[2];
    }
    const SubCoef &C = R.sc[s];
    T h = helm_face<ORD>(WH, dz, k, C.nuadt);
    if (R.implicit) h += (T) C.nuadt * lxy;
    if (s > 0) h += (T) C.gamdt * f + (T) C.zetdt * g[q];
    else h += (T) C.gamdt * f;
    h -= (T) R.liftg[((size_t) s * 4 + 2) * nz + k];
    if (s < 2) fo[q] = f;                     /* the last substep's tendency is never read */
    ro[q] = R.implicit ? h : h * ((T) 0.5 * (DZ(k + 1) + DZ(k)));   /* times the face volume (diffusew.c:205) */
}

template <int ORD, bool SC, typename T>
__global__ __launch_bounds__(256) void k_grhs_w(const RunDev R, const int s,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ c, const T *__restrict__ g, const double *__restrict__ scal,
    T *__restrict__ fo, T *__restrict__ ro, double *__restrict__ cflpart, double cfldt)
{
    rhsw_body<ORD, SC, T>(R, s, u, v, w, c, g, scal, fo, ro, cflpart, cfldt, blk_native());
}

/* ------------------------------------------------------------ c */

/* flux_quick + advect_scalar (scalar.c:79-162), laplacexy with nu / Pr,
 * then the scalar's implicit right-hand side (chanfast.c:540-571) */
template <int ORD, bool SC, typename T>
static __device__ __forceinline__ void rhsc_body(const RunDev &R, const int s,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ c, const T *__restrict__ g, const double *__restrict__ scal,
    T *__restrict__ fo, T *__restrict__ ro, double *__restrict__ cflpart, double cfldt, const Blk b)
{
    (void) cflpart; (void) cfldt;
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = b.x * 64 + threadIdx.x, j = b.y * 4 + threadIdx.y, k = b.z;
    if (i >= nx || j >= ny) return;
    const size_t P = (size_t) nx * ny;
    const double *dz = R.dzt + CF_NGT;
    const T A1 = (T) A1d, A2 = (T) A2d, B1 = (T) B1d, B2 = (T) B2d;
    (void) A1; (void) A2; (void) B1; (void) B2;
    auto DZ = [&](int kk) -> T { return (T) dz[kk]; };
    const T dx = (T) R.g.dx, dy = (T) R.g.dy;
    const size_t cc = (size_t) j * nx + i, q = (size_t) k * P + cc;
    const size_t row = rowof(k, j, nx, P);
    T Cz[7], CX[7], CY[7];
    cell_col7<ORD>(c, P, cc, k, nz, R.cl.cell, R.cbot, R.ctop, Cz);
#pragma unroll
    for (int t = 0; t < 7; t++) {
        CX[t] = (t == 0 || t == 6) && ORD == 2 ? (T) 0 : c[row + wrapm(i - 3 + t, nx)];
        CY[t] = (t == 0 || t == 6) && ORD == 2 ? (T) 0 : c[rowof(k, wrapm(j - 3 + t, ny), nx, P) + i];
    }
    const T qk = (T) -0.125;
    /* QUICK face value upwinded by the sign of the face volume flux */
    auto quick = [&](T fl, T cm, T c0, T c1, T c2) -> T {
        T cf = (T) 0.5 * (c0 + c1);
        cf += fl > (T) 0 ? qk * (c1 - (T) 2.0 * c0 + cm) : qk * (c2 - (T) 2.0 * c1 + c0);
        return fl * cf;
    };
    const int im = wrapm(i - 1, nx), jm = wrapm(j - 1, ny);
    const T fu0 = dy * DZ(k) * u[row + i], fum = dy * DZ(k) * u[row + im];
    const T fv0 = dx * DZ(k) * v[q], fvm = dx * DZ(k) * v[rowof(k, jm, nx, P) + i];
    const T fw0 = dx * dy * wadv(w, P, cc, k, nz), fwm = dx * dy * wadv(w, P, cc, k - 1, nz);
    const T mu0 = quick(fu0, CX[2], CX[3], CX[4], CX[5]), mum = quick(fum, CX[1], CX[2], CX[3], CX[4]);
    const T mv0 = quick(fv0, CY[2], CY[3], CY[4], CY[5]), mvm = quick(fvm, CY[1], CY[2], CY[3], CY[4]);
    const T mw0 = quick(fw0, Cz[2], Cz[3], Cz[4], Cz[5]), mwm = quick(fwm, Cz[1], Cz[2], Cz[3], Cz[4]);
    T f = -((mu0 - mum) + (mv0 - mvm) + (mw0 - mwm)) / (dx * dy * DZ(k));
    const T lxy = lapxy<ORD>(CX, CY, dx, dy, (T) 1);
    if (!R.implicit) f += (T) R.kap * lxy;
    const SubCoef &C = R.sc[s];
    T h = helm_cell<ORD>(Cz, dz, k, C.kadt);
    if (R.implicit) h += (T) C.kadt * lxy;
    if (s > 0) h += (T) C.gamdt * f + (T) C.zetdt * (g[q] + (T) scal[SC_MC0 + (s + 2) % 3]);
    else h += (T) C.gamdt * f;
    h -= (T) R.liftg[((size_t) s * 4 + 3) * nz + k];
    if (s < 2) fo[q] = f;                     /* the last substep's tendency is never read */
    ro[q] = R.implicit ? h : h * DZ(k);
}

template <int ORD, bool SC, typename T>
__global__ __launch_bounds__(256) void k_grhs_c(const RunDev R, const int s,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ c, const T *__restrict__ g, const double *__restrict__ scal,
    T *__restrict__ fo, T *__restrict__ ro, double *__restrict__ cflpart, double cfldt)
{
    rhsc_body<ORD, SC, T>(R, s, u, v, w, c, g, scal, fo, ro, cflpart, cfldt, blk_native());
}

/* ------------------------------------------------------------ solve */

/* order-4 wall-reduced z solves (solve_helmholtz1d_u4 / _w4 with
 * heptdag.c:5-61), constant coefficients from the host, one thread per
 * (column, field), in place.  Fields 0 u, 1 v, 2 w (nz-1 unknowns, top
 * wall := 0), 3 c. */
template <typename T>
static __device__ __forceinline__ void solve4_body(const RunDev &R, const int s,
    T *__restrict__ ru, T *__restrict__ rv, T *__restrict__ rw, T *__restrict__ rc, const Blk b)
{
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t P = (size_t) nx * ny;
    const size_t col = (size_t) b.x * blockDim.x + threadIdx.x;
    const int f = b.y;
    if (col >= P || f > 3) return;
    T *x = f == 0 ? ru : (f == 1 ? rv : (f == 2 ? rw : rc));
    const int type = f == 2 ? 2 : (f == 3 ? 1 : 0);
    const HeptTab HT = R.hp[s][type];
    const int n = HT.n;
    const double *A = HT.t, *AL = A + n, *BE = AL + n, *ID = BE + n, *Pp = ID + n, *Qq = Pp + n, *Rr = Qq + n;
    T *xc = x + col;
    T s1 = 0, s2 = 0, s3 = 0;
    constexpr int BT = 8;
    for (int i0 = 0; i0 < n; i0 += BT) {
        T hb[BT];
#pragma unroll
        for (int t = 0; t < BT; t++) hb[t] = i0 + t < n ? xc[(size_t) (i0 + t) * P] : (T) 0;
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const int i = i0 + t;
            if (i < n) {
                const T si = (hb[t] - (T) A[i] * s3 - (T) AL[i] * s2 - (T) BE[i] * s1) * (T) ID[i];
                xc[(size_t) i * P] = si;
                s3 = s2; s2 = s1; s1 = si;
            }
        }
    }
    T x1 = 0, x2 = 0, x3 = 0;
    for (int i0 = n - 1; i0 >= 0; i0 -= BT) {
        T sb[BT];
#pragma unroll
        for (int t = 0; t < BT; t++) sb[t] = i0 - t >= 0 ? xc[(size_t) (i0 - t) * P] : (T) 0;
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const int i = i0 - t;
            if (i >= 0) {
                const T xi = sb[t] - (T) Pp[i] * x1 - (T) Qq[i] * x2 - (T) Rr[i] * x3;
                xc[(size_t) i * P] = xi;
                x3 = x2; x2 = x1; x1 = xi;
            }
        }
    }
    if (f == 2) xc[(size_t) (nz - 1) * P] = (T) 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gsolve4(const RunDev R, const int s,
    T *__restrict__ ru, T *__restrict__ rv, T *__restrict__ rw, T *__restrict__ rc)
{
    solve4_body<T>(R, s, ru, rv, rw, rc, blk_native());
}

template __global__ void k_gsolve4<double>(const RunDev, const int, double *, double *, double *, double *);
template __global__ void k_gsolve4<float>(const RunDev, const int, float *, float *, float *, float *);

/* ------------------------------------------------------------ constraints */

/* interior plane sums of u* (kind 0, blockIdx.z = 0) and c* (kind 1) for
 * get_xy_avg (get.c:281-307); deterministic fixed-order trees.
 * part[(kind * nz + k) * nb + b] */
template <typename T>
static __device__ __forceinline__ void plane_body(const RunDev &R, const T *__restrict__ q0,
                                                const T *__restrict__ q1, double *__restrict__ part, int nb, const Blk b)
{
    __shared__ double red[4];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t P = (size_t) nx * ny;
    const int lev = b.y, kind = b.z, bb = b.x;
    const T *q = kind == 0 ? q0 : q1;
    const size_t chunk = (P + nb - 1) / nb;
    const size_t lo = (size_t) bb * chunk, hi = lo + chunk < P ? lo + chunk : P;
    const T *pl = q + (size_t) lev * P;
    double acc = 0.0;                             /* FP64 accumulation for either storage */
    /* unrolled so that a thread's loads are in flight together */
#pragma unroll 8
    for (size_t e = lo + threadIdx.x; e < hi; e += 256) acc += (double) pl[e];
    const double tot = block_sum256(acc, red);    /* DPP wave sums, then the four in order */
    if (threadIdx.x == 0) part[((size_t) kind * nz + lev) * nb + bb] = tot;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gplane(const RunDev R, const T *__restrict__ q0,
                                                const T *__restrict__ q1, double *__restrict__ part, int nb)
{
    plane_body<T>(R, q0, q1, part, nb, blk_native());
}

template __global__ void k_gplane<double>(const RunDev, const double *, const double *, double *, int);
template __global__ void k_gplane<float>(const RunDev, const float *, const float *, double *, int);

/* mass-flux (kind 0, chanfast.c:485-530) and scalar-mean (kind 1,
 * :572-617) constraints: plane averages, their wall ghosts from the
 * closures, bulk value, wall gradients of order 2 or 4 (:489-494,
 * :1413-1418), centreline; forcing from the unit-forcing response. */
static __device__ __forceinline__ void mass_body(const RunDev &R, const int s, const double *__restrict__ part, int nb,
                                               double *__restrict__ scal, double *__restrict__ logbuf, const Blk b)
{
    extern __shared__ double avg[];               /* nz + 6: level k at k + 3 */
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz, ord = R.g.order;
    const int kind = b.x;
    const double *dz = R.dzt + CF_NGT;
    const double scale = 1.0 / ((double) nx * (double) ny);
    for (int l = threadIdx.x; l < nz; l += blockDim.x) {
        double acc = 0.0;
        for (int b = 0; b < nb; b++) acc += part[((size_t) kind * nz + l) * nb + b];
        avg[l + 3] = acc * scale;
    }
    __syncthreads();
    double *A = avg + 3;
    /* the volume integral and the height as fixed-order block sums */
    __shared__ double shv[4], shl[4];
    double pv = 0.0, pl = 0.0;
    for (int k = threadIdx.x; k < nz; k += blockDim.x) { pv += A[k] * dz[k]; pl += dz[k]; }
    const double volsum = block_sum256(pv, shv), lzsum = block_sum256(pl, shl);
    if (threadIdx.x != 0) return;
    const double bcb = kind == 0 ? R.ubc : R.cbot, bct = kind == 0 ? R.ubc : R.ctop;
    for (int m = 0; m < 3; m++) {
        const Clos1 &B = R.cl.cell[0], &T = R.cl.cell[1];
        const double b1 = nz > 1 ? A[1] : 0.0, b2 = nz > 2 ? A[2] : 0.0;
        const double t1 = nz > 1 ? A[nz - 2] : 0.0, t2 = nz > 2 ? A[nz - 3] : 0.0;
        A[-1 - m] = B.G[m][0] * A[0] + B.G[m][1] * b1 + B.G[m][2] * b2 + B.B[m] * bcb;
        A[nz + m] = T.G[m][0] * A[nz - 1] + T.G[m][1] * t1 + T.G[m][2] * t2 + T.B[m] * bct;
    }
    double vol = volsum / lzsum;
    const double kap = kind == 0 ? R.nu : R.kap;
    double nb_, nt_;
    if (ord == 2) {
        nb_ = A[0] - A[-1];
        nt_ = A[nz] - A[nz - 1];
    } else {
        nb_ = A1d * (A[0] - A[-1]) + A2d * (A[1] - A[-2]);
        nt_ = A1d * (A[nz] - A[nz - 1]) + A2d * (A[nz + 1] - A[nz - 2]);
    }
    double bot = kap * (nb_ / (0.5 * (dz[0] + dz[-1])));
    double top = kap * (nt_ / (0.5 * (dz[nz] + dz[nz - 1])));
    double cl = nz % 2 == 0 ? 0.5 * (A[nz / 2 - 1] + A[nz / 2]) : A[nz / 2];
    const double *s1 = R.stat1 + (size_t) (s * 2 + kind) * 4;   /* vol1, bot1, top1, cl1 */
    double m = R.cmode[kind] == 1 ? (R.ctarget[kind] - vol) / s1[0] : R.cfixed[kind];
    vol += s1[0] * m;
    bot += s1[1] * m;
    top += s1[2] * m;
    cl += s1[3] * m;
    scal[(kind == 0 ? SC_M0 : SC_MC0) + s] = m;
    if (s == 2) {
        long idx = (long) scal[SC_STEP] % R.logcap;
        double *lg = logbuf + (size_t) idx * 13 + (kind == 0 ? 3 : 8);
        lg[0] = m; lg[1] = bot; lg[2] = top; lg[3] = vol; lg[4] = cl;
    }
}

__global__ __launch_bounds__(256) void k_gmass(const RunDev R, const int s, const double *__restrict__ part, int nb,
                                               double *__restrict__ scal, double *__restrict__ logbuf)
{
    mass_body(R, s, part, nb, scal, logbuf, blk_native());
}

/* ------------------------------------------------------------ Poisson, order 4 */

/* pressure_poisson4 (pressure.c:357-444): per (kx, ky) the wall-reduced
 * heptadiagonal system with the 4th-order modified wavenumbers on the
 * diagonal, factored in the recurrence of heptdag.c:5-61; the U factors
 * go to S (3 per level), the forward sweep in place in Y.  Real and
 * imaginary parts share the factors. */
template <typename Ar, typename C>
static __device__ __forceinline__ void pois4_mode(const Pois4Tab &T, C *__restrict__ y, double *__restrict__ S,
                                                  size_t M, int nz, double lxd, double lyd, int variant)
{
    /* the rows go through registers in blocks of BT: every load of a block
     * is in flight before the first row's recurrence needs it (one wave per
     * 64 modes leaves no other waves to hide a per-row load latency) */
    constexpr int BT = 8;
    const Ar lx = (Ar) lxd, ly = (Ar) lyd;
    const double *bd = T.band[variant];
    double *Sp = S, *Sq = S + (size_t) nz * M, *Sr = S + (size_t) 2 * nz * M;
    Ar p1 = 0, p2 = 0, p3 = 0, q1 = 0, q2 = 0, q3 = 0, r1 = 0, r2 = 0, r3 = 0;
    Ar s1x = 0, s2x = 0, s3x = 0, s1y = 0, s2y = 0, s3y = 0;
    for (int i0 = 0; i0 < nz; i0 += BT) {
        C rb[BT];
#pragma unroll
        for (int t = 0; t < BT; t++) rb[t] = y[(size_t) min(i0 + t, nz - 1) * M];
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const int i = i0 + t;
            if (i < nz) {
                const Ar a = (Ar) bd[i], b = (Ar) bd[nz + i], c = (Ar) bd[2 * nz + i];
                const Ar d = (Ar) bd[3 * nz + i] - lx * (Ar) T.ex[i] - ly * (Ar) T.ey[i];
                const Ar e = (Ar) bd[4 * nz + i], f = (Ar) bd[5 * nz + i], g = (Ar) bd[6 * nz + i];
                const Ar alp = b - a * p3;
                const Ar bet = c - a * q3 - alp * p2;
                const Ar den = d - a * r3 - alp * q2 - bet * p1;
                const Ar iden = (Ar) 1 / den;
                const Ar pi = (e - alp * r2 - bet * q1) * iden;
                const Ar qi = (f - bet * r1) * iden;
                const Ar ri = g * iden;
                const C rh = rb[t];
                const Ar dv = (Ar) T.dvol[i];
                const Ar six = ((Ar) rh.x * dv - a * s3x - alp * s2x - bet * s1x) * iden;
                const Ar siy = ((Ar) rh.y * dv - a * s3y - alp * s2y - bet * s1y) * iden;
                C o;
                o.x = six; o.y = siy;
                y[(size_t) i * M] = o;
                Sp[(size_t) i * M] = pi; Sq[(size_t) i * M] = qi; Sr[(size_t) i * M] = ri;
                p3 = p2; p2 = p1; p1 = pi;
                q3 = q2; q2 = q1; q1 = qi;
                r3 = r2; r2 = r1; r1 = ri;
                s3x = s2x; s2x = s1x; s1x = six;
                s3y = s2y; s2y = s1y; s1y = siy;
            }
        }
    }
    Ar x1x = 0, x2x = 0, x3x = 0, x1y = 0, x2y = 0, x3y = 0;
    for (int i0 = nz - 1; i0 >= 0; i0 -= BT) {
        C sb[BT];
        double pb[BT], qb[BT], rbv[BT];
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const size_t o = (size_t) max(i0 - t, 0) * M;
            sb[t] = y[o]; pb[t] = Sp[o]; qb[t] = Sq[o]; rbv[t] = Sr[o];
        }
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const int i = i0 - t;
            if (i >= 0) {
                const Ar xix = (Ar) sb[t].x - (Ar) pb[t] * x1x - (Ar) qb[t] * x2x - (Ar) rbv[t] * x3x;
                const Ar xiy = (Ar) sb[t].y - (Ar) pb[t] * x1y - (Ar) qb[t] * x2y - (Ar) rbv[t] * x3y;
                C o;
                o.x = xix; o.y = xiy;
                y[(size_t) i * M] = o;
                x3x = x2x; x2x = x1x; x1x = xix;
                x3y = x2y; x2y = x1y; x1y = xiy;
            }
        }
    }
}

template <typename T>
static __device__ __forceinline__ void pois4_body(const RunDev &R, const Pois4Tab T_,
                                                  typename Cplx<T>::type *__restrict__ Y, double *__restrict__ S, const Blk b)
{
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const size_t M = (size_t) ny * nkx;
    const size_t qm = (size_t) b.x * blockDim.x + threadIdx.x;
    if (qm >= M) return;
    const int kx = (int) (qm / ny), ky = (int) (qm % ny);
    const int variant = (kx == 0 && ky == 0) ? 1 : 0;
    const bool low = kx <= CF_F64_KLOW && (ky <= CF_F64_KLOW || ny - ky <= CF_F64_KLOW);
    if (sizeof(T) == sizeof(double) || low)
        pois4_mode<double>(T_, Y + qm, S + qm, M, nz, R.lamx[kx], R.lamy[ky], variant);
    else
        pois4_mode<T>(T_, Y + qm, S + qm, M, nz, R.lamx[kx], R.lamy[ky], variant);
}

template <typename T>
__global__ __launch_bounds__(256) void k_poisson4(const RunDev R, const Pois4Tab T_,
                                                  typename Cplx<T>::type *__restrict__ Y, double *__restrict__ S)
{
    pois4_body<T>(R, T_, Y, S, blk_native());
}

template __global__ void k_poisson4<double>(const RunDev, const Pois4Tab, double2 *, double *);
template __global__ void k_poisson4<float>(const RunDev, const Pois4Tab, float2 *, double *);

/* The same solve split at the middle row m = nz / 2 ("twisted"), for
 * nz >= 6: wave 0 of a 128-thread workgroup eliminates rows 0..m-1 of 64
 * modes from the top (the recurrence above), wave 1 rows nz-1..m from the
 * bottom (the same recurrence on the mirrored bands), so each dependent
 * chain is half as long and twice as many waves run.  Each side's last
 * three relations,
 *   top:    x_R + p x_{R+1} + q x_{R+2} + r x_{R+3} = s   (R = m-3..m-1)
 *   bottom: x_R + p x_{R-1} + q x_{R-2} + r x_{R-3} = s   (R = m..m+2),
 * couple the six unknowns x_{m-3..m+2}; both waves solve that 6x6 system
 * (unit diagonal, no pivoting) and back-substitute their own half.  Every
 * row is still eliminated once, from the rows on its side only, as in the
 * sequential elimination; results differ from it by round-off. */
template <typename Ar, typename C>
static __device__ __forceinline__ void p4t_sweep(const Pois4Tab &T, C *__restrict__ y, double *__restrict__ S,
                                                 size_t M, int nz, double lxd, double lyd, int variant, int side,
                                                 double *__restrict__ xme)
{
    constexpr int BT = 8;
    const int m = nz / 2, cnt = side == 0 ? m : nz - m;
    const Ar lx = (Ar) lxd, ly = (Ar) lyd;
    const double *bd = T.band[variant];
    /* band offsets -3..+3 as seen from this side (mirrored for the bottom) */
    const double *ba = bd + (side == 0 ? 0 : 6) * nz, *bb = bd + (side == 0 ? 1 : 5) * nz, *bc = bd + (side == 0 ? 2 : 4) * nz;
    const double *be = bd + (side == 0 ? 4 : 2) * nz, *bf = bd + (side == 0 ? 5 : 1) * nz, *bg = bd + (side == 0 ? 6 : 0) * nz;
    const double *bdd = bd + 3 * nz;
    double *Sp = S, *Sq = S + (size_t) nz * M, *Sr = S + (size_t) 2 * nz * M;
    Ar p1 = 0, p2 = 0, p3 = 0, q1 = 0, q2 = 0, q3 = 0, r1 = 0, r2 = 0, r3 = 0;
    Ar s1x = 0, s2x = 0, s3x = 0, s1y = 0, s2y = 0, s3y = 0;
    for (int j0 = 0; j0 < cnt; j0 += BT) {
        C rb[BT];
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const int j = min(j0 + t, cnt - 1);
            rb[t] = y[(size_t) (side == 0 ? j : nz - 1 - j) * M];
        }
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const int j = j0 + t;
            if (j < cnt) {
                const int i = side == 0 ? j : nz - 1 - j;
                const Ar a = (Ar) ba[i], b = (Ar) bb[i], c = (Ar) bc[i];
                const Ar d = (Ar) bdd[i] - lx * (Ar) T.ex[i] - ly * (Ar) T.ey[i];
                const Ar e = (Ar) be[i], f = (Ar) bf[i], g = (Ar) bg[i];
                const Ar alp = b - a * p3;
                const Ar bet = c - a * q3 - alp * p2;
                const Ar den = d - a * r3 - alp * q2 - bet * p1;
                const Ar iden = (Ar) 1 / den;
                const Ar pi = (e - alp * r2 - bet * q1) * iden;
                const Ar qi = (f - bet * r1) * iden;
                const Ar ri = g * iden;
                const Ar dv = (Ar) T.dvol[i];
                const Ar six = ((Ar) rb[t].x * dv - a * s3x - alp * s2x - bet * s1x) * iden;
                const Ar siy = ((Ar) rb[t].y * dv - a * s3y - alp * s2y - bet * s1y) * iden;
                C o;
                o.x = six; o.y = siy;
                y[(size_t) i * M] = o;
                Sp[(size_t) i * M] = pi; Sq[(size_t) i * M] = qi; Sr[(size_t) i * M] = ri;
                p3 = p2; p2 = p1; p1 = pi;
                q3 = q2; q2 = q1; q1 = qi;
                r3 = r2; r2 = r1; r1 = ri;
                s3x = s2x; s2x = s1x; s1x = six;
                s3y = s2y; s2y = s1y; s1y = siy;
            }
        }
    }
    /* the relations of local rows cnt-1, cnt-2, cnt-3 (five values each) */
    xme[0 * 64] = (double) p1; xme[1 * 64] = (double) q1; xme[2 * 64] = (double) r1;
    xme[3 * 64] = (double) s1x; xme[4 * 64] = (double) s1y;
    xme[5 * 64] = (double) p2; xme[6 * 64] = (double) q2; xme[7 * 64] = (double) r2;
    xme[8 * 64] = (double) s2x; xme[9 * 64] = (double) s2y;
    xme[10 * 64] = (double) p3; xme[11 * 64] = (double) q3; xme[12 * 64] = (double) r3;
    xme[13 * 64] = (double) s3x; xme[14 * 64] = (double) s3y;
}

template <typename Ar, typename C>
static __device__ __forceinline__ void p4t_finish(C *__restrict__ y, const double *__restrict__ S, size_t M, int nz,
                                                  int side, const double *__restrict__ xt, const double *__restrict__ xb)
{
    constexpr int BT = 8;
    const int m = nz / 2, cnt = side == 0 ? m : nz - m;
    /* the 6x6 system over z_u = x_{m-3+u}: rows 0..2 the top relations of
     * R = m-3, m-2, m-1 (exported as local rows cnt-3, cnt-2, cnt-1 of the
     * top), rows 3..5 the bottom relations of R = m, m+1, m+2 */
    Ar A[6][6], bx[6], by[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
#pragma unroll
        for (int c = 0; c < 6; c++) A[r][c] = (Ar) 0;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        /* top: relation of R = m-1-k is the export slot k (local row cnt-1-k) */
        const int r = 2 - k;                      /* matrix row and unknown index of R */
        A[r][r] = (Ar) 1;
        A[r][r + 1] = (Ar) xt[(5 * k + 0) * 64];
        A[r][r + 2] = (Ar) xt[(5 * k + 1) * 64];
        A[r][r + 3] = (Ar) xt[(5 * k + 2) * 64];
        bx[r] = (Ar) xt[(5 * k + 3) * 64];
        by[r] = (Ar) xt[(5 * k + 4) * 64];
        /* bottom: relation of R = m+k is the export slot k (local row cnt_b-1-k) */
        const int rb = 3 + k;
        A[rb][rb] = (Ar) 1;
        A[rb][rb - 1] = (Ar) xb[(5 * k + 0) * 64];
        A[rb][rb - 2] = (Ar) xb[(5 * k + 1) * 64];
        A[rb][rb - 3] = (Ar) xb[(5 * k + 2) * 64];
        bx[rb] = (Ar) xb[(5 * k + 3) * 64];
        by[rb] = (Ar) xb[(5 * k + 4) * 64];
    }
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const Ar ip = (Ar) 1 / A[c][c];
#pragma unroll
        for (int r = c + 1; r < 6; r++) {
            const Ar f = A[r][c] * ip;
#pragma unroll
            for (int k = c + 1; k < 6; k++) A[r][k] -= f * A[c][k];
            bx[r] -= f * bx[c];
            by[r] -= f * by[c];
        }
    }
    Ar zx[6], zy[6];
#pragma unroll
    for (int r = 5; r >= 0; r--) {
        Ar vx = bx[r], vy = by[r];
#pragma unroll
        for (int k = r + 1; k < 6; k++) { vx -= A[r][k] * zx[k]; vy -= A[r][k] * zy[k]; }
        const Ar ip = (Ar) 1 / A[r][r];
        zx[r] = vx * ip;
        zy[r] = vy * ip;
    }
    const double *Sp = S, *Sq = S + (size_t) nz * M, *Sr = S + (size_t) 2 * nz * M;
    /* own three unknowns of the six, then back-substitution of the rest of
     * this side, moving away from the middle */
    Ar x1x, x2x, x3x, x1y, x2y, x3y;
    if (side == 0) {
#pragma unroll
        for (int u = 0; u < 3; u++) {
            C o; o.x = zx[u]; o.y = zy[u];
            y[(size_t) (m - 3 + u) * M] = o;
        }
        x1x = zx[0]; x2x = zx[1]; x3x = zx[2]; x1y = zy[0]; x2y = zy[1]; x3y = zy[2];
    } else {
#pragma unroll
        for (int u = 3; u < 6; u++) {
            C o; o.x = zx[u]; o.y = zy[u];
            y[(size_t) (m - 3 + u) * M] = o;
        }
        x1x = zx[5]; x2x = zx[4]; x3x = zx[3]; x1y = zy[5]; x2y = zy[4]; x3y = zy[3];
    }
    for (int j0 = cnt - 4; j0 >= 0; j0 -= BT) {
        C sb[BT];
        double pb[BT], qb[BT], rbv[BT];
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const int j = max(j0 - t, 0);
            const size_t o = (size_t) (side == 0 ? j : nz - 1 - j) * M;
            sb[t] = y[o]; pb[t] = Sp[o]; qb[t] = Sq[o]; rbv[t] = Sr[o];
        }
#pragma unroll
        for (int t = 0; t < BT; t++) {
            const int j = j0 - t;
            if (j >= 0) {
                const Ar xix = (Ar) sb[t].x - (Ar) pb[t] * x1x - (Ar) qb[t] * x2x - (Ar) rbv[t] * x3x;
                const Ar xiy = (Ar) sb[t].y - (Ar) pb[t] * x1y - (Ar) qb[t] * x2y - (Ar) rbv[t] * x3y;
                C o;
                o.x = xix; o.y = xiy;
                y[(size_t) (side == 0 ? j : nz - 1 - j) * M] = o;
                x3x = x2x; x2x = x1x; x1x = xix;
                x3y = x2y; x2y = x1y; x1y = xiy;
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(128) void k_poisson4t(const RunDev R, const Pois4Tab T_,
                                                   typename Cplx<T>::type *__restrict__ Y, double *__restrict__ S)
{
    typedef typename Cplx<T>::type C;
    __shared__ double xch[2][15 * 64];
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const size_t M = (size_t) ny * nkx;
    const int lane = threadIdx.x & 63;
    const int side = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   /* 0 top, 1 bottom */
    const size_t q0 = (size_t) blockIdx.x * 64 + lane;
    const bool act = q0 < M;
    const size_t qm = act ? q0 : M - 1;           /* idle lanes repeat the last mode, and never write */
    const int kx = (int) (qm / ny), ky = (int) (qm % ny);
    const int variant = (kx == 0 && ky == 0) ? 1 : 0;
    const bool low = kx <= CF_F64_KLOW && (ky <= CF_F64_KLOW || ny - ky <= CF_F64_KLOW);
    const bool wide = sizeof(T) == sizeof(double) || low;
    /* lanes past the last mode take part in the barrier only */
    double *xme = &xch[side][lane];
    if (act) {
        if (wide) p4t_sweep<double, C>(T_, Y + qm, S + qm, M, nz, R.lamx[kx], R.lamy[ky], variant, side, xme);
        else p4t_sweep<T, C>(T_, Y + qm, S + qm, M, nz, R.lamx[kx], R.lamy[ky], variant, side, xme);
    }
    __syncthreads();
    if (!act) return;
    const double *xt = &xch[0][lane], *xb = &xch[1][lane];
    if (wide) p4t_finish<double, C>(Y + qm, S + qm, M, nz, side, xt, xb);
    else p4t_finish<T, C>(Y + qm, S + qm, M, nz, side, xt, xb);
}

template __global__ void k_poisson4t<double>(const RunDev, const Pois4Tab, double2 *, double *);
template __global__ void k_poisson4t<float>(const RunDev, const Pois4Tab, float2 *, double *);

/* ------------------------------------------------------------ projection */

/* project2/4 (pressure.c:120-195) of the mass-flux corrected solve
 * output into the state: u = (u* + h1 m) - dt dp/dx, v, w likewise (the
 * top-wall face := 0), c = c* + h1c mc.  For order 4 the first pressure
 * ghost equals the adjacent interior value (the copy row of the closure,
 * pressure.c:534-632), which w at the first and last interior face uses. */
/* (us, u), (vs, v), (ws, w) may be the same arrays (in-place projection):
 * each thread reads and writes only its own point of them */
template <int ORD, typename T>
static __device__ __forceinline__ void proj_body(const RunDev &R, const T *us,
    const T *vs, const T *ws, const T *cs,
    const double *__restrict__ h1u, const double *__restrict__ mu, const double *__restrict__ h1c,
    const double *__restrict__ mc, const T *__restrict__ p, double dtp,
    T *u, T *v, T *w, T *c, const Blk b)
{
    const T A1 = (T) A1d, A2 = (T) A2d;
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = b.x * 64 + threadIdx.x;
    const int j = b.y * 4 + threadIdx.y;
    const int k = b.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const size_t P = (size_t) nx * ny;
    const size_t row = rowof(k, j, nx, P), q = row + i;
    const double *rdzh = R.rdzh + CF_NGT;
    const T m = mu ? (T) *mu : (T) 0;
    const T p0 = p[q];
    const int ip = wrapm(i + 1, nx), jp = wrapm(j + 1, ny);
    T gx, gy;
    if (ORD == 2) {
        gx = p[row + ip] - p0;
        gy = p[rowof(k, jp, nx, P) + i] - p0;
    } else {
        const int im = wrapm(i - 1, nx), ip2 = wrapm(i + 2, nx), jm = wrapm(j - 1, ny), jp2 = wrapm(j + 2, ny);
        gx = A1 * (p[row + ip] - p0) + A2 * (p[row + ip2] - p[row + im]);
        gy = A1 * (p[rowof(k, jp, nx, P) + i] - p0) + A2 * (p[rowof(k, jp2, nx, P) + i] - p[rowof(k, jm, nx, P) + i]);
    }
    u[q] = (us[q] + (h1u ? (T) h1u[k + 1] * m : (T) 0)) - gx * (T) (dtp * R.rdx);
    v[q] = vs[q] - gy * (T) (dtp * R.rdy);
    if (k < nz - 1) {
        const T p1 = p[q + P];
        T gz;
        if (ORD == 2) gz = p1 - p0;
        else {
            const T pm = k > 0 ? p[q - P] : p0;
            const T p2 = k + 2 < nz ? p[q + 2 * P] : p1;
            gz = A1 * (p1 - p0) + A2 * (p2 - pm);
        }
        w[q] = ws[q] - gz * (T) (dtp * rdzh[k]);
    } else {
        w[q] = (T) 0;
    }
    if (c) c[q] = cs[q] + (h1c ? (T) h1c[k + 1] * (mc ? (T) *mc : (T) 0) : (T) 0);
}

template <int ORD, typename T>
__global__ __launch_bounds__(256) void k_gproject(const RunDev R, const T *us,
    const T *vs, const T *ws, const T *cs,
    const double *__restrict__ h1u, const double *__restrict__ mu, const double *__restrict__ h1c,
    const double *__restrict__ mc, const T *__restrict__ p, double dtp,
    T *u, T *v, T *w, T *c)
{
    proj_body<ORD, T>(R, us, vs, ws, cs, h1u, mu, h1c, mc, p, dtp, u, v, w, c, blk_native());
}

template <typename T> struct GPick {
    template <int ORD, bool SC> static void fill(GSel<T> &S)
    {
        S.u = k_grhs_u<ORD, SC, T>;
        S.v = k_grhs_v<ORD, SC, T>;
        S.w = k_grhs_w<ORD, SC, T>;
        S.c = SC ? k_grhs_c<ORD, true, T> : nullptr;
        S.proj = k_gproject<ORD, T>;
    }
};

template <typename T>
GSel<T> pick_general(int order, int scalar)
{
    GSel<T> S;
    if (order == 4) {
        if (scalar) GPick<T>::template fill<4, true>(S);
        else GPick<T>::template fill<4, false>(S);
    } else {
        if (scalar) GPick<T>::template fill<2, true>(S);
        else GPick<T>::template fill<2, false>(S);
    }
    return S;
}

template GSel<double> pick_general<double>(int, int);
template GSel<float> pick_general<float>(int, int);

/* ------------------------------------------------------------ batched launches
 * Runs of one (order, scalar, precision) with any grid sizes and
 * coefficients go out together: each launch below is persistent over the
 * virtual blocks of all runs (batch_loop), each run reading its own
 * parameter block and buffers from the GArgs table.  Substep index s in
 * 0..2; s == 3 is a stand-alone projection of the state (dt = dt_adhoc). */

template <int ORD, bool SC, typename T, int F>
__global__ __launch_bounds__(256) void kb_rhs(const GArgs<T> *__restrict__ A, const unsigned *__restrict__ vs,
                                              int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = A[r];
        T *const *fnew = (s % 2 == 0) ? a.fa : a.fb;
        T *const *gold = (s % 2 == 0) ? a.fb : a.fa;
        const Blk b = blk_of(a, ST_CELL, lb);
        if (F == 0)
            rhsu_body<ORD, SC, T>(a.R, s, a.u, a.v, a.w, a.c, gold[0], a.scal, fnew[0], a.ru,
                                  s == 0 ? a.cflpart : (double *) nullptr, a.dt, b);
        else if (F == 1)
            rhsv_body<ORD, SC, T>(a.R, s, a.u, a.v, a.w, a.c, gold[1], a.scal, fnew[1], a.rv, nullptr, a.dt, b);
        else if (F == 2)
            rhsw_body<ORD, SC, T>(a.R, s, a.u, a.v, a.w, a.c, gold[2], a.scal, fnew[2], a.rw, nullptr, a.dt, b);
        else
            rhsc_body<ORD, SC, T>(a.R, s, a.u, a.v, a.w, a.c, gold[3], a.scal, fnew[3], a.rc, nullptr, a.dt, b);
    });
}

template <int ORD, typename T>
__global__ __launch_bounds__(256) void kb_proj(const GArgs<T> *__restrict__ A, const unsigned *__restrict__ vs,
                                               int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = A[r];
        const Blk b = blk_of(a, ST_CELL, lb);
        if (s < 3) {
            const int nz = a.R.g.nz;
            const double *h1u = a.R.h1 + (size_t) (s * 2 + 0) * (nz + 2);
            const double *h1c = a.R.h1 + (size_t) (s * 2 + 1) * (nz + 2);
            const bool sc = a.R.scalar != 0;
            proj_body<ORD, T>(a.R, a.ru, a.rv, a.rw, a.rc, h1u, a.scal + SC_M0 + s, sc ? h1c : nullptr,
                              a.scal + SC_MC0 + s, a.p, a.R.sc[s].abdt, a.u, a.v, a.w, sc ? a.c : nullptr, b);
        } else {
            proj_body<ORD, T>(a.R, a.u, a.v, a.w, nullptr, nullptr, nullptr, nullptr, nullptr, a.p, a.dt_adhoc,
                              a.u, a.v, a.w, nullptr, b);
        }
    });
}

template <typename T>
__global__ __launch_bounds__(256) void kb_solve4(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        solve4_body<T>(a.R, s, a.ru, a.rv, a.rw, a.rc, blk_of(a, ST_SOLVE, lb));
    });
}

template <typename T>
__global__ __launch_bounds__(256) void kb_plane(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    (void) s;
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        plane_body<T>(a.R, a.ru, a.rc, a.part, a.nb_plane, blk_of(a, ST_PLANE, lb));
    });
}

template <typename T>
__global__ __launch_bounds__(256) void kb_mass(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        mass_body(a.R, s, a.part, a.nb_plane, a.scal, a.logbuf, blk_of(a, ST_MASS, lb));
    });
}

template <typename T>
__global__ __launch_bounds__(256) void kb_poisson4(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        pois4_body<T>(a.R, a.pt4[s], a.X, a.S4, blk_of(a, ST_MODE, lb));
    });
}

#define CF_INST_B(T)                                                                                              \
    template __global__ void kb_solve4<T>(const GArgs<T> *, const unsigned *, int, int);                          \
    template __global__ void kb_plane<T>(const GArgs<T> *, const unsigned *, int, int);                           \
    template __global__ void kb_mass<T>(const GArgs<T> *, const unsigned *, int, int);                            \
    template __global__ void kb_poisson4<T>(const GArgs<T> *, const unsigned *, int, int);
CF_INST_B(double)
CF_INST_B(float)
#undef CF_INST_B

template <typename T> struct BPick {
    template <int ORD, bool SC> static void fill(BSel<T> &S)
    {
        S.u = kb_rhs<ORD, SC, T, 0>;
        S.v = kb_rhs<ORD, SC, T, 1>;
        S.w = kb_rhs<ORD, SC, T, 2>;
        S.c = SC ? kb_rhs<ORD, true, T, 3> : nullptr;
        S.proj = kb_proj<ORD, T>;
        S.divx = pick_batched_divx<T>(ORD);
    }
};

template <typename T>
BSel<T> pick_batched(int order, int scalar)
{
    BSel<T> S;
    if (order == 4) {
        if (scalar) BPick<T>::template fill<4, true>(S);
        else BPick<T>::template fill<4, false>(S);
    } else {
        if (scalar) BPick<T>::template fill<2, true>(S);
        else BPick<T>::template fill<2, false>(S);
    }
    return S;
}

template BSel<double> pick_batched<double>(int, int);
template BSel<float> pick_batched<float>(int, int);
