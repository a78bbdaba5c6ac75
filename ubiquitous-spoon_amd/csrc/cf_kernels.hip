/*
 * cf_kernels.hip -- HIP kernels of one RK3 substep of the channel DNS on
 * gfx950, FP64.  Reference semantics per kernel are cited beside it; the
 * device layout is described in cf_dev.h.  Every kernel bounds-checks its
 * indices against the Grid it is given; the host checks launch shapes.
 */
#include "cf_dev.h"
#include "cf_kernels.h"

#include <math.h>

/* ---------------------------------------------------------------------
 * io layout [nx][ny][nz] <-> device layout [nz][ny][nx]: for each y the
 * (x, z) plane is transposed through a 32x33 LDS tile (bit-exact moves).
 * Replaces the pencil redistributions of transpose.c:409-448 on one GPU.
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_io_to_dev(const double *__restrict__ io, double *__restrict__ dev,
                                                   int nx, int ny, int nz)
{
    __shared__ double tile[32][33];
    const int k0 = blockIdx.x * 32, i0 = blockIdx.y * 32, j = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;       /* 32 x 8 */
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + r, k = k0 + tx;
        if (i < nx && k < nz) tile[r][tx] = io[((size_t) i * ny + j) * nz + k];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int k = k0 + r, i = i0 + tx;
        if (i < nx && k < nz) dev[((size_t) k * ny + j) * nx + i] = tile[tx][r];
    }
}

__global__ __launch_bounds__(256) void k_dev_to_io(const double *__restrict__ dev, double *__restrict__ io,
                                                   int nx, int ny, int nz)
{
    __shared__ double tile[32][33];
    const int i0 = blockIdx.x * 32, k0 = blockIdx.y * 32, j = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        int k = k0 + r, i = i0 + tx;
        if (i < nx && k < nz) tile[r][tx] = dev[((size_t) k * ny + j) * nx + i];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + r, k = k0 + tx;
        if (i < nx && k < nz) io[((size_t) i * ny + j) * nz + k] = tile[tx][r];
    }
}

/* ---------------------------------------------------------------------
 * Explicit right-hand side and implicit-solve right-hand side of one
 * substep, 2nd order, one thread per cell (i, j, k).  Fuses, per field:
 *   wall ghosts        apply_bc_u2 / apply_bc_w2   (diffuseu.c:37-69, diffusew.c:37-69)
 *   mass fluxes        massflux{x,y,z}2            (advect.c:6-83)
 *   advection          advect{u,v,w}2              (advect.c:286-463)
 *   x-y diffusion      add_laplacexy_operator_*2   (diffuseu.c:286-313, diffusew.c:290-315)
 *   z-Helmholtz        helmholtz1d_operator_*2     (diffuseu.c:107-132, diffusew.c:107-133)
 *   RK3 combination    h += gam dt f + zet dt g, h -= lift (chanfast.c:465-477, :619-631)
 * Writes f (the new explicit tendency) and r = h * dvol, the right-hand
 * side of the tridiagonal solve (diffuseu.c:202, diffusew.c:207).
 * g_eff = g + m_prev folds the previous substep's mass-flux forcing into
 * the zeta term, as `fu += mdpdx` does (chanfast.c:519-522).
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_rhs2(const RunDev R, const int s,
    const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
    const double *__restrict__ gu, const double *__restrict__ gv, const double *__restrict__ gw,
    const double *__restrict__ scal,
    double *__restrict__ fu, double *__restrict__ fv, double *__restrict__ fw,
    double *__restrict__ ru, double *__restrict__ rv, double *__restrict__ rw)
{
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const size_t P = (size_t) nx * ny;
    const int ip = i + 1 < nx ? i + 1 : 0, im = i > 0 ? i - 1 : nx - 1;
    const int jp = j + 1 < ny ? j + 1 : 0, jm = j > 0 ? j - 1 : ny - 1;
    const double dx = R.g.dx, dy = R.g.dy;
    const double *dz = R.dzt + CF_NGT;
    const double ub = R.ubc, vb = R.vbc;

#define AT(a, ii, jj, kk) (a)[(size_t) (kk) * P + (size_t) (jj) * nx + (ii)]
    /* wall ghosts: 2nd-order Dirichlet closure gives q_g = 2 bc - q_1 */
    auto U = [&](int ii, int jj, int kk) -> double {
        if (kk < 0) return 2.0 * ub - AT(u, ii, jj, 0);
        if (kk >= nz) return 2.0 * ub - AT(u, ii, jj, nz - 1);
        return AT(u, ii, jj, kk);
    };
    auto V = [&](int ii, int jj, int kk) -> double {
        if (kk < 0) return 2.0 * vb - AT(v, ii, jj, 0);
        if (kk >= nz) return 2.0 * vb - AT(v, ii, jj, nz - 1);
        return AT(v, ii, jj, kk);
    };
    /* w: both wall faces carry the Dirichlet value 0 */
    auto W = [&](int ii, int jj, int kk) -> double {
        return (kk < 0 || kk >= nz - 1) ? 0.0 : AT(w, ii, jj, kk);
    };

    const double dzk = dz[k], dzkp = dz[k + 1];
    const double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
    const double dzm1h = 0.5 * (dz[k - 1] + dz[k]);
    const SubCoef C = R.sc[s];
    const double nu = R.nu;
    const size_t q = (size_t) k * P + (size_t) j * nx + i;

    /* ---- u ---- */
    const double u0 = U(i, j, k), uip = U(ip, j, k), uim = U(im, j, k);
    const double ujp = U(i, jp, k), ujm = U(i, jm, k), ukp = U(i, j, k + 1), ukm = U(i, j, k - 1);
    const double v0 = V(i, j, k), vip = V(ip, j, k), vim = V(im, j, k);
    const double vjp = V(i, jp, k), vjm = V(i, jm, k), vkp = V(i, j, k + 1), vkm = V(i, j, k - 1);
    const double vipjm = V(ip, jm, k);
    const double w0 = W(i, j, k), wip = W(ip, j, k), wkm = W(i, j, k - 1), wipkm = W(ip, j, k - 1);
    const double wjp = W(i, jp, k), wjpkm = W(i, jp, k - 1);
    {
        double mxp = 0.5 * (dy * dzk * u0 + dy * dzk * uip);
        double mxm = 0.5 * (dy * dzk * uim + dy * dzk * u0);
        double myp = 0.5 * (dx * dzk * v0 + dx * dzk * vip);
        double mym = 0.5 * (dx * dzk * vjm + dx * dzk * vipjm);
        double mzp = 0.5 * (dx * dy * w0 + dx * dy * wip);
        double mzm = 0.5 * (dx * dy * wkm + dx * dy * wipkm);
        double adv = -((mxp * 0.5 * (u0 + uip) - mxm * 0.5 * (u0 + uim))
                     + (myp * 0.5 * (u0 + ujp) - mym * 0.5 * (u0 + ujm))
                     + (mzp * 0.5 * (u0 + ukp) - mzm * 0.5 * (u0 + ukm))) / (dx * dy * dzk);
        double f = adv + nu * (((uip - u0) / dx - (u0 - uim) / dx) / dx
                             + ((ujp - u0) / dy - (u0 - ujm) / dy) / dy);
        double h = u0 + C.nuadt * (((ukp - u0) / dzp1h - (u0 - ukm) / dzm1h) / dzk);
        double g = gu[q] + scal[R.m_prev_slot[s]];
        h += C.gamdt * f + C.zetdt * g;
        if (k == 0) h -= R.lift[s][0];
        if (k == nz - 1) h -= R.lift[s][1];
        fu[q] = f;
        ru[q] = h * dzk;
    }
    /* ---- v ---- */
    {
        const double uimv = U(im, j, k), uimjpv = U(im, jp, k);
        double mxp = 0.5 * (dy * dzk * u0 + dy * dzk * ujp);
        double mxm = 0.5 * (dy * dzk * uimv + dy * dzk * uimjpv);
        double myp = 0.5 * (dx * dzk * v0 + dx * dzk * vjp);
        double mym = 0.5 * (dx * dzk * vjm + dx * dzk * v0);
        double mzp = 0.5 * (dx * dy * w0 + dx * dy * wjp);
        double mzm = 0.5 * (dx * dy * wkm + dx * dy * wjpkm);
        double adv = -((mxp * 0.5 * (v0 + vip) - mxm * 0.5 * (v0 + vim))
                     + (myp * 0.5 * (v0 + vjp) - mym * 0.5 * (v0 + vjm))
                     + (mzp * 0.5 * (v0 + vkp) - mzm * 0.5 * (v0 + vkm))) / (dx * dy * dzk);
        double f = adv + nu * (((vip - v0) / dx - (v0 - vim) / dx) / dx
                             + ((vjp - v0) / dy - (v0 - vjm) / dy) / dy);
        double h = v0 + C.nuadt * (((vkp - v0) / dzp1h - (v0 - vkm) / dzm1h) / dzk);
        h += C.gamdt * f + C.zetdt * gv[q];
        if (k == 0) h -= R.lift[s][2];
        if (k == nz - 1) h -= R.lift[s][3];
        fv[q] = f;
        rv[q] = h * dzk;
    }
    /* ---- w (faces k < nz - 1; the top wall face carries no equation) ---- */
    if (k < nz - 1) {
        const double ukp1 = U(i, j, k + 1), uimkp = U(im, j, k + 1);
        const double vjmkp = V(i, jm, k + 1);
        const double wim = W(im, j, k), wjm = W(i, jm, k), wkp = W(i, j, k + 1);
        double mxp = 0.5 * (dy * dzk * u0 + dy * dzkp * ukp1);
        double mxm = 0.5 * (dy * dzk * uim + dy * dzkp * uimkp);
        double myp = 0.5 * (dx * dzk * v0 + dx * dzkp * vkp);
        double mym = 0.5 * (dx * dzk * vjm + dx * dzkp * vjmkp);
        double mzp = 0.5 * (dx * dy * w0 + dx * dy * wkp);
        double mzm = 0.5 * (dx * dy * wkm + dx * dy * w0);
        double adv = -((mxp * 0.5 * (w0 + wip) - mxm * 0.5 * (w0 + wim))
                     + (myp * 0.5 * (w0 + wjp) - mym * 0.5 * (w0 + wjm))
                     + (mzp * 0.5 * (w0 + wkp) - mzm * 0.5 * (w0 + wkm))) / (dx * dy * dzp1h);
        double f = adv + nu * (((wip - w0) / dx - (w0 - wim) / dx) / dx
                             + ((wjp - w0) / dy - (w0 - wjm) / dy) / dy);
        double h = w0 + C.nuadt * (((wkp - w0) / dzkp - (w0 - wkm) / dzk) / dzp1h);
        h += C.gamdt * f + C.zetdt * gw[q];
        fw[q] = f;
        rw[q] = h * dzp1h;
    } else {
        fw[q] = 0.0;
        rw[q] = 0.0;
    }
#undef AT
}

/* ---------------------------------------------------------------------
 * Batched constant-coefficient Thomas solve, one thread per (column,
 * field), in place: the right-hand side r becomes the solution.  The
 * factors come from the wall-reduced system of solve_helmholtz1d_u2 /
 * _w2 (diffuseu.c:173-219, diffusew.c:175-223 with tridag.c:4-26), built
 * on the host once per substep since they do not depend on the column.
 * Fields: 0 u (nz rows), 1 v (nz rows), 2 w (nz-1 rows; top wall := 0).
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_thomas(const RunDev R, const int s,
    double *__restrict__ ru, double *__restrict__ rv, double *__restrict__ rw)
{
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t P = (size_t) nx * ny;
    const size_t col = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (col >= P || f > 2) return;
    double *x = f == 0 ? ru : (f == 1 ? rv : rw);
    const int type = f == 2 ? 2 : 0;
    const int n = f == 2 ? nz - 1 : nz;
    const double *a = R.th + ((size_t) (s * 4 + type) * 3 + 0) * nz;
    const double *ib = R.th + ((size_t) (s * 4 + type) * 3 + 1) * nz;
    const double *gm = R.th + ((size_t) (s * 4 + type) * 3 + 2) * nz;
    double *xc = x + col;
    double d = xc[0] * ib[0];
    xc[0] = d;
    int k = 1;
    for (; k + 8 <= n; k += 8) {
        double rr[8];
#pragma unroll
        for (int t = 0; t < 8; t++) rr[t] = xc[(size_t) (k + t) * P];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            d = (rr[t] - a[k + t] * d) * ib[k + t];
            xc[(size_t) (k + t) * P] = d;
        }
    }
    for (; k < n; k++) {
        d = (xc[(size_t) k * P] - a[k] * d) * ib[k];
        xc[(size_t) k * P] = d;
    }
    k = n - 2;
    for (; k - 7 >= 0; k -= 8) {
        double rr[8];
#pragma unroll
        for (int t = 0; t < 8; t++) rr[t] = xc[(size_t) (k - t) * P];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            d = rr[t] - gm[k - t + 1] * d;
            xc[(size_t) (k - t) * P] = d;
        }
    }
    for (; k >= 0; k--) {
        d = xc[(size_t) k * P] - gm[k + 1] * d;
        xc[(size_t) k * P] = d;
    }
    if (f == 2) xc[(size_t) (nz - 1) * P] = 0.0;   /* ghost_fill_top2, diffusew.c:683-688 */
}

/* ---------------------------------------------------------------------
 * Plane sums for get_xy_avg (get.c:281-307) of the solved u: levels
 * -1 and nz are the wall ghosts 2 bc - u (apply_bc_u2 after the solve,
 * chanfast.c:482-483).  Deterministic: block b sums a fixed chunk of the
 * plane in a fixed tree order; the final sum over blocks is sequential.
 * part[(k + 1) * nb + b].
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_plane_sum(const RunDev R, const double *__restrict__ q,
                                                   double bc, double *__restrict__ part, int nb)
{
    __shared__ double red[256];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t P = (size_t) nx * ny;
    const int lev = blockIdx.y - 1;              /* -1 .. nz */
    const int b = blockIdx.x;
    if (lev > nz || b >= nb) return;
    const size_t chunk = (P + nb - 1) / nb;
    const size_t lo = (size_t) b * chunk, hi = lo + chunk < P ? lo + chunk : P;
    const int src = lev < 0 ? 0 : (lev >= nz ? nz - 1 : lev);
    const bool ghost = lev < 0 || lev >= nz;
    const double *pl = q + (size_t) src * P;
    double acc = 0.0;
    for (size_t e = lo + threadIdx.x; e < hi; e += 256) {
        double val = pl[e];
        acc += ghost ? (2.0 * bc - val) : val;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t) (lev + 1) * nb + b] = red[0];
}

/* ---------------------------------------------------------------------
 * Mass-flux constraint (chanfast.c:485-530): plane averages -> bulk
 * velocity (get_z_avg, get.c:490-502), wall shear, centreline; forcing
 * mdpdx = (Ub - uframe - U0) / U1 with U1 from the column-independent
 * unit-forcing solution h1 (host-precomputed per substep).  Writes the
 * forcing into scal[slot]; on the last substep writes the pstat record.
 * One block.
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_mass_finalize(const RunDev R, const int s, const double *__restrict__ part,
                                                       int nb, double *__restrict__ scal, double *__restrict__ logbuf,
                                                       int kind)
{
    extern __shared__ double avg[];               /* nz + 2 */
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const double *dz = R.dzt + CF_NGT;
    const double scale = 1.0 / ((double) nx * (double) ny);
    for (int l = threadIdx.x; l < nz + 2; l += blockDim.x) {
        double acc = 0.0;
        for (int b = 0; b < nb; b++) acc += part[(size_t) l * nb + b];
        avg[l] = acc * scale;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double *A = avg + 1;                    /* A[-1 .. nz] */
    double vol = 0.0, lz = 0.0;
    for (int k = 0; k < nz; k++) { vol += A[k] * dz[k]; lz += dz[k]; }
    vol = vol / lz;
    const double kap = kind == 0 ? R.nu : R.nu / R.prandtl;
    double bot = kap * (A[0] - A[-1]) / (0.5 * (dz[0] + dz[-1]));
    double top = kap * (A[nz] - A[nz - 1]) / (0.5 * (dz[nz] + dz[nz - 1]));
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

/* ---------------------------------------------------------------------
 * CFL at the start of a step (get_cfl, get.c:390-417): per-block max,
 * then a one-block finish that writes the cfl.dat record and bumps the
 * step counter at the end of the step (k_step_end).
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_cfl_part(const RunDev R, const double *__restrict__ u,
                                                  const double *__restrict__ v, const double *__restrict__ w,
                                                  double dt, double *__restrict__ part, int nb)
{
    __shared__ double red[256];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t N = (size_t) nx * ny * nz, P = (size_t) nx * ny;
    const double *dz = R.dzt + CF_NGT;
    double m = 0.0;
    for (size_t e = (size_t) blockIdx.x * 256 + threadIdx.x; e < N; e += (size_t) nb * 256) {
        int k = (int) (e / P);
        double dzph = 0.5 * (dz[k] + dz[k + 1]);
        double c;
        if (m < (c = fabs(u[e]) * dt / R.g.dx)) m = c;
        if (m < (c = fabs(v[e]) * dt / R.g.dy)) m = c;
        if (m < (c = fabs(w[e]) * dt / dzph)) m = c;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_cfl_finish(const RunDev R, const double *__restrict__ part, int nb, double dt,
                             double *__restrict__ scal, double *__restrict__ logbuf)
{
    if (threadIdx.x != 0) return;
    double m = 0.0;
    for (int b = 0; b < nb; b++) m = fmax(m, part[b]);
    long idx = (long) scal[SC_STEP] % R.logcap;
    double *lg = logbuf + (size_t) idx * 13;
    for (int t = 0; t < 13; t++) lg[t] = 0.0;
    lg[0] = scal[SC_STEP] + 1.0;
    lg[1] = dt;
    lg[2] = m;
    scal[SC_CFL] = m;
}

__global__ void k_step_end(double *__restrict__ scal)
{
    if (threadIdx.x == 0) scal[SC_STEP] += 1.0;
}

/* ---------------------------------------------------------------------
 * divergence2 (pressure.c:40-61) of the mass-flux-corrected velocity
 * (u + h1 mdpdx, chanfast.c:519-522) fused with the forward real-to-
 * complex x FFT scaled 1/nx (fft.c:161-177).  Two real lines (j, k) are
 * packed into one complex FFT and separated afterwards.
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_div_fftx(const RunDev R, const double *__restrict__ us,
    const double *__restrict__ vs, const double *__restrict__ ws, const double *__restrict__ h1,
    const double *__restrict__ mptr, double2 *__restrict__ X)
{
    extern __shared__ double2 sb[];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const int ld = nx + 1;
    const int nl = R.fx_pairs;
    const long NL = (long) ny * nz;
    const size_t P = (size_t) nx * ny;
    const double *dz = R.dzt + CF_NGT;
    const double m = mptr ? *mptr : 0.0;
    const double dx = R.g.dx, dy = R.g.dy;
    const long pair0 = (long) blockIdx.x * nl;
    auto divg = [&](long line, int i) -> double {
        if (line >= NL) return 0.0;
        int k = (int) (line / ny), j = (int) (line - (long) k * ny);
        int im = i > 0 ? i - 1 : nx - 1, jm = j > 0 ? j - 1 : ny - 1;
        size_t q = (size_t) k * P + (size_t) j * nx;
        double hm = h1 ? h1[k + 1] * m : 0.0;
        double u1 = us[q + i] + hm, u0 = us[q + im] + hm;
        double v1 = vs[q + i], v0 = vs[(size_t) k * P + (size_t) jm * nx + i];
        double w1 = k < nz - 1 ? ws[q + i] : 0.0;
        double w0 = k > 0 ? ws[q - P + i] : 0.0;
        return (u1 - u0) / dx + (v1 - v0) / dy + (w1 - w0) / dz[k];
    };
    for (int e = threadIdx.x; e < nl * nx; e += blockDim.x) {
        int l = e / nx, i = e - l * nx;
        long pr = pair0 + l;
        sb[l * ld + i] = make_double2(divg(2 * pr, i), divg(2 * pr + 1, i));
    }
    __syncthreads();
    lds_fft<8>(sb, nx, ld, nl, R.radx, R.nradx, R.twx, 1.0);
    const double sc = 0.5 / (double) nx;
    for (int e = threadIdx.x; e < nl * nkx; e += blockDim.x) {
        int l = e / nkx, kx = e - l * nkx;
        long la = 2 * (pair0 + l), lb = la + 1;
        if (la >= NL) continue;
        double2 Z = sb[l * ld + kx], Wc = sb[l * ld + (kx == 0 ? 0 : nx - kx)];
        X[(size_t) la * nkx + kx] = make_double2((Z.x + Wc.x) * sc, (Z.y - Wc.y) * sc);
        if (lb < NL) X[(size_t) lb * nkx + kx] = make_double2((Z.y + Wc.y) * sc, (Wc.x - Z.x) * sc);
    }
}

/* plain forward r2c x FFT of a real device-layout field (no divergence),
 * scaled 1/nx: the fftx_d2c stage of FFT3D_FWD on its own */
__global__ __launch_bounds__(256) void k_fftx_fwd_plain(const RunDev R, const double *__restrict__ a,
                                                        double2 *__restrict__ X)
{
    extern __shared__ double2 sb[];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const int ld = nx + 1;
    const int nl = R.fx_pairs;
    const long NL = (long) ny * nz;
    const long pair0 = (long) blockIdx.x * nl;
    for (int e = threadIdx.x; e < nl * nx; e += blockDim.x) {
        int l = e / nx, i = e - l * nx;
        long la = 2 * (pair0 + l), lb = la + 1;
        double va = la < NL ? a[(size_t) la * nx + i] : 0.0;
        double vb = lb < NL ? a[(size_t) lb * nx + i] : 0.0;
        sb[l * ld + i] = make_double2(va, vb);
    }
    __syncthreads();
    lds_fft<8>(sb, nx, ld, nl, R.radx, R.nradx, R.twx, 1.0);
    const double sc = 0.5 / (double) nx;
    for (int e = threadIdx.x; e < nl * nkx; e += blockDim.x) {
        int l = e / nkx, kx = e - l * nkx;
        long la = 2 * (pair0 + l), lb = la + 1;
        if (la >= NL) continue;
        double2 Z = sb[l * ld + kx], Wc = sb[l * ld + (kx == 0 ? 0 : nx - kx)];
        X[(size_t) la * nkx + kx] = make_double2((Z.x + Wc.x) * sc, (Z.y - Wc.y) * sc);
        if (lb < NL) X[(size_t) lb * nkx + kx] = make_double2((Z.y + Wc.y) * sc, (Wc.x - Z.x) * sc);
    }
}

/* ---------------------------------------------------------------------
 * Complex y FFT of X[nz][ny][nkx] in place, forward scaled 1/ny
 * (fft.c:186-202) or unscaled inverse (fft.c:205-208).  Block = TK
 * consecutive kx columns at one level.
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_ffty(const RunDev R, double2 *__restrict__ X, int fwd)
{
    extern __shared__ double2 sb[];
    const int ny = R.g.ny, nkx = R.g.nkx;
    const int TK = R.fy_cols, ld = ny + 1;
    const int k = blockIdx.y, kx0 = blockIdx.x * TK;
    double2 *base = X + (size_t) k * ny * nkx;
    for (int e = threadIdx.x; e < TK * ny; e += blockDim.x) {
        int j = e / TK, t = e - j * TK;
        int kx = kx0 + t;
        sb[t * ld + j] = kx < nkx ? base[(size_t) j * nkx + kx] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    lds_fft<8>(sb, ny, ld, TK, R.rady, R.nrady, R.twy, fwd ? 1.0 : -1.0);
    const double sc = fwd ? 1.0 / (double) ny : 1.0;
    for (int e = threadIdx.x; e < TK * ny; e += blockDim.x) {
        int j = e / TK, t = e - j * TK;
        int kx = kx0 + t;
        if (kx < nkx) {
            double2 val = sb[t * ld + j];
            base[(size_t) j * nkx + kx] = fwd ? cscale(val, sc) : val;
        }
    }
}

/* ---------------------------------------------------------------------
 * pressure_poisson2 (pressure.c:293-355): one thread per (kx, ky) mode,
 * z-tridiagonal with modified wavenumbers lam = 2 - 2 cos, dvol
 * symmetrisation, Neumann walls except the (0,0) mode whose bottom row
 * is the Dirichlet-mean closure; real and imaginary parts share the
 * matrix.  Thomas with the gam factors kept in G for the back sweep.
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_poisson2(const RunDev R, const PoisTab T,
                                                  double2 *__restrict__ Y, double *__restrict__ G)
{
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const size_t M = (size_t) ny * nkx;
    const size_t qm = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (qm >= M) return;
    const int kx = (int) (qm % nkx), ky = (int) (qm / nkx);
    const double lx = R.lamx[kx], ly = R.lamy[ky];
    const bool zero = kx == 0 && ky == 0;
    double2 *y = Y + qm;
    double *g = G + qm;
    double b = T.bz[0] - lx * T.ex[0] - ly * T.ey[0];
    b -= zero ? T.db_dir : T.db_neu;
    if (nz == 1) b -= T.db_top;
    double bet = b;
    double2 r = cscale(y[0], T.dvol[0]);
    double2 d = make_double2(r.x / bet, r.y / bet);
    y[0] = d;
    for (int k = 1; k < nz; k++) {
        double bk = T.bz[k] - lx * T.ex[k] - ly * T.ey[k];
        if (k == nz - 1) bk -= T.db_top;
        double gam = T.c[k - 1] / bet;
        bet = bk - T.a[k] * gam;
        g[(size_t) k * M] = gam;
        double2 rk = cscale(y[(size_t) k * M], T.dvol[k]);
        d = make_double2((rk.x - T.a[k] * d.x) / bet, (rk.y - T.a[k] * d.y) / bet);
        y[(size_t) k * M] = d;
    }
    for (int k = nz - 2; k >= 0; k--) {
        double gk = g[(size_t) (k + 1) * M];
        double2 yk = y[(size_t) k * M];
        d = make_double2(yk.x - gk * d.x, yk.y - gk * d.y);
        y[(size_t) k * M] = d;
    }
}

/* ---------------------------------------------------------------------
 * Inverse complex-to-real x FFT (fft.c:180-183, FFTW c2r semantics: the
 * imaginary parts of the kx = 0 and kx = nx/2 terms are ignored), two
 * real lines per complex transform, writing p.
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_fftx_inv(const RunDev R, const double2 *__restrict__ X,
                                                  double *__restrict__ p)
{
    extern __shared__ double2 sb[];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const int ld = nx + 1;
    const int nl = R.fx_pairs;
    const long NL = (long) ny * nz;
    const long pair0 = (long) blockIdx.x * nl;
    const int half = nx / 2;
    auto spec = [&](long line, int kx) -> double2 {
        if (line >= NL) return make_double2(0.0, 0.0);
        const double2 *row = X + (size_t) line * nkx;
        if (kx <= half) {
            double2 a = row[kx];
            if (kx == 0 || (nx % 2 == 0 && kx == half)) a.y = 0.0;
            return a;
        }
        double2 a = row[nx - kx];
        return make_double2(a.x, -a.y);
    };
    for (int e = threadIdx.x; e < nl * nx; e += blockDim.x) {
        int l = e / nx, kx = e - l * nx;
        long la = 2 * (pair0 + l);
        double2 A = spec(la, kx), B = spec(la + 1, kx);
        sb[l * ld + kx] = make_double2(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    lds_fft<8>(sb, nx, ld, nl, R.radx, R.nradx, R.twx, -1.0);
    for (int e = threadIdx.x; e < nl * nx; e += blockDim.x) {
        int l = e / nx, i = e - l * nx;
        long la = 2 * (pair0 + l), lb = la + 1;
        double2 z = sb[l * ld + i];
        if (la < NL) p[(size_t) la * nx + i] = z.x;
        if (lb < NL) p[(size_t) lb * nx + i] = z.y;
    }
}

/* ---------------------------------------------------------------------
 * project2 (pressure.c:120-152) fused with the mass-flux correction:
 * u = (u* + h1 mdpdx) - (p_{i+1} - p_i) dt / dx, v, w likewise.  The
 * top-wall face is left at its Dirichlet value 0 (the reference subtracts
 * the Neumann-ghost difference there, which is round-off).
 * ------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_project2(const RunDev R, const double *__restrict__ us,
    const double *__restrict__ vs, const double *__restrict__ ws, const double *__restrict__ h1,
    const double *__restrict__ mptr, const double *__restrict__ p, double dtp,
    double *__restrict__ u, double *__restrict__ v, double *__restrict__ w)
{
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const size_t P = (size_t) nx * ny;
    const int ip = i + 1 < nx ? i + 1 : 0, jp = j + 1 < ny ? j + 1 : 0;
    const size_t q = (size_t) k * P + (size_t) j * nx + i;
    const double *dz = R.dzt + CF_NGT;
    const double m = mptr ? *mptr : 0.0;
    const double p0 = p[q];
    double uc = us[q] + (h1 ? h1[k + 1] * m : 0.0);
    u[q] = uc - (p[(size_t) k * P + (size_t) j * nx + ip] - p0) * dtp / R.g.dx;
    v[q] = vs[q] - (p[(size_t) k * P + (size_t) jp * nx + i] - p0) * dtp / R.g.dy;
    if (k < nz - 1) {
        double dzp1h = 0.5 * (dz[k + 1] + dz[k]);
        w[q] = ws[q] - (p[q + P] - p0) * dtp / dzp1h;
    } else {
        w[q] = 0.0;
    }
}

/* max |div| over interior cells (case1 check, cases.c:473-486) */
__global__ __launch_bounds__(256) void k_maxdiv_part(const RunDev R, const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, double *__restrict__ part, int nb)
{
    __shared__ double red[256];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t N = (size_t) nx * ny * nz, P = (size_t) nx * ny;
    const double *dz = R.dzt + CF_NGT;
    double mx = 0.0;
    for (size_t e = (size_t) blockIdx.x * 256 + threadIdx.x; e < N; e += (size_t) nb * 256) {
        int k = (int) (e / P);
        int rem = (int) (e - (size_t) k * P);
        int j = rem / nx, i = rem - j * nx;
        int im = i > 0 ? i - 1 : nx - 1, jm = j > 0 ? j - 1 : ny - 1;
        size_t row = (size_t) k * P + (size_t) j * nx;
        double w1 = k < nz - 1 ? w[e] : 0.0;
        double w0 = k > 0 ? w[e - P] : 0.0;
        double d = (u[e] - u[row + im]) / R.g.dx + (v[e] - v[(size_t) k * P + (size_t) jm * nx + i]) / R.g.dy
                 + (w1 - w0) / dz[k];
        mx = fmax(mx, fabs(d));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
