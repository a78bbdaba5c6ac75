/*
 * cf_rhs4.hip -- the explicit stage of an order-4 substep in one launch:
 * the tendency f and the implicit right-hand side h of u, v, w and (with
 * the scalar) c.  chanfast4 (chanfast.c:1319-1470): advectu4 / v4 / w4
 * with massflux*4 (advect.c:6-734), laplacexy (diffuseu.c:286-359),
 * add_buoyancy4 (scalar.c:6-76), flux_quick + advect_scalar
 * (scalar.c:79-162), helmholtz1d_operator_u4 / w4 (diffuseu.c:107-170,
 * diffusew.c:107-172) and the solve's right-hand side; with `implicit`
 * (channel4) the x-y diffusion goes into h instead of f.
 *
 * A workgroup owns a 32 x 8 tile of columns and marches up a chunk of
 * levels:
 *   - each thread keeps its own column of u, v, c (levels k-3..k+3) and w
 *     (faces k-3..k+3) in registers, wall ghosts formed from them;
 *   - per level the tile's planes u(k), v(k), c(k) and w at faces k and
 *     k+1, each with the halo the stencils reach (a 3-wide cross plus the
 *     2 x 2 corner blocks the cross-interpolated fluxes touch), are staged
 *     in LDS;
 *   - the z-interpolated face fluxes of u and v (massfluxz) are formed in
 *     the column that owns them and shared through LDS; halo columns keep
 *     a 4-level window of their own;
 *   - the x- and y-interpolated fluxes of w (massfluxx / massfluxy) are
 *     formed once per face and kept as 4-face register windows.
 * Every stencil neighbour is then an LDS read at a constant offset, and
 * every division is a multiplication by a host-computed reciprocal.
 */
#include "cf_kernels.h"

namespace {

constexpr double A1d = 1.125 / 1.0;      /* fourth.c:3-6 */
constexpr double A2d = -0.125 / 3.0;
constexpr double B1d = 1.125 / 2.0;
constexpr double B2d = -0.125 / 2.0;

constexpr int TX = CF_R4_X, TY = CF_R4_Y, H = 3;
constexpr int PX = TX + 2 * H, PY = TY + 2 * H;      /* staged plane with halo */
constexpr int UZX = TX + 3;                          /* uz: x in [-2, TX] */
constexpr int VZY = TY + 3;                          /* vz: y in [-2, TY] */

/* halo element t (0..255) of a staged plane -> (x, y) relative to the tile */
__device__ __forceinline__ void halo_xy(int t, int &x, int &y)
{
    if (t < 3 * TX) { y = t / TX - 3; x = t % TX; return; }                 /* rows -3..-1 */
    if (t < 6 * TX) { const int e = t - 3 * TX; y = TY + e / TX; x = e % TX; return; }   /* rows TY..TY+2 */
    if (t < 6 * TX + 6 * TY) {                                              /* columns -3..-1, TX..TX+2 */
        const int e = t - 6 * TX, c = e / TY;
        x = c < 3 ? c - 3 : TX + c - 3;
        y = e % TY;
        return;
    }
    const int e = t - 6 * TX - 6 * TY;                                      /* corners */
    const int cx = e & 3, cy = e >> 2;
    x = cx < 2 ? cx - 2 : TX + cx - 2;
    y = cy < 2 ? cy - 2 : TY + cy - 2;
}
static_assert(6 * TX + 6 * TY + 16 == 256, "one halo element per thread");

/* ghosts of a cell column window q[t] = level k-3+t (raw, levels inside
 * [0, nz)) from the closure: ghost m of a wall = G[m] . (x0, x1, x2) + B[m] bc.
 * k is uniform over the workgroup; the cases keep every register index a
 * compile-time constant. */
template <int KB, bool BC, typename T>
__device__ __forceinline__ void ghost_bot(const T q[7], T o[7], const Clos1 &Bc, double bc)
{
    const T x0 = q[3 - KB], x1 = q[4 - KB], x2 = q[5 - KB];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const int t = 2 - m - KB;
        if (t >= 0) o[t] = (T) Bc.G[m][0] * x0 + (T) Bc.G[m][1] * x1 + (T) Bc.G[m][2] * x2 + (BC ? (T) (Bc.B[m] * bc) : (T) 0);
    }
}
/* KT = index of the last unknown (level nz-1, or face nz-2) minus 3 */
template <int KT, bool BC, typename T>
__device__ __forceinline__ void ghost_top(const T q[7], T o[7], const Clos1 &Tc, double bc)
{
    constexpr int b = KT + 3;
    const T x0 = q[b], x1 = q[b - 1], x2 = q[b - 2];
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const int t = b + 1 + m;
        if (t <= 6) o[t] = (T) Tc.G[m][0] * x0 + (T) Tc.G[m][1] * x1 + (T) Tc.G[m][2] * x2 + (BC ? (T) (Tc.B[m] * bc) : (T) 0);
    }
}

template <typename T>
__device__ __forceinline__ void cell_ghosts(const T q[7], T o[7], int k, int nz, const Clos1 &Bc, const Clos1 &Tc,
                                            double bcb, double bct)
{
#pragma unroll
    for (int t = 0; t < 7; t++) o[t] = q[t];
    if (k == 0) ghost_bot<0, true>(q, o, Bc, bcb);
    else if (k == 1) ghost_bot<1, true>(q, o, Bc, bcb);
    else if (k == 2) ghost_bot<2, true>(q, o, Bc, bcb);
    const int kt = nz - 1 - k;
    if (kt == 0) ghost_top<0, true>(q, o, Tc, bct);
    else if (kt == 1) ghost_top<1, true>(q, o, Tc, bct);
    else if (kt == 2) ghost_top<2, true>(q, o, Tc, bct);
}

/* the level-k form of cell_ghosts for a 4-level window q[t] = level k-1+t
 * (face fluxes at faces 0..nz-2 reach one ghost at most) */
template <typename T>
__device__ __forceinline__ void cell_ghosts4(const T q[4], T o[4], int k, int nz, const Clos1 &Bc, const Clos1 &Tc,
                                             double bcb, double bct)
{
#pragma unroll
    for (int t = 0; t < 4; t++) o[t] = q[t];
    if (k == 0) o[0] = (T) Bc.G[0][0] * q[1] + (T) Bc.G[0][1] * q[2] + (T) Bc.G[0][2] * q[3] + (T) (Bc.B[0] * bcb);
    if (k == nz - 2) o[3] = (T) Tc.G[0][0] * q[2] + (T) Tc.G[0][1] * q[1] + (T) Tc.G[0][2] * q[0] + (T) (Tc.B[0] * bct);
}

/* w as the advection sees it from a raw face window q[t] = face k-3+t
 * (k <= nz-2): wall faces -1 and nz-1 are 0, ghost faces mirror oddly
 * (apply_bc_w and the ghost rewrite of massflux*4, advect.c:100-131) */
template <typename T>
__device__ __forceinline__ void wadv_window(const T q[7], T o[7], int k, int nz)
{
#pragma unroll
    for (int t = 0; t < 7; t++) o[t] = q[t];
    if (k == 0) { o[0] = -q[4]; o[1] = -q[3]; o[2] = (T) 0; }
    else if (k == 1) { o[0] = -q[2]; o[1] = (T) 0; }
    else if (k == 2) { o[0] = (T) 0; }
    const int kt = nz - 2 - k;
    if (kt == 0) { o[4] = (T) 0; o[5] = -q[3]; o[6] = -q[2]; }
    else if (kt == 1) { o[5] = (T) 0; o[6] = -q[4]; }
    else if (kt == 2) { o[6] = (T) 0; }
}

/* w for the implicit operator (k <= nz-2): closure ghosts (apply_bc_w4
 * re-applied before helmholtz1d_operator_w4) */
template <typename T>
__device__ __forceinline__ void wclos_window(const T q[7], T o[7], int k, int nz, const Clos1 &Bc, const Clos1 &Tc)
{
#pragma unroll
    for (int t = 0; t < 7; t++) {
        const int f = k - 3 + t;
        o[t] = (f >= 0 && f <= nz - 2) ? q[t] : (T) 0;
    }
    if (k == 0) ghost_bot<0, false>(q, o, Bc, 0.0);
    else if (k == 1) ghost_bot<1, false>(q, o, Bc, 0.0);
    else if (k == 2) ghost_bot<2, false>(q, o, Bc, 0.0);
    const int kt = nz - 2 - k;
    if (kt == 0) ghost_top<0, false>(q, o, Tc, 0.0);
    else if (kt == 1) ghost_top<1, false>(q, o, Tc, 0.0);
    else if (kt == 2) ghost_top<2, false>(q, o, Tc, 0.0);
}

template <typename T>
__device__ __forceinline__ T wall4(T gm2, T gm1, T x0, T x1)
{
    return (T) (0.5 * A2d) * (gm2 + x1) + (T) (0.5 * (A1d + 2.0 * A2d)) * (gm1 + x0);
}

/* sum_{+-} A1 (q(+1/2) - q(-1/2)) + A2 (q(+3/2) - q(-3/2)) of a 7-point
 * line, the 4th-order second difference without its 1/h^2 */
template <typename T>
__device__ __forceinline__ T d2line(const T X[7])
{
    const T A1 = (T) A1d, A2 = (T) A2d;
    const T p3 = A1 * (X[5] - X[4]) + A2 * (X[6] - X[3]);
    const T p1 = A1 * (X[4] - X[3]) + A2 * (X[5] - X[2]);
    const T m1 = A1 * (X[3] - X[2]) + A2 * (X[4] - X[1]);
    const T m3 = A1 * (X[2] - X[1]) + A2 * (X[3] - X[0]);
    return A1 * (p1 - m1) + A2 * (p3 - m3);
}

/* loads and stores at a 32-bit byte offset from a scalar base: the
 * hardware adds the base to the 32-bit vector offset, with no 64-bit
 * address arithmetic (arrays below 2^32 bytes, checked at creation) */
template <typename T>
__device__ __forceinline__ T ldo(gp<const T> b, unsigned off) { return *(gp<const T>) ((gp<const char>) b + off); }
template <typename T>
__device__ __forceinline__ void sto(gp<T> b, unsigned off, T v) { *(gp<T>) ((gp<char>) b + off) = v; }

}  // namespace


namespace {

/* 4th-order cell Helmholtz operator q + nudt L_z q from ghosted levels
 * k-3..k+3 (helmholtz1d_operator_u4, diffuseu.c:107-170), with the
 * reciprocals of the half-level distances at k+1/2, k-1/2, k+3/2, k-3/2 */
template <typename T>
__device__ __forceinline__ T helm4_cell(const T Q[7], T nudt, T rp1, T rm1, T rp3, T rm3, T rdzk)
{
    const T A1 = (T) A1d, A2 = (T) A2d;
    const T wp3h = A1 * (Q[5] - Q[4]) + A2 * (Q[6] - Q[3]);
    const T wp1h = A1 * (Q[4] - Q[3]) + A2 * (Q[5] - Q[2]);
    const T wm1h = A1 * (Q[3] - Q[2]) + A2 * (Q[4] - Q[1]);
    const T wm3h = A1 * (Q[2] - Q[1]) + A2 * (Q[3] - Q[0]);
    return Q[3] + nudt * ((A1 * (wp1h * rp1 - wm1h * rm1) + A2 * (wp3h * rp3 - wm3h * rm3)) * rdzk);
}

/* the tile of a workgroup: lb -> (tile x, tile y, z-chunk) */
struct Tile4 {
    int i0, j0, k0, k1, lx, ly, tid;
    __device__ __forceinline__ Tile4(int nx, int ny, int nz, int nch)
    {
        const int ntx = nx / TX, nty = ny / TY;
        const unsigned lb = xcd_block(blockIdx.x, gridDim.x);
        const int tx = (int) (lb % (unsigned) ntx);
        const unsigned rest = lb / (unsigned) ntx;
        const int ty = (int) (rest % (unsigned) nty);
        const int chunk = (int) (rest / (unsigned) nty);
        k0 = (int) (((long) chunk * nz) / nch);
        k1 = (int) (((long) (chunk + 1) * nz) / nch);
        tid = threadIdx.x; lx = tid % TX; ly = tid / TX;
        i0 = tx * TX; j0 = ty * TY;
    }
    /* plane offset of column (x, y) of the tile frame, periodic */
    __device__ __forceinline__ int col(int x, int y, int nx, int ny) const
    {
        int ii = i0 + x, jj = j0 + y;
        ii = ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii);
        jj = jj < 0 ? jj + ny : (jj >= ny ? jj - ny : jj);
        return jj * nx + ii;
    }
};

}  // namespace

/* ------------------------------------------------------------ u and v */

/* advectu4 / advectv4 with massfluxx4 / massfluxy4, laplacexy, buoyancy,
 * helmholtz1d_operator_u4 and the u / v right-hand sides
 * (chanfast.c:1380-1420); CFL partials when cflpart is set */
template <bool SC, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CF_R4_WAVES, 8))) void k_rhs4_uv(
    const RunDev R, const int s,
    const T *__restrict__ u_, const T *__restrict__ v_, const T *__restrict__ w_, const T *__restrict__ c_,
    const T *__restrict__ gu_, const T *__restrict__ gv_, const double *__restrict__ scal_,
    T *__restrict__ fu_, T *__restrict__ fv_, T *__restrict__ ru_, T *__restrict__ rv_,
    double *__restrict__ cflpart_, double cfldt, int nch)
{
    const gp<const T> u = (gp<const T>) u_, v = (gp<const T>) v_, w = (gp<const T>) w_, c = (gp<const T>) c_;
    const gp<const T> gu = (gp<const T>) gu_, gv = (gp<const T>) gv_;
    const gp<T> fu = (gp<T>) fu_, fv = (gp<T>) fv_, ru = (gp<T>) ru_, rv = (gp<T>) rv_;
    const cp<double> scal = (cp<double>) scal_;

    __shared__ T Pu[PY][PX], Pv[PY][PX], Pw[PY][PX];
    __shared__ T Pc[SC ? PY : 1][SC ? PX : 1];
    __shared__ double red[256];

    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int P = nx * ny;
    const Tile4 tl(nx, ny, nz, nch);
    const int lx = tl.lx, ly = tl.ly, tid = tl.tid, k0 = tl.k0, k1 = tl.k1;
    const unsigned oc = (unsigned) tl.col(lx, ly, nx, ny) * (unsigned) sizeof(T);
    int hx, hy;
    halo_xy(tid, hx, hy);
    const unsigned oh = (unsigned) tl.col(hx, hy, nx, ny) * (unsigned) sizeof(T);
    const unsigned Pb = (unsigned) P * (unsigned) sizeof(T);

    const T dx = (T) R.g.dx, dy = (T) R.g.dy;
    const T rdx2 = (T) (R.rdx * R.rdx), rdy2 = (T) (R.rdy * R.rdy);
    const cp<double> dz = (cp<double>) R.dzt + CF_NGT;
    const cp<double> rdz = (cp<double>) R.rdz + CF_NGT, rdzh = (cp<double>) R.rdzh + CF_NGT;
    const cp<double> lift = (cp<double>) R.liftg + (size_t) s * 4 * nz;
    const T A1 = (T) A1d, A2 = (T) A2d, B1 = (T) B1d, B2 = (T) B2d;
    const SubCoef &C = R.sc[s];
    const T gamdt = (T) C.gamdt, zetdt = (T) C.zetdt, nuadt = (T) C.nuadt;
    const T nu = (T) R.nu;
    const bool impl = R.implicit != 0;
    const T mfu = s > 0 ? (T) scal[R.m_prev_slot[s]] : (T) 0;
    const T bgu = SC ? (T) R.betg[0] : (T) 0, bgv = SC ? (T) R.betg[1] : (T) 0;
    const Clos1 &Cb = R.cl.cell[0], &Ct = R.cl.cell[1];
    const T h5 = (T) 0.5;

    auto ld_cell = [&](gp<const T> q, int L, unsigned o) -> T { return (L >= 0 && L < nz) ? ldo(q, (unsigned) L * Pb + o) : (T) 0; };
    auto ld_face = [&](int L, unsigned o) -> T { return (L >= 0 && L <= nz - 2) ? ldo(w, (unsigned) L * Pb + o) : (T) 0; };

    /* x- and y-interpolated w at faces k-3..k (before a level's update) */
    T WXw[4], WYw[4];
    WXw[0] = (T) 0; WYw[0] = (T) 0;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const int f = k0 - 2 + e;
        T wxv = (T) 0, wyv = (T) 0;
        if (f >= 0 && f <= nz - 2) {                   /* uniform */
            Pw[ly + H][lx + H] = ldo(w, (unsigned) f * Pb + oc);
            Pw[hy + H][hx + H] = ldo(w, (unsigned) f * Pb + oh);
            __syncthreads();
            wxv = B1 * (Pw[ly + H][lx + H] + Pw[ly + H][lx + H + 1]) + B2 * (Pw[ly + H][lx + H - 1] + Pw[ly + H][lx + H + 2]);
            wyv = B1 * (Pw[ly + H][lx + H] + Pw[ly + H + 1][lx + H]) + B2 * (Pw[ly + H - 1][lx + H] + Pw[ly + H + 2][lx + H]);
            __syncthreads();
        }
        WXw[e + 1] = wxv; WYw[e + 1] = wyv;
    }
    if (k0 == 0) { WXw[1] = -WXw[3]; WYw[1] = -WYw[3]; }     /* face -2 = -face 0; face -1 = 0 */

    T Uc[7], Vc[7];                                     /* levels k-3..k+3 of the own column */
#pragma unroll
    for (int t = 0; t < 7; t++) { Uc[t] = ld_cell(u, k0 - 3 + t, oc); Vc[t] = ld_cell(v, k0 - 3 + t, oc); }
    T hu = ld_cell(u, k0, oh), hv = ld_cell(v, k0, oh), hc = SC ? ld_cell(c, k0, oh) : (T) 0;
    T cc = SC ? ld_cell(c, k0, oc) : (T) 0;
    T wo = ld_face(k0 + 1, oc), hw = ld_face(k0 + 1, oh);
    const bool want_cfl = cflpart_ != nullptr;
    double cflv = 0.0;

    for (int k = k0; k < k1; k++) {
        Pu[ly + H][lx + H] = Uc[3];
        Pv[ly + H][lx + H] = Vc[3];
        Pw[ly + H][lx + H] = wo;
        Pu[hy + H][hx + H] = hu;
        Pv[hy + H][hx + H] = hv;
        Pw[hy + H][hx + H] = hw;
        if (SC) { Pc[ly + H][lx + H] = cc; Pc[hy + H][hx + H] = hc; }
        /* the next level's loads, in flight across this level's work */
        const T nU = ld_cell(u, k + 4, oc), nV = ld_cell(v, k + 4, oc);
        const T nhu = ld_cell(u, k + 1, oh), nhv = ld_cell(v, k + 1, oh);
        const T ncc = SC ? ld_cell(c, k + 1, oc) : (T) 0, nhc = SC ? ld_cell(c, k + 1, oh) : (T) 0;
        const T nwo = ld_face(k + 2, oc), nhw = ld_face(k + 2, oh);
        const unsigned q = (unsigned) k * Pb + oc;
        const T gu_q = s > 0 ? ldo(gu, q) : (T) 0, gv_q = s > 0 ? ldo(gv, q) : (T) 0;
        const T wq = want_cfl && k <= nz - 2 ? ldo(w, q) : (T) 0;
        __syncthreads();

        /* x/y-interpolated w at face k+1 */
        {
            T ax, ay;
            if (k + 1 <= nz - 2) {
                ax = B1 * (Pw[ly + H][lx + H] + Pw[ly + H][lx + H + 1]) + B2 * (Pw[ly + H][lx + H - 1] + Pw[ly + H][lx + H + 2]);
                ay = B1 * (Pw[ly + H][lx + H] + Pw[ly + H + 1][lx + H]) + B2 * (Pw[ly + H - 1][lx + H] + Pw[ly + H + 2][lx + H]);
            } else if (k + 1 == nz - 1) {
                ax = (T) 0; ay = (T) 0;
            } else {                                   /* face nz = -face nz-2 */
                ax = -WXw[2]; ay = -WYw[2];
            }
            WXw[0] = WXw[1]; WXw[1] = WXw[2]; WXw[2] = WXw[3]; WXw[3] = ax;
            WYw[0] = WYw[1]; WYw[1] = WYw[2]; WYw[2] = WYw[3]; WYw[3] = ay;
        }
        const T dzk = (T) dz[k];
        const T rden = (T) (R.rdxdy * rdz[k]);
        const T rp1 = (T) rdzh[k], rm1 = (T) rdzh[k - 1], rp3 = (T) rdzh[k + 1], rm3 = (T) rdzh[k - 2];
        const T rdzk = (T) rdz[k];
        const T sy = dy * dzk, sx = dx * dzk, sxy = dx * dy;

        /* ---- u */
        {
            T Ug[7];
            cell_ghosts(Uc, Ug, k, nz, Cb, Ct, R.ubc, R.ubc);
            T UX[7], UY[7];
#pragma unroll
            for (int t = 0; t < 7; t++) { UX[t] = Pu[ly + H][lx + t]; UY[t] = Pu[ly + t][lx + H]; }
            const T Q0 = Ug[3];
            auto MU = [&](int d) -> T { return sy * (B1 * (UX[3 + d] + UX[4 + d]) + B2 * (UX[2 + d] + UX[5 + d])); };
            auto MV = [&](int e) -> T {
                return sx * (B1 * (Pv[ly + H + e][lx + H] + Pv[ly + H + e][lx + H + 1])
                           + B2 * (Pv[ly + H + e][lx + H - 1] + Pv[ly + H + e][lx + H + 2]));
            };
            const T xp = MU(0) * h5 * (Q0 + UX[4]), xm = MU(-1) * h5 * (Q0 + UX[2]);
            const T xp3 = MU(1) * h5 * (Q0 + UX[6]), xm3 = MU(-2) * h5 * (Q0 + UX[0]);
            const T yp = MV(0) * h5 * (Q0 + UY[4]), ym = MV(-1) * h5 * (Q0 + UY[2]);
            const T yp3 = MV(1) * h5 * (Q0 + UY[6]), ym3 = MV(-2) * h5 * (Q0 + UY[0]);
            const T mw0 = sxy * WXw[2], mwm = sxy * WXw[1], mwp = sxy * WXw[3], mwm2 = sxy * WXw[0];
            T zp = mw0 * h5 * (Q0 + Ug[4]), zm = mwm * h5 * (Q0 + Ug[2]);
            T zp3 = mwp * h5 * (Q0 + Ug[6]), zm3 = mwm2 * h5 * (Q0 + Ug[0]);
            if (k == nz - 1) {
                const T bct = wall4(Ug[5], Ug[4], Ug[3], Ug[2]);
                zp = mw0 * bct;
                zp3 = -mwm * h5 * (Ug[1] + Ug[4]) + (T) 2.0 * mw0 * bct;
            }
            if (k == nz - 2) zp3 = mwp * wall4(Ug[6], Ug[5], Ug[4], Ug[3]);
            if (k == 0) {
                const T bcb = wall4(Ug[1], Ug[2], Ug[3], Ug[4]);
                zm = mwm * bcb;
                zm3 = -mw0 * h5 * (Ug[5] + Ug[2]) + (T) 2.0 * mwm * bcb;
            }
            if (k == 1) zm3 = mwm2 * wall4(Ug[0], Ug[1], Ug[2], Ug[3]);
            T f = -((A1 * (xp - xm) + A2 * (xp3 - xm3)) + (A1 * (yp - ym) + A2 * (yp3 - ym3))
                    + (A1 * (zp - zm) + A2 * (zp3 - zm3))) * rden;
            const T lxy = d2line(UX) * rdx2 + d2line(UY) * rdy2;
            if (!impl) f += nu * lxy;
            if (SC) f += (B1 * (Pc[ly + H][lx + H] + Pc[ly + H][lx + H + 1]) + B2 * (Pc[ly + H][lx + H - 1] + Pc[ly + H][lx + H + 2])) * bgu;
            T h = helm4_cell(Ug, nuadt, rp1, rm1, rp3, rm3, rdzk);
            if (impl) h += nuadt * lxy;
            h += gamdt * f;
            if (s > 0) h += zetdt * (gu_q + mfu);
            h -= (T) lift[0 * nz + k];
            if (s < 2) sto(fu, q, f);
            sto(ru, q, impl ? h : h * dzk);
            if (want_cfl)
                cflv = fmax(cflv, fmax(fmax(fabs((double) Q0) * cfldt * R.rdx, fabs((double) Vc[3]) * cfldt * R.rdy),
                                       fabs((double) wq) * cfldt * rdzh[k]));
        }
        /* ---- v */
        {
            T Vg[7];
            cell_ghosts(Vc, Vg, k, nz, Cb, Ct, R.vbc, R.vbc);
            T VX[7], VY[7];
#pragma unroll
            for (int t = 0; t < 7; t++) { VX[t] = Pv[ly + H][lx + t]; VY[t] = Pv[ly + t][lx + H]; }
            const T Q0 = Vg[3];
            auto MU = [&](int d) -> T {
                return sy * (B1 * (Pu[ly + H][lx + H + d] + Pu[ly + H + 1][lx + H + d])
                           + B2 * (Pu[ly + H - 1][lx + H + d] + Pu[ly + H + 2][lx + H + d]));
            };
            auto MV = [&](int e) -> T { return sx * (B1 * (VY[3 + e] + VY[4 + e]) + B2 * (VY[2 + e] + VY[5 + e])); };
            const T xp = MU(0) * h5 * (Q0 + VX[4]), xm = MU(-1) * h5 * (Q0 + VX[2]);
            const T xp3 = MU(1) * h5 * (Q0 + VX[6]), xm3 = MU(-2) * h5 * (Q0 + VX[0]);
            const T yp = MV(0) * h5 * (Q0 + VY[4]), ym = MV(-1) * h5 * (Q0 + VY[2]);
            const T yp3 = MV(1) * h5 * (Q0 + VY[6]), ym3 = MV(-2) * h5 * (Q0 + VY[0]);
            const T mw0 = sxy * WYw[2], mwm = sxy * WYw[1], mwp = sxy * WYw[3], mwm2 = sxy * WYw[0];
            T zp = mw0 * h5 * (Q0 + Vg[4]), zm = mwm * h5 * (Q0 + Vg[2]);
            T zp3 = mwp * h5 * (Q0 + Vg[6]), zm3 = mwm2 * h5 * (Q0 + Vg[0]);
            if (k == nz - 1) {
                const T bct = wall4(Vg[5], Vg[4], Vg[3], Vg[2]);
                zp = mw0 * bct;
                zp3 = -mwm * h5 * (Vg[1] + Vg[4]) + (T) 2.0 * mw0 * bct;
            }
            if (k == nz - 2) zp3 = mwp * wall4(Vg[6], Vg[5], Vg[4], Vg[3]);
            if (k == 0) {
                const T bcb = wall4(Vg[1], Vg[2], Vg[3], Vg[4]);
                zm = mwm * bcb;
                zm3 = -mw0 * h5 * (Vg[5] + Vg[2]) + (T) 2.0 * mwm * bcb;
            }
            if (k == 1) zm3 = mwm2 * wall4(Vg[0], Vg[1], Vg[2], Vg[3]);
            T f = -((A1 * (xp - xm) + A2 * (xp3 - xm3)) + (A1 * (yp - ym) + A2 * (yp3 - ym3))
                    + (A1 * (zp - zm) + A2 * (zp3 - zm3))) * rden;
            const T lxy = d2line(VX) * rdx2 + d2line(VY) * rdy2;
            if (!impl) f += nu * lxy;
            if (SC) f += (B1 * (Pc[ly + H][lx + H] + Pc[ly + H + 1][lx + H]) + B2 * (Pc[ly + H - 1][lx + H] + Pc[ly + H + 2][lx + H])) * bgv;
            T h = helm4_cell(Vg, nuadt, rp1, rm1, rp3, rm3, rdzk);
            if (impl) h += nuadt * lxy;
            h += gamdt * f;
            if (s > 0) h += zetdt * gv_q;
            h -= (T) lift[1 * nz + k];
            if (s < 2) sto(fv, q, f);
            sto(rv, q, impl ? h : h * dzk);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 6; t++) { Uc[t] = Uc[t + 1]; Vc[t] = Vc[t + 1]; }
        Uc[6] = nU; Vc[6] = nV;
        hu = nhu; hv = nhv; hc = nhc; cc = ncc; wo = nwo; hw = nhw;
    }
    if (want_cfl) {
        red[tid] = cflv;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (tid < st) red[tid] = fmax(red[tid], red[tid + st]);
            __syncthreads();
        }
        if (tid == 0) ((gp<double>) cflpart_)[blockIdx.x] = red[0];
    }
}

/* ------------------------------------------------------------ w and c */

/* advectw4 with massfluxz4, laplacexy_w, buoyancy,
 * helmholtz1d_operator_w4 and the w right-hand side; with the scalar,
 * flux_quick + advect_scalar, laplacexy with nu / Pr and the scalar's
 * right-hand side (chanfast.c:1420-1470, :1495-1530) */
template <bool SC, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CF_R4_WAVES, 8))) void k_rhs4_wc(
    const RunDev R, const int s,
    const T *__restrict__ u_, const T *__restrict__ v_, const T *__restrict__ w_, const T *__restrict__ c_,
    const T *__restrict__ gw_, const T *__restrict__ gc_, const double *__restrict__ scal_,
    T *__restrict__ fw_, T *__restrict__ fc_, T *__restrict__ rw_, T *__restrict__ rc_, int nch)
{
    const gp<const T> u = (gp<const T>) u_, v = (gp<const T>) v_, w = (gp<const T>) w_, c = (gp<const T>) c_;
    const gp<const T> gw = (gp<const T>) gw_, gc = (gp<const T>) gc_;
    const gp<T> fw = (gp<T>) fw_, fc = (gp<T>) fc_, rw = (gp<T>) rw_, rc = (gp<T>) rc_;
    const cp<double> scal = (cp<double>) scal_;

    __shared__ T Pw[PY][PX];
    __shared__ T Pc[SC ? PY : 1][SC ? PX : 1];
    __shared__ T Puz[TY][UZX], Pvz[VZY][TX];

    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int P = nx * ny;
    const Tile4 tl(nx, ny, nz, nch);
    const int lx = tl.lx, ly = tl.ly, tid = tl.tid, k0 = tl.k0, k1 = tl.k1;
    const unsigned oc = (unsigned) tl.col(lx, ly, nx, ny) * (unsigned) sizeof(T);
    int hx, hy;
    halo_xy(tid, hx, hy);
    const unsigned oh = (unsigned) tl.col(hx, hy, nx, ny) * (unsigned) sizeof(T);
    const unsigned Pb = (unsigned) P * (unsigned) sizeof(T);
    /* a 4-level window for a halo column of the shared face fluxes: threads
     * 0..3TX-1 the vz rows y = -2, -1, TY; the next 3TY threads the uz
     * columns x = -2, -1, TX */
    const int zkind = tid < 3 * TX ? 1 : (tid < 3 * TX + 3 * TY ? 0 : -1);      /* 0 u, 1 v, -1 none */
    int zx = 0, zy = 0;
    if (zkind == 1) { zx = tid % TX; const int r = tid / TX; zy = r < 2 ? r - 2 : TY; }
    else if (zkind == 0) { const int e = tid - 3 * TX, q = e / TY; zx = q < 2 ? q - 2 : TX; zy = e % TY; }
    const unsigned oz = (unsigned) tl.col(zx, zy, nx, ny) * (unsigned) sizeof(T);
    const gp<const T> zsrc = zkind == 1 ? v : u;
    const unsigned oxm = (unsigned) tl.col(lx - 1, ly, nx, ny) * (unsigned) sizeof(T);
    const unsigned oym = (unsigned) tl.col(lx, ly - 1, nx, ny) * (unsigned) sizeof(T);

    const T dx = (T) R.g.dx, dy = (T) R.g.dy;
    const T rdx2 = (T) (R.rdx * R.rdx), rdy2 = (T) (R.rdy * R.rdy);
    const cp<double> dz = (cp<double>) R.dzt + CF_NGT;
    const cp<double> rdz = (cp<double>) R.rdz + CF_NGT, rdzh = (cp<double>) R.rdzh + CF_NGT;
    const cp<double> lift = (cp<double>) R.liftg + (size_t) s * 4 * nz;
    const T A1 = (T) A1d, A2 = (T) A2d, B1 = (T) B1d, B2 = (T) B2d;
    const SubCoef &C = R.sc[s];
    const T gamdt = (T) C.gamdt, zetdt = (T) C.zetdt, nuadt = (T) C.nuadt, kadt = (T) C.kadt;
    const T nu = (T) R.nu, kap = (T) R.kap;
    const bool impl = R.implicit != 0;
    const T mfc = (SC && s > 0) ? (T) scal[SC_MC0 + (s + 2) % 3] : (T) 0;
    const T bgw = SC ? (T) R.betg[2] : (T) 0;
    const Clos1 &Cb = R.cl.cell[0], &Ct = R.cl.cell[1], &Fb = R.cl.face[0], &Ft = R.cl.face[1];
    const T h5 = (T) 0.5;

    auto ld_cell = [&](gp<const T> q, int L, unsigned o) -> T { return (L >= 0 && L < nz) ? ldo(q, (unsigned) L * Pb + o) : (T) 0; };
    auto ld_face = [&](int L, unsigned o) -> T { return (L >= 0 && L <= nz - 2) ? ldo(w, (unsigned) L * Pb + o) : (T) 0; };

    T Wc[7], Cc[7];                                     /* faces / levels k-3..k+3 of the own column */
    T U4[4], V4[4], Zh[4];                              /* levels k-1..k+2: own column, halo column */
#pragma unroll
    for (int t = 0; t < 7; t++) {
        Wc[t] = ld_face(k0 - 3 + t, oc);
        Cc[t] = SC ? ld_cell(c, k0 - 3 + t, oc) : (T) 0;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        U4[t] = ld_cell(u, k0 - 1 + t, oc);
        V4[t] = ld_cell(v, k0 - 1 + t, oc);
        Zh[t] = zkind >= 0 ? ld_cell(zsrc, k0 - 1 + t, oz) : (T) 0;
    }
    T hw = ld_face(k0, oh), hc = SC ? ld_cell(c, k0, oh) : (T) 0;

    for (int k = k0; k < k1; k++) {
        const bool wface = k <= nz - 2;                 /* w unknown at face k (uniform) */
        const T dzk = (T) dz[k];
        Pw[ly + H][lx + H] = Wc[3];
        Pw[hy + H][hx + H] = hw;
        if (SC) { Pc[ly + H][lx + H] = Cc[3]; Pc[hy + H][hx + H] = hc; }
        if (wface) {
            /* massfluxz of u and v at face k */
            const T dzm = (T) dz[k - 1], dzp = (T) dz[k + 1], dzp2 = (T) dz[k + 2];
            T Ug[4], Vg[4];
            cell_ghosts4(U4, Ug, k, nz, Cb, Ct, R.ubc, R.ubc);
            cell_ghosts4(V4, Vg, k, nz, Cb, Ct, R.vbc, R.vbc);
            Puz[ly][lx + 2] = dy * (B1 * (dzk * Ug[1] + dzp * Ug[2]) + B2 * (dzm * Ug[0] + dzp2 * Ug[3]));
            Pvz[ly + 2][lx] = dx * (B1 * (dzk * Vg[1] + dzp * Vg[2]) + B2 * (dzm * Vg[0] + dzp2 * Vg[3]));
            if (zkind >= 0) {
                const double bc = zkind == 1 ? R.vbc : R.ubc;
                T Zg[4];
                cell_ghosts4(Zh, Zg, k, nz, Cb, Ct, bc, bc);
                const T val = (zkind == 1 ? dx : dy) * (B1 * (dzk * Zg[1] + dzp * Zg[2]) + B2 * (dzm * Zg[0] + dzp2 * Zg[3]));
                if (zkind == 1) Pvz[zy + 2][zx] = val;
                else Puz[zy][zx + 2] = val;
            }
        }
        /* the next level's loads, in flight across this level's work */
        const T nW = ld_face(k + 4, oc), nC = SC ? ld_cell(c, k + 4, oc) : (T) 0;
        const T nU = ld_cell(u, k + 3, oc), nV = ld_cell(v, k + 3, oc);
        const T nZ = zkind >= 0 ? ld_cell(zsrc, k + 3, oz) : (T) 0;
        const T nhw = ld_face(k + 1, oh), nhc = SC ? ld_cell(c, k + 1, oh) : (T) 0;
        const unsigned q = (unsigned) k * Pb + oc;
        const T gw_q = (s > 0 && wface) ? ldo(gw, q) : (T) 0;
        const T gc_q = (SC && s > 0) ? ldo(gc, q) : (T) 0;
        const T uxm = SC ? ldo(u, (unsigned) k * Pb + oxm) : (T) 0, vym = SC ? ldo(v, (unsigned) k * Pb + oym) : (T) 0;
        __syncthreads();

        /* ---- w at face k */
        if (wface) {
            T WA[7], WH[7];
            wadv_window(Wc, WA, k, nz);
            wclos_window(Wc, WH, k, nz, Fb, Ft);
            T WX[7], WY[7];
#pragma unroll
            for (int t = 0; t < 7; t++) { WX[t] = Pw[ly + H][lx + t]; WY[t] = Pw[ly + t][lx + H]; }
            const T Q0 = WA[3];
            const T sxy = dx * dy;
            auto MUZ = [&](int d) -> T { return Puz[ly][lx + 2 + d]; };
            auto MVZ = [&](int e) -> T { return Pvz[ly + 2 + e][lx]; };
            auto MWZ = [&](int e) -> T { return sxy * (B1 * (WA[3 + e] + WA[4 + e]) + B2 * (WA[2 + e] + WA[5 + e])); };
            const T xp = MUZ(0) * h5 * (Q0 + WX[4]), xm = MUZ(-1) * h5 * (Q0 + WX[2]);
            const T xp3 = MUZ(1) * h5 * (Q0 + WX[6]), xm3 = MUZ(-2) * h5 * (Q0 + WX[0]);
            const T yp = MVZ(0) * h5 * (Q0 + WY[4]), ym = MVZ(-1) * h5 * (Q0 + WY[2]);
            const T yp3 = MVZ(1) * h5 * (Q0 + WY[6]), ym3 = MVZ(-2) * h5 * (Q0 + WY[0]);
            const T zp = MWZ(0) * h5 * (Q0 + WA[4]), zm = MWZ(-1) * h5 * (Q0 + WA[2]);
            const T zp3 = MWZ(1) * h5 * (Q0 + WA[6]), zm3 = MWZ(-2) * h5 * (Q0 + WA[0]);
            const T rfh = (T) rdzh[k];
            T f = -((A1 * (xp - xm) + A2 * (xp3 - xm3)) + (A1 * (yp - ym) + A2 * (yp3 - ym3))
                    + (A1 * (zp - zm) + A2 * (zp3 - zm3))) * ((T) R.rdxdy * rfh);
            const T lxy = d2line(WX) * rdx2 + d2line(WY) * rdy2;
            if (!impl) f += nu * lxy;
            if (SC) {
                T Cg[7];
                cell_ghosts(Cc, Cg, k, nz, Cb, Ct, R.cbot, R.ctop);
                f += (B1 * (Cg[3] + Cg[4]) + B2 * (Cg[2] + Cg[5])) * bgw;
            }
            /* helmholtz1d_operator_w4: face distances dz(k+1), dz(k), dz(k+2), dz(k-1), face volume dzh(k) */
            const T wp3h = A1 * (WH[5] - WH[4]) + A2 * (WH[6] - WH[3]);
            const T wp1h = A1 * (WH[4] - WH[3]) + A2 * (WH[5] - WH[2]);
            const T wm1h = A1 * (WH[3] - WH[2]) + A2 * (WH[4] - WH[1]);
            const T wm3h = A1 * (WH[2] - WH[1]) + A2 * (WH[3] - WH[0]);
            T h = WH[3] + nuadt * ((A1 * (wp1h * (T) rdz[k + 1] - wm1h * (T) rdz[k])
                                    + A2 * (wp3h * (T) rdz[k + 2] - wm3h * (T) rdz[k - 1])) * rfh);
            if (impl) h += nuadt * lxy;
            h += gamdt * f;
            if (s > 0) h += zetdt * gw_q;
            h -= (T) lift[2 * nz + k];
            if (s < 2) sto(fw, q, f);
            sto(rw, q, impl ? h : h * (h5 * ((T) dz[k + 1] + dzk)));
        } else {                                       /* top wall face: f = 0 (advect.c:301), no unknown */
            if (s < 2) sto(fw, q, (T) 0);
            sto(rw, q, (T) 0);
        }
        /* ---- c */
        if (SC) {
            T Cg[7];
            cell_ghosts(Cc, Cg, k, nz, Cb, Ct, R.cbot, R.ctop);
            T CX[7], CY[7];
#pragma unroll
            for (int t = 0; t < 7; t++) { CX[t] = Pc[ly + H][lx + t]; CY[t] = Pc[ly + t][lx + H]; }
            const T qk = (T) -0.125;
            auto quick = [&](T fl, T cm, T c0, T c1, T c2) -> T {
                T cf = h5 * (c0 + c1);
                cf += fl > (T) 0 ? qk * (c1 - (T) 2.0 * c0 + cm) : qk * (c2 - (T) 2.0 * c1 + c0);
                return fl * cf;
            };
            const T fu0 = dy * dzk * U4[1], fum = dy * dzk * uxm;
            const T fv0 = dx * dzk * V4[1], fvm = dx * dzk * vym;
            const T wk = k <= nz - 2 ? Wc[3] : (T) 0, wkm = k >= 1 ? Wc[2] : (T) 0;
            const T fw0 = dx * dy * wk, fwm = dx * dy * wkm;
            const T mu0 = quick(fu0, CX[2], CX[3], CX[4], CX[5]), mum = quick(fum, CX[1], CX[2], CX[3], CX[4]);
            const T mv0 = quick(fv0, CY[2], CY[3], CY[4], CY[5]), mvm = quick(fvm, CY[1], CY[2], CY[3], CY[4]);
            const T mw0 = quick(fw0, Cg[2], Cg[3], Cg[4], Cg[5]), mwm = quick(fwm, Cg[1], Cg[2], Cg[3], Cg[4]);
            T f = -((mu0 - mum) + (mv0 - mvm) + (mw0 - mwm)) * (T) (R.rdxdy * rdz[k]);
            const T lxy = d2line(CX) * rdx2 + d2line(CY) * rdy2;
            if (!impl) f += kap * lxy;
            T h = helm4_cell(Cg, kadt, (T) rdzh[k], (T) rdzh[k - 1], (T) rdzh[k + 1], (T) rdzh[k - 2], (T) rdz[k]);
            if (impl) h += kadt * lxy;
            h += gamdt * f;
            if (s > 0) h += zetdt * (gc_q + mfc);
            h -= (T) lift[3 * nz + k];
            if (s < 2) sto(fc, q, f);
            sto(rc, q, impl ? h : h * dzk);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 6; t++) { Wc[t] = Wc[t + 1]; Cc[t] = Cc[t + 1]; }
        Wc[6] = nW; Cc[6] = nC;
#pragma unroll
        for (int t = 0; t < 3; t++) { U4[t] = U4[t + 1]; V4[t] = V4[t + 1]; Zh[t] = Zh[t + 1]; }
        U4[3] = nU; V4[3] = nV; Zh[3] = nZ;
        hw = nhw; hc = nhc;
    }
}

#define R4_INST(SC, T)                                                                                           \
    template __global__ void k_rhs4_uv<SC, T>(const RunDev, const int, const T *, const T *, const T *, const T *, \
        const T *, const T *, const double *, T *, T *, T *, T *, double *, double, int);                         \
    template __global__ void k_rhs4_wc<SC, T>(const RunDev, const int, const T *, const T *, const T *, const T *, \
        const T *, const T *, const double *, T *, T *, T *, T *, int);
R4_INST(false, double)
R4_INST(true, double)
R4_INST(false, float)
R4_INST(true, float)
