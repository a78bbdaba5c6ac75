/*
 * cf_kernels.hip -- HIP kernels of one RK3 substep of the channel DNS on
 * gfx950: the FP64 order-2 fast path, and the solves, FFT passes and
 * transposes the general path shares (templated on the storage type T,
 * double or float).  Reference semantics per kernel are cited beside it; the
 * device layout is described in cf_dev.h.  Every kernel bounds-checks its
 * indices against the Grid it is given; the host checks launch shapes.
 */
#include "cf_dev.h"
#include "cf_kernels.h"

#include <math.h>

static constexpr double A1_ = 1.125 / 1.0;      /* fourth.c:3-6 */
static constexpr double A2_ = -0.125 / 3.0;

/* ---------------------------------------------------------------------
 * io layout [nx][ny][nz] <-> device layout [nz][ny][nx]: for each y the
 * (x, z) plane is transposed through a 32x33 LDS tile (bit-exact moves;
 * a float device field is the io double rounded to nearest).
 * Replaces the pencil redistributions of transpose.c:409-448 on one GPU.
 * ------------------------------------------------------------------- */
template <typename T>
__global__ __launch_bounds__(256) void k_io_to_dev(const double *__restrict__ io, T *__restrict__ dev,
                                                   int nx, int ny, int nz, const double *__restrict__ off)
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
        if (i < nx && k < nz) dev[((size_t) k * ny + j) * nx + i] = (T) (off ? tile[tx][r] - off[k] : tile[tx][r]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_dev_to_io(const T *__restrict__ dev, double *__restrict__ io,
                                                   int nx, int ny, int nz, const double *__restrict__ off)
{
    __shared__ double tile[32][33];
    const int i0 = blockIdx.x * 32, k0 = blockIdx.y * 32, j = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        int k = k0 + r, i = i0 + tx;
        if (i < nx && k < nz) tile[r][tx] = off ? (double) dev[((size_t) k * ny + j) * nx + i] + off[k]
                                                : (double) dev[((size_t) k * ny + j) * nx + i];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + r, k = k0 + tx;
        if (i < nx && k < nz) io[((size_t) i * ny + j) * nz + k] = tile[tx][r];
    }
}

template __global__ void k_io_to_dev<double>(const double *, double *, int, int, int, const double *);
template __global__ void k_io_to_dev<float>(const double *, float *, int, int, int, const double *);
template __global__ void k_dev_to_io<double>(const double *, double *, int, int, int, const double *);
template __global__ void k_dev_to_io<float>(const float *, double *, int, int, int, const double *);

/* ---------------------------------------------------------------------
 * Explicit and implicit-solve right-hand sides of one substep, 2nd order,
 * 2.5-D blocked and with the
 * previous substep's projection (project2, pressure.c:120-152, and the
 * mass-flux correction u += h1 mdpdx, chanfast.c:519-522) applied on the
 * fly, so the projected velocity is never stored.  A workgroup owns a
 * 32 x 8 column tile and marches levels [k0, k1); a four-slot LDS ring
 * holds the projected u, v, w of the tile plus a one-cell halo (periodic
 * wrap in x, y; wall ghosts 2 bc - q_1 for u, v and 0 for w).  Level k + 3
 * is fetched into registers before level k is computed and level k + 2 is
 * built into the ring right after it, with one barrier per level.  Substep 0 also
 * reduces the CFL number of the projected state (get_cfl, get.c:390-417).
 * ------------------------------------------------------------------- */
#define RT_X CF_RT_X
#define RT_Y CF_RT_Y
#define RT_HX (RT_X + 2)
#define RT_HY (RT_Y + 2)
#define RT_PX (RT_X + 3)
#define RT_PY (RT_Y + 3)
#define RT_NA (RT_HX * RT_HY)      /* 396 halo-tile cells */
#define RT_NP (RT_PX * RT_PY)      /* 469 pressure cells (one more in +x, +y) */

/* lb: the tile's block index within the run (ntx nty nchunk of them) */
template <typename T>
static __device__ __forceinline__ void rhs2t_body(const RunDev &R, const int s, const ProjSrcT<T> &S,
    const T *__restrict__ gu_, const T *__restrict__ gv_, const T *__restrict__ gw_,
    const double *__restrict__ scal_,
    T *__restrict__ fu_, T *__restrict__ fv_, T *__restrict__ fw_,
    T *__restrict__ ru_, T *__restrict__ rv_, T *__restrict__ rw_,
    double *__restrict__ cflpart_, double cfldt, int kc, const unsigned lb)
{
    const gp<const T> gu = (gp<const T>) gu_, gv = (gp<const T>) gv_, gw = (gp<const T>) gw_;
    const gp<T> fu = (gp<T>) fu_, fv = (gp<T>) fv_, fw = (gp<T>) fw_;
    const gp<T> ru = (gp<T>) ru_, rv = (gp<T>) rv_, rw = (gp<T>) rw_;
    const gp<double> cflpart = (gp<double>) cflpart_;
    const cp<double> scal = (cp<double>) scal_;
    const gp<const T> Sus = (gp<const T>) S.us, Svs = (gp<const T>) S.vs, Sws = (gp<const T>) S.ws;
    const gp<const T> Sp = (gp<const T>) S.p;
    const cp<double> Sh1 = (cp<double>) S.h1, Sm = (cp<double>) S.m;
    __shared__ __align__(16) T SU[4][RT_HY][RT_HX], SV[4][RT_HY][RT_HX], SW[4][RT_HY][RT_HX];
    __shared__ T SP[2][2][RT_PY][RT_PX];          /* [buffer][p(L), p(L+1)] */
    double *red = (double *) &SU[0][0][0];        /* CFL reduction after the last level (>= 256 doubles) */
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t P = (size_t) nx * ny;
    const int ntx = (nx + RT_X - 1) / RT_X, nty = (ny + RT_Y - 1) / RT_Y;
    const int tx = (int) (lb % (unsigned) ntx);
    const unsigned rest = lb / (unsigned) ntx;
    const int ty = (int) (rest % (unsigned) nty);
    const int chunk = (int) (rest / (unsigned) nty);
    const int i0 = tx * RT_X, j0 = ty * RT_Y;
    const int k0 = chunk * kc;
    if (k0 >= nz) return;                         /* uniform over the block */
    const int k1 = min(k0 + kc, nz);
    const int tid = threadIdx.x;
    const int lx = tid % RT_X, ly = tid / RT_X;
    const int i = i0 + lx, j = j0 + ly;
    const bool act = i < nx && j < ny;

    const T m = Sm ? (T) *Sm : (T) 0;
    const T cx = (T) (S.dtp * R.rdx), cy = (T) (S.dtp * R.rdy);
    const cp<double> dz = (cp<double>) R.dzt + CF_NGT;
    const cp<double> rdz = (cp<double>) R.rdz + CF_NGT, rdzh = (cp<double>) R.rdzh + CF_NGT;
    const T ub = (T) R.ubc, vb = (T) R.vbc;
    /* FP32 storage holds u - U(z) (R.uoff): the ring then holds that
     * deviation, whose wall ghost is its negative (U's ghosts carry the
     * wall value); full values are formed only where the advection needs
     * them, and U's share of the implicit right-hand side is added per
     * level.  FP64 storage holds u itself. */
    constexpr bool OFF = sizeof(T) == 4;
    const cp<double> uof = (cp<double>) (OFF ? R.uoff : nullptr);
    const T ubs = (OFF && uof) ? (T) 0 : ub;
    auto F = [&](T v, T o) -> T { if constexpr (OFF) return v + o; else return v; };

    /* element e of the halo tile / pressure tile -> global column offset */
    auto colo = [&](int e, int w) -> size_t {
        const int r = e / w, c = e - r * w;
        int ii = i0 - 1 + c, jj = j0 - 1 + r;
        ii %= nx; if (ii < 0) ii += nx;
        jj %= ny; if (jj < 0) jj += ny;
        return (size_t) jj * nx + ii;
    };
    /* element offsets as 32-bit byte offsets from each array's (scalar)
     * base: the loads take base + 32-bit vector offset directly, with no
     * 64-bit address arithmetic (N sizeof(T) < 2^32, checked at creation) */
    unsigned oa[2], op[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int e = tid + 256 * t;
        oa[t] = (unsigned) ((e < RT_NA ? colo(e, RT_HX) : 0) * sizeof(T));
        op[t] = (unsigned) ((e < RT_NP ? colo(e, RT_PX) : 0) * sizeof(T));
    }
    const unsigned Pb = (unsigned) (P * sizeof(T));
    auto ldb = [](gp<const T> base, unsigned off) -> T { return *(gp<const T>) ((gp<const char>) base + off); };
    auto stb = [](gp<T> base, unsigned off, T v) { *(gp<T>) ((gp<char>) base + off) = v; };
    /* a register set holds one fetched level L of the tile: u*, v*, w* on
     * the halo tile and, two levels ahead, p(L + 2) on the pressure tile
     * (clamped to the walls), so the pressure a level stages was read a
     * full level earlier */
    struct RSet { T u[2], v[2], w[2], pn[2]; };
    auto wok_of = [&](int L) { return L >= 0 && L <= nz - 2; };
    auto clampz = [&](int L) { return L < 0 ? 0 : (L >= nz ? nz - 1 : L); };
    auto fetch = [&](int L, RSet &X) {
        const int Lc = clampz(L), L2 = clampz(L + 2);
        const bool wok = wok_of(L);
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int e = tid + 256 * t;
            if (e < RT_NA) {
                X.u[t] = ldb(Sus, (unsigned) Lc * Pb + oa[t]);
                X.v[t] = ldb(Svs, (unsigned) Lc * Pb + oa[t]);
                X.w[t] = wok ? ldb(Sws, (unsigned) L * Pb + oa[t]) : (T) 0;
            }
            if (e < RT_NP) X.pn[t] = ldb(Sp, (unsigned) L2 * Pb + op[t]);
        }
    };
    /* pressure of level L into buffer (L - k0 + 1) & 1: p(clamp(L)) and,
     * where w has a face at L, p(L + 1) */
    auto stage_p = [&](int L, const T *pa, const T *pb) {
        const int b = (L - k0 + 1) & 1;
        const bool wok = wok_of(L);
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int e = tid + 256 * t;
            if (e < RT_NP) {
                const int r = e / RT_PX, c = e - r * RT_PX;
                SP[b][0][r][c] = pa[t];
                SP[b][1][r][c] = wok ? pb[t] : (T) 0;
            }
        }
    };
    auto pdirect = [&](int L, T *out) {
        const int Lc = clampz(L);
#pragma unroll
        for (int t = 0; t < 2; t++) out[t] = tid + 256 * t < RT_NP ? ldb(Sp, (unsigned) Lc * Pb + op[t]) : (T) 0;
    };
    /* projected u, v, w of level L into ring slot (L - k0 + 1) & 3 */
    auto build = [&](int L, const RSet &X) {
        const int sl = (L - k0 + 1) & 3, b = (L - k0 + 1) & 1;
        const int Lc = L < 0 ? 0 : (L >= nz ? nz - 1 : L);
        const bool wok = wok_of(L);
        const bool ghost = L < 0 || L >= nz;
        const T hm = Sh1 ? (T) Sh1[Lc + 1] * m : (T) 0;
        const T cz = wok ? (T) (S.dtp * rdzh[L]) : (T) 0;
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int e = tid + 256 * t;
            if (e < RT_NA) {
                const int r = e / RT_HX, c = e - r * RT_HX;
                const T p0 = SP[b][0][r][c];
                const T uc = X.u[t] + hm;
                T up = uc - (SP[b][0][r][c + 1] - p0) * cx;
                T vp = X.v[t] - (SP[b][0][r + 1][c] - p0) * cy;
                if (ghost) { up = (T) 2.0 * ubs - up; vp = (T) 2.0 * vb - vp; }
                SU[sl][r][c] = up;
                SV[sl][r][c] = vp;
                SW[sl][r][c] = wok ? X.w[t] - (SP[b][1][r][c] - p0) * cz : (T) 0;
            }
        }
    };

    /* prologue: levels k0 - 1, k0, k0 + 1 built into the ring, the
     * pressure of k0 + 2 staged, set A holding level k0 + 1 (p(k0 + 3))
     * and set B level k0 + 2 (p(k0 + 4)) */
    RSet A, B;
    T carry[2], pd0[2], pd1[2];
    fetch(k0 - 1, A);
    fetch(k0, B);
    pdirect(k0 - 1, pd0);
    pdirect(k0, pd1);
    stage_p(k0 - 1, pd0, pd1);
    stage_p(k0, pd1, A.pn);                       /* A.pn = p(k0 + 1) */
    __syncthreads();
    build(k0 - 1, A);
    build(k0, B);
    carry[0] = A.pn[0]; carry[1] = A.pn[1];       /* p(k0 + 1) */
    fetch(k0 + 1, A);
    __syncthreads();                              /* buffer of k0 + 1 was read by build(k0 - 1) */
    stage_p(k0 + 1, carry, B.pn);                 /* B.pn = p(k0 + 2) */
    __syncthreads();
    build(k0 + 1, A);
    carry[0] = B.pn[0]; carry[1] = B.pn[1];       /* p(k0 + 2) */
    fetch(k0 + 2, B);
    stage_p(k0 + 2, carry, A.pn);                 /* its buffer was last read by build(k0); A.pn = p(k0 + 3) */
    __syncthreads();

    const SubCoef C = R.sc[s];
    const T nu = (T) R.nu;
    const T rdx = (T) R.rdx, rdy = (T) R.rdy;
    const T rdx2 = rdx * rdx, rdy2 = rdy * rdy;
    const T gmprev = (T) scal[R.m_prev_slot[s]];
    const T nuadt = (T) C.nuadt, gamdt = (T) C.gamdt, zetdt = (T) C.zetdt;
    const int r0 = ly + 1, c0 = lx + 1;
    double cflm = 0.0;
    /* previous tendencies of this point, one level ahead */
    const size_t qc = (size_t) j * nx + i;
    const unsigned qcb = (unsigned) (qc * sizeof(T));
    /* gu == nullptr: the substep's zeta is 0 (no previous tendency read);
     * fu == nullptr: this substep's tendency is never read (the next
     * step's first substep has zeta = 0), so it is not stored */
    const bool useg = gu != nullptr, keepf = fu != nullptr;
    T gnu = 0, gnv = 0, gnw = 0;
    if (act && useg) {
        const unsigned o = (unsigned) k0 * Pb + qcb;
        gnu = ldb(gu, o); gnv = ldb(gv, o); gnw = ldb(gw, o);
    }

    /* level k: register set Y holds level k + 2 (its pressure staged), X
     * (level k + 1, holding p(k + 3)) receives level k + 3.  Four ring slots
     * and two pressure buffers let build(k + 2) and stage(k + 3) run right
     * after the stencil of level k without a barrier between them: one
     * barrier per level. */
    auto level = [&](const int k, RSet &X, RSet &Y) {
        const bool more = k + 2 <= k1, more3 = k + 3 <= k1;
        T pk3[2] = { X.pn[0], X.pn[1] };          /* p(k + 3) */
        const T gcu = gnu, gcv = gnv, gcw = gnw;
        if (act && useg && k + 1 < k1) {
            const unsigned o = (unsigned) (k + 1) * Pb + qcb;
            gnu = ldb(gu, o); gnv = ldb(gv, o); gnw = ldb(gw, o);
        }
        if (more3) fetch(k + 3, X);
        const int sm = (k - k0) & 3, s0 = (k - k0 + 1) & 3, sp = (k - k0 + 2) & 3;
        if (act) {
            const unsigned qb = (unsigned) k * Pb + qcb;
            const T dzk = (T) dz[k], dzkp = (T) dz[k + 1];
            const T rzk = (T) rdz[k], rzkp = (T) rdz[k + 1];
            const T rzp = (T) rdzh[k], rzm = (T) rdzh[k - 1];
            T Uk = 0, Ukp = 0, Ukm = 0, Uh = 0;    /* U at levels k, k + 1, k - 1; U's right-hand-side share */
            if (OFF && uof) {
                Uk = (T) uof[k + 1]; Ukp = (T) uof[k + 2]; Ukm = (T) uof[k];
                Uh = (T) uof[(size_t) nz + 2 + (size_t) s * nz + k];
            }
#define LU(sl, dr, dc) SU[sl][r0 + (dr)][c0 + (dc)]
#define LV(sl, dr, dc) SV[sl][r0 + (dr)][c0 + (dc)]
#define LW(sl, dr, dc) SW[sl][r0 + (dr)][c0 + (dc)]
/* keeps each component's LDS reads next to their use (register pressure) */
#define SEQ() asm volatile("" ::: "memory")
            if (cflpart) {
                double c;
                if (cflm < (c = fabs(F(LU(s0, 0, 0), Uk)) * cfldt * rdx)) cflm = c;
                if (cflm < (c = fabs(LV(s0, 0, 0)) * cfldt * rdy)) cflm = c;
                if (cflm < (c = fabs(LW(s0, 0, 0)) * cfldt * rzp)) cflm = c;
            }
            {   /* u */
                const T u0 = LU(s0, 0, 0), uip = LU(s0, 0, 1), uim = LU(s0, 0, -1);
                const T ujp = LU(s0, 1, 0), ujm = LU(s0, -1, 0), ukp = LU(sp, 0, 0), ukm = LU(sm, 0, 0);
                const T v0 = LV(s0, 0, 0), vip = LV(s0, 0, 1), vjm = LV(s0, -1, 0), vipjm = LV(s0, -1, 1);
                const T w0 = LW(s0, 0, 0), wip = LW(s0, 0, 1), wkm = LW(sm, 0, 0), wipkm = LW(sm, 0, 1);
                const T u0f = F(u0, Uk), uipf = F(uip, Uk), uimf = F(uim, Uk);
                const T ujpf = F(ujp, Uk), ujmf = F(ujm, Uk), ukpf = F(ukp, Ukp), ukmf = F(ukm, Ukm);
                /* the face mass fluxes (dy dz, dx dz, dx dy times the face
                 * velocity) times the face value of u, over the cell volume:
                 * the cell-size factors cancel to 1/dx, 1/dy, 1/dz */
                const T sxp = u0f + uipf, sxm = u0f + uimf;
                T adv = -(T) 0.25 * ((sxp * sxp - sxm * sxm) * rdx
                                     + ((v0 + vip) * (u0f + ujpf) - (vjm + vipjm) * (u0f + ujmf)) * rdy
                                     + ((w0 + wip) * (u0f + ukpf) - (wkm + wipkm) * (u0f + ukmf)) * rzk);
                T f = adv + nu * (((uip - u0) - (u0 - uim)) * rdx2 + ((ujp - u0) - (u0 - ujm)) * rdy2);
                T h = u0 + nuadt * (((ukp - u0) * rzp - (u0 - ukm) * rzm) * rzk);
                if constexpr (OFF) h += Uh;
                T g = gcu + gmprev;
                h += gamdt * f + zetdt * g;
                if (k == 0) h -= (T) R.lift[s][0];
                if (k == nz - 1) h -= (T) R.lift[s][1];
                if (keepf) stb(fu, qb, f);
                stb(ru, qb, h * dzk);
            }
            SEQ();
            {   /* v */
                const T u0 = LU(s0, 0, 0), ujp = LU(s0, 1, 0), uim = LU(s0, 0, -1), uimjp = LU(s0, 1, -1);
                const T v0 = LV(s0, 0, 0), vip = LV(s0, 0, 1), vim = LV(s0, 0, -1);
                const T vjp = LV(s0, 1, 0), vjm = LV(s0, -1, 0), vkp = LV(sp, 0, 0), vkm = LV(sm, 0, 0);
                const T w0 = LW(s0, 0, 0), wjp = LW(s0, 1, 0), wkm = LW(sm, 0, 0), wjpkm = LW(sm, 1, 0);
                const T syp = v0 + vjp, sym = vjm + v0;
                T adv = -(T) 0.25 * (((F(u0, Uk) + F(ujp, Uk)) * (v0 + vip) - (F(uim, Uk) + F(uimjp, Uk)) * (v0 + vim)) * rdx
                                     + (syp * syp - sym * sym) * rdy
                                     + ((w0 + wjp) * (v0 + vkp) - (wkm + wjpkm) * (v0 + vkm)) * rzk);
                T f = adv + nu * (((vip - v0) - (v0 - vim)) * rdx2 + ((vjp - v0) - (v0 - vjm)) * rdy2);
                T h = v0 + nuadt * (((vkp - v0) * rzp - (v0 - vkm) * rzm) * rzk);
                h += gamdt * f + zetdt * gcv;
                if (k == 0) h -= (T) R.lift[s][2];
                if (k == nz - 1) h -= (T) R.lift[s][3];
                if (keepf) stb(fv, qb, f);
                stb(rv, qb, h * dzk);
            }
            SEQ();
            if (k < nz - 1) {   /* w on the interior faces */
                const T u0 = LU(s0, 0, 0), uim = LU(s0, 0, -1), ukp = LU(sp, 0, 0), uimkp = LU(sp, 0, -1);
                const T v0 = LV(s0, 0, 0), vjm = LV(s0, -1, 0), vkp = LV(sp, 0, 0), vjmkp = LV(sp, -1, 0);
                const T w0 = LW(s0, 0, 0), wip = LW(s0, 0, 1), wim = LW(s0, 0, -1);
                const T wjp = LW(s0, 1, 0), wjm = LW(s0, -1, 0), wkp = LW(sp, 0, 0), wkm = LW(sm, 0, 0);
                const T dzp1h = (T) 0.5 * (dz[k + 1] + dz[k]);
                const T szp = w0 + wkp, szm = wkm + w0;
                T adv = -(T) 0.25 * rzp * (((dzk * F(u0, Uk) + dzkp * F(ukp, Ukp)) * (w0 + wip)
                                            - (dzk * F(uim, Uk) + dzkp * F(uimkp, Ukp)) * (w0 + wim)) * rdx
                                           + ((dzk * v0 + dzkp * vkp) * (w0 + wjp) - (dzk * vjm + dzkp * vjmkp) * (w0 + wjm)) * rdy
                                           + (szp * szp - szm * szm));
                T f = adv + nu * (((wip - w0) - (w0 - wim)) * rdx2 + ((wjp - w0) - (w0 - wjm)) * rdy2);
                T h = w0 + nuadt * (((wkp - w0) * rzkp - (w0 - wkm) * rzk) * rzp);
                h += gamdt * f + zetdt * gcw;
                if (keepf) stb(fw, qb, f);
                stb(rw, qb, h * dzp1h);
            } else {
                if (keepf) stb(fw, qb, (T) 0);
                stb(rw, qb, (T) 0);
            }
#undef SEQ
#undef LU
#undef LV
#undef LW
        }
        if (more) build(k + 2, Y);
        if (more3) stage_p(k + 3, pk3, Y.pn);     /* Y.pn = p(k + 4) */
        __syncthreads();
    };
    for (int k = k0; k < k1; k += 2) {
        level(k, A, B);
        if (k + 1 < k1) level(k + 1, B, A);
    }
    if (cflpart) {
        red[tid] = cflm;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
            __syncthreads();
        }
        if (tid == 0) cflpart[lb] = red[0];
    }
}

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CF_RHS_WAVES, 8))) void k_rhs2t(const RunDev R, const int s, const ProjSrcT<T> S,
    const T *__restrict__ gu, const T *__restrict__ gv, const T *__restrict__ gw,
    const double *__restrict__ scal,
    T *__restrict__ fu, T *__restrict__ fv, T *__restrict__ fw,
    T *__restrict__ ru, T *__restrict__ rv, T *__restrict__ rw,
    double *__restrict__ cflpart, double cfldt, int kc)
{
    rhs2t_body<T>(R, s, S, gu, gv, gw, scal, fu, fv, fw, ru, rv, rw, cflpart, cfldt, kc, xcd_block(blockIdx.x, gridDim.x));
}
template __global__ void k_rhs2t<double>(const RunDev, const int, const ProjSrcT<double>, const double *, const double *,
    const double *, const double *, double *, double *, double *, double *, double *, double *, double *, double, int);
template __global__ void k_rhs2t<float>(const RunDev, const int, const ProjSrcT<float>, const float *, const float *,
    const float *, const double *, float *, float *, float *, float *, float *, float *, double *, double, int);

/* ---------------------------------------------------------------------
 * Batched constant-coefficient Thomas solve, one thread per (column,
 * field), in place: the right-hand side r becomes the solution.  The
 * factors come from the wall-reduced system of solve_helmholtz1d_u2 /
 * _w2 (diffuseu.c:173-219, diffusew.c:175-223 with tridag.c:4-26), built
 * on the host once per substep since they do not depend on the column.
 * Fields: 0 u (nz rows), 1 v (nz rows), 2 w (nz-1 rows; top wall := 0),
 * 3 the scalar c (nz rows, diffusivity nu / Pr).
 * ------------------------------------------------------------------- */
template <typename T>
static __device__ __forceinline__ void thomas_body(const RunDev &R, const int s,
    T *__restrict__ ru, T *__restrict__ rv, T *__restrict__ rw, T *__restrict__ rc, const Blk b)
{
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t P = (size_t) nx * ny;
    const size_t col = (size_t) b.x * blockDim.x + threadIdx.x;
    const int f = b.y;
    if (col >= P || f > 3) return;
    T *x = f == 0 ? ru : (f == 1 ? rv : (f == 2 ? rw : rc));
    const int type = f == 2 ? 2 : (f == 3 ? 1 : 0);
    const int n = f == 2 ? nz - 1 : nz;
    const double *a = R.th + ((size_t) (s * 4 + type) * 3 + 0) * nz;
    const double *ib = R.th + ((size_t) (s * 4 + type) * 3 + 1) * nz;
    const double *gm = R.th + ((size_t) (s * 4 + type) * 3 + 2) * nz;
    T *xc = x + col;
    T d = xc[0] * (T) ib[0];
    xc[0] = d;
    int k = 1;
    for (; k + 8 <= n; k += 8) {
        T rr[8];
#pragma unroll
        for (int t = 0; t < 8; t++) rr[t] = xc[(size_t) (k + t) * P];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            d = (rr[t] - (T) a[k + t] * d) * (T) ib[k + t];
            xc[(size_t) (k + t) * P] = d;
        }
    }
    for (; k < n; k++) {
        d = (xc[(size_t) k * P] - (T) a[k] * d) * (T) ib[k];
        xc[(size_t) k * P] = d;
    }
    k = n - 2;
    for (; k - 7 >= 0; k -= 8) {
        T rr[8];
#pragma unroll
        for (int t = 0; t < 8; t++) rr[t] = xc[(size_t) (k - t) * P];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            d = rr[t] - (T) gm[k - t + 1] * d;
            xc[(size_t) (k - t) * P] = d;
        }
    }
    for (; k >= 0; k--) {
        d = xc[(size_t) k * P] - (T) gm[k + 1] * d;
        xc[(size_t) k * P] = d;
    }
    if (f == 2) xc[(size_t) (nz - 1) * P] = (T) 0;   /* ghost_fill_top2, diffusew.c:683-688 */
}

template <typename T>
__global__ __launch_bounds__(256) void k_thomas(const RunDev R, const int s,
    T *__restrict__ ru, T *__restrict__ rv, T *__restrict__ rw, T *__restrict__ rc)
{
    thomas_body<T>(R, s, ru, rv, rw, rc, blk_native());
}

template __global__ void k_thomas<double>(const RunDev, const int, double *, double *, double *, double *);
template __global__ void k_thomas<float>(const RunDev, const int, float *, float *, float *, float *);

/* ---------------------------------------------------------------------
 * The same constant-coefficient tridiagonal solves as k_thomas, as a
 * partitioned ("spike") solve held in registers: a workgroup owns 64
 * columns of one field and S waves, wave c owning rows [c L, c L + L).
 * Each lane solves its chunk with the chunk-local Thomas factors (y),
 * publishes the chunk's first and last values, and recovers the two
 * neighbouring interface unknowns from the precomputed inverse of the
 * 2S x 2S interface system; x = y - E_{c-1} v - F_{c+1} w with the
 * column-independent spike vectors v, w.  Every value is read once and
 * written once (Thomas in global memory reads and writes it twice).
 * Same system as solve_helmholtz1d_u2 / _w2 (diffuseu.c:173-219,
 * diffusew.c:175-223); differs from sequential Thomas by round-off.
 * ------------------------------------------------------------------- */
/* plane != null (field 0 only): also write, per level, the sum of the
 * solved u* over the workgroup's 64 columns to plane[level * b.nx + b.x],
 * the first stage of get_xy_avg (get.c:281-307) */
template <typename T, int LM>
static __device__ __forceinline__ void spike_body(const RunDev &R, const int s,
    T *__restrict__ ru, T *__restrict__ rv, T *__restrict__ rw, T *__restrict__ rc, const Blk b,
    double *__restrict__ plane = nullptr)
{
    __shared__ T itf[2 * CF_SPIKE_SMAX][64];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t P = (size_t) nx * ny;
    const int f = b.y;
    if (f > 3) return;
    const int type = f == 2 ? 2 : (f == 3 ? 1 : 0);
    const SpikeTab ST = R.sp[s][type];
    const int lane = threadIdx.x & 63;
    /* wave-uniform chunk index, so the table reads below are scalar loads;
     * the tables hold LM rows per chunk, identity past its end,
     * and every read of the column is at a valid (clamped) row, so the
     * sweeps are straight-line code whose table reads can all be hoisted */
    const int chunk = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool live = chunk < ST.S;
    const size_t col = (size_t) b.x * 64 + lane;
    const bool act = col < P;
    const gp<T> x = (gp<T>) (f == 0 ? ru : (f == 1 ? rv : (f == 2 ? rw : rc)));
    const gp<double> pl = (gp<double>) plane;
    const int lo = chunk * ST.L;
    const int Lc = live ? min(ST.L, ST.n - lo) : 0;     /* rows of this chunk */
    const gp<T> xc = x + (act ? col : P - 1);
    const size_t tb = (size_t) chunk * LM;
    const cp<double> ta = (cp<double>) ST.a + tb, tib = (cp<double>) ST.ib + tb;
    const cp<double> tgam = (cp<double>) ST.gam + tb, tvl = (cp<double>) ST.vl + tb;
    const cp<double> twr = (cp<double>) ST.wr + tb, tminv = (cp<double>) ST.minv;
    T y[LM];
    if (live) {                                   /* wave-uniform */
#pragma unroll
        for (int t = 0; t < LM; t++) y[t] = xc[(size_t) (lo + min(t, Lc - 1)) * P];
#pragma unroll
        for (int t = 0; t < LM; t++) y[t] = (act && t < Lc) ? y[t] : (T) 0;
        T d = 0;
#pragma unroll
        for (int t = 0; t < LM; t++) {
            d = (y[t] - (T) ta[t] * d) * (T) tib[t];
            y[t] = d;
        }
        T yl = 0;
        d = 0;
#pragma unroll
        for (int t = LM - 1; t >= 0; t--) {
            d = t < LM - 1 ? y[t] - (T) tgam[t + 1] * d : y[t];
            yl = t == Lc - 1 ? d : yl;
            y[t] = d;
        }
        itf[2 * chunk][lane] = y[0];
        itf[2 * chunk + 1][lane] = yl;
    }
    __syncthreads();
    if (!live) return;                            /* wave-uniform */
    const bool sums = plane != nullptr && f == 0;
    if (!act && !sums) return;
    const int m = 2 * ST.S;
    constexpr int MP = 2 * CF_SPIKE_SMAX;
    T E = 0, F = 0;
    if (chunk > 0) {
        const cp<double> row = tminv + (size_t) (2 * (chunk - 1) + 1) * MP;
#pragma unroll
        for (int q = 0; q < MP; q++) E += (T) row[q] * (q < m ? itf[q][lane] : (T) 0);
    }
    if (chunk < ST.S - 1) {
        const cp<double> row = tminv + (size_t) (2 * (chunk + 1)) * MP;
#pragma unroll
        for (int q = 0; q < MP; q++) F += (T) row[q] * (q < m ? itf[q][lane] : (T) 0);
    }
    double sv[LM];
#pragma unroll
    for (int t = 0; t < LM; t++) {
        sv[t] = 0.0;
        if (t < Lc) {                             /* wave-uniform */
            const T val = y[t] - E * (T) tvl[t] - F * (T) twr[t];
            if (act) xc[(size_t) (lo + t) * P] = val;
            sv[t] = act ? (double) val : 0.0;
        }
    }
    if (sums) {                                   /* wave-uniform; every lane takes part */
        const double tot = wave_sums<LM>(sv);
        const int t = lane >> 1;
        if ((lane & 1) == 0 && t < Lc) pl[(size_t) (lo + t) * b.nx + b.x] = tot;
    }
    if (act && f == 2 && chunk == ST.S - 1) xc[(size_t) (nz - 1) * P] = (T) 0;   /* top wall face */
}

template <typename T, int LM>
__global__ __launch_bounds__(512) void k_spike(const RunDev R, const int s,
    T *__restrict__ ru, T *__restrict__ rv, T *__restrict__ rw, T *__restrict__ rc)
{
    spike_body<T, LM>(R, s, ru, rv, rw, rc, blk_native());
}

/* the fast path's solve: k_spike with the plane sums of u* fused in
 * (accumulated in FP64 whatever the storage) */
#ifndef CF_SPIKE_WPE
#define CF_SPIKE_WPE 1
#endif
template <typename T, int LM>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(CF_SPIKE_WPE, 8))) void k_spike_sums(const RunDev R, const int s, T *__restrict__ ru,
    T *__restrict__ rv, T *__restrict__ rw, double *__restrict__ plane)
{
    spike_body<T, LM>(R, s, ru, rv, rw, nullptr, blk_native(), plane);
}

template <typename T>
k_spike_t<T> pick_spike(int L)
{
    switch (L) {
#define X_(N) case N: return k_spike<T, N>;
    CF_SPIKE_LSET(X_)
#undef X_
    default: return nullptr;
    }
}
template k_spike_t<double> pick_spike<double>(int);
template k_spike_t<float> pick_spike<float>(int);

template <typename T>
k_spike_sums_t<T> pick_spike_sums(int L)
{
    switch (L) {
#define X_(N) case N: return k_spike_sums<T, N>;
    CF_SPIKE_LSET(X_)
#undef X_
    default: return nullptr;
    }
}
template k_spike_sums_t<double> pick_spike_sums<double>(int);
template k_spike_sums_t<float> pick_spike_sums<float>(int);

/* second stage of the plane sums: level blockIdx.x, nbx partials in a
 * fixed tree */
static __device__ __forceinline__ void level_sum_body(const double *__restrict__ plane, int nbx,
                                                      double *__restrict__ lsum, const int l)
{
    __shared__ double red[4];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nbx; q += 256) acc += plane[(size_t) l * nbx + q];
    const double tot = block_sum256(acc, red);
    if (threadIdx.x == 0) lsum[l] = tot;
}

__global__ __launch_bounds__(256) void k_level_sum(const double *__restrict__ plane, int nbx,
                                                   double *__restrict__ lsum)
{
    level_sum_body(plane, nbx, lsum, blockIdx.x);
}




/* ---------------------------------------------------------------------
 * Mass-flux constraint (chanfast.c:485-530): plane averages -> bulk
 * velocity (get_z_avg, get.c:490-502), wall shear, centreline; forcing
 * mdpdx = (Ub - uframe - U0) / U1 with U1 from the column-independent
 * unit-forcing solution h1 (host-precomputed per substep).  Input: the
 * per-level plane sums of the solved field (k_spike_sums, k_level_sum);
 * the wall-ghost averages are 2 bc - the adjacent level (apply_bc_u2).
 * Writes the forcing into scal[slot]; on the last substep writes the
 * pstat record.  One block of 256 threads.
 * ------------------------------------------------------------------- */
static __device__ __forceinline__ void mass_finalize_body(const RunDev &R, const int s, const double *__restrict__ lsum,
                                                          double *__restrict__ scal, double *__restrict__ logbuf, int kind)
{
    extern __shared__ __align__(16) unsigned char smem_[];
    double *avg = (double *) smem_;               /* nz + 2 */
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const double *dz = R.dzt + CF_NGT;
    const double scale = 1.0 / ((double) nx * (double) ny);
    /* volume integral and height: per-thread partials over k = t, t + 256,
     * ..., then a fixed tree (deterministic, the same in every launch) */
    __shared__ double rv[4], rl[4];
    double pv = 0.0, pl = 0.0;
    const double *uo = kind == 0 && R.uoff ? R.uoff + 1 : nullptr;   /* stored u is u - U(z) */
    for (int l = threadIdx.x; l < nz; l += 256) {
        const double a = uo ? lsum[l] * scale + uo[l] : lsum[l] * scale;
        avg[l + 1] = a;
        pv += a * dz[l];
        pl += dz[l];
    }
    const double volsum = block_sum256(pv, rv), hsum = block_sum256(pl, rl);
    if (threadIdx.x != 0) return;
    double *A = avg + 1;                          /* A[-1 .. nz] */
    const double bc = kind == 0 ? R.ubc : R.cbot;
    A[-1] = 2.0 * bc - A[0];                      /* apply_bc_u2 ghosts of the solved field */
    A[nz] = 2.0 * (kind == 0 ? R.ubc : R.ctop) - A[nz - 1];
    double vol = volsum / hsum;
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

/* The second stage of the plane sums and the mass-flux constraint in one
 * launch: each workgroup sums its level (level_sum_body), and the last
 * workgroup to finish -- counted with a device-scope ticket after a
 * release fence -- runs mass_finalize_body on all levels and resets the
 * ticket for the next launch.  The sums are in the same fixed order as the
 * two-launch form; only which workgroup finishes is a race, and it does
 * not enter the arithmetic.  step_end: also count the completed step
 * (the last substep). */
__global__ __launch_bounds__(256) void k_level_mass(const RunDev R, const int s, const double *__restrict__ plane,
                                                    int nbx, double *__restrict__ lsum, double *__restrict__ scal,
                                                    double *__restrict__ logbuf, int step_end)
{
    level_sum_body(plane, nbx, lsum, blockIdx.x);
    __shared__ int last;
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned long long *ticket = (unsigned long long *) (scal + SC_TICKET);
        last = atomicAdd(ticket, 1ull) == (unsigned long long) (gridDim.x - 1);
        __threadfence();
    }
    __syncthreads();
    if (!last) return;
    mass_finalize_body(R, s, lsum, scal, logbuf, 0);
    if (threadIdx.x == 0) {
        *(unsigned long long *) (scal + SC_TICKET) = 0ull;
        if (step_end) scal[SC_STEP] += 1.0;
    }
}

__global__ __launch_bounds__(256) void k_mass_finalize(const RunDev R, const int s, const double *__restrict__ lsum,
                                                       double *__restrict__ scal, double *__restrict__ logbuf, int kind)
{
    mass_finalize_body(R, s, lsum, scal, logbuf, kind);
}

/* ---------------------------------------------------------------------
 * CFL at the start of a step (get_cfl, get.c:390-417): per-block max,
 * then a one-block finish that writes the cfl.dat record and bumps the
 * step counter at the end of the step (k_step_end).
 * ------------------------------------------------------------------- */
static __device__ __forceinline__ void cfl_body(const RunDev &R, const double *__restrict__ part, int nb,
                                                    double dt, double *__restrict__ scal, double *__restrict__ logbuf, const Blk b)
{
    __shared__ double red[4];
    double m = 0.0;                               /* the block maxima are >= 0 */
#pragma unroll 4
    for (int b = threadIdx.x; b < nb; b += 256) m = fmax(m, part[b]);
    const double mx = block_max256_nonneg(m, red);
    if (threadIdx.x != 0) return;
    long idx = (long) scal[SC_STEP] % R.logcap;
    double *lg = logbuf + (size_t) idx * 13;
    for (int t = 0; t < 13; t++) lg[t] = 0.0;
    lg[0] = scal[SC_STEP] + 1.0;
    lg[1] = dt;
    lg[2] = R.cfl_log > 0.0 ? R.cfl_log : mx;
    scal[SC_CFL] = lg[2];
}

__global__ __launch_bounds__(256) void k_cfl_finish(const RunDev R, const double *__restrict__ part, int nb,
                                                    double dt, double *__restrict__ scal, double *__restrict__ logbuf)
{
    cfl_body(R, part, nb, dt, scal, logbuf, blk_native());
}

/* cflmode 1 (chanfast.c:385-391): get_cfl (get.c:390-417) of the plain
 * fields (u, v, w) with the current dt, before the step that rescales dt;
 * one workgroup of 64 x 4 per (x tile, y tile, level), its max into
 * part[block] */
template <typename T>
__global__ __launch_bounds__(256) void k_cfl_plain(const RunDev R, const T *__restrict__ u, const T *__restrict__ v,
                                                   const T *__restrict__ w, double dt, double *__restrict__ part)
{
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    double m = 0.0;
    if (i < nx && j < ny && k < nz) {
        const size_t q = (size_t) k * nx * ny + (size_t) j * nx + i;
        const double *dz = R.dzt + CF_NGT;
        const double dzph = 0.5 * (dz[k] + dz[k + 1]);
        double c;
        const double uq = R.uoff ? (double) u[q] + R.uoff[k + 1] : (double) u[q];
        if (m < (c = fabs(uq) * dt / R.g.dx)) m = c;
        if (m < (c = fabs((double) v[q]) * dt / R.g.dy)) m = c;
        if (m < (c = fabs((double) w[q]) * dt / dzph)) m = c;
    }
    __shared__ double red[256];
    const int tid = threadIdx.y * 64 + threadIdx.x;
    red[tid] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] = fmax(red[tid], red[tid + s]);
        __syncthreads();
    }
    if (tid == 0) part[((size_t) blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}
template __global__ void k_cfl_plain<double>(const RunDev, const double *, const double *, const double *, double,
                                             double *);
template __global__ void k_cfl_plain<float>(const RunDev, const float *, const float *, const float *, double,
                                            double *);

/* the max of nb partials into scal[SC_CFL0] (a fixed tree) */
__global__ __launch_bounds__(256) void k_cfl_reduce(const double *__restrict__ part, int nb, double *__restrict__ scal)
{
    __shared__ double red[256];
    double m = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) m = fmax(m, part[b]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) scal[SC_CFL0] = red[0];
}

__global__ void k_step_end(double *__restrict__ scal)
{
    if (threadIdx.x == 0) scal[SC_STEP] += 1.0;
}

/* ---------------------------------------------------------------------
 * FFT passes.  Spectral layout X[nz][nx/2+1][ny] (ky fastest), so y lines
 * are contiguous.  x passes pack two real lines (j, j+1 at one level)
 * into one complex transform.  NX / NY > 0 selects a compile-time plan,
 * 0 the runtime one.  T is the storage and arithmetic precision; C its
 * complex type; the twiddles come from host libm in FP64 (rounded for T
 * = float).
 * ------------------------------------------------------------------- */

template <typename C> static __device__ __forceinline__ const C *twsel(const double2 *d, const float2 *f);
template <> __device__ __forceinline__ const double2 *twsel<double2>(const double2 *d, const float2 *) { return d; }
template <> __device__ __forceinline__ const float2 *twsel<float2>(const double2 *, const float2 *f) { return f; }

/* twiddles of an n-point pass into LDS after the line data (tws), in the
 * order the plan reads them (tw_src): the first one per thread is read
 * into a register with the data gather (tw_first) and stored after it
 * (tw_rest stores it and copies any others), so the passes read twiddles
 * from LDS without an extra round trip to memory */
template <int N, typename C>
static __device__ __forceinline__ C tw_first(cp<cv<C>> tw, int n)
{
    const int t = (int) threadIdx.x, cnt = tw_count<N>(n);
    return cx_of<C>(tw[tw_src<N>(t < cnt ? t : cnt - 1)]);
}
template <int N, typename C>
static __device__ __forceinline__ void tw_rest(C *tws, C first, cp<cv<C>> tw, int n)
{
    const int t = (int) threadIdx.x, cnt = tw_count<N>(n);
    if (t < cnt) tws[t] = first;
    for (int q = t + (int) blockDim.x; q < cnt; q += (int) blockDim.x) tws[q] = cx_of<C>(tw[tw_src<N>(q)]);
}

#define CF_FX_UNR 3            /* elements per thread gathered together by the divergence + x-FFT */
#define CF_FXI_UNR 5           /* ... and by the inverse x-FFT (nl (nx/2 + 1) <= 1152 for nx >= 16) */

/* unpack two packed real-line spectra and store kx-major: for each kx the
 * block's 2 nl lines are consecutive ky positions (coalesced) */
/* the half spectra of a workgroup's line pairs to the ky-fastest layout:
 * lane (kx, j) writes one value, so the 2 nl consecutive lines of a kx
 * go out as one contiguous run of 16 (or 8) bytes per lane */
template <typename C>
static __device__ __forceinline__ void store_half_spectra(const C *sb, int ld, int nl, int nx, int nkx,
                                                          int ny, long NL, long pair0, double scd, gp<cv<C>> X,
                                                          bool swz = true)
{
    typedef typename Real<C>::type F;
    const F sc = (F) scd;
    const bool on = swz && fsw_on(nx);
    const int nj = 2 * nl;
    const double rny = 1.0 / (double) ny;
    for (int e = threadIdx.x; e < nj * nkx; e += blockDim.x) {
        const int kx = e / nj, jj = e - kx * nj;
        const int l = jj >> 1;
        const bool second = (jj & 1) != 0;
        const long line = 2 * (pair0 + l) + (second ? 1 : 0);
        if (line >= NL) continue;
        const C Z = sb[l * ld + fsw(kx, on)], Wc = sb[l * ld + fsw(kx == 0 ? 0 : nx - kx, on)];
        const int k = qdiv((int) line, ny, rny), j = (int) line - k * ny;
        const C val = second ? mkc<C>((Z.y + Wc.y) * sc, (Wc.x - Z.x) * sc) : mkc<C>((Z.x + Wc.x) * sc, (Z.y - Wc.y) * sc);
        X[((size_t) k * nkx + kx) * ny + j] = cv_of(val);
    }
}

template <int N, typename C>
static __device__ __forceinline__ void dft_regs(C *a, double sgd);

/* twiddles w^(m k1) of a two-pass length n = P1 P2 as a P1 x P2 table (k1
 * rows) after a workgroup's lines */
template <int N, typename C>
static __device__ __forceinline__ void tw2_stage(C *tw, cp<cv<C>> twg)
{
    constexpr int P2 = N / fft2_p1(N);
    for (int q = (int) threadIdx.x; q < N; q += (int) blockDim.x) {
        const int k1 = q / P2, m = q - k1 * P2;
        tw[q] = cx_of<C>(twg[k1 * m]);
    }
}

/* pass 1 of a two-pass transform on registers a[0..P1) (column n2 of a
 * line, natural order in): DFT-P1, twiddle w^(n2 k1) (conjugated for sg
 * = -1), row k1 of the line's LDS image at row[k1 (P2 + 1)] */
template <int P1, int P2, typename C>
static __device__ __forceinline__ void pass1_rows(C *a, C *row, const C *tw, int n2, double sgd)
{
    typedef typename Real<C>::type F;
    dft_regs<P1>(a, sgd);
#pragma unroll
    for (int k1 = 0; k1 < P1; k1++) {
        C v = a[k1];
        if (k1 > 0) {
            C w = tw[k1 * P2 + n2];
            w.y *= (F) sgd;
            v = cmul(v, w);
        }
        row[k1 * (P2 + 1)] = v;
    }
}

/* the order-2 divergence + forward x FFT of lengths NX = P1 P2 in two
 * register passes (ny even: a workgroup's line pairs are rows j, j + 1 of
 * one level).  Pass 1: the lane of (pair, column n2) forms the divergence
 * of both rows at x = P2 n1 + n2, n1 < P1, straight from memory (the x - 1
 * neighbours are reloads that hit the cache), takes the DFT-P1 and writes
 * the pair's LDS rows; pass 2 (pair, k1) takes the DFT-P2 of column k1 and
 * writes the transform in natural order over the rows, which
 * store_half_spectra unpacks.  Same sums as divx_body's paired path. */
template <int NX, typename T>
static __device__ __forceinline__ void divx2_body(const RunDev &R, const T *__restrict__ us_,
    const T *__restrict__ vs_, const T *__restrict__ ws_, const double *__restrict__ h1_,
    const double *__restrict__ mptr_, typename Cplx<T>::type *__restrict__ X_, const Blk b)
{
    typedef typename Cplx<T>::type C;
    typedef typename std::conditional<sizeof(T) == 4, double, T>::type D;
    constexpr int P1 = fx2_p1(NX), P2 = NX / P1, LS = fx_ld(NX, sizeof(T)), nl = fx_pairs_of(NX, sizeof(T));
    const gp<const T> us = (gp<const T>) us_, vs = (gp<const T>) vs_, ws = (gp<const T>) ws_;
    const cp<double> h1 = (cp<double>) h1_, mptr = (cp<double>) mptr_;
    const gp<cv<C>> X = (gp<cv<C>>) X_;
    extern __shared__ __align__(16) unsigned char smem_[];
    C *sb = (C *) smem_;
    C *tw = sb + nl * LS;
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const int NL = ny * nz;
    const size_t P = (size_t) NX * ny;
    const double rny = 1.0 / (double) ny;
    const cp<double> rdz = (cp<double>) R.rdz + CF_NGT;
    const T m = mptr ? (T) *mptr : (T) 0;
    const D rdx = (D) (T) R.rdx, rdy = (D) (T) R.rdy;
    const int pair0 = (int) xcd_block(b.x, b.nx) * nl;
    const int t1 = (int) threadIdx.x / P2, n2 = (int) threadIdx.x - t1 * P2;
    C a[P1];
    if (t1 < nl) {
        const int la = 2 * (pair0 + t1);
        const bool ok = la < NL;
        const int lc = ok ? la : NL - 2;
        const int k = qdiv(lc, ny, rny), j = lc - k * ny;
        const int jm = j > 0 ? j - 1 : ny - 1;
        const size_t qa = (size_t) k * P + (size_t) j * NX, qb = qa + NX, qm = (size_t) k * P + (size_t) jm * NX;
        const size_t qam = k > 0 ? qa - P : qa, qbm = k > 0 ? qb - P : qb;
        const D hm = h1 ? (D) ((T) h1[k + 1] * m) : (D) 0;
        const D rz = (D) ((T) rdz[k]);
        const bool top = k < nz - 1, bot = k > 0;
#pragma unroll
        for (int n1 = 0; n1 < P1; n1++) {
            const int x = P2 * n1 + n2, xm = x > 0 ? x - 1 : NX - 1;
            const D ua = (D) us[qa + x] + hm, ub = (D) us[qb + x] + hm;
            const D uam = (D) us[qa + xm] + hm, ubm = (D) us[qb + xm] + hm;
            const D va = vs[qa + x], vb = vs[qb + x], vam = vs[qm + x];
            const D war = ws[qa + x], wbr = ws[qb + x], wamr = ws[qam + x], wbmr = ws[qbm + x];
            const D wa = top ? war : (D) 0, wb = top ? wbr : (D) 0;
            const D wam = bot ? wamr : (D) 0, wbm = bot ? wbmr : (D) 0;
            const D da = (ua - uam) * rdx + (va - vam) * rdy + (wa - wam) * rz;
            const D db = (ub - ubm) * rdx + (vb - va) * rdy + (wb - wbm) * rz;
            a[n1] = mkc<C>(ok ? (T) da : (T) 0, ok ? (T) db : (T) 0);
            /* at most four columns' loads in flight: a compiler memory
             * barrier keeps the next group's loads below this one */
            if ((n1 & 3) == 3) __asm__ volatile("" ::: "memory");
        }
    }
    tw2_stage<NX, C>(tw, (cp<cv<C>>) twsel<C>(R.twx, R.twx32));
    __syncthreads();
    if (t1 < nl) pass1_rows<P1, P2>(a, sb + t1 * LS + n2, tw, n2, 1.0);
    __syncthreads();
    const int t2 = (int) threadIdx.x / P1, k1 = (int) threadIdx.x - t2 * P1;
    C u[P2];
    if (t2 < nl) {
        const C *col = sb + t2 * LS + k1 * (P2 + 1);
#pragma unroll
        for (int q = 0; q < P2; q++) u[q] = col[q];
        dft_regs<P2>(u, 1.0);
    }
    __syncthreads();
    if (t2 < nl) {
#pragma unroll
        for (int k2 = 0; k2 < P2; k2++) sb[t2 * LS + k1 + P1 * k2] = u[k2];
    }
    __syncthreads();
    store_half_spectra<C>(sb, LS, nl, NX, nkx, ny, (long) NL, (long) pair0, 0.5 / (double) NX, X, false);
}

/* divergence2 / divergence4 (pressure.c:40-118) of the mass-flux-corrected
 * velocity (u + h1 mdpdx, chanfast.c:519-522) fused with the forward r2c x
 * FFT scaled 1/nx (fft.c:161-177).  Order 4 sees the w ghost faces
 * mirrored oddly about the (zero) wall faces, as divergence4 sets them. */
template <int NX, int ORD, typename T, int UNRX = 0>
static __device__ __forceinline__ void divx_body(const RunDev &R, const T *__restrict__ us_,
    const T *__restrict__ vs_, const T *__restrict__ ws_, const double *__restrict__ h1_,
    const double *__restrict__ mptr_, typename Cplx<T>::type *__restrict__ X_, const Blk b)
{
    if constexpr (NX > 0 && ORD == 2 && divx2_on(NX, sizeof(T))) {
        if ((R.g.ny & 1) == 0) {                  /* workgroup-uniform */
            divx2_body<NX, T>(R, us_, vs_, ws_, h1_, mptr_, X_, b);
            return;
        }
    }
    typedef typename Cplx<T>::type C;
    const gp<const T> us = (gp<const T>) us_, vs = (gp<const T>) vs_, ws = (gp<const T>) ws_;
    const cp<double> h1 = (cp<double>) h1_, mptr = (cp<double>) mptr_;
    const gp<cv<C>> X = (gp<cv<C>>) X_;
    extern __shared__ __align__(16) unsigned char smem_[];
    C *sb = (C *) smem_;
    const int nx = NX > 0 ? NX : R.g.nx;
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const int ld = fft_ld(nx);
    /* compiled lengths run only on runs of element size sizeof(T), whose
     * pair count is fx_pairs_of(NX, sizeof(T)) */
    const int nl = NX > 0 ? fx_pairs_of(NX, sizeof(T)) : R.fx_pairs;
    const int NL = ny * nz;
    const size_t P = (size_t) nx * ny;
    const double rny = 1.0 / (double) ny;
    const cp<double> rdz = (cp<double>) R.rdz + CF_NGT;
    const T m = mptr ? (T) *mptr : (T) 0;
    const T rdx = (T) R.rdx, rdy = (T) R.rdy;
    const int pair0 = (int) xcd_block(b.x, b.nx) * nl;
    /* every read is at a valid (clamped) address and the boundary values
     * are applied by selects, so the reads of a group of elements issue
     * together without branches */
    /* FP32 storage: the divergence is a small sum of large nearly
     * cancelling differences, so it is formed in FP64 from the stored
     * values and rounded once (the pressure then corrects the velocity by
     * a rounding of the divergence, not of its terms) */
    typedef typename std::conditional<sizeof(T) == 4, double, T>::type D;
    auto divg = [&](int line, int i) -> T {
        const bool ok = line < NL;
        const int lc = ok ? line : NL - 1;
        const int k = qdiv(lc, ny, rny), j = lc - k * ny;
        const int im = i > 0 ? i - 1 : nx - 1, jm = j > 0 ? j - 1 : ny - 1;
        const size_t q = (size_t) k * P + (size_t) j * nx;
        const D hm = h1 ? (D) ((T) h1[k + 1] * m) : (D) 0;
        const D u1 = (D) us[q + i] + hm, u0 = (D) us[q + im] + hm;
        const D v1 = vs[q + i], v0 = vs[(size_t) k * P + (size_t) jm * nx + i];
        const D w1r = ws[q + i], w0r = ws[(k > 0 ? q - P : q) + i];
        const D w1 = k < nz - 1 ? w1r : (D) 0;
        const D w0 = k > 0 ? w0r : (D) 0;
        D d;
        if (ORD == 2) {
            d = (u1 - u0) * (D) rdx + (v1 - v0) * (D) rdy + (w1 - w0) * (D) ((T) rdz[k]);
        } else {
            const D a1 = (D) (T) A1_, a2 = (D) (T) A2_;
            const int ip = i + 1 < nx ? i + 1 : i + 1 - nx, im2 = i - 2 >= 0 ? i - 2 : i - 2 + nx;
            const int jp = j + 1 < ny ? j + 1 : j + 1 - ny, jm2 = j - 2 >= 0 ? j - 2 : j - 2 + ny;
            const D u2 = (D) us[q + ip] + hm, um = (D) us[q + im2] + hm;
            const D v2 = vs[(size_t) k * P + (size_t) jp * nx + i], vm = vs[(size_t) k * P + (size_t) jm2 * nx + i];
            /* w at faces k+1 and k-2: walls at -1 and nz-1 are 0, ghosts odd */
            const D w2r = ws[(k <= nz - 3 ? q + P : (k == nz - 2 ? q : q - P)) + i];
            const D w2 = k <= nz - 3 ? w2r : (k == nz - 2 ? (D) 0 : -w2r);
            const D wmr = ws[(k >= 2 ? q - 2 * P : q) + i];
            const D wm = k >= 2 ? wmr : (k == 1 ? (D) 0 : -wmr);
            d = (a1 * (u1 - u0) + a2 * (u2 - um)) * (D) rdx + (a1 * (v1 - v0) + a2 * (v2 - vm)) * (D) rdy
              + (a1 * (w1 - w0) + a2 * (w2 - wm)) * (D) ((T) rdz[k]);
        }
        return ok ? (T) d : (T) 0;
    };
    /* compiled order-2 lengths whose first pass is radix 4 with one
     * butterfly per thread (n = 256, 4 line pairs): the lane of (pair l,
     * i < n / 4) forms the divergence at x = i + m n / 4, m < 4 -- the
     * first pass's inputs -- and takes that butterfly in registers, so the
     * values go to LDS once, as the first pass's outputs (one LDS round
     * trip and two barriers fewer); the other passes follow as usual */
    constexpr int NQ = NX > 0 ? NX / 4 : 1;
    if constexpr (NX > 0 && ORD == 2 && ct_rad_at(NX, 0) == 4 && fx_pairs_of(NX, sizeof(T)) * NQ == 256 && NQ == 64) {
        if ((ny & 1) == 0) {                      /* workgroup-uniform */
            const cp<cv<C>> twg0 = (cp<cv<C>>) twsel<C>(R.twx, R.twx32);
            C *tws0 = sb + nl * ld;
            const C tw0f = tw_first<NX, C>(twg0, nx);
            const int l = (int) threadIdx.x >> 6, i = (int) threadIdx.x & 63;
            const int la = 2 * (pair0 + l);
            const bool ok = la < NL;
            const int lc = ok ? la : NL - 2;
            const int k = qdiv(lc, ny, rny), j = lc - k * ny;
            const int jm = j > 0 ? j - 1 : ny - 1;
            const size_t qa = (size_t) k * P + (size_t) j * nx, qb = qa + nx, qm = (size_t) k * P + (size_t) jm * nx;
            const D hm = h1 ? (D) ((T) h1[k + 1] * m) : (D) 0;
            const D rz = (D) ((T) rdz[k]);
            C z[4];
#pragma unroll
            for (int g = 0; g < 4; g += 2) {
#pragma unroll
                for (int mm = g; mm < g + 2; mm++) {
                    const int x = i + NQ * mm;
                    const D ua = (D) us[qa + x] + hm, ub = (D) us[qb + x] + hm;
                    D uam = __shfl_up(ua, 1, 64), ubm = __shfl_up(ub, 1, 64);
                    if (i == 0) {
                        const int xm = x > 0 ? x - 1 : nx - 1;
                        uam = (D) us[qa + xm] + hm;
                        ubm = (D) us[qb + xm] + hm;
                    }
                    const D va = vs[qa + x], vb = vs[qb + x], vam = vs[qm + x];
                    const D war = ws[qa + x], wbr = ws[qb + x];
                    const D wamr = ws[(k > 0 ? qa - P : qa) + x], wbmr = ws[(k > 0 ? qb - P : qb) + x];
                    const D wa = k < nz - 1 ? war : (D) 0, wb = k < nz - 1 ? wbr : (D) 0;
                    const D wam = k > 0 ? wamr : (D) 0, wbm = k > 0 ? wbmr : (D) 0;
                    const D da = (ua - uam) * (D) rdx + (va - vam) * (D) rdy + (wa - wam) * rz;
                    const D db = (ub - ubm) * (D) rdx + (vb - va) * (D) rdy + (wb - wbm) * rz;
                    z[mm] = mkc<C>(ok ? (T) da : (T) 0, ok ? (T) db : (T) 0);
                }
            }
            dft_small<4>(z, 1.0);
            constexpr bool on = (NX & 63) == 0;
#pragma unroll
            for (int q = 0; q < 4; q++) sb[l * ld + fsw(4 * i + q, on)] = z[q];
            tw_rest<NX, C>(tws0, tw0f, twg0, nx);
            __syncthreads();
            lds_fft_ct<NX, 8, 1, 4, 3>(sb, ld, nl, (const C *) tws0, 1.0);
            store_half_spectra<C>(sb, ld, nl, nx, nkx, ny, (long) NL, (long) pair0, 0.5 / (double) nx, X);
            return;
        }
    }
    /* groups of CF_FX_UNR elements per thread: all reads of a group are
     * in flight before the first LDS write (blockDim.x = 256) */
    constexpr int UNR = UNRX > 0 ? UNRX
                      : (ORD == 2 ? (sizeof(T) == 4 ? CF_FX_UNR : 2) : (sizeof(T) == 4 ? 2 : 1));   /* fit 128 VGPRs */
    const cp<cv<C>> twg = (cp<cv<C>>) twsel<C>(R.twx, R.twx32);
    C *tws = sb + nl * ld;
    const C tw0 = tw_first<NX, C>(twg, nx);
    const int tot = nl * nx;
    for (int e0 = 0; e0 < tot; e0 += UNR * 256) {
        C res[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int e = e0 + (int) threadIdx.x + u * 256;
            const int ec = e < tot ? e : tot - 1;
            const int l = ec / nx, i = ec - l * nx;
            const int pr = pair0 + l;
            if (ORD == 2 && (nx & 63) == 0 && (ny & 1) == 0) {
                /* the pair's lines are rows j, j + 1 of one level: the
                 * second line's v(j - 1) is the first line's v, and
                 * u(i - 1) comes from the neighbouring lane */
                const int la = 2 * pr;
                const bool ok = la < NL;
                const int lc = ok ? la : NL - 2;
                const int k = qdiv(lc, ny, rny), j = lc - k * ny;
                const int jm = j > 0 ? j - 1 : ny - 1;
                const size_t qa = (size_t) k * P + (size_t) j * nx, qb = qa + nx, qm = (size_t) k * P + (size_t) jm * nx;
                const D hm = h1 ? (D) ((T) h1[k + 1] * m) : (D) 0;
                const D ua = (D) us[qa + i] + hm, ub = (D) us[qb + i] + hm;
                D uam = __shfl_up(ua, 1, 64), ubm = __shfl_up(ub, 1, 64);
                if ((threadIdx.x & 63) == 0) {
                    const int im = i > 0 ? i - 1 : nx - 1;
                    uam = (D) us[qa + im] + hm;
                    ubm = (D) us[qb + im] + hm;
                }
                const D va = vs[qa + i], vb = vs[qb + i], vam = vs[qm + i];
                const D war = ws[qa + i], wbr = ws[qb + i];
                const D wamr = ws[(k > 0 ? qa - P : qa) + i], wbmr = ws[(k > 0 ? qb - P : qb) + i];
                const D wa = k < nz - 1 ? war : (D) 0, wb = k < nz - 1 ? wbr : (D) 0;
                const D wam = k > 0 ? wamr : (D) 0, wbm = k > 0 ? wbmr : (D) 0;
                const D rz = (D) ((T) rdz[k]);
                const D da = (ua - uam) * (D) rdx + (va - vam) * (D) rdy + (wa - wam) * rz;
                const D db = (ub - ubm) * (D) rdx + (vb - va) * (D) rdy + (wb - wbm) * rz;
                res[u] = mkc<C>(ok ? (T) da : (T) 0, ok ? (T) db : (T) 0);
            } else {
                res[u] = mkc<C>(divg(2 * pr, i), divg(2 * pr + 1, i));
            }
        }
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int e = e0 + (int) threadIdx.x + u * 256;
            if (e < tot) {
                const int l = e / nx, i = e - l * nx;
                sb[l * ld + fsw(i, fsw_on(nx))] = res[u];
            }
        }
    }
    tw_rest<NX, C>(tws, tw0, twg, nx);
    __syncthreads();
    fft_lines<NX, 8>(sb, nx, ld, nl, (cp<int>) R.radx, R.nradx, (const C *) tws, 1.0);
    store_half_spectra<C>(sb, ld, nl, nx, nkx, ny, (long) NL, (long) pair0, 0.5 / (double) nx, X);
}

template <int NX, int ORD, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((ORD == 2 && divx2_on(NX, sizeof(T))) ? 2 : 4, 8))) void k_div_fftx(const RunDev R, const T *__restrict__ us,
    const T *__restrict__ vs, const T *__restrict__ ws, const double *__restrict__ h1,
    const double *__restrict__ mptr, typename Cplx<T>::type *__restrict__ X)
{
    divx_body<NX, ORD, T>(R, us, vs, ws, h1, mptr, X, blk_native());
}

/* the fftx_d2c stage of FFT3D_FWD on its own (no divergence), scaled 1/nx */
template <int NX>
__global__ __launch_bounds__(256) void k_fftx_fwd_plain(const RunDev R, const double *__restrict__ a,
                                                        double2 *__restrict__ X)
{
    extern __shared__ __align__(16) unsigned char smem_[];
    double2 *sb = (double2 *) smem_;
    const int nx = NX > 0 ? NX : R.g.nx;
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const int ld = fft_ld(nx);
    const int nl = R.fx_pairs;
    const long NL = (long) ny * nz;
    const long pair0 = (long) blockIdx.x * nl;
    const cp<cv<double2>> twg = (cp<cv<double2>>) R.twx;
    double2 *tws = sb + nl * ld;
    const double2 tw0 = tw_first<NX, double2>(twg, nx);
    for (int e = threadIdx.x; e < nl * nx; e += blockDim.x) {
        int l = e / nx, i = e - l * nx;
        long la = 2 * (pair0 + l), lb = la + 1;
        double va = la < NL ? a[(size_t) la * nx + i] : 0.0;
        double vb = lb < NL ? a[(size_t) lb * nx + i] : 0.0;
        sb[l * ld + fsw(i, fsw_on(nx))] = make_double2(va, vb);
    }
    tw_rest<NX, double2>(tws, tw0, twg, nx);
    __syncthreads();
    fft_lines<NX, 8>(sb, nx, ld, nl, (cp<int>) R.radx, R.nradx, (const double2 *) tws, 1.0);
    store_half_spectra<double2>(sb, ld, nl, nx, nkx, ny, NL, pair0, 0.5 / (double) nx, (gp<cv<double2>>) X);
}

/* DFT of N = 8 or 16 points in registers, natural order in and out:
 * four-step over N = 4 (N / 4), n = (N / 4) n1 + n2, k = k1 + 4 k2 --
 * DFT-4 over n1 for each n2, twiddle w_N^(n2 k1), DFT-(N / 4) over n2 for
 * each k1.  w_N^j = cos(2 pi j / N) - i sg sin(2 pi j / N), the constants
 * below being cos and sin of 2 pi j / 16. */
template <int N, typename C>
static __device__ __forceinline__ void dft_regs(C *a, double sgd)
{
    typedef typename Real<C>::type F;
    if constexpr (N <= 5) {
        dft_small<N>(a, sgd);
    } else {
        static_assert(N == 8 || N == 16, "dft_regs: 8 or 16 points");
        constexpr double c16[16] = { 1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                     0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                     -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                                     0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613 };
        constexpr double s16[16] = { 0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                                     1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                     0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                     -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173 };
        constexpr int Q = N / 4;
        const F sg = (F) sgd;
        C y[N];
#pragma unroll
        for (int n2 = 0; n2 < Q; n2++) {
            C t[4];
#pragma unroll
            for (int n1 = 0; n1 < 4; n1++) t[n1] = a[Q * n1 + n2];
            dft_small<4>(t, sgd);
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) {
                const int j = (n2 * k1) * (16 / N);       /* w_N^(n2 k1) = w_16^j */
                C v = t[k1];
                if (j != 0) {
                    const F c = (F) c16[j], sn = sg * (F) s16[j];
                    v = mkc<C>(v.x * c + v.y * sn, v.y * c - v.x * sn);
                }
                y[k1 * Q + n2] = v;
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            C u[Q];
#pragma unroll
            for (int n2 = 0; n2 < Q; n2++) u[n2] = y[k1 * Q + n2];
            if constexpr (Q == 2) dft_small<2>(u, sgd);
            else dft_small<4>(u, sgd);
#pragma unroll
            for (int k2 = 0; k2 < Q; k2++) a[k1 + 4 * k2] = u[k2];
        }
    }
}

/* y FFT of length NY = P1 P2 (P1, P2 <= 16) in two register passes, n =
 * P2 n1 + n2, k = k1 + P1 k2: pass 1 reads x[P2 n1 + n2] of one line
 * straight from memory (lanes along n2), takes the DFT-P1 over n1,
 * multiplies by w^(n2 k1) and writes row k1 of the line's LDS image
 * (rows of P2 + 1, so the column reads of pass 2 fall on distinct bank
 * groups); pass 2 reads column k1, takes the DFT-P2 over n2 and writes
 * X[k1 + P1 k2] straight to memory (lanes along k1).  One LDS round trip
 * per element instead of one per radix-4 pass plus the staging copies. */
template <int NY, typename T>
static __device__ __forceinline__ void ffty2_body(const RunDev &R, typename Cplx<T>::type *__restrict__ X_, int fwd, const Blk b)
{
    typedef typename Cplx<T>::type C;
    constexpr int P1 = fy2_p1(NY), P2 = NY / P1, RL = P2 + 1, LL = P1 * RL;
    const gp<cv<C>> X = (gp<cv<C>>) X_;
    extern __shared__ __align__(16) unsigned char smem_[];
    C *sb = (C *) smem_;
    const long nlines = (long) R.g.nz * R.g.nkx;
    const int TK = R.fy_cols;
    const long L0 = (long) b.x * TK;
    const int nl = (int) (nlines - L0 < TK ? nlines - L0 : TK);
    const gp<cv<C>> base = X + (size_t) L0 * NY;
    const double sgd = fwd ? 1.0 : -1.0;
    /* twiddles w^(n2 k1) as a P1 x P2 table after the lines */
    C *tw = sb + TK * LL;
    const cp<cv<C>> twg = (cp<cv<C>>) twsel<C>(R.twy, R.twy32);
    typedef typename Real<C>::type F;
    const int t1 = (int) threadIdx.x / P2, n2 = (int) threadIdx.x - t1 * P2;
    const bool act1 = t1 < nl;
    C a[P1];
    if (t1 < TK) {
        const gp<cv<C>> src = base + (size_t) (act1 ? t1 : 0) * NY + n2;
#pragma unroll
        for (int n1 = 0; n1 < P1; n1++) a[n1] = cx_of<C>(src[P2 * n1]);
    }
    for (int q = (int) threadIdx.x; q < NY; q += (int) blockDim.x) {
        const int k1 = q / P2, m = q - k1 * P2;
        tw[q] = cx_of<C>(twg[k1 * m]);
    }
    __syncthreads();
    if (t1 < TK) {
        dft_regs<P1>(a, sgd);
        C *row = sb + t1 * LL + n2;
#pragma unroll
        for (int k1 = 0; k1 < P1; k1++) {
            C v = act1 ? a[k1] : mkc<C>((F) 0, (F) 0);
            if (k1 > 0) {
                C w = tw[k1 * P2 + n2];
                w.y *= (F) sgd;
                v = cmul(v, w);
            }
            row[k1 * RL] = v;
        }
    }
    __syncthreads();
    const int t2 = (int) threadIdx.x / P1, k1 = (int) threadIdx.x - t2 * P1;
    if (t2 < nl) {
        C u[P2];
        const C *col = sb + t2 * LL + k1 * RL;
#pragma unroll
        for (int m = 0; m < P2; m++) u[m] = col[m];
        dft_regs<P2>(u, sgd);
        const F sc = fwd ? (F) (1.0 / (double) NY) : (F) 1;
        const gp<cv<C>> dst = base + (size_t) t2 * NY + k1;
#pragma unroll
        for (int k2 = 0; k2 < P2; k2++) dst[P1 * k2] = cv_of(fwd ? cscale(u[k2], sc) : u[k2]);
    }
}

/* complex y FFT in place, forward scaled 1/ny (fft.c:186-202) or unscaled
 * inverse (fft.c:205-208); a block transforms TK contiguous lines */
template <int NY, typename T>
static __device__ __forceinline__ void ffty_body(const RunDev &R, typename Cplx<T>::type *__restrict__ X_, int fwd, const Blk b)
{
    if constexpr (NY > 0 && fy2_p1(NY) > 0) {
        ffty2_body<NY, T>(R, X_, fwd, b);
        return;
    }
    typedef typename Cplx<T>::type C;
    const gp<cv<C>> X = (gp<cv<C>>) X_;
    extern __shared__ __align__(16) unsigned char smem_[];
    C *sb = (C *) smem_;
    const int ny = NY > 0 ? NY : R.g.ny;
    const long nlines = (long) R.g.nz * R.g.nkx;
    const int TK = R.fy_cols, ld = fft_ld(ny);
    const long L0 = (long) b.x * TK;
    const int nl = (int) (nlines - L0 < TK ? nlines - L0 : TK);
    const gp<cv<C>> base = X + (size_t) L0 * ny;
    /* all reads of a thread in flight before its first LDS write
     * (TK ny <= 8 x 256, blockDim.x = 256) */
    const cp<cv<C>> twg = (cp<cv<C>>) twsel<C>(R.twy, R.twy32);
    C *tws = sb + TK * ld;
    const C tw0 = tw_first<NY, C>(twg, ny);
    const int tot = TK * ny, totv = nl * ny;
    for (int e0 = 0; e0 < tot; e0 += 8 * 256) {
        C v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + (int) threadIdx.x + u * 256;
            const C x = cx_of<C>(base[e < totv ? e : totv - 1]);
            v[u] = e < totv ? x : mkc<C>((T) 0, (T) 0);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int e = e0 + (int) threadIdx.x + u * 256;
            if (e < tot) {
                const int t = e / ny, j = e - t * ny;
                sb[t * ld + fsw(j, fsw_on(ny))] = v[u];
            }
        }
    }
    tw_rest<NY, C>(tws, tw0, twg, ny);
    __syncthreads();
    fft_lines<NY, 8>(sb, ny, ld, TK, (cp<int>) R.rady, R.nrady, (const C *) tws, fwd ? 1.0 : -1.0);
    const T sc = fwd ? (T) (1.0 / (double) ny) : (T) 1;
    for (int e = threadIdx.x; e < nl * ny; e += blockDim.x) {
        int t = e / ny, j = e - t * ny;
        C val = sb[t * ld + fsw(j, fsw_on(ny))];
        base[e] = cv_of(fwd ? cscale(val, sc) : val);
    }
}

template <int NY, typename T>
__global__ __launch_bounds__(256) void k_ffty(const RunDev R, typename Cplx<T>::type *__restrict__ X, int fwd)
{
    ffty_body<NY, T>(R, X, fwd, blk_native());
}

/* the inverse x FFT of lengths NX = P1 P2 in two register passes: the
 * pairs' packed spectra are unpacked into LDS in natural order as in
 * fftxi_body; pass 1 (pair, n2) reads column n2, takes the inverse DFT-P1
 * and writes the pair's rows; pass 2 (pair, k1) takes the inverse DFT-P2
 * of column k1 and writes x = k1 + P1 k2 of both real lines straight to p
 * (lanes along k1). */
template <int NX, typename T>
static __device__ __forceinline__ void fftxi2_body(const RunDev &R, const typename Cplx<T>::type *__restrict__ X_,
                                                   T *__restrict__ p_, const Blk b)
{
    typedef typename Cplx<T>::type C;
    constexpr int P1 = fx2_p1(NX), P2 = NX / P1, LS = fx_ld(NX, sizeof(T)), nl = fx_pairs_of(NX, sizeof(T)), half = NX / 2;
    const gp<const cv<C>> X = (gp<const cv<C>>) X_;
    const gp<T> p = (gp<T>) p_;
    extern __shared__ __align__(16) unsigned char smem_[];
    C *sb = (C *) smem_;
    C *tw = sb + nl * LS;
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const int NLi = ny * nz;
    const int pair0 = (int) xcd_block(b.x, b.nx) * nl;
    const double rny = 1.0 / (double) ny;
    const int tot = nl * nkx;
    for (int e0 = 0; e0 < tot; e0 += CF_FXI_UNR * 256) {
        C A[CF_FXI_UNR], B[CF_FXI_UNR];
#pragma unroll
        for (int u = 0; u < CF_FXI_UNR; u++) {
            const int e = e0 + (int) threadIdx.x + u * 256;
            const int ec = e < tot ? e : tot - 1;
            const int kx = ec / nl, l = ec - kx * nl;
            const int la = 2 * (pair0 + l), lb = la + 1;
            const int lac = la < NLi ? la : NLi - 1;
            const int ka = qdiv(lac, ny, rny), ja = lac - ka * ny;
            const bool bnext = lb < NLi, wrap = ja + 1 == ny;
            const int kb = bnext ? (wrap ? ka + 1 : ka) : ka, jb = bnext ? (wrap ? 0 : ja + 1) : ja;
            const C av = cx_of<C>(X[((size_t) ka * nkx + kx) * ny + ja]);
            const C bv = cx_of<C>(X[((size_t) kb * nkx + kx) * ny + jb]);
            A[u] = la < NLi ? av : mkc<C>((T) 0, (T) 0);
            B[u] = lb < NLi ? bv : mkc<C>((T) 0, (T) 0);
        }
#pragma unroll
        for (int u = 0; u < CF_FXI_UNR; u++) {
            const int e = e0 + (int) threadIdx.x + u * 256;
            if (e < tot) {
                const int kx = e / nl, l = e - kx * nl;
                C a = A[u], bq = B[u];
                if (kx == 0 || kx == half) { a.y = (T) 0; bq.y = (T) 0; }
                sb[l * LS + kx] = mkc<C>(a.x - bq.y, a.y + bq.x);
                if (kx > 0 && kx < NX - kx) sb[l * LS + NX - kx] = mkc<C>(a.x + bq.y, bq.x - a.y);
            }
        }
    }
    tw2_stage<NX, C>(tw, (cp<cv<C>>) twsel<C>(R.twx, R.twx32));
    __syncthreads();
    const int t1 = (int) threadIdx.x / P2, n2 = (int) threadIdx.x - t1 * P2;
    C a[P1];
    if (t1 < nl) {
#pragma unroll
        for (int n1 = 0; n1 < P1; n1++) a[n1] = sb[t1 * LS + P2 * n1 + n2];
    }
    __syncthreads();
    if (t1 < nl) pass1_rows<P1, P2>(a, sb + t1 * LS + n2, tw, n2, -1.0);
    __syncthreads();
    const int t2 = (int) threadIdx.x / P1, k1 = (int) threadIdx.x - t2 * P1;
    if (t2 < nl) {
        C u[P2];
        const C *col = sb + t2 * LS + k1 * (P2 + 1);
#pragma unroll
        for (int q = 0; q < P2; q++) u[q] = col[q];
        dft_regs<P2>(u, -1.0);
        const int la = 2 * (pair0 + t2), lb = la + 1;
        const gp<T> pa = p + (size_t) la * NX + k1, pb = pa + NX;
#pragma unroll
        for (int k2 = 0; k2 < P2; k2++) {
            if (la < NLi) pa[P1 * k2] = u[k2].x;
            if (lb < NLi) pb[P1 * k2] = u[k2].y;
        }
    }
}

/* inverse complex-to-real x FFT (fft.c:180-183; FFTW c2r semantics: the
 * imaginary parts of the kx = 0 and kx = nx/2 terms are ignored), two real
 * lines per complex transform, writing p */
template <int NX, typename T>
static __device__ __forceinline__ void fftxi_body(const RunDev &R, const typename Cplx<T>::type *__restrict__ X_,
                                                  T *__restrict__ p_, const Blk b)
{
    /* a run's pair count is fx_pairs_of(nx, its element size), which for
     * the lengths compiled here in two passes is the same in FP32 and FP64 */
    if constexpr (NX > 0 && fxi2_on(NX, sizeof(T))) {
        fftxi2_body<NX, T>(R, X_, p_, b);
        return;
    }
    typedef typename Cplx<T>::type C;
    const gp<const cv<C>> X = (gp<const cv<C>>) X_;
    const gp<T> p = (gp<T>) p_;
    extern __shared__ __align__(16) unsigned char smem_[];
    C *sb = (C *) smem_;
    const int nx = NX > 0 ? NX : R.g.nx;
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const int ld = fft_ld(nx);
    const int nl = NX > 0 ? fx_pairs_of(NX, sizeof(T)) : R.fx_pairs;
    const double rnl = 1.0 / (double) nl;
    const long NL = (long) ny * nz;
    const long pair0 = (long) xcd_block(b.x, b.nx) * nl;
    const int half = nx / 2;
    const double rny = 1.0 / (double) ny;
    /* groups of CF_FXI_UNR elements per thread, reads at clamped indices,
     * all in flight before the first LDS write (blockDim.x = 256) */
    const cp<cv<C>> twg = (cp<cv<C>>) twsel<C>(R.twx, R.twx32);
    C *tws = sb + nl * ld;
    const C tw0 = tw_first<NX, C>(twg, nx);
    const int tot = nl * nkx, NLi = (int) NL;
    for (int e0 = 0; e0 < tot; e0 += CF_FXI_UNR * 256) {
        C A[CF_FXI_UNR], B[CF_FXI_UNR];
#pragma unroll
        for (int u = 0; u < CF_FXI_UNR; u++) {
            const int e = e0 + (int) threadIdx.x + u * 256;
            const int ec = e < tot ? e : tot - 1;
            const int kx = NX > 0 ? ec / nl : qdiv(ec, nl, rnl), l = ec - kx * nl;
            const int la = 2 * ((int) pair0 + l), lb = la + 1;
            const int lac = la < NLi ? la : NLi - 1;
            const int ka = qdiv(lac, ny, rny), ja = lac - ka * ny;
            /* line lb = la + 1 follows la (the clamped la when lb is past
             * the end: an in-bounds read whose value is dropped) */
            const bool bnext = lb < NLi, wrap = ja + 1 == ny;
            const int kb = bnext ? (wrap ? ka + 1 : ka) : ka, jb = bnext ? (wrap ? 0 : ja + 1) : ja;
            const C a = cx_of<C>(X[((size_t) ka * nkx + kx) * ny + ja]);
            const C bq = cx_of<C>(X[((size_t) kb * nkx + kx) * ny + jb]);
            A[u] = la < NLi ? a : mkc<C>((T) 0, (T) 0);
            B[u] = lb < NLi ? bq : mkc<C>((T) 0, (T) 0);
        }
#pragma unroll
        for (int u = 0; u < CF_FXI_UNR; u++) {
            const int e = e0 + (int) threadIdx.x + u * 256;
            if (e < tot) {
                const int kx = NX > 0 ? e / nl : qdiv(e, nl, rnl), l = e - kx * nl;
                C a = A[u], bq = B[u];
                if (kx == 0 || (nx % 2 == 0 && kx == half)) { a.y = (T) 0; bq.y = (T) 0; }
                sb[l * ld + fsw(kx, fsw_on(nx))] = mkc<C>(a.x - bq.y, a.y + bq.x);
                if (kx > 0 && kx < nx - kx) sb[l * ld + fsw(nx - kx, fsw_on(nx))] = mkc<C>(a.x + bq.y, bq.x - a.y);
            }
        }
    }
    tw_rest<NX, C>(tws, tw0, twg, nx);
    __syncthreads();
    fft_lines<NX, 8>(sb, nx, ld, nl, (cp<int>) R.radx, R.nradx, (const C *) tws, -1.0);
    for (int e = threadIdx.x; e < nl * nx; e += blockDim.x) {
        int l = e / nx, i = e - l * nx;
        long la = 2 * (pair0 + l), lb = la + 1;
        C z = sb[l * ld + fsw(i, fsw_on(nx))];
        if (la < NL) p[(size_t) la * nx + i] = z.x;
        if (lb < NL) p[(size_t) lb * nx + i] = z.y;
    }
}

template <int NX, typename T>
__global__ __launch_bounds__(256) void k_fftx_inv(const RunDev R, const typename Cplx<T>::type *__restrict__ X,
                                                  T *__restrict__ p)
{
    fftxi_body<NX, T>(R, X, p, blk_native());
}

/* host-side selectors of the compile-time instances */
#define CF_FFT_SIZES(X) X(16) X(20) X(24) X(30) X(32) X(40) X(48) X(60) X(64) X(80) X(96) \
    X(120) X(128) X(160) X(192) X(240) X(256) X(320) X(384) X(480) X(512)

template <typename T, int ORD>
static typename KSel<T>::div_t pick_div_ord(int n)
{
    switch (n) {
#define X_(N) case N: return k_div_fftx<N, ORD, T>;
    CF_FFT_SIZES(X_)
#undef X_
    default: return k_div_fftx<0, ORD, T>;
    }
}

template <typename T>
typename KSel<T>::div_t pick_div_fftx(int n, int order)
{
    return order == 4 ? pick_div_ord<T, 4>(n) : pick_div_ord<T, 2>(n);
}

k_fftx_plain_t pick_fftx_plain(int n)
{
    switch (n) {
#define X_(N) case N: return k_fftx_fwd_plain<N>;
    CF_FFT_SIZES(X_)
#undef X_
    default: return k_fftx_fwd_plain<0>;
    }
}

template <typename T>
typename KSel<T>::ffty_t pick_ffty(int n)
{
    switch (n) {
#define X_(N) case N: return k_ffty<N, T>;
    CF_FFT_SIZES(X_)
#undef X_
    default: return k_ffty<0, T>;
    }
}

template <typename T>
typename KSel<T>::fftxi_t pick_fftx_inv(int n)
{
    switch (n) {
#define X_(N) case N: return k_fftx_inv<N, T>;
    CF_FFT_SIZES(X_)
#undef X_
    default: return k_fftx_inv<0, T>;
    }
}

template KSel<double>::div_t pick_div_fftx<double>(int, int);
template KSel<float>::div_t pick_div_fftx<float>(int, int);
template KSel<double>::ffty_t pick_ffty<double>(int);
template KSel<float>::ffty_t pick_ffty<float>(int);
template KSel<double>::fftxi_t pick_fftx_inv<double>(int);
template KSel<float>::fftxi_t pick_fftx_inv<float>(int);

/* ---------------------------------------------------------------------
 * pressure_poisson2 (pressure.c:293-355): one thread per (kx, ky) mode,
 * z-tridiagonal with modified wavenumbers lam = 2 - 2 cos, dvol
 * symmetrisation, Neumann walls except the (0,0) mode whose bottom row
 * is the Dirichlet-mean closure; real and imaginary parts share the
 * matrix.  Thomas with the gam factors kept in G for the back sweep.
 * Arithmetic in A: FP64 always for T = double; for T = float the modes
 * with |kx|, |ky| <= CF_F64_KLOW are still solved in FP64 (the mean
 * pressure and the large-scale modes carry the wall-normal pressure
 * balance), the rest in FP32.
 * ------------------------------------------------------------------- */
/* One mode's Thomas solve in two passes over Y without storing the
 * factors: the forward sweep keeps only the recurrence state (d, 1/bet)
 * entering every chunk of CF_P2_CH levels; the backward pass walks the
 * chunks from the top, recomputes each chunk's forward sweep from its
 * checkpoint into registers and back-substitutes.  HBM traffic is two
 * reads and one write of Y (48 B per complex value) instead of writing
 * the forward sweep and the factors and reading both back (80 B).  G holds
 * the checkpoints: level slot 3 (c - 1) + {0, 1, 2} for chunk c >= 1. */
#define CF_P2_CH 16
template <typename A, typename C>
static __device__ __forceinline__ void pois2_mode(const PoisTab &T, C *__restrict__ y, double *__restrict__ g,
                                                  size_t M, int nz, double lxd, double lyd, bool zero)
{
    const A lx = (A) lxd, ly = (A) lyd;
    const int nch = (nz + CF_P2_CH - 1) / CF_P2_CH;
    /* one level of the forward recurrence: (d, ibet) of level k-1 -> level k */
    auto fwd = [&](int k, C rk, A &dx_, A &dy_, A &ibet, A &gam) {
        A b = (A) T.bz[k] - lx * (A) T.ex[k] - ly * (A) T.ey[k];
        if (k == 0) b -= (A) (zero ? T.db_dir : T.db_neu);
        if (k == nz - 1) b -= (A) T.db_top;
        A bet = b;
        gam = 0;
        if (k > 0) {
            gam = (A) T.c[k - 1] * ibet;              /* c_{k-1} / bet_{k-1} */
            bet = b - (A) T.a[k] * gam;
        }
        ibet = (A) 1 / bet;
        const A dv = (A) T.dvol[k];
        const A ak = k > 0 ? (A) T.a[k] : (A) 0;
        dx_ = ((A) rk.x * dv - ak * dx_) * ibet;
        dy_ = ((A) rk.y * dv - ak * dy_) * ibet;
    };
    {   /* pass 1: forward sweep, checkpoints only */
        A dx_ = 0, dy_ = 0, ibet = 0, gam;
        for (int c = 0; c < nch; c++) {
            const int k0 = c * CF_P2_CH;
            if (c > 0) {
                g[(size_t) (3 * (c - 1) + 0) * M] = (double) dx_;
                g[(size_t) (3 * (c - 1) + 1) * M] = (double) dy_;
                g[(size_t) (3 * (c - 1) + 2) * M] = (double) ibet;
            }
            C rb[CF_P2_CH];
#pragma unroll
            for (int t = 0; t < CF_P2_CH; t++)
                if (k0 + t < nz) rb[t] = y[(size_t) (k0 + t) * M];
#pragma unroll
            for (int t = 0; t < CF_P2_CH; t++)
                if (k0 + t < nz) fwd(k0 + t, rb[t], dx_, dy_, ibet, gam);
        }
    }
    /* pass 2: chunks from the top, recompute then back-substitute */
    A xnx = 0, xny = 0, gnext = 0;
    for (int c = nch - 1; c >= 0; c--) {
        const int k0 = c * CF_P2_CH;
        A dx_ = 0, dy_ = 0, ibet = 0;
        if (c > 0) {
            dx_ = (A) g[(size_t) (3 * (c - 1) + 0) * M];
            dy_ = (A) g[(size_t) (3 * (c - 1) + 1) * M];
            ibet = (A) g[(size_t) (3 * (c - 1) + 2) * M];
        }
        C rb[CF_P2_CH];
        A dxs[CF_P2_CH], dys[CF_P2_CH], gs[CF_P2_CH];
#pragma unroll
        for (int t = 0; t < CF_P2_CH; t++)
            if (k0 + t < nz) rb[t] = y[(size_t) (k0 + t) * M];
#pragma unroll
        for (int t = 0; t < CF_P2_CH; t++) {
            if (k0 + t < nz) {
                A gam;
                fwd(k0 + t, rb[t], dx_, dy_, ibet, gam);
                dxs[t] = dx_; dys[t] = dy_; gs[t] = gam;
            }
        }
#pragma unroll
        for (int t = CF_P2_CH - 1; t >= 0; t--) {
            const int k = k0 + t;
            if (k < nz) {
                if (k < nz - 1) {
                    xnx = dxs[t] - gnext * xnx;
                    xny = dys[t] - gnext * xny;
                } else {
                    xnx = dxs[t];
                    xny = dys[t];
                }
                gnext = gs[t];
                C o;
                o.x = xnx; o.y = xny;
                y[(size_t) k * M] = o;
            }
        }
    }
}

/* Partitioned form of the same solve, for nz <= 16 * 32.  A workgroup of
 * CF_P2S_THREADS threads owns CF_P2S_MW modes; thread t takes part t & 1
 * (real or imaginary: both parts share the matrix), mode (t >> 1) mod
 * CF_P2S_MW of the group and z-chunk t / (2 CF_P2S_MW) of L rows (L a
 * compile-time choice >= nz / 16).  Rows
 * above nz are identity rows (x = 0), so every chunk has L - 1 interior
 * rows and a separator (its last row).  Each thread holds its chunk in
 * registers and solves the interior as a Dirichlet problem in the two
 * neighbouring separators (x = y + P X_below + Q X_above, Thomas in
 * registers); the separators' Schur complement (tridiagonal, one unknown
 * per chunk) is solved in LDS by parallel cyclic reduction, every chunk's
 * thread taking part, then every thread writes its rows.  Every value is read once and written once (32 B
 * per complex FP64 value), and the 16-fold parallelism over z hides the
 * latency of the level recurrences, which bounds the one-thread-per-mode
 * sweep at these mode counts.  Differs from sequential Thomas by
 * round-off. */
/* occupancy target by chunk length: the largest without register spills */
#ifndef CF_P2S_WAVES
#define CF_P2S_WAVES(L) ((L) == 0 ? 2 : ((L) <= 13 ? 4 : ((L) <= 20 ? 3 : 1)))
#endif

template <int L, typename A, typename T>
static __device__ __forceinline__ void pois2s_run(const PoisTab &T_, typename Real<typename Cplx<T>::type>::type *__restrict__ yr,
                                                  size_t M, int nz, int nch, size_t idx, bool act,
                                                  double lxd, double lyd, bool zero, double (*sh)[CF_P2S_NC][CF_P2S_GW],
                                                  double (*xs)[CF_P2S_GW], double *tl)
{
    typedef typename Real<typename Cplx<T>::type>::type S;
    const int g = threadIdx.x & (CF_P2S_GW - 1), chunk = threadIdx.x / CF_P2S_GW;
    const A lx = (A) lxd, ly = (A) lyd;
    const bool live = chunk < nch;
    const int lo = chunk * L;
    const bool ld = live && act;
    const size_t stride = 2 * M;
    const gp<S> col = (gp<S>) yr + idx;
    const int ktop = nz - 1;
    /* the column's rows (clamped, valid reads), then the z tables into LDS
     * (tl: a, c, bz, ex, ey, dvol, nz each) with the same round trip */
    S raw[L];
#pragma unroll
    for (int t = 0; t < L; t++) raw[t] = col[(size_t) min(lo + t, ktop) * stride];
    {
        const cp<double> src[6] = { (cp<double>) T_.a, (cp<double>) T_.c, (cp<double>) T_.bz,
                                    (cp<double>) T_.ex, (cp<double>) T_.ey, (cp<double>) T_.dvol };
#pragma unroll
        for (int w = 0; w < 6; w++)
            for (int k = (int) threadIdx.x; k < nz; k += (int) blockDim.x) tl[w * nz + k] = src[w][k];
    }
    __syncthreads();
    const double *ta = tl, *tc = tl + nz, *tbz = tl + 2 * nz, *tex = tl + 3 * nz, *tey = tl + 4 * nz,
                 *tdv = tl + 5 * nz;
    /* row coefficients, identity beyond the top wall; every read is at a
     * clamped (valid) level and the padding is applied by selects */
    const A dbot = (A) (zero ? T_.db_dir : T_.db_neu), dtop = (A) T_.db_top;
    auto cb = [&](int k) {
        const int kc = min(k, ktop);
        A b = (A) tbz[kc] - lx * (A) tex[kc] - ly * (A) tey[kc];
        b -= k == 0 ? dbot : (A) 0;
        b -= k == ktop ? dtop : (A) 0;
        return k > ktop ? (A) 1 : b;
    };
    auto ca = [&](int k) { const A v = (A) ta[min(k, ktop)]; return (k > 0 && k <= ktop) ? v : (A) 0; };
    auto cc = [&](int k) { const A v = (A) tc[min(k, ktop)]; return k < ktop ? v : (A) 0; };
    A d[L - 1], ph[L - 1], gm[L - 1];
    A ds = 0;
#pragma unroll
    for (int t = 0; t < L; t++) {
        const int k = lo + t;
        const A v = (A) raw[t] * (A) tdv[min(k, ktop)];
        const A r = (ld && k <= ktop) ? v : (A) 0;
        if (t < L - 1) d[t] = r;
        else ds = r;
    }
    /* forward sweep of the interior; the separator below enters through
     * a_lo, carried as ph */
    A ibet = 0, dp = 0, pp = 0;
#pragma unroll
    for (int t = 0; t < L - 1; t++) {
        const int k = lo + t;
        A b = cb(k), gam = 0;
        if (t > 0) {
            gam = cc(k - 1) * ibet;
            b -= ca(k) * gam;
        }
        ibet = (A) 1 / b;
        if (t == 0) { dp = d[t] * ibet; pp = -ca(k) * ibet; }
        else { dp = (d[t] - ca(k) * dp) * ibet; pp = -ca(k) * pp * ibet; }
        d[t] = dp; ph[t] = pp; gm[t] = gam;
    }
    /* back substitution of the particular (y), below (P) and above (Q)
     * responses, in place: d <- y, ph <- P, gm <- Q */
    A yv = d[L - 2], pv = ph[L - 2], qv = -cc(lo + L - 2) * ibet;
    const A yl = yv, pl = pv, ql = qv;
    A gnext = gm[L - 2];
    gm[L - 2] = qv;
#pragma unroll
    for (int t = L - 3; t >= 0; t--) {
        yv = d[t] - gnext * yv;
        pv = ph[t] - gnext * pv;
        qv = -gnext * qv;
        gnext = gm[t];
        d[t] = yv; ph[t] = pv; gm[t] = qv;
    }
    if (live) {
        sh[0][chunk][g] = (double) yv; sh[1][chunk][g] = (double) pv; sh[2][chunk][g] = (double) qv;
        sh[3][chunk][g] = (double) yl; sh[4][chunk][g] = (double) pl; sh[5][chunk][g] = (double) ql;
        sh[6][chunk][g] = (double) ds;
    }
    __syncthreads();
    /* separator s_j = (j + 1) L - 1 above chunk j:
     * a x_{s-1} + b x_s + c x_{s+1} = d_s with x_{s-1} chunk j's top
     * interior row and x_{s+1} chunk j + 1's bottom one, which gives the
     * tridiagonal system sub_j X_{j-1} + dia_j X_j + sup_j X_{j+1} = rhs_j
     * over the separators; each chunk's lanes form their separator's row */
    A sub = 0, dia = 1, sup = 0, rhs = 0;
    if (live) {
        const int j = chunk, sj = (j + 1) * L - 1;
        const bool nxt = j < nch - 1;
        const A as = ca(sj), cs = cc(sj);
        sub = j > 0 ? as * (A) sh[4][j][g] : (A) 0;
        dia = cb(sj) + as * (A) sh[5][j][g] + (nxt ? cs * (A) sh[1][j + 1][g] : (A) 0);
        sup = nxt ? cs * (A) sh[2][j + 1][g] : (A) 0;
        rhs = (A) sh[6][j][g] - as * (A) sh[3][j][g] - (nxt ? cs * (A) sh[0][j + 1][g] : (A) 0);
    }
    /* parallel cyclic reduction over the chunks (log2 nch steps, every
     * chunk's lanes at once): row j takes rows j - d and j + d out of its
     * equation; sh[0..3] hold the rows of the current step */
    for (int d = 1; d < nch; d *= 2) {
        __syncthreads();
        if (live) {
            sh[0][chunk][g] = (double) sub; sh[1][chunk][g] = (double) dia;
            sh[2][chunk][g] = (double) sup; sh[3][chunk][g] = (double) rhs;
        }
        __syncthreads();
        if (live) {
            A nsub = 0, nsup = 0, ndia = dia, nrhs = rhs;
            if (chunk - d >= 0) {
                const A al = -sub / (A) sh[1][chunk - d][g];
                nsub = al * (A) sh[0][chunk - d][g];
                ndia += al * (A) sh[2][chunk - d][g];
                nrhs += al * (A) sh[3][chunk - d][g];
            }
            if (chunk + d < nch) {
                const A ga = -sup / (A) sh[1][chunk + d][g];
                nsup = ga * (A) sh[2][chunk + d][g];
                ndia += ga * (A) sh[0][chunk + d][g];
                nrhs += ga * (A) sh[3][chunk + d][g];
            }
            sub = nsub; sup = nsup; dia = ndia; rhs = nrhs;
        }
    }
    if (live) xs[chunk][g] = (double) (rhs / dia);
    __syncthreads();
    if (!ld) return;
    const A XB = chunk > 0 ? (A) xs[chunk - 1][g] : (A) 0;
    const A XA = (A) xs[chunk][g];
    const gp<S> out = col + (size_t) lo * stride;
    if (lo + L - 1 <= ktop) {                     /* a full chunk */
#pragma unroll
        for (int t = 0; t < L - 1; t++) out[(size_t) t * stride] = (S) (d[t] + ph[t] * XB + gm[t] * XA);
        out[(size_t) (L - 1) * stride] = (S) XA;
    } else {
#pragma unroll
        for (int t = 0; t < L - 1; t++)
            if (lo + t <= ktop) out[(size_t) t * stride] = (S) (d[t] + ph[t] * XB + gm[t] * XA);
    }
}

/* L = pois2_L(nz) (or larger): the partitioned solve with chunks of L
 * rows; L = 0: one thread per mode.  FP32 storage solves the modes with
 * |kx|, |ky| <= CF_F64_KLOW in FP64 (the islands); ISL selects what a
 * launch covers: 0 the storage-precision modes only (island workgroups
 * return at once), 1 the islands only, 2 both.  Splitting them keeps the
 * FP64 path's registers out of the FP32 kernel. */
template <typename T, int L, int ISL = 2>
static __device__ __forceinline__ void pois2_body(const RunDev &R, const PoisTab T_,
                                                  typename Cplx<T>::type *__restrict__ Y, double *__restrict__ G, const Blk b)
{
    const int ny = R.g.ny, nz = R.g.nz, nkx = R.g.nkx;
    const size_t M = (size_t) ny * nkx;
    constexpr bool f64 = sizeof(T) == sizeof(double);
    if (L > 0) {
        __shared__ double sh[7][CF_P2S_NC][CF_P2S_GW];
        __shared__ double xs[CF_P2S_NC][CF_P2S_GW];
        const int nch = (nz + L - 1) / (L > 0 ? L : 1);
        /* FP32 rows are half a cache line per workgroup: with CF_P2S_XCD the
         * workgroups holding neighbouring modes share an XCD (and its L2;
         * FP32 step 1.267 -> 1.243 ms at 256x193x256); the island launch
         * covers only the leading workgroups and keeps the plain order */
        const unsigned bx = (CF_P2S_XCD && ISL != 1) ? xcd_block(b.x, b.nx) : b.x;
        const size_t m0 = (size_t) bx * CF_P2S_MW;
        const size_t mode = m0 + ((threadIdx.x >> 1) & (CF_P2S_MW - 1));
        const bool act = mode < M;
        const size_t mc = act ? mode : M - 1;
        const int kx = (int) (mc / ny), ky = (int) (mc % ny);
        const bool zero = kx == 0 && ky == 0;
        /* FP32 storage: workgroups holding any mode with |kx|, |ky| <=
         * CF_F64_KLOW solve in FP64 (a workgroup-uniform choice) */
        bool low = false;
        if (!f64) {
            const size_t m1 = min(m0 + CF_P2S_MW, M) - 1;
            for (size_t q = m0; q <= m1; q++) {
                const int qx = (int) (q / ny), qy = (int) (q % ny);
                low = low || (qx <= CF_F64_KLOW && (qy <= CF_F64_KLOW || ny - qy <= CF_F64_KLOW));
            }
        }
        if (!f64 && ((ISL == 0 && low) || (ISL == 1 && !low))) return;   /* workgroup-uniform */
        typename Real<typename Cplx<T>::type>::type *yr = (typename Real<typename Cplx<T>::type>::type *) Y;
        extern __shared__ __align__(16) unsigned char smem_[];
        double *tl = (double *) smem_;            /* 6 nz doubles (pois2_lds) */
        const size_t idx = 2 * mc + (threadIdx.x & 1);
        const double lx = ((cp<double>) R.lamx)[kx], ly = ((cp<double>) R.lamy)[ky];
        constexpr int LL = L > 1 ? L : 2;
        if constexpr (f64 || ISL == 1) {
            pois2s_run<LL, double, T>(T_, yr, M, nz, nch, idx, act, lx, ly, zero, sh, xs, tl);
        } else if constexpr (ISL == 0) {
            pois2s_run<LL, T, T>(T_, yr, M, nz, nch, idx, act, lx, ly, zero, sh, xs, tl);
        } else {
            if (low)
                pois2s_run<LL, double, T>(T_, yr, M, nz, nch, idx, act, lx, ly, zero, sh, xs, tl);
            else
                pois2s_run<LL, T, T>(T_, yr, M, nz, nch, idx, act, lx, ly, zero, sh, xs, tl);
        }
        return;
    }
    const size_t qm = (size_t) b.x * blockDim.x + threadIdx.x;
    if (qm >= M) return;
    const int kx = (int) (qm / ny), ky = (int) (qm % ny);     /* layout [kx][ky] per level */
    const bool zero = kx == 0 && ky == 0;
    const bool low = kx <= CF_F64_KLOW && (ky <= CF_F64_KLOW || ny - ky <= CF_F64_KLOW);
    if (!f64 && ((ISL == 0 && low) || (ISL == 1 && !low))) return;
    if (f64 || low)
        pois2_mode<double>(T_, Y + qm, G + qm, M, nz, R.lamx[kx], R.lamy[ky], zero);
    else
        pois2_mode<T>(T_, Y + qm, G + qm, M, nz, R.lamx[kx], R.lamy[ky], zero);
}

template <typename T, int L, int ISL>
__global__ __launch_bounds__(CF_P2S_THREADS) __attribute__((amdgpu_waves_per_eu(CF_P2S_WAVES(L), 8))) void k_poisson2(const RunDev R, const PoisTab T_,
                                                  typename Cplx<T>::type *__restrict__ Y, double *__restrict__ G)
{
    pois2_body<T, L, ISL>(R, T_, Y, G, blk_native());
}

/* FP64: the one solve; FP32: the storage-precision modes (the islands are
 * a launch of their own, pick_poisson2_islands) */
template <typename T>
typename KSel<T>::pois2_t pick_poisson2(int L)
{
    constexpr int I = sizeof(T) == sizeof(double) ? 2 : 0;
    switch (L) {
#define X_(N) case N: return k_poisson2<T, N, I>;
    CF_P2S_LSET(X_)
#undef X_
    default: return k_poisson2<T, 0, I>;
    }
}
template KSel<double>::pois2_t pick_poisson2<double>(int);
template KSel<float>::pois2_t pick_poisson2<float>(int);

KSel<float>::pois2_t pick_poisson2_islands(int L)
{
    switch (L) {
#define X_(N) case N: return k_poisson2<float, N, 1>;
    CF_P2S_LSET(X_)
#undef X_
    default: return k_poisson2<float, 0, 1>;
    }
}

/* ---------------------------------------------------------------------
 * project2 (pressure.c:120-152) fused with the mass-flux correction:
 * u = (u* + h1 mdpdx) - (p_{i+1} - p_i) dt / dx, v, w likewise.  The
 * top-wall face is left at its Dirichlet value 0 (the reference subtracts
 * the Neumann-ghost difference there, which is round-off).
 * ------------------------------------------------------------------- */
template <typename T>
__global__ __launch_bounds__(256) void k_project2(const RunDev R, const T *__restrict__ us,
    const T *__restrict__ vs, const T *__restrict__ ws, const double *__restrict__ h1,
    const double *__restrict__ mptr, const T *__restrict__ p, double dtp,
    T *__restrict__ u, T *__restrict__ v, T *__restrict__ w)
{
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const size_t P = (size_t) nx * ny;
    const int ip = i + 1 < nx ? i + 1 : 0, jp = j + 1 < ny ? j + 1 : 0;
    const size_t q = (size_t) k * P + (size_t) j * nx + i;
    const double *rdzh = R.rdzh + CF_NGT;
    const T m = mptr ? (T) *mptr : (T) 0;
    const T p0 = p[q];
    T uc = us[q] + (h1 ? (T) h1[k + 1] * m : (T) 0);
    u[q] = uc - (p[(size_t) k * P + (size_t) j * nx + ip] - p0) * (T) (dtp * R.rdx);
    v[q] = vs[q] - (p[(size_t) k * P + (size_t) jp * nx + i] - p0) * (T) (dtp * R.rdy);
    if (k < nz - 1) {
        w[q] = ws[q] - (p[q + P] - p0) * (T) (dtp * rdzh[k]);
    } else {
        w[q] = (T) 0;
    }
}
template __global__ void k_project2<double>(const RunDev, const double *, const double *, const double *, const double *,
    const double *, const double *, double, double *, double *, double *);
template __global__ void k_project2<float>(const RunDev, const float *, const float *, const float *, const double *,
    const double *, const float *, double, float *, float *, float *);

/* max |div| over interior cells (case1 check, cases.c:473-486), the
 * divergence of the order in force, evaluated in FP64 whatever the
 * storage precision */
template <typename T>
__global__ __launch_bounds__(256) void k_maxdiv_part(const RunDev R, const T *__restrict__ u,
    const T *__restrict__ v, const T *__restrict__ w, double *__restrict__ part, int nb)
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
        double w1 = k < nz - 1 ? (double) w[e] : 0.0;
        double w0 = k > 0 ? (double) w[e - P] : 0.0;
        double d;
        if (R.g.order == 2) {
            d = ((double) u[e] - (double) u[row + im]) / R.g.dx
              + ((double) v[e] - (double) v[(size_t) k * P + (size_t) jm * nx + i]) / R.g.dy
              + (w1 - w0) / dz[k];
        } else {                                   /* divergence4, pressure.c:63-118 */
            int ip = i + 1 < nx ? i + 1 : i + 1 - nx, im2 = i - 2 >= 0 ? i - 2 : i - 2 + nx;
            int jp = j + 1 < ny ? j + 1 : j + 1 - ny, jm2 = j - 2 >= 0 ? j - 2 : j - 2 + ny;
            double w2 = k + 1 <= nz - 2 ? (double) w[e + P] : (k + 1 == nz - 1 ? 0.0 : -(double) w[e - P]);
            double wm = k - 2 >= 0 ? (double) w[e - 2 * P] : (k - 2 == -1 ? 0.0 : -(double) w[e]);
            double um = u[row + im], u2 = u[row + ip], um2 = u[row + im2];
            double vm = v[(size_t) k * P + (size_t) jm * nx + i], v2 = v[(size_t) k * P + (size_t) jp * nx + i];
            double vm2 = v[(size_t) k * P + (size_t) jm2 * nx + i];
            d = (A1_ * ((double) u[e] - um) + A2_ * (u2 - um2)) / R.g.dx
              + (A1_ * ((double) v[e] - vm) + A2_ * (v2 - vm2)) / R.g.dy
              + (A1_ * (w1 - w0) + A2_ * (w2 - wm)) / dz[k];
        }
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

template __global__ void k_maxdiv_part<double>(const RunDev, const double *, const double *, const double *, double *, int);
template __global__ void k_maxdiv_part<float>(const RunDev, const float *, const float *, const float *, double *, int);

/* ------------------------------------------------------------ batched launches
 * (see cf_general.hip): the shared solves, FFT passes, order-2 Poisson and
 * the per-step bookkeeping over all runs of a batch; FFTs use the runtime
 * plan since the runs' lengths differ. */

/* NX / NY > 0: a launch over the runs of one FFT length with its plan
 * compiled in; 0: any lengths, runtime plan.  Where it fits without
 * spills, the occupancy target caps them at 128 VGPRs (4 waves per SIMD). */
/* compiled lengths gather 2 elements per thread in FP32 (not 3) so that
 * the batch loop's extra registers still fit 4 waves per SIMD */
template <int NX, int ORD, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((NX > 0 && ORD == 2) ? (divx2_on(NX, sizeof(T)) ? 2 : 4) : 1, 8))) void kb_divx(
    const GArgs<T> *__restrict__ A, const unsigned *__restrict__ vs, int nrun, int s)
{
    constexpr int U = (NX > 0 && ORD == 2 && sizeof(T) == 4) ? 2 : 0;
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = A[r];
        const Blk b = blk_of(a, ST_FFTX, lb);
        if (s < 3)
            divx_body<NX, ORD, T, U>(a.R, a.ru, a.rv, a.rw, a.R.h1 + (size_t) (s * 2) * (a.R.g.nz + 2),
                                     a.scal + SC_M0 + s, a.X, b);
        else
            divx_body<NX, ORD, T, U>(a.R, a.u, a.v, a.w, nullptr, nullptr, a.X, b);
    });
}

template <typename T>
typename BSel<T>::divx_t_ pick_batched_divx(int order)
{
    return order == 4 ? kb_divx<0, 4, T> : kb_divx<0, 2, T>;
}

template <int NY, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NY > 0 ? 4 : 1, 8))) void kb_ffty(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int fwd)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        ffty_body<NY, T>(a.R, a.X, fwd, blk_of(a, ST_FFTY, lb));
    });
}

template <int NX, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((NX > 0 && NX <= 128) ? (fxi2_on(NX, sizeof(T)) && sizeof(T) == 8 ? 2 : 4) : 1, 8))) void kb_fftxi(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    (void) s;
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        fftxi_body<NX, T>(a.R, a.X, a.p, blk_of(a, ST_FFTX, lb));
    });
}

/* per-length batched FFT launches (n = bfft_key of the group's length).
 * The divergence + x-FFT has its length compiled in only where that fits
 * 4 waves per SIMD without spills inside the batch loop (order 2, n <= 128
 * in FP32, n <= 64 in FP64); other groups take the runtime plan. */
template <typename T>
typename BSel<T>::kb_t pick_batched_divx_n(int order, int n)
{
    constexpr int lim = sizeof(T) == 4 ? 128 : 64;
    if (order == 2) {
        switch (n) {
        case 32: return kb_divx<32, 2, T>;
        case 48: return kb_divx<48, 2, T>;
        case 64: return kb_divx<64, 2, T>;
        case 96: if constexpr (lim >= 96) return kb_divx<96, 2, T>; break;
        case 128: if constexpr (lim >= 128) return kb_divx<128, 2, T>; break;
        default: break;
        }
    }
    return pick_batched_divx<T>(order);
}
template <typename T>
typename BSel<T>::kb_t pick_batched_ffty(int n)
{
    switch (n) {
#define X_(N) case N: return kb_ffty<N, T>;
    CF_BFFT_SIZES(X_)
#undef X_
    default: return kb_ffty<0, T>;
    }
}
template <typename T>
typename BSel<T>::kb_t pick_batched_fftxi(int n)
{
    switch (n) {
#define X_(N) case N: return kb_fftxi<N, T>;
    CF_BFFT_SIZES(X_)
#undef X_
    default: return kb_fftxi<0, T>;
    }
}

template <typename T, int L, int ISL>
__global__ __launch_bounds__(CF_P2S_THREADS) void kb_poisson(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        pois2_body<T, L, ISL>(a.R, a.pt[s], a.X, a.G, blk_of(a, ST_MODE, lb));
    });
}

template <typename T>
typename BSel<T>::kb_t pick_batched_poisson(int L)
{
    constexpr int I = sizeof(T) == sizeof(double) ? 2 : 0;
    switch (L) {
#define X_(N) case N: return kb_poisson<T, N, I>;
    CF_P2S_LSET(X_)
#undef X_
    default: return kb_poisson<T, 0, I>;
    }
}

BSel<float>::kb_t pick_batched_poisson_islands(int L)
{
    switch (L) {
#define X_(N) case N: return kb_poisson<float, N, 1>;
    CF_P2S_LSET(X_)
#undef X_
    default: return kb_poisson<float, 0, 1>;
    }
}

template <typename T, int LM>
__global__ __launch_bounds__(512) void kb_spike(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        spike_body<T, LM>(a.R, s, a.ru, a.rv, a.rw, a.rc, blk_of(a, ST_SOLVE, lb));
    });
}

template <typename T>
typename BSel<T>::kb_t pick_batched_spike(int L)
{
    switch (L) {
#define X_(N) case N: return kb_spike<T, N>;
    CF_SPIKE_LSET(X_)
#undef X_
    default: return nullptr;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void kb_thomas(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        thomas_body<T>(a.R, s, a.ru, a.rv, a.rw, a.rc, blk_of(a, ST_SOLVE, lb));
    });
}

template <typename T>
__global__ __launch_bounds__(256) void kb_cfl(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    (void) s;
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        cfl_body(a.R, a.cflpart, a.gblocks, a.dt, a.scal, a.logbuf, blk_of(a, ST_ONE, lb));
    });
}

template <typename T>
__global__ __launch_bounds__(256) void kb_step_end(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    (void) s;
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        (void) lb;
        if (threadIdx.x == 0) a_[r].scal[SC_STEP] += 1.0;
    });
}

/* ---- batched order-2 fast path (FP64 or FP32 storage): the single-run
 * kernels' bodies over the runs of a batch.  The argument table in use maps
 * u, v, w to the substep's input state set and ru, rv, rw to its output
 * set. */
template <typename T>
__global__ __launch_bounds__(256) void kb_rhs2t(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        const int ps = (s + 2) % 3;               /* the substep whose projection is applied */
        ProjSrcT<T> S;
        S.us = a.u; S.vs = a.v; S.ws = a.w; S.p = a.p;
        S.h1 = a.R.h1 + (size_t) (ps * 2) * (a.R.g.nz + 2);
        S.m = a.scal + SC_M0 + ps;
        S.dtp = a.R.sc[ps].abdt;
        T *const *fnew = (s % 2 == 0) ? a.fa : a.fb;
        T *const *gold = (s % 2 == 0) ? a.fb : a.fa;
        const bool useg = a.R.sc[s].zetdt != 0.0, keepf = s < 2;
        rhs2t_body<T>(a.R, s, S, useg ? gold[0] : nullptr, gold[1], gold[2], a.scal,
                      keepf ? fnew[0] : nullptr, fnew[1], fnew[2], a.ru, a.rv, a.rw,
                      s == 0 ? a.cflpart : nullptr, a.dt, a.rhs_kc, lb);
    });
}

template <typename T, int LM>
__global__ __launch_bounds__(512) void kb_spike_sums(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        spike_body<T, LM>(a.R, s, a.ru, a.rv, a.rw, nullptr, blk_of(a, ST_SOLVE, lb), a.part);
    });
}

template <typename T>
typename BSel<T>::kb_t pick_batched_spike_sums(int L)
{
    switch (L) {
#define X_(N) case N: return kb_spike_sums<T, N>;
    CF_SPIKE_LSET(X_)
#undef X_
    default: return nullptr;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void kb_level_sum(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    (void) s;
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        const GArgs<T> &a = a_[r];
        level_sum_body(a.part, (int) a.dims[ST_SOLVE][0], a.lsum, (int) lb);
    });
}

template <typename T>
__global__ __launch_bounds__(256) void kb_mass_fast(const GArgs<T> *__restrict__ a_, const unsigned *__restrict__ vs, int nrun, int s)
{
    batch_loop(vs, nrun, [&](int r, unsigned lb) {
        (void) lb;
        const GArgs<T> &a = a_[r];
        mass_finalize_body(a.R, s, a.lsum, a.scal, a.logbuf, 0);
    });
}

#define CF_INST_B(T)                                                                                              \
    template typename BSel<T>::kb_t pick_batched_divx_n<T>(int, int);                                             \
    template typename BSel<T>::kb_t pick_batched_ffty<T>(int);                                                     \
    template typename BSel<T>::kb_t pick_batched_fftxi<T>(int);                                                    \
    template typename BSel<T>::kb_t pick_batched_poisson<T>(int);                                                  \
    template typename BSel<T>::kb_t pick_batched_spike<T>(int);                                                    \
    template __global__ void kb_thomas<T>(const GArgs<T> *, const unsigned *, int, int);                          \
    template __global__ void kb_cfl<T>(const GArgs<T> *, const unsigned *, int, int);                             \
    template __global__ void kb_step_end<T>(const GArgs<T> *, const unsigned *, int, int);                        \
    template typename BSel<T>::divx_t_ pick_batched_divx<T>(int);                                                  \
    template __global__ void kb_rhs2t<T>(const GArgs<T> *, const unsigned *, int, int);                           \
    template typename BSel<T>::kb_t pick_batched_spike_sums<T>(int);                                               \
    template __global__ void kb_level_sum<T>(const GArgs<T> *, const unsigned *, int, int);                       \
    template __global__ void kb_mass_fast<T>(const GArgs<T> *, const unsigned *, int, int);
CF_INST_B(double)
CF_INST_B(float)
#undef CF_INST_B
