/*
 * cf_dev.h -- device-side types and helpers of the chanflow HIP library
 * (gfx950 / CDNA4, wave64).
 *
 * Device layout of every field is z-slowest: q = (k * ny + j) * nx + i
 * over interior cells only (i < nx, j < ny, k < nz).  Periodic x and y
 * neighbours are taken by index wrap inside the kernels, wall ghosts are
 * formed on the fly from the boundary closures, so no halo copies exist.
 * Cell fields (u, v, c, p) hold the nz cells; w holds the nz faces above
 * them, the last being the top wall (the bottom wall face is implicit 0).
 * The spectral array holds [nz][ny][nx/2+1] complex values.
 */
#ifndef CF_DEV_H
#define CF_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CF_NGT 3   /* offset of level 0 in the z metric tables (3 ghosts below) */

struct Grid {
    int nx, ny, nz, nkx;
    int order;
    double dx, dy;
};

/* per-substep coefficients (chanfast.c:459-463) */
struct SubCoef {
    double nuadt, nubdt, gamdt, zetdt, abdt;
    double kadt, kbdt;            /* scalar: (nu/Pr) alpha dt, (nu/Pr) beta dt */
};

/* Thomas factors of one constant-coefficient tridiagonal system:
 * d_0 = r_0 ibet_0, d_j = (r_j - a_j d_{j-1}) ibet_j, x_j = d_j - gam_{j+1} x_{j+1} */
struct Thomas {
    const double *a, *ibet, *gam;
};

/* real <-> complex type pairs of the two storage precisions */
/* A pointer into global memory, typed so.  Pointers read out of memory
 * (a batch's argument table, a run's parameter block) are generic, and
 * accesses through them compile to flat instructions, which also hold the
 * LDS wait counter and never take the scalar path; held as gp<T> every use
 * compiles to global (or, when uniform and read-only, scalar) loads. */
template <typename P>
using gp = __attribute__((address_space(1))) P *;
/* read-only tables (written by the host before the launch, never by a
 * kernel): the constant address space lets uniform reads compile to scalar
 * loads even inside a persistent loop whose stores the compiler cannot
 * prove disjoint from them */
template <typename P>
using cp = __attribute__((address_space(4))) const P *;

template <typename T> struct Cplx;
template <> struct Cplx<double> { typedef double2 type; };
template <> struct Cplx<float> { typedef float2 type; };
template <typename C> struct Real;
template <> struct Real<double2> { typedef double type; };
template <> struct Real<float2> { typedef float type; };

template <typename C> static __device__ __forceinline__ C mkc(typename Real<C>::type x, typename Real<C>::type y)
{
    C c;
    c.x = x;
    c.y = y;
    return c;
}
/* a complex element as clang's native two-lane vector: loads and stores of
 * it go through address-space-typed pointers (gp, cp), which the HIP vector
 * types' constructors and assignments (generic references) cannot */
template <typename C> using cv = typename Real<C>::type __attribute__((ext_vector_type(2)));
template <typename C> static __device__ __forceinline__ C cx_of(const cv<C> v) { return mkc<C>(v.x, v.y); }
template <typename C> static __device__ __forceinline__ cv<C> cv_of(const C c)
{
    cv<C> v;
    v.x = c.x;
    v.y = c.y;
    return v;
}
template <typename C> static __device__ __forceinline__ C cadd(C a, C b) { return mkc<C>(a.x + b.x, a.y + b.y); }
template <typename C> static __device__ __forceinline__ C csub(C a, C b) { return mkc<C>(a.x - b.x, a.y - b.y); }
template <typename C> static __device__ __forceinline__ C cmul(C a, C b)
{
    return mkc<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C, typename S> static __device__ __forceinline__ C cscale(C a, S s)
{
    typedef typename Real<C>::type R;
    return mkc<C>(a.x * (R) s, a.y * (R) s);
}
/* multiply by -i*sg (sg = +1 forward, -1 inverse) */
template <typename C, typename S> static __device__ __forceinline__ C cmul_mi(C a, S sg)
{
    typedef typename Real<C>::type R;
    return mkc<C>((R) sg * a.y, -(R) sg * a.x);
}

static __device__ __forceinline__ int wrapm(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* ------------------------------------------------------------------ FFT
 * In-place mixed-radix Stockham transform of `nlines` lines of length n
 * held in LDS (line l at buf[l * ld], ld >= n), after the radix-r pass formulation
 * of Stockham autosort (input i + m n/r, output (i - k) r + k + q p with
 * k = i mod p).  Radices 2, 3, 4, 5.  tw[j] = exp(-2 pi i j / n) from host
 * libm.  sg = +1 forward (exp(-i ...)), -1 inverse; unnormalised.  Each
 * stage reads all of a thread's butterfly inputs into registers, barriers,
 * then writes, so one LDS buffer suffices; MAXE bounds the elements per
 * thread (nlines * n <= MAXE * blockDim.x). */
template <int R, typename C>
static __device__ __forceinline__ void dft_small(C *a, double sgd)
{
    typedef typename Real<C>::type F;
    const F sg = (F) sgd;
    if (R == 2) {
        C t = a[1];
        a[1] = csub(a[0], t);
        a[0] = cadd(a[0], t);
    } else if (R == 4) {
        C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        C t2 = cadd(a[1], a[3]), t3 = cmul_mi(csub(a[1], a[3]), sg);
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    } else if (R == 3) {
        const F C3 = (F) -0.5, S3 = (F) 0.86602540378443864676;       /* cos, sin of 2 pi / 3 */
        C t = cadd(a[1], a[2]);
        C d = cmul_mi(csub(a[1], a[2]), sg * S3);
        C m0 = mkc<C>(a[0].x + C3 * t.x, a[0].y + C3 * t.y);
        a[0] = cadd(a[0], t);
        a[1] = cadd(m0, d);
        a[2] = csub(m0, d);
    } else {  /* R == 5 */
        const F C51 = (F) 0.30901699437494742410, C52 = (F) -0.80901699437494742410;
        const F S51 = (F) 0.95105651629515357212, S52 = (F) 0.58778525229247312917;
        C t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
        C d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
        C m1 = mkc<C>(a[0].x + C51 * t1.x + C52 * t2.x, a[0].y + C51 * t1.y + C52 * t2.y);
        C m2 = mkc<C>(a[0].x + C52 * t1.x + C51 * t2.x, a[0].y + C52 * t1.y + C51 * t2.y);
        C s1 = cmul_mi(mkc<C>(S51 * d1.x + S52 * d2.x, S51 * d1.y + S52 * d2.y), sg);
        C s2 = cmul_mi(mkc<C>(S52 * d1.x - S51 * d2.x, S52 * d1.y - S51 * d2.y), sg);
        a[0] = cadd(a[0], cadd(t1, t2));
        a[1] = cadd(m1, s1);
        a[4] = csub(m1, s1);
        a[2] = cadd(m2, s2);
        a[3] = csub(m2, s2);
    }
}

/* LDS slot of element i of a transform line whose length is a multiple of
 * 64 (on): within each 64-element block, bits 0-1 and 2-3 of i are XORed
 * with bits 4-5.  With 16-byte elements the strided writes of the first
 * two radix-4 passes (stride 4, and blocks of 4 at stride 16) then fall on
 * 16 distinct bank groups per 16 lanes instead of 4, while contiguous runs
 * stay conflict-free.  A permutation within each block, so any length
 * that is a multiple of 64 stays inside its line; other lengths keep the
 * plain order. */
static __device__ __forceinline__ int fsw(int i, bool on)
{
    return on ? i ^ (((i >> 4) & 3) * 5) : i;
}
static __host__ __device__ __forceinline__ bool fsw_on(int n) { return (n & 63) == 0; }

template <int R, int MAXE, typename C>
static __device__ __forceinline__ void fft_stage(C *buf, int n, int ld, int nlines, int p,
                                                 const C *__restrict__ tw, double sg)
{
    typedef typename Real<C>::type F;
    constexpr int NB = (MAXE + R - 1) / R;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int T = n / R;
    const int nb = nlines * T;
    const int tstride = n / (p * R);
    const bool on = fsw_on(n);
    C y[NB][R];
#pragma unroll
    for (int it = 0; it < NB; it++) {
        const int b = tid + it * nt;
        if (b < nb) {
            const int line = b / T, i = b - line * T;
            const int k = i % p;
            const C *src = buf + line * ld;
#pragma unroll
            for (int m = 0; m < R; m++) {
                C x = src[fsw(i + m * T, on)];
                if (m > 0 && k > 0) {
                    C w = tw[(k * m * tstride) % n];
                    w.y *= (F) sg;
                    x = cmul(x, w);
                }
                y[it][m] = x;
            }
            dft_small<R>(y[it], sg);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NB; it++) {
        const int b = tid + it * nt;
        if (b < nb) {
            const int line = b / T, i = b - line * T;
            const int k = i % p;
            C *dstl = buf + line * ld;
            const int d0 = (i - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) dstl[fsw(d0 + q * p, on)] = y[it][q];
        }
    }
    __syncthreads();
}

template <int MAXE, typename C>
static __device__ void lds_fft(C *buf, int n, int ld, int nlines, cp<int> rad, int nrad,
                               const C *__restrict__ tw, double sg)
{
    int p = 1;
    for (int st = 0; st < nrad; st++) {
        const int r = rad[st];
        if (r == 4) fft_stage<4, MAXE>(buf, n, ld, nlines, p, tw, sg);
        else if (r == 2) fft_stage<2, MAXE>(buf, n, ld, nlines, p, tw, sg);
        else if (r == 3) fft_stage<3, MAXE>(buf, n, ld, nlines, p, tw, sg);
        else fft_stage<5, MAXE>(buf, n, ld, nlines, p, tw, sg);
        p *= r;
    }
}

/* ---- compile-time plans: the same radix order as the host's factor_radix
 * (4s first, then 2, 3, 5), so butterfly index math folds to constants */
constexpr int ct_rad_count(int n)
{
    int c = 0;
    while (n % 4 == 0 && n > 4) { n /= 4; c++; }
    if (n == 4) return c + 1;
    while (n % 2 == 0) { n /= 2; c++; }
    while (n % 3 == 0) { n /= 3; c++; }
    while (n % 5 == 0) { n /= 5; c++; }
    return c;
}

constexpr int ct_rad_at(int n, int idx)
{
    int c = 0;
    while (n % 4 == 0 && n > 4) { if (c == idx) return 4; n /= 4; c++; }
    if (n == 4) return 4;
    while (n % 2 == 0) { if (c == idx) return 2; n /= 2; c++; }
    while (n % 3 == 0) { if (c == idx) return 3; n /= 3; c++; }
    while (n % 5 == 0) { if (c == idx) return 5; n /= 5; c++; }
    return 1;
}

/* Twiddles of a compiled plan in stage order: stage s (radix R, span P)
 * keeps w^(k m N / (P R)), k < P, 1 <= m < R, at OFF_s + (m - 1) P + k, so
 * the lanes of a pass (consecutive k) read consecutive slots instead of a
 * stride of m N / (P R) slots, which put up to four lanes of a 16-lane
 * group on one LDS bank group; the stages fill N - 1 slots.  A run-time
 * plan (N = 0) keeps the plain table w^j, j < n. */
template <int N>
static __device__ __forceinline__ int tw_count(int n) { return N > 0 ? N - 1 : n; }
template <int N>
static __device__ __forceinline__ int tw_src(int q)
{
    if constexpr (N > 0) {
        int res = 0, P = 1;
        bool found = false;
#pragma unroll
        for (int s = 0; s < ct_rad_count(N); s++) {
            const int R = ct_rad_at(N, s), cnt = P * (R - 1);
            if (!found && q < cnt) {
                const int m = q / P + 1, k = q - (m - 1) * P;
                res = k * m * (N / (P * R));
                found = true;
            }
            if (!found) q -= cnt;
            P *= R;
        }
        return res;
    } else {
        return q;
    }
}

template <int N, int R, int P, int OFF, int MAXE, typename C>
static __device__ __forceinline__ void fft_stage_ct(C *buf, int ld, int nlines,
                                                    const C *__restrict__ tw, double sg)
{
    typedef typename Real<C>::type F;
    constexpr int NB = (MAXE + R - 1) / R;
    constexpr int T = N / R;
    constexpr bool on = (N & 63) == 0;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nb = nlines * T;
    C y[NB][R];
#pragma unroll
    for (int it = 0; it < NB; it++) {
        const int b = tid + it * nt;
        if (b < nb) {
            const int line = b / T, i = b - line * T;
            const int k = i % P;
            const C *src = buf + line * ld;
#pragma unroll
            for (int m = 0; m < R; m++) {
                C x = src[fsw(i + m * T, on)];
                if (m > 0 && P > 1) {
                    C w = tw[OFF + (m - 1) * P + k];
                    w.y *= (F) sg;
                    x = cmul(x, w);
                }
                y[it][m] = x;
            }
            dft_small<R>(y[it], sg);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NB; it++) {
        const int b = tid + it * nt;
        if (b < nb) {
            const int line = b / T, i = b - line * T;
            const int k = i % P;
            C *dstl = buf + line * ld;
            const int d0 = (i - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; q++) dstl[fsw(d0 + q * P, on)] = y[it][q];
        }
    }
    __syncthreads();
}

template <int N, int MAXE, int ST = 0, int P = 1, int OFF = 0, typename C>
static __device__ __forceinline__ void lds_fft_ct(C *buf, int ld, int nlines,
                                                  const C *__restrict__ tw, double sg)
{
    if constexpr (ST < ct_rad_count(N)) {
        constexpr int R = ct_rad_at(N, ST);
        fft_stage_ct<N, R, P, OFF, MAXE>(buf, ld, nlines, tw, sg);
        lds_fft_ct<N, MAXE, ST + 1, P * R, OFF + P * (R - 1)>(buf, ld, nlines, tw, sg);
    }
}

/* N > 0: compile-time plan (twiddles in stage order); N == 0: runtime
 * plan (any 2^a 3^b 5^c, plain twiddle table) */
template <int N, int MAXE, typename C>
static __device__ __forceinline__ void fft_lines(C *buf, int n, int ld, int nlines, cp<int> rad,
                                                 int nrad, const C *__restrict__ tw, double sg)
{
    if constexpr (N > 0) lds_fft_ct<N, MAXE>(buf, ld, nlines, tw, sg);
    else lds_fft<MAXE>(buf, n, ld, nlines, rad, nrad, tw, sg);
}

template <int CTRL, int ROWS>
static __device__ __forceinline__ double dpp_f64(double x)
{
    /* DPP moves the two 32-bit halves; lanes outside ROWS read 0 */
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}

/* sum over the 64 lanes in a fixed order (the result does not depend on
 * scheduling), returned by every lane: pairs, quads, half rows and rows by
 * DPP permutes within each row of 16 lanes, then rows 0 + 1 and 2 + 3 and
 * the two halves by row broadcasts; lane 63 ends with the total */
static __device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0xB1, 0xf>(v);        /* quad_perm [1, 0, 3, 2] */
    v += dpp_f64<0x4E, 0xf>(v);        /* quad_perm [2, 3, 0, 1] */
    v += dpp_f64<0x141, 0xf>(v);       /* row_half_mirror */
    v += dpp_f64<0x140, 0xf>(v);       /* row_mirror */
    v += dpp_f64<0x142, 0xa>(v);       /* row_bcast:15 into rows 1 and 3 */
    v += dpp_f64<0x143, 0xc>(v);       /* row_bcast:31 into rows 2 and 3 */
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

/* the maximum over a 64-lane wave of non-negative values (lanes a DPP
 * step does not reach read 0, which leaves such a maximum unchanged),
 * and over a 256-thread workgroup (sh: 4 doubles of LDS, one barrier) */
static __device__ __forceinline__ double wave_max_nonneg(double v)
{
    v = fmax(v, dpp_f64<0xB1, 0xf>(v));
    v = fmax(v, dpp_f64<0x4E, 0xf>(v));
    v = fmax(v, dpp_f64<0x141, 0xf>(v));
    v = fmax(v, dpp_f64<0x140, 0xf>(v));
    v = fmax(v, dpp_f64<0x142, 0xa>(v));
    v = fmax(v, dpp_f64<0x143, 0xc>(v));
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}
static __device__ __forceinline__ double block_max256_nonneg(double v, double *sh)
{
    const double w = wave_max_nonneg(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = w;
    __syncthreads();
    return fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

/* the sum of one value per thread over a 256-thread workgroup: a DPP sum
 * per wave, then the four wave sums in order (one barrier, fixed order);
 * every thread returns the total.  sh: 4 doubles of LDS. */
static __device__ __forceinline__ double block_sum256(double v, double *sh)
{
    const double w = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = w;
    __syncthreads();
    return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

/* the sums over a wave's 64 lanes of each of a lane's LM <= 32 values
 * v[t], by recursive halving: at the exchange over lane bit m = 32, 16,
 * ..., 2 a lane keeps the half of its values that bit selects, adds its
 * partner's copy of that half and passes on the other half; a last
 * exchange over bit 0 completes the sums.  Lanes 2t and 2t + 1 return the
 * sum of v[t] (t < 32), each the same fixed tree over the 64 lanes.  32
 * exchanges in place of 6 per value. */
template <int LM>
static __device__ __forceinline__ double wave_sums(const double *v)
{
    static_assert(LM <= 32, "wave_sums: at most 32 values per lane");
    double a[32];
#pragma unroll
    for (int t = 0; t < 32; t++) a[t] = t < LM ? v[t] : 0.0;
    const int lane = (int) threadIdx.x & 63;
#pragma unroll
    for (int st = 0; st < 5; st++) {
        const int m = 32 >> st, half = 16 >> st;
        const bool up = (lane & m) != 0;
#pragma unroll
        for (int i = 0; i < half; i++) {
            const double send = up ? a[i] : a[i + half];
            const double keep = up ? a[i + half] : a[i];
            a[i] = keep + __shfl_xor(send, m, 64);
        }
    }
    return a[0] + __shfl_xor(a[0], 1, 64);
}

/* n / d for 0 <= n < 2^31 and 1 <= d < 2^20, given rd = 1 / d correctly
 * rounded to double: n rd differs from n / d by at most (n / d) 2^-52 <
 * 2^-21, less than the distance 1 / d from a non-integer n / d to the
 * nearest integer, so its truncation is the quotient; when d divides n it
 * may be one less, which the remainder test corrects.  A few instructions
 * in place of the integer division sequence when d is a run-time value. */
static __device__ __forceinline__ int qdiv(int n, int d, double rd)
{
    const int q = (int) ((double) n * rd);
    return n - q * d >= d ? q + 1 : q;
}

/* Workgroups are dispatched round-robin over the 8 XCDs (each with its own
 * L2); renumber so that consecutive logical blocks share an XCD and the
 * halo rows they exchange stay in one L2. */
static __device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblocks)
{
    const unsigned per = nblocks / 8u;
    if (per == 0 || b >= per * 8u) return b;
    return (b % 8u) * per + b / 8u;
}

#endif
