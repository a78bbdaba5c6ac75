/*
 * cf_kernels.h -- kernel parameter blocks and kernel prototypes shared
 * by cf_kernels.hip (definitions) and cf_run.hip (launches).
 */
#ifndef CF_KERNELS_H
#define CF_KERNELS_H

#include "cf_dev.h"
#include "cf_tables.h"

/* slots of the per-run device scalar block */
enum {
    SC_M0 = 0,      /* mass-flux forcing of substeps 0..2 */
    SC_MC0 = 3,     /* scalar-mean forcing of substeps 0..2 */
    SC_STEP = 6,    /* completed steps */
    SC_CFL = 7,     /* cfl at the start of the current step */
    SC_ZERO = 8,    /* always 0 */
    SC_CFL0 = 9,    /* cflmode 1: cfl of the state at the start of a step, with the previous dt */
    SC_TICKET = 10, /* k_level_mass: workgroups done, as a 64-bit counter (0 between launches) */
    SC_N = 16
};

#define CF_MAX_RAD 16
#ifndef CF_RHS_WAVES
#define CF_RHS_WAVES 3          /* occupancy target (waves per SIMD) of k_rhs2t */
#endif
/* xy tile of the fast-path RHS (one thread per column, 256 threads; 32 x 8 measured faster
 * than 64 x 4 and 16 x 16 at 256x193x256) */
#define CF_RT_X 32
#define CF_RT_Y 8
#ifndef CF_RHS_KC
#define CF_RHS_KC 16             /* target levels per z-chunk of k_rhs2t (16: 264 us, 24: 270 us, 32: 266 us at 256x193x256) */
#endif
#ifndef CF_RHS_KC_BATCH
#define CF_RHS_KC_BATCH 24       /* ... of kb_rhs2t, whose launches cover many runs */
#endif
/* xy tile of the fused order-4 RHS (k_rhs4t); nx and ny must be multiples */
#define CF_R4_X 32
#define CF_R4_Y 8
#ifndef CF_R4_KC
#define CF_R4_KC 8               /* target levels per z-chunk of the order-4 RHS (8: 153 us, 16: 164 us, 32: 198 us per launch at 192x129x192) */
#endif
#ifndef CF_R4_WAVES
#define CF_R4_WAVES 2            /* occupancy target (waves per SIMD) of the order-4 RHS */
#endif
#define CF_SPIKE_LMAX 32        /* rows per chunk held in registers */
#define CF_SPIKE_SMAX 8         /* chunks (waves) per workgroup */
#define CF_F64_KLOW 4           /* FP32 runs solve Poisson modes |kx|, |ky| <= this in FP64 */

/* Partitioned (spike) solve of one constant-coefficient tridiagonal
 * system split into S chunks of L rows: per-chunk Thomas factors, the
 * left/right spike vectors and the inverse of the 2S x 2S interface
 * system, all column-independent (host-built). */
struct SpikeTab {
    int n, S, L;
    const double *a, *ib, *gam, *vl, *wr, *minv;
};

/* order-4 partitioned z solve of one system over n unknowns, S chunks of L
 * rows (cf_spike4.hip): f = 7 S L chunk-local heptdag factors
 * (a | al | be | iden | p | q | r), vw = 6 S L spikes (three left, three
 * right), minv = the inverse interface system, 6 CF_SPIKE4_SMAX square */
struct Spike4Tab {
    int n, S, L;
    const double *f, *vw, *minv;
};

/* ghost closures of cell (Dirichlet) and face fields at both walls */
struct Clos {
    Clos1 cell[2], face[2];     /* [0] bottom, [1] top */
};

/* factors of one wall-reduced order-4 z system over n unknowns:
 * t[0..7n) = a | al | be | iden | p | q | r (cf_tables.h, hept_factor) */
struct HeptTab {
    int n;
    const double *t;
};

/* order-4 Poisson z-tables for one dt: the wall-reduced bands at zero
 * horizontal wavenumber for the Neumann bottom (band[0]) and the
 * Dirichlet-mean bottom of the (0,0) mode (band[1]), 7 nz each (column
 * offsets -3..+3, level fastest), the diagonal shifts per unit modified
 * wavenumber and the volumes (pressure.c:357-444) */
struct Pois4Tab {
    const double *band[2];
    const double *ex, *ey, *dvol;
};

/* Everything a substep kernel needs that is fixed for the run: passed by
 * value (a few hundred bytes of kernel arguments). */
struct RunDev {
    Grid g;
    const double *dzt;          /* nz + 6: dzt[k + 3] = dz of level k (mirrored ghosts) */
    SubCoef sc[3];
    double nu, prandtl;
    double ubc, vbc, cbot, ctop;
    double lift[3][4];          /* boundary lift of u (bot, top) and v (bot, top), per substep */
    int m_prev_slot[3];         /* scalar slot holding the previous substep's forcing */
    const double *th;           /* Thomas factors [s][type 0..3][a, ibet, gam][nz] */
    SpikeTab sp[3][3];          /* [s][type 0 cell-nu, 1 cell-kappa, 2 face-nu] */
    const double *rdz;          /* 1 / dz, level-indexed like dzt (offset CF_NGT) */
    const double *rdzh;         /* 1 / (0.5 (dz_k + dz_k+1)), offset CF_NGT */
    double rdx, rdy, rdxdy;
    const double *stat1;        /* [s][kind][vol1, bot1, top1, cl1] */
    const double *h1;           /* [s][kind][nz + 2] unit-forcing profile, level k at k + 1 */
    int cmode[2];
    double ctarget[2], cfixed[2];
    int logcap;
    double cfl_log;             /* > 0: the cfl.dat value of the step (cflmode 1 logs want_cfl) */
    /* FFT plans */
    int radx[CF_MAX_RAD], nradx, rady[CF_MAX_RAD], nrady;
    const double2 *twx, *twy;
    const float2 *twx32, *twy32;  /* the same twiddles rounded to FP32 */
    const double *lamx, *lamy;  /* modified wavenumbers 2 - 2 cos(2 pi k / n) */
    int fx_pairs;               /* complex x lines per block */
    int fy_cols;                /* kx columns per y-FFT block */
    /* general path (order 4 and/or the scalar) */
    Clos cl;
    int scalar;
    double betg[3];
    double kap;                 /* scalar diffusivity nu / Pr */
    const double *liftg;        /* [s][field u, v, w, c][nz] boundary-lift profiles */
    HeptTab hp[3][3];           /* order-4 implicit z factors [s][type 0 cell-nu, 1 cell-kappa, 2 face-nu] */
    Spike4Tab s4[3][3];         /* ... as partitioned-solve tables (n = 0: none) */
    int implicit;               /* 1: every diffusion term implicit (channel2/4): x-y diffusion in h, not f */
    /* FP32 storage on the fused path: u is stored as u - U(z), U a per-level
     * FP64 profile, so storage rounding scales with the deviation from U
     * instead of with u.  Layout: [nz + 2] U with the wall ghosts at 0 and
     * nz + 1, then [3][nz] U's share of each substep's implicit right-hand
     * side.  Null: no offset (FP64, or the general path). */
    const double *uoff;
};

/* A workgroup's coordinates in one run's launch grid: blockIdx / gridDim
 * of a single-run launch, or decoded from a batched launch's virtual block. */
struct Blk {
    unsigned x, y, z, nx, ny, nz;
};
static __device__ __forceinline__ Blk blk_native()
{
    Blk b;
    b.x = blockIdx.x; b.y = blockIdx.y; b.z = blockIdx.z;
    b.nx = gridDim.x; b.ny = gridDim.y; b.nz = gridDim.z;
    return b;
}

/* launch stages of a batched step; each run's grid per stage is dims[stage] */
enum { ST_CELL = 0, ST_SOLVE, ST_PLANE, ST_MASS, ST_ONE, ST_FFTX, ST_FFTY, ST_MODE, ST_RHS2, ST_LEVEL, CF_NST };
#define CF_BATCH_MAX 256        /* runs per batched launch */

/* Where a substep's projected velocity comes from: the previous
 * substep's solve output, its pressure, mass-flux profile and forcing
 * (h1, m may be null: no correction) and its projection dt. */
template <typename T> struct ProjSrcT {
    const T *us, *vs, *ws, *p;               /* state of storage type T */
    const double *h1, *m;
    double dtp;
};
typedef ProjSrcT<double> ProjSrc;

/* z-tables of the Poisson operator for one dt (pressure.c:318-339) */
struct PoisTab {
    const double *a, *c, *bz, *ex, *ey, *dvol;
    double db_neu, db_dir, db_top;
};

/* z-solve chunk lengths compiled for the spike kernels: a run's tables
 * hold L rows per chunk, L the smallest of these with CF_SPIKE_SMAX L >= nz
 * (0: nz beyond CF_SPIKE_SMAX CF_SPIKE_LMAX, Thomas instead) */
#define CF_SPIKE_LSET(X) X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(20) X(24) X(28) X(32)
static inline int spike_L(int nz)
{
    static const int set[] = { 4, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32 };
    /* six chunks where a compiled length gives exactly six (measured faster
     * than seven at nz = 65 and 129; nz = 97 would round to five, slower) */
    const int need6 = (nz + 5) / 6;
    for (int l : set)
        if (l >= need6) {
            if ((nz + l - 1) / l == 6) return l;
            break;
        }
    const int need = (nz + CF_SPIKE_SMAX - 1) / CF_SPIKE_SMAX;
    for (int l : set)
        if (l >= need) return l;
    return 0;
}

/* order-4 spike solve: at most CF_SPIKE4_SMAX chunks of L rows, L a
 * compiled length; one L serves the cell (nz) and face (nz - 1) systems,
 * and every chunk but a system's last has L >= 6 rows, the last >= 3, so
 * a chunk's three-row couplings reach only its neighbours (0: none fits,
 * the sequential sweep instead) */
#define CF_SPIKE4_SMAX 8
#define CF_SPIKE4_LSET(X) X(6) X(8) X(10) X(12) X(14) X(16) X(18) X(20) X(24) X(28) X(32)
static inline int spike4_L(int nz)
{
    static const int set[] = { 6, 8, 10, 12, 14, 16, 18, 20, 24, 28, 32 };
    auto fits = [](int n, int l) {
        const int S = (n + l - 1) / l;
        return S <= CF_SPIKE4_SMAX && (S == 1 || n - (S - 1) * l >= 3);
    };
    for (int l : set)
        if (fits(nz, l) && fits(nz - 1, l)) return l;
    return 0;
}

/* LDS of the FFT kernels: lines of n elements at stride fft_ld(n), then
 * the n twiddles of the passes, staged once per workgroup */
#ifndef CF_FFT_PAD
#define CF_FFT_PAD 1
#endif
static inline __host__ __device__ int fft_ld(int n) { return n + CF_FFT_PAD; }
/* FFT lengths done in two register passes, n = P1 P2 (P1 = fft2_p1(n),
 * 0 for the Stockham passes): the y FFT (CF_FY2), the inverse x FFT and
 * the order-2 divergence + x FFT (CF_FX2) where they measured faster than
 * the Stockham passes -- the inverse x FFT in FP32 and for lengths up to
 * 128 in FP64, the divergence in FP32 at 256 (esz: bytes per stored real) */
#ifndef CF_FY2
#define CF_FY2 1
#endif
#ifndef CF_FX2
#define CF_FX2 1
#endif
static constexpr __host__ __device__ int fft2_p1(int n) { return n == 256 ? 16 : (n == 128 || n == 64 ? 8 : 0); }
static constexpr __host__ __device__ int fy2_p1(int n) { return CF_FY2 ? fft2_p1(n) : 0; }
static constexpr __host__ __device__ int fx2_p1(int n) { return CF_FX2 ? fft2_p1(n) : 0; }
static constexpr __host__ __device__ bool fxi2_on(int n, int esz) { return fx2_p1(n) > 0 && (esz == 4 || n <= 128); }
static constexpr __host__ __device__ bool divx2_on(int n, int esz) { return fx2_p1(n) > 0 && esz == 4 && n == 256; }
/* y lines sit in LDS as P1 rows of P2 + 1 */
static inline __host__ __device__ int fy_ld(int n) { return fy2_p1(n) ? n + fy2_p1(n) : fft_ld(n); }
/* x line pairs: P1 rows of P2 + 1 and 2 more, so the unpacking writes of
 * neighbouring pairs (lanes along the pair) fall on distinct bank groups */
static constexpr __host__ __device__ int fx_ld(int n, int esz)
{
    return fxi2_on(n, esz) ? n + fx2_p1(n) + 2 : n + CF_FFT_PAD;
}
/* line pairs per x-FFT workgroup: CF_FX_ELEMS complex elements (at least
 * one pair); 2048 where the two-pass forms run, so that a pass has 128 or
 * 256 (pair, column) tasks.  The Stockham kernels take the pair count
 * from the run (RunDev::fx_pairs); the two-pass ones compile it in, and
 * for their lengths it does not depend on the element size, except at 256,
 * whose two-pass forms are FP32 only (FP32 runs launch FP64 x FFTs only
 * through the Stockham kernels there). */
#ifndef CF_FX_ELEMS
#define CF_FX_ELEMS 1024
#endif
static constexpr __host__ __device__ int fx_pairs_of(int nx, int esz)
{
    return fxi2_on(nx, esz) ? 2048 / nx : (CF_FX_ELEMS / nx > 0 ? CF_FX_ELEMS / nx : 1);
}

#ifndef CF_P2S_NC
#define CF_P2S_NC 16
#endif
#ifndef CF_P2S_MW
#define CF_P2S_MW 8
#endif
#ifndef CF_P2S_XCD
#define CF_P2S_XCD 1
#endif
/* order-2 Poisson: the partitioned solve (at most 16 z-chunks of L rows
 * per mode, 8 modes per 256-thread workgroup), L a compiled chunk length
 * >= nz / 16, covers nz <= 512; beyond that one thread per mode (256
 * modes per workgroup). */
#define CF_P2S_LSET(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
                       X(20) X(24) X(28) X(32)
static inline int pois2_L_of(int need)
{
    static const int set[] = { 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 20, 24, 28, 32 };
    for (int l : set)
        if (l >= need) return l;
    return 0;
}
/* chunk length for nz (0: one thread per mode) */
static inline int pois2_L(int nz) { return pois2_L_of((nz + CF_P2S_NC - 1) / CF_P2S_NC); }
/* dynamic LDS of the partitioned solve: its six z tables */
static inline size_t pois2_lds(int L, int nz) { return L > 0 ? 6 * sizeof(double) * (size_t) nz : 0; }
/* workgroup shape: CF_P2S_NC z-chunks x CF_P2S_MW modes x 2 parts; a
 * z-level of the group is 2 CF_P2S_MW contiguous reals (128 B in FP64;
 * 16 modes per workgroup measured slower at 256x193x256) */
#define CF_P2S_GW (2 * CF_P2S_MW)
#define CF_P2S_THREADS (CF_P2S_NC * CF_P2S_GW)
/* modes per workgroup: CF_P2S_MW partitioned, or one per thread (L = 0) */
static inline size_t pois2_modes_per_block(int L) { return L > 0 ? CF_P2S_MW : CF_P2S_THREADS; }
static inline unsigned pois2_blocks(int L, size_t M)
{
    const size_t per = pois2_modes_per_block(L);
    return (unsigned) ((M + per - 1) / per);
}
/* FP32 storage: the leading workgroups of the order-2 Poisson grid that
 * can hold an FP64 island mode (kx <= CF_F64_KLOW: modes below
 * (CF_F64_KLOW + 1) ny in the kx-major order) */
static inline unsigned pois2_island_blocks(int L, int ny, size_t M)
{
    const size_t per = pois2_modes_per_block(L);
    const size_t n = ((size_t) (CF_F64_KLOW + 1) * (size_t) ny + per - 1) / per;
    const unsigned all = pois2_blocks(L, M);
    return n < all ? (unsigned) n : all;
}

/* One run of a batched launch: its parameter block, buffers (storage
 * precision T), Poisson tables and per-stage grid shapes.  An array of
 * these lives in device memory; batched kernels read the entry of the run
 * a virtual block belongs to. */
template <typename T> struct GArgs {
    RunDev R;
    T *u, *v, *w, *c, *ru, *rv, *rw, *rc, *fa[4], *fb[4], *p;
    typename Cplx<T>::type *X;
    double *G, *S4, *scal, *logbuf, *part, *cflpart, *lsum;
    double dt;                  /* step */
    double dt_adhoc;            /* dt of a stand-alone projection (substep index 3) */
    int nb_plane, gblocks;      /* gblocks: CFL partials of the run's RHS launch */
    int rhs_kc;                 /* levels per z-chunk of the fast-path RHS */
    PoisTab pt[4];              /* substeps 0..2, stand-alone projection */
    Pois4Tab pt4[4];
    unsigned dims[CF_NST][3];
};

/* Persistent loop of a batched launch: vstart[0..nrun] is the prefix of the
 * runs' virtual block counts for this stage; each workgroup walks virtual
 * blocks vb = blockIdx.x, + gridDim.x, ..., finds the run by binary search
 * and calls fn(run, local block).  A barrier separates iterations (bodies
 * reuse their LDS). */
template <typename F>
static __device__ __forceinline__ void batch_loop(const unsigned *__restrict__ vstart, int nrun, F &&fn)
{
    __shared__ unsigned lvs[CF_BATCH_MAX + 1];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int nt = blockDim.x * blockDim.y;
    for (int t = tid; t <= nrun; t += nt) lvs[t] = vstart[t];
    __syncthreads();
    /* every thread searches for the same virtual block, so the run and the
     * local block are workgroup-uniform: readfirstlane keeps them (and the
     * argument-table loads they index) in scalar registers */
    const unsigned total = (unsigned) __builtin_amdgcn_readfirstlane((int) lvs[nrun]);
    for (unsigned vb = blockIdx.x; vb < total; vb += gridDim.x) {
        int lo = 0, hi = nrun - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            const unsigned v = (unsigned) __builtin_amdgcn_readfirstlane((int) lvs[mid]);
            if (v <= vb) lo = mid; else hi = mid - 1;
        }
        const unsigned lb = vb - (unsigned) __builtin_amdgcn_readfirstlane((int) lvs[lo]);
        fn(lo, lb);
        __syncthreads();
    }
}

/* local block index -> Blk in a run's grid of the given stage */
template <typename T>
static __device__ __forceinline__ Blk blk_of(const GArgs<T> &a, int stage, unsigned lb)
{
    Blk b;
    b.nx = a.dims[stage][0]; b.ny = a.dims[stage][1]; b.nz = a.dims[stage][2];
    b.x = lb % b.nx;
    const unsigned r = lb / b.nx;
    b.y = r % b.ny;
    b.z = r / b.ny;
    return b;
}

/* off (may be null): a per-level profile subtracted on the way in and
 * added on the way out */
template <typename T> __global__ void k_io_to_dev(const double *io, T *dev, int nx, int ny, int nz, const double *off);
template <typename T> __global__ void k_dev_to_io(const T *dev, double *io, int nx, int ny, int nz, const double *off);
template <typename T>
__global__ void k_rhs2t(const RunDev R, const int s, const ProjSrcT<T> S, const T *gu, const T *gv, const T *gw,
                        const double *scal, T *fu, T *fv, T *fw, T *ru, T *rv, T *rw, double *cflpart, double cfldt,
                        int kc);
template <typename T> __global__ void k_thomas(const RunDev R, const int s, T *ru, T *rv, T *rw, T *rc);
template <typename T> using k_spike_t = void (*)(const RunDev, const int, T *, T *, T *, T *);
template <typename T> k_spike_t<T> pick_spike(int L);
template <typename T> k_spike_t<T> pick_spike4(int L);
template <typename T> using k_spike_sums_t = void (*)(const RunDev, const int, T *, T *, T *, double *);
template <typename T> k_spike_sums_t<T> pick_spike_sums(int L);
__global__ void k_level_sum(const double *plane, int nbx, double *lsum);
__global__ void k_mass_finalize(const RunDev R, const int s, const double *lsum, double *scal, double *logbuf,
                                int kind);
template <typename T>
__global__ void k_cfl_plain(const RunDev R, const T *u, const T *v, const T *w, double dt, double *part);
__global__ void k_cfl_reduce(const double *part, int nb, double *scal);
__global__ void k_level_mass(const RunDev R, const int s, const double *plane, int nbx, double *lsum, double *scal,
                             double *logbuf, int step_end);
__global__ void k_cfl_finish(const RunDev R, const double *part, int nb, double dt,
                             double *scal, double *logbuf);
__global__ void k_step_end(double *scal);
/* kernel pointer types of the FFT passes for storage precision T */
template <typename T> struct KSel {
    typedef typename Cplx<T>::type C;
    typedef void (*div_t)(const RunDev, const T *, const T *, const T *, const double *, const double *, C *);
    typedef void (*ffty_t)(const RunDev, C *, int);
    typedef void (*fftxi_t)(const RunDev, const C *, T *);
    typedef void (*pois2_t)(const RunDev, const PoisTab, C *, double *);
};
typedef void (*k_fftx_plain_t)(const RunDev, const double *, double2 *);
template <typename T> typename KSel<T>::div_t pick_div_fftx(int nx, int order);
k_fftx_plain_t pick_fftx_plain(int nx);
template <typename T> typename KSel<T>::ffty_t pick_ffty(int ny);
template <typename T> typename KSel<T>::fftxi_t pick_fftx_inv(int nx);
template <typename T> typename KSel<T>::pois2_t pick_poisson2(int lmax);
KSel<float>::pois2_t pick_poisson2_islands(int lmax);
template <typename T>
__global__ void k_project2(const RunDev R, const T *us, const T *vs, const T *ws, const double *h1, const double *mptr,
                           const T *p, double dtp, T *u, T *v, T *w);
template <typename T>
__global__ void k_maxdiv_part(const RunDev R, const T *u, const T *v, const T *w, double *part, int nb);

/* general path (cf_general.hip), storage and arithmetic precision T */
template <typename T> struct GSel {
    typedef void (*rhs_t)(const RunDev, const int, const T *, const T *, const T *, const T *, const T *,
                          const double *, T *, T *, double *, double);
    typedef void (*proj_t)(const RunDev, const T *, const T *, const T *, const T *, const double *,
                           const double *, const double *, const double *, const T *, double, T *, T *, T *, T *);
    rhs_t u, v, w, c;
    proj_t proj;
};
template <typename T> GSel<T> pick_general(int order, int scalar);

/* the order-4 explicit stage (cf_rhs4.hip): u and v in one launch, w and c
 * in another; grid (nx / CF_R4_X) * (ny / CF_R4_Y) * nch blocks of 256,
 * nch z-chunks */
template <bool SC, typename T>
__global__ void k_rhs4_uv(const RunDev R, const int s, const T *u, const T *v, const T *w, const T *c,
    const T *gu, const T *gv, const double *scal, T *fu, T *fv, T *ru, T *rv, double *cflpart, double cfldt, int nch);
template <bool SC, typename T>
__global__ void k_rhs4_wc(const RunDev R, const int s, const T *u, const T *v, const T *w, const T *c,
    const T *gw, const T *gc, const double *scal, T *fw, T *fc, T *rw, T *rc, int nch);

/* batched launches (cf_general.hip, cf_kernels.hip): a = device array of
 * nrun GArgs, vs = that stage's virtual-block prefix (nrun + 1) */
template <typename T> struct BSel {
    typedef void (*divx_t_)(const GArgs<T> *, const unsigned *, int, int);
    typedef void (*rhs_t)(const GArgs<T> *, const unsigned *, int, int);
    rhs_t u, v, w, c;
    rhs_t proj;
    divx_t_ divx;
    typedef void (*kb_t)(const GArgs<T> *, const unsigned *, int, int);
};
template <typename T> BSel<T> pick_batched(int order, int scalar);
template <typename T> typename BSel<T>::divx_t_ pick_batched_divx(int order);
template <typename T> __global__ void kb_poisson4(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
template <typename T> __global__ void kb_solve4(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
template <typename T> typename BSel<T>::kb_t pick_batched_spike(int L);
template <typename T> __global__ void kb_thomas(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
template <typename T> __global__ void kb_plane(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
template <typename T> __global__ void kb_mass(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
template <typename T> __global__ void kb_cfl(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
template <typename T> __global__ void kb_step_end(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
/* batched order-2 fast path (FP64 or FP32 storage) */
template <typename T> __global__ void kb_rhs2t(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
template <typename T> typename BSel<T>::kb_t pick_batched_spike_sums(int L);
template <typename T> __global__ void kb_level_sum(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
template <typename T> __global__ void kb_mass_fast(const GArgs<T> *a, const unsigned *vs, int nrun, int s);
/* batched FFT passes: one launch per group of runs of one length; lengths
 * in CF_BFFT_SIZES have their plan compiled in, the others share a launch
 * with the runtime plan (key 0) */
#define CF_BFFT_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
static inline int bfft_key(int n)
{
    switch (n) {
#define X_(N) case N: return N;
    CF_BFFT_SIZES(X_)
#undef X_
    default: return 0;
    }
}
template <typename T> typename BSel<T>::kb_t pick_batched_divx_n(int order, int n);
template <typename T> typename BSel<T>::kb_t pick_batched_ffty(int n);
template <typename T> typename BSel<T>::kb_t pick_batched_fftxi(int n);
template <typename T> typename BSel<T>::kb_t pick_batched_poisson(int L);
BSel<float>::kb_t pick_batched_poisson_islands(int L);
template <typename T> __global__ void k_gsolve4(const RunDev R, const int s, T *ru, T *rv, T *rw, T *rc);
template <typename T> __global__ void k_gplane(const RunDev R, const T *q0, const T *q1, double *part, int nb);
__global__ void k_gmass(const RunDev R, const int s, const double *part, int nb, double *scal, double *logbuf);
template <typename T>
__global__ void k_poisson4(const RunDev R, const Pois4Tab T_, typename Cplx<T>::type *Y, double *S);
/* the twisted form (nz >= 6): grid ceil(M / 64) blocks of 128 */
template <typename T>
__global__ void k_poisson4t(const RunDev R, const Pois4Tab T_, typename Cplx<T>::type *Y, double *S);

/* plane statistics (cf_stats.hip) */
template <typename T, int MODE>
__global__ void k_stat_plane(const T *a, const T *b, const double *ma, const double *mb, double en, int P, int nbl,
                             double *part);
__global__ void k_stat_finish(const double *part, int nbl, int nz, double scale, double *out, const double *off);
__global__ void k_stat_spec(const double2 *A, const double2 *B, int nx, int ny, int nkx, int nz, double *specx,
                            double *specy);
template <typename T> __global__ void k_stat_widen(const T *in, double *out, size_t n);
#define CF_STAT_NBL 64          /* workgroups (partials) per level of a plane statistic */

#endif
