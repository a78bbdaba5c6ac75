/*
 * cf_spike4.hip -- the order-4 implicit z solves (solve_helmholtz1d_u4 /
 * _w4 with heptdag.c:5-61, diffuseu.c:173-283, diffusew.c:175-287) as a
 * partitioned ("spike") solve held in registers.
 *
 * The systems are heptadiagonal with coefficients that depend on z only,
 * so everything but the right-hand side is built on the host once per
 * (dt, substep): the rows are cut into S chunks of L; each chunk has its
 * own heptdag factors (couplings across the chunk edges dropped), three
 * left spikes V = A_c^-1 B_left and three right spikes W = A_c^-1 B_right,
 * and the interface system over the first three and last three unknowns
 * of every chunk is inverted.  A workgroup owns 64 columns of one field
 * and S waves, wave c owning chunk c:
 *   - each lane solves its chunk's rows in registers (y = A_c^-1 h),
 *   - publishes the chunk's first and last three values,
 *   - forms the three unknowns below and the three above its chunk from
 *     the inverse interface system,
 *   - and writes x = y - V x_below - W x_above.
 * Every value is read once and written once (the sequential sweep reads
 * and writes it twice).  Same system as k_gsolve4; differs by round-off.
 */
#include "cf_kernels.h"

template <typename T, int LM>
__global__ __launch_bounds__(64 * CF_SPIKE4_SMAX) void k_spike4(const RunDev R, const int s,
    T *__restrict__ ru, T *__restrict__ rv, T *__restrict__ rw, T *__restrict__ rc)
{
    constexpr int MP = 6 * CF_SPIKE4_SMAX;
    __shared__ T itf[MP][64];
    const int nx = R.g.nx, ny = R.g.ny, nz = R.g.nz;
    const size_t P = (size_t) nx * ny;
    const int f = blockIdx.y;
    const int type = f == 2 ? 2 : (f == 3 ? 1 : 0);
    const Spike4Tab ST = R.s4[s][type];
    const int lane = threadIdx.x & 63;
    /* wave-uniform chunk index: the table reads below are scalar loads;
     * the tables hold LM rows per chunk, identity past the system's end */
    const int chunk = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool live = chunk < ST.S;
    const size_t col = (size_t) blockIdx.x * 64 + lane;
    const bool act = col < P;
    const gp<T> x = (gp<T>) (f == 0 ? ru : (f == 1 ? rv : (f == 2 ? rw : rc)));
    const int lo = chunk * LM;
    const int Lc = live ? min(LM, ST.n - lo) : 0;     /* rows of this chunk */
    const gp<T> xc = x + (act ? col : P - 1);
    const size_t SL = (size_t) ST.S * LM, tb = (size_t) chunk * LM;
    const cp<double> fA = (cp<double>) ST.f + tb, fAL = fA + SL, fBE = fA + 2 * SL, fID = fA + 3 * SL;
    const cp<double> fP = fA + 4 * SL, fQ = fA + 5 * SL, fR = fA + 6 * SL;
    const cp<double> v0 = (cp<double>) ST.vw + tb, v1 = v0 + SL, v2 = v0 + 2 * SL;
    const cp<double> w0 = v0 + 3 * SL, w1 = v0 + 4 * SL, w2 = v0 + 5 * SL;
    T y[LM];
    if (live) {                                   /* wave-uniform */
#pragma unroll
        for (int t = 0; t < LM; t++) y[t] = xc[(size_t) (lo + min(t, Lc - 1)) * P];
#pragma unroll
        for (int t = 0; t < LM; t++) y[t] = (act && t < Lc) ? y[t] : (T) 0;
        T s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
        for (int t = 0; t < LM; t++) {
            const T v = (y[t] - (T) fA[t] * s3 - (T) fAL[t] * s2 - (T) fBE[t] * s1) * (T) fID[t];
            y[t] = v;
            s3 = s2; s2 = s1; s1 = v;
        }
        T x1 = 0, x2 = 0, x3 = 0, l0 = 0, l1 = 0, l2 = 0;
#pragma unroll
        for (int t = LM - 1; t >= 0; t--) {
            const T v = y[t] - (T) fP[t] * x1 - (T) fQ[t] * x2 - (T) fR[t] * x3;
            y[t] = v;
            l0 = t == Lc - 3 ? v : l0;
            l1 = t == Lc - 2 ? v : l1;
            l2 = t == Lc - 1 ? v : l2;
            x3 = x2; x2 = x1; x1 = v;
        }
        itf[6 * chunk + 0][lane] = y[0];
        itf[6 * chunk + 1][lane] = y[1];
        itf[6 * chunk + 2][lane] = y[2];
        itf[6 * chunk + 3][lane] = l0;
        itf[6 * chunk + 4][lane] = l1;
        itf[6 * chunk + 5][lane] = l2;
    }
    __syncthreads();
    if (!live || !act) return;                    /* no barrier follows */
    const int m = 6 * ST.S;
    const cp<double> minv = (cp<double>) ST.minv;
    T E0 = 0, E1 = 0, E2 = 0, F0 = 0, F1 = 0, F2 = 0;
    const bool below = chunk > 0, above = chunk < ST.S - 1;
    const cp<double> re = minv + (size_t) (6 * (chunk - 1) + 3) * MP;   /* rows of the unknowns below */
    const cp<double> rf = minv + (size_t) (6 * (chunk + 1)) * MP;       /* ... and above */
#pragma unroll
    for (int q = 0; q < MP; q++) {
        if (q < m) {                              /* wave-uniform */
            const T v = itf[q][lane];
            if (below) {
                E0 += (T) re[q] * v; E1 += (T) re[MP + q] * v; E2 += (T) re[2 * MP + q] * v;
            }
            if (above) {
                F0 += (T) rf[q] * v; F1 += (T) rf[MP + q] * v; F2 += (T) rf[2 * MP + q] * v;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < LM; t++) {
        if (t < Lc) {                             /* wave-uniform */
            const T val = y[t] - ((T) v0[t] * E0 + (T) v1[t] * E1 + (T) v2[t] * E2)
                               - ((T) w0[t] * F0 + (T) w1[t] * F1 + (T) w2[t] * F2);
            xc[(size_t) (lo + t) * P] = val;
        }
    }
    if (f == 2 && chunk == ST.S - 1) xc[(size_t) (nz - 1) * P] = (T) 0;   /* top wall face */
}

template <typename T>
k_spike_t<T> pick_spike4(int L)
{
    switch (L) {
#define X_(N) case N: return k_spike4<T, N>;
        CF_SPIKE4_LSET(X_)
#undef X_
    default: return nullptr;
    }
}
template k_spike_t<double> pick_spike4<double>(int);
template k_spike_t<float> pick_spike4<float>(int);
