/*
 * cf_stats.hip -- plane statistics of a run's fields on the device: the
 * x-y means, correlations and central moments of get.c:100-174 and
 * get_xy_avg (get.c:281-307), and the one-dimensional cross-spectra of
 * get_xy_spec (get.c:5-97), at every interior level.
 *
 * Fields are in the device layout [nz][ny][nx]: a level is one contiguous
 * plane.  Plane sums are two-stage and in a fixed order (per workgroup a
 * strided thread sum and a tree, then the workgroups' partials in index
 * order), so a statistic is bitwise repeatable.
 */
#include "cf_kernels.h"

#include <hip/hip_runtime.h>

/* per (level, workgroup) partial of a plane sum: MODE 0 sum of a,
 * 1 sum of (a - ma)(b - mb), 2 sum of (a - ma)^en */
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_stat_plane(const T *a, const T *b, const double *ma, const double *mb,
                                                    double en, int P, int nbl, double *part)
{
    const int k = blockIdx.y, bl = blockIdx.x, tid = threadIdx.x;
    const T *ak = a + (size_t) k * P;
    const T *bk = MODE == 1 ? b + (size_t) k * P : nullptr;
    const double mak = MODE != 0 ? ma[k] : 0.0;
    const double mbk = MODE == 1 ? mb[k] : 0.0;
    const int per = (P + nbl - 1) / nbl;
    const int lo = bl * per, hi = lo + per < P ? lo + per : P;
    double s = 0.0;
    for (int t = lo + tid; t < hi; t += 256) {
        const double x = (double) ak[t];
        if (MODE == 0) s += x;
        else if (MODE == 1) s += (x - mak) * ((double) bk[t] - mbk);
        else s += pow(x - mak, en);
    }
    __shared__ double red[256];
    red[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) part[(size_t) k * nbl + bl] = red[0];
}

/* out[k] = scale * (the level's partials summed in index order), plus
 * off[k] when given (a field stored relative to a per-level profile) */
__global__ void k_stat_finish(const double *part, int nbl, int nz, double scale, double *out, const double *off)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nz) return;
    double s = 0.0;
    for (int b = 0; b < nbl; b++) s += part[(size_t) k * nbl + b];
    out[k] = off ? s * scale + off[k] : s * scale;
}

/* one-dimensional cross-spectra of one level (workgroup = level): A, B the
 * FFT3D spectra [nz][nkx][ny]; specx[ix][k] sums conj-products over ky
 * without the (0, 0) mode; specy[t][k], t <= ny/2, sums over kx the mode
 * ky = t and the mirror of ky = ny - t (imaginary part negated) in the
 * columns where get.c:46-84 takes them */
__global__ void __launch_bounds__(256) k_stat_spec(const double2 *A, const double2 *B, int nx, int ny, int nkx,
                                                   int nz, double *specx, double *specy)
{
    const int k = blockIdx.x, tid = threadIdx.x;
    const double2 *a = A + (size_t) k * nkx * ny, *b = B + (size_t) k * nkx * ny;
    __shared__ double rr[256], ri[256];
    for (int ix = 0; ix < nkx; ix++) {
        double sr = 0.0, si = 0.0;
        for (int iy = tid; iy < ny; iy += 256) {
            if (ix == 0 && iy == 0) continue;
            const double2 u = a[(size_t) ix * ny + iy], v = b[(size_t) ix * ny + iy];
            sr += u.x * v.x + u.y * v.y;
            si += u.y * v.x - u.x * v.y;
        }
        rr[tid] = sr;
        ri[tid] = si;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) { rr[tid] += rr[tid + w]; ri[tid] += ri[tid + w]; }
            __syncthreads();
        }
        if (tid == 0) {
            specx[((size_t) ix * nz + k) * 2] = rr[0];
            specx[((size_t) ix * nz + k) * 2 + 1] = ri[0];
        }
        __syncthreads();
    }
    for (int t = tid; t <= ny / 2; t += 256) {
        const int iyc = (ny - t) % ny;      /* the ky whose mirror is t */
        double sr = 0.0, si = 0.0;
        for (int ix = 0; ix < nkx; ix++) {
            const bool nyq = ix == nx / 2 && nx % 2 == 0;
            if (ix != 0 || t > 0) {
                const double2 u = a[(size_t) ix * ny + t], v = b[(size_t) ix * ny + t];
                sr += u.x * v.x + u.y * v.y;
                si += u.y * v.x - u.x * v.y;
            }
            if (ix != 0 && !nyq) {
                const double2 u = a[(size_t) ix * ny + iyc], v = b[(size_t) ix * ny + iyc];
                sr += u.x * v.x + u.y * v.y;
                si -= u.y * v.x - u.x * v.y;
            }
        }
        specy[((size_t) t * nz + k) * 2] = sr;
        specy[((size_t) t * nz + k) * 2 + 1] = si;
    }
}

/* FP32 storage to FP64 for the spectra */
template <typename T>
__global__ void k_stat_widen(const T *in, double *out, size_t n)
{
    const size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double) in[i];
}

template __global__ void k_stat_plane<double, 0>(const double *, const double *, const double *, const double *,
                                                 double, int, int, double *);
template __global__ void k_stat_plane<double, 1>(const double *, const double *, const double *, const double *,
                                                 double, int, int, double *);
template __global__ void k_stat_plane<double, 2>(const double *, const double *, const double *, const double *,
                                                 double, int, int, double *);
template __global__ void k_stat_plane<float, 0>(const float *, const float *, const double *, const double *,
                                                double, int, int, double *);
template __global__ void k_stat_plane<float, 1>(const float *, const float *, const double *, const double *,
                                                double, int, int, double *);
template __global__ void k_stat_plane<float, 2>(const float *, const float *, const double *, const double *,
                                                double, int, int, double *);
template __global__ void k_stat_widen<float>(const float *, double *, size_t);
