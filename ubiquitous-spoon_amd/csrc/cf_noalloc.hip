/*
 * cf_noalloc.hip -- cf_run_xy_stats_noalloc: the plane statistics of
 * cf_run_xy_stats (kinds CF_STAT_AVG, CF_STAT_CORR, CF_STAT_MOMENT) with
 * their work space in a fixed device array instead of a per-call
 * allocation, for callers that cannot allocate device memory.
 *
 * The work space (CF_STAT_NBL partial sums per level, two mean profiles
 * and the result profile) lives in `scratch`, one copy per GPU for the
 * life of the process.  Its device address is looked up once per device,
 * with that device current, and only that address is handed to kernels and
 * copies.  Every call on a GPU shares the GPU's array: calls that overlap
 * on one GPU race.  No value carries over from one call to the next (every
 * partial, mean and result is written before it is read), so nothing needs
 * clearing.
 */
#include "chanflow.h"
#include "cf_internal_host.h"
#include "cf_kernels.h"

#include <hip/hip_runtime.h>

#include <mutex>

#define CF_NOALLOC_NZ_MAX 2048
#define CF_NOALLOC_DOUBLES ((size_t) (CF_STAT_NBL + 3) * CF_NOALLOC_NZ_MAX)
#define CF_NOALLOC_DEVICES 64

static __device__ __align__(256) double scratch[CF_NOALLOC_DOUBLES];

static std::mutex g_lock;
static double *g_ptr[CF_NOALLOC_DEVICES];

/* the device address of `scratch` on `device`, looked up on first use with
 * that device current */
static cf_status scratch_ptr(int device, double **out)
{
    if (device < 0 || device >= CF_NOALLOC_DEVICES) {
        cf_set_error("cf_run_xy_stats_noalloc: device %d out of range", device);
        return CF_ERR_ARG;
    }
    std::lock_guard<std::mutex> hold(g_lock);
    if (!g_ptr[device]) {
        int prev = 0;
        if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) {
            cf_set_error("cf_run_xy_stats_noalloc: cannot select device %d", device);
            return CF_ERR_DEVICE;
        }
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void) hipSetDevice(prev);
        if (e != hipSuccess || !p) {
            cf_set_error("cf_run_xy_stats_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
            return CF_ERR_DEVICE;
        }
        g_ptr[device] = (double *) p;
    }
    *out = g_ptr[device];
    return CF_OK;
}

extern "C" cf_status cf_run_xy_stats_noalloc(cf_run *run, int kind, int field_a, int field_b, double en, double *out)
{
    if (!run || !out) { cf_set_error("cf_run_xy_stats_noalloc: null run or output"); return CF_ERR_ARG; }
    if (kind != CF_STAT_AVG && kind != CF_STAT_CORR && kind != CF_STAT_MOMENT) {
        cf_set_error("cf_run_xy_stats_noalloc: kind %d (AVG, CORR or MOMENT only)", kind);
        return CF_ERR_ARG;
    }
    int nz = 0, device = 0;
    void *stream = nullptr;
    cf_status st = cf_run_stat_plane_into(run, kind, field_a, field_b, en, nullptr, nullptr, nullptr, &nz, &stream,
                                          &device);
    if (st != CF_OK) return st;
    if (nz > CF_NOALLOC_NZ_MAX) {
        cf_set_error("cf_run_xy_stats_noalloc: nz = %d exceeds the fixed work space (%d levels)", nz,
                     CF_NOALLOC_NZ_MAX);
        return CF_ERR_ARG;
    }
    double *ws = nullptr;
    if ((st = scratch_ptr(device, &ws)) != CF_OK) return st;
    double *part = ws, *means = ws + (size_t) CF_STAT_NBL * nz, *res = means + 2 * (size_t) nz;
    st = cf_run_stat_plane_into(run, kind, field_a, field_b, en, part, means, res, nullptr, nullptr, nullptr);
    if (st != CF_OK) return st;
    const hipStream_t s = (hipStream_t) stream;
    if (hipMemcpyAsync(out, res, sizeof(double) * (size_t) nz, hipMemcpyDeviceToHost, s) != hipSuccess
        || hipStreamSynchronize(s) != hipSuccess) {
        cf_set_error("cf_run_xy_stats_noalloc: result copy failed");
        return CF_ERR_DEVICE;
    }
    return CF_OK;
}
