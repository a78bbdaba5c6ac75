/* cf_internal_host.h -- internal helpers shared by the host C files and
 * the HIP runtime of the chanflow library (not part of the public ABI). */
#ifndef CF_INTERNAL_HOST_H
#define CF_INTERNAL_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* record a printf-style message for cf_last_error() on this thread */
void cf_set_error(const char *fmt, ...);

/* a plane statistic (CF_STAT_AVG, _CORR, _MOMENT) of a run into res[nz]
 * with caller-provided device work space, enqueued on the run's stream
 * (returned as an opaque hipStream_t); part == NULL only reports nz, the
 * stream and the device */
struct cf_run;
cf_status cf_run_stat_plane_into(struct cf_run *r, int kind, int fa, int fb, double en, double *part,
                                 double *means, double *res, int *nz, void **stream, int *device);

#ifdef __cplusplus
}
#endif

#endif
