/* cf_internal_host.h -- internal helpers shared by the host C files and
 * the HIP runtime of the chanflow library (not part of the public ABI). */
#ifndef CF_INTERNAL_HOST_H
#define CF_INTERNAL_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* record a printf-style message for cf_last_error() on this thread */
void cf_set_error(const char *fmt, ...);

#ifdef __cplusplus
}
#endif

#endif
