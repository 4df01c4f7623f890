/*
 * copss_ggem_noalloc.h -- an entry point of the GGEM particle path that
 * allocates no device memory.
 */
#ifndef COPSS_GGEM_NOALLOC_H
#define COPSS_GGEM_NOALLOC_H

#include "copss_ggem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ggem_local_velocity_points with its staging (points, element ids,
 * results, an error flag) in one fixed 4 MiB device array per GPU instead
 * of the context's growable pools: the call itself allocates nothing.  No
 * counterpart in the reference; for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid it).  At most 65536
 * points per call (GGEM_ERR_CAPACITY beyond).  Same arguments, results
 * and errors as ggem_local_velocity_points.  All calls on one GPU share
 * the array: calls that overlap on the same GPU race. */
int ggem_local_velocity_points_noalloc(ggem_ctx* ctx, int64_t m, const double* pts, const int32_t* elem,
                                       double* u);

#ifdef __cplusplus
}
#endif
#endif
