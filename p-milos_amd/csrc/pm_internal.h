// pm_internal.h -- declarations shared by the library's own sources (not
// part of the C ABI).
#ifndef PM_INTERNAL_H
#define PM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stddef.h>

struct pm_ctx;

// Queues one inversion pass of npix pixels on stream st with the given
// work-queue counters (pm_counters_bytes() of device memory) and fallback
// counter: launches on different streams need their own counters.  share
// (0, 1] caps the workgroups at that fraction of what the GPU holds at once.
// reset: zero the counters on st first (a one-workgroup kernel); false when
// the caller has zeroed them already (pm_reset_launch).
int pm_invert_launch(pm_ctx *ctx, const float *d_stokes, const float *d_slight, int npix,
                     double *d_models, float *d_chi2, int *d_iters, int *d_accepted,
                     float *d_fit, int *d_counters, unsigned long long *d_fallbacks,
                     hipStream_t st, double share, bool reset = true);
// Queues the zeroing of one set of work-queue counters and its fallback
// counter on st.
int pm_reset_launch(int *d_counters, unsigned long long *d_fallbacks, hipStream_t st);
int pm_ctx_nlambda(const pm_ctx *ctx);
// Estimated work per pixel and iteration, relative (for sizing and ordering
// chunks of work).
double pm_ctx_cost(const pm_ctx *ctx);
int pm_ctx_device(const pm_ctx *ctx);
size_t pm_counters_bytes();

#endif
