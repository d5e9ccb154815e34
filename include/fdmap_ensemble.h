/*
 * fdmap_ensemble.h -- batched ensembles and launch tuning for the MI355X
 * time-stepping library (include/fdmap.h).
 *
 * These entry points have no counterpart in the reference, which runs one
 * problem per process.  They serve callers that advance many independent
 * problems together (rupture ensembles) and the autotuner that picks launch
 * shapes per problem size.
 */
#ifndef FDMAP_ENSEMBLE_H
#define FDMAP_ENSEMBLE_H

#include "fdmap.h"

#ifdef __cplusplus
extern "C" {
#endif

/* launch knobs of one context */
#define FDM_TUNE_TILE 0    /* interior tile shape: 0 = 64x4, 1 = 64x8, 2 = 128x4
                              (SBP4 in a uniform medium without dissipation;
                              SBP6 there: 0 or 1, default 1; else 0 only) */
#define FDM_TUNE_SWIZZLE 1 /* XCD-contiguous tile order: 0 or 1 */
#define FDM_TUNE_NT_STORE 2 /* non-temporal stores of the interior kernel: 0 or 1
                               (default: 1 when the 23 field, rate and metric
                               planes exceed 256 MB) */
#define FDM_TUNE_METRIC_F32 3 /* interior tiles read the metrics as FP32 deltas
                                 from a per-block base (FP64 plane-strain
                                 contexts, 64x4 and 64x8 tiles): 0 or 1.  fdm_create turns it
                                 on where the encoding is exact for all but
                                 1/64 of the metric values (those are read
                                 from the FP64 planes); value 1 fails where
                                 it did not. */

/* Set a launch knob of ctx.  Results do not depend on it.  Set knobs
   before putting ctx into an ensemble (its graphs keep the shapes they were
   captured with). */
int fdm_tune(fdm_ctx* ctx, int32_t knob, int32_t value);

typedef struct fdm_ensemble fdm_ensemble;

/* Group n contexts (each an independent problem with its own storage) so
   that they advance together: every RK stage of all members is one interior
   launch per tile shape in use, one edge launch and one band launch over
   the members' merged work lists, and every step is one hipGraph launch,
   captured on first use.  Members must have no point source, no
   manufactured solution and no load ramp, use the tile kernels, and share
   the precision, RK order and device; at each step they must share the
   current buffer parity.  The contexts stay owned by the caller and must
   outlive the ensemble. */
int fdm_ensemble_create(fdm_ctx* const* members, int32_t n, fdm_ensemble** out);

/* nsteps steps of every member k, step n starting at t[k] + (n-1)*dt[k] for
   n = n_first .. n_first+nsteps-1: the same arithmetic as fdm_step on each
   member.  Times must be >= 0.  Returns before the device finishes; the
   members' own calls (fdm_get_*) are ordered after it. */
int fdm_ensemble_step(fdm_ensemble* e, const double* t, const double* dt, int64_t n_first,
                      int32_t nsteps);

/* Wait for the ensemble's queued steps. */
int fdm_ensemble_sync(fdm_ensemble* e);

/* Release the ensemble (not its members). */
void fdm_ensemble_destroy(fdm_ensemble* e);

/* Message of the last failed ensemble call (e may be NULL: last failed
   fdm_ensemble_create, see also fdm_last_error(NULL)). */
const char* fdm_ensemble_last_error(const fdm_ensemble* e);

#ifdef __cplusplus
}
#endif

#endif /* FDMAP_ENSEMBLE_H */
