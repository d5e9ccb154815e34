/* fdmap_adjoint.h -- discrete adjoint of the time-stepping path (C ABI).
 *
 * The reference has no adjoint: these entry points have no counterpart in
 * it.  They serve callers that need gradients of a waveform misfit with
 * respect to fault parameters or initial fields (inversion, sensitivity
 * studies), e.g. the torch.autograd.Function of fdmap_github_amd.adjoint.
 *
 * The forward pass is the library's own time stepping (fdm_step) on a
 * context, recording vx, vy at receiver grid points after every step.  The
 * backward pass runs the exact transpose of every stage of the discrete
 * scheme (LSRK(5,4) stages, SBP4 split-form rates, SAT penalties, boundary
 * characteristic hats, interface hats and the rate-and-state friction solve
 * through the implicit function theorem at its converged root) and returns
 * the gradient of the scalar whose derivative with respect to the recorded
 * receiver values is given.
 *
 * Trajectory storage: the backward pass needs, per stage, the slip velocity,
 * state variable and normal stress of every fault node (the bulk operator is
 * affine, so no bulk state is kept): 3 x 8 bytes per fault node and stage.
 * The forward pass stores checkpoints of the full state every `interval`
 * steps and keeps the traces of the last interval; the backward pass
 * recomputes every other interval's steps from its checkpoint to regenerate
 * its traces, then sweeps the interval backwards.
 *   interval > 0   : that interval (clamped to nsteps; nsteps keeps the whole
 *                    trace and recomputes nothing)
 *   interval < 0   : ceil(sqrt(nsteps)) (square-root checkpointing)
 *   interval = 0   : the whole trace when it needs at most FDM_ADJ_TRACE_MB
 *                    MiB (environment, default 2048), else ceil(sqrt(nsteps))
 *
 * Supported problems: FP64 contexts without plasticity, every block at least
 * 12 points wide in both directions, interfaces locked or frictional with a
 * regularised rate-and-state law (RSL, RFL, RAL), no normal-stress
 * regularisation and no locked region.  Point sources, manufactured-solution
 * forcing and load ramps are additive in time and allowed.
 */
#ifndef FDMAP_ADJOINT_H
#define FDMAP_ADJOINT_H

#include <stdint.h>

#include "fdmap.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdm_adjoint fdm_adjoint;

/* Gradient outputs of fdm_adjoint_backward; concatenated over interfaces in
 * interface order (node_off of each), length = total interface nodes. */
enum {
  FDM_GRAD_S0 = 0, /* background shear stress S0base (stress file)     */
  FDM_GRAD_A = 1,  /* rate-and-state a                                */
  FDM_GRAD_B = 2,  /* rate-and-state b                                */
  FDM_GRAD_PSI0 = 3, /* state variable at the start of the forward pass */
  FDM_GRAD_NFIELDS = 4
};

/* Adjoint workspace for `nsteps` steps of ctx with nrec receivers at global
 * grid points (rec_i[r], rec_j[r]).  ctx must outlive it. */
int fdm_adjoint_create(fdm_ctx* ctx, int32_t nrec, const int32_t* rec_i, const int32_t* rec_j,
                       int32_t nsteps, int32_t interval, fdm_adjoint** out);

/* Number of interface nodes (length of each gradient array) and the
 * checkpoint interval in use. */
int fdm_adjoint_info(const fdm_adjoint* adj, int32_t* nodes, int32_t* interval);

/* Overwrite friction parameters of all interface nodes (concatenated as the
 * gradients): field FDM_GRAD_S0, FDM_GRAD_A or FDM_GRAD_B. */
int fdm_adjoint_set_param(fdm_adjoint* adj, int32_t field, const double* values);

/* Run nsteps steps of the context from its current state, step n at time
 * t + (n-1) dt as fdm_step, storing checkpoints; rec (may be NULL) receives
 * [nsteps][nrec][2] values of vx, vy after each step. */
int fdm_adjoint_forward(fdm_adjoint* adj, double t, double dt, double* rec);

/* Backward pass of the most recent forward pass.  rec_bar: [nsteps][nrec][2]
 * derivative of the scalar with respect to rec.  grad (may be NULL):
 * [FDM_GRAD_NFIELDS][nodes] accumulated gradients; gF (may be NULL):
 * [6][ny][nx] gradient with respect to the fields at the start of the
 * forward pass (the szz plane is zero: nothing depends on szz but szz).
 * Leaves the context at the end of the first checkpoint interval when it
 * recomputes, at the end of the forward pass otherwise. */
int fdm_adjoint_backward(fdm_adjoint* adj, const double* rec_bar, double* grad, double* gF);

/* Device time of the most recent passes in ms: forward, backward
 * recomputation, backward sweep. */
int fdm_adjoint_timing(const fdm_adjoint* adj, double* fwd_ms, double* recompute_ms,
                       double* sweep_ms);

const char* fdm_adjoint_last_error(const fdm_adjoint* adj);
void fdm_adjoint_destroy(fdm_adjoint* adj);

#ifdef __cplusplus
}
#endif

#endif /* FDMAP_ADJOINT_H */
