/*
 * gecco_hip_solver.h -- amplitude solver of the MI355X engine: diagonal
 * preconditioner and Pulay DIIS on device-resident amplitude vectors.
 *
 * Replaces the in-core route of GeCCo's DIIS optimiser for one amplitude
 * vector (solve/optc_macit.f:229,392,398):
 *   - the preconditioned step  r = g / D with |D| <= 1e-10 -> 1e-10
 *     (baserout/diavc.f:1-38, called from solve/optc_diis_sbsp_add.f:155),
 *   - the subspace vectors r and v = t - r
 *     (solve/optc_diis_sbsp_add.f:150-205),
 *   - the subspace matrix B(i,j) = <r_i|r_j>, oldest vector dropped when the
 *     subspace is full (solve/optc_update_redsp1.f),
 *   - the weights from [B -1; -1 0] (w, l) = (0, -1)
 *     (solve/optc_diis_extr.f:66-105),
 *   - the new amplitudes t = sum_i w_i v_i (solve/optc_expand_vec.f),
 *   - the gradient norm ||g|| (type_xret = 1 accumulation,
 *     contract/contr_op1op2_wmaps_c.f:2060-2061, solve/optc_macit.f).
 * The diagonal D of a list (contract/dia4op_ev.f:300-365, one-electron
 * part) is gecco_hip_list_diagonal in gecco_hip_model.h.
 *
 * All vectors are device pointers of n doubles; errors as in gecco_hip.h.
 */
#ifndef GECCO_HIP_SOLVER_H
#define GECCO_HIP_SOLVER_H

#include <stdint.h>
#include "gecco_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gecco_hip_diis gecco_hip_diis;

/* Subspace of at most maxsub (1..16) vectors of length n (device memory for
 * 2*maxsub vectors is allocated here). */
int gecco_hip_diis_create(int64_t n, int64_t maxsub, gecco_hip_diis **out);
int gecco_hip_diis_destroy(gecco_hip_diis *d);
/* Empty the subspace. */
int gecco_hip_diis_reset(gecco_hip_diis *d);

/* One iteration for the vector function g(t) = diag*t + rest (rest: g
 * without its diagonal part, as the formula computes it).  Writes ||g|| to
 * *gnorm and, if energy_dev is given, the device scalar energy_dev[0] to
 * *energy (read in the same synchronisation).  If ||g|| < tol, *converged =
 * 1 and t is unchanged; otherwise t is replaced by the DIIS extrapolation.
 * Work is queued on `stream` (hipStream_t, may be 0); returns after the
 * stream has drained the reduction. */
int gecco_hip_diis_step(gecco_hip_diis *d, double *t, const double *rest, const double *diag,
                        const double *energy_dev, double tol, double *gnorm, double *energy,
                        int64_t *converged, void *stream);

/* The same step without any synchronisation: the subspace bookkeeping (slot
 * rotation, the new row of B, the augmented solve) runs on the device, so
 * every step queues the same four launches with the same arguments and an
 * iteration can be captured and replayed as a graph.  Step i (0-based since
 * create/reset) writes its record into a ring of 64 in pinned host memory;
 * read it with gecco_hip_diis_status once the stream has passed the step.
 * A converged step (||g|| < tol) or a singular system leaves t unchanged.
 * A subspace is either synchronous or asynchronous until reset.  No
 * counterpart in the reference, whose solver is host code between
 * contractions (solve/optc_macit.f); it serves callers that keep the whole
 * iteration on the GPU. */
int gecco_hip_diis_step_async(gecco_hip_diis *d, double *t, const double *rest, const double *diag,
                              const double *energy_dev, double tol, void *stream);
/* Record of asynchronous step `step`: ||g||, energy_dev[0] at that step, and
 * converged = 1 (met tol), 0 (extrapolated) or -1 (singular system).  Fails
 * with GECCO_HIP_ESTATE if the step has not completed or its record was
 * overwritten (more than 64 steps ago). */
int gecco_hip_diis_status(const gecco_hip_diis *d, int64_t step, double *gnorm, double *energy,
                          int64_t *converged);
/* Device int32 that an asynchronous step sets to 1 when it meets tol
 * (converged) and leaves 0 otherwise (NULL: none; the caller keeps the
 * memory alive and clears it together with reset).  While the flag is set,
 * later asynchronous steps do no subspace work and leave t unchanged; each
 * repeats the converged record (||g|| of the converging step).  With plan
 * skip flags (gecco_hip_plan_set_skip) on the same memory, an instance's
 * contractions and DIIS steps stop once it has converged.  No counterpart
 * in the reference. */
int gecco_hip_diis_set_flag(gecco_hip_diis *d, int32_t *flag);
/* Device vector scale[n] (NULL: none, the default): later steps use
 * r = scale * g / D in place of r = g / D -- a per-amplitude factor on the
 * preconditioned step, e.g. the learned preconditioner's s = 1 + y
 * (gecco_hip_mlp_scale, gecco_hip_mlp.h).  ||g|| and the convergence test
 * are unchanged.  The caller keeps the memory alive and fills it before
 * each step.  No counterpart in the reference (baserout/diavc.f divides by
 * D only). */
int gecco_hip_diis_set_scale(gecco_hip_diis *d, const double *scale);

/* Many subspaces stepped together (independent problems, e.g. batched CC
 * instances): member k steps t[k] with rest[k], diag[k] and energy[k]
 * (energy may be NULL) exactly as gecco_hip_diis_step_async would, but one
 * step of the whole group is three launches (all members' fused passes, one
 * workgroup per member for the small solve, all expansions).  Each member
 * keeps its own records (gecco_hip_diis_status), flag and scale; set those
 * before creating the group, and recreate it after a member's reset.  No
 * counterpart in the reference. */
typedef struct gecco_hip_diis_group gecco_hip_diis_group;
int gecco_hip_diis_group_create(gecco_hip_diis *const *d, int64_t count, double *const *t,
                                const double *const *rest, const double *const *diag,
                                const double *const *energy, gecco_hip_diis_group **out);
int gecco_hip_diis_group_step_async(gecco_hip_diis_group *g, double tol, void *stream);
int gecco_hip_diis_group_destroy(gecco_hip_diis_group *g);

/* Weights of the last step (w[0..m-1], oldest vector first) and the
 * current subspace matrix (b[m*m], row-major), for checks. */
int gecco_hip_diis_weights(const gecco_hip_diis *d, double *w, int64_t nw, int64_t *m);
int gecco_hip_diis_matrix(const gecco_hip_diis *d, double *b, int64_t nb, int64_t *m);

#ifdef __cplusplus
}
#endif
#endif /* GECCO_HIP_SOLVER_H */
