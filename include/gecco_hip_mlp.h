/* Learned per-amplitude preconditioner of the MI355X engine (C ABI).
 *
 * No counterpart in the reference: the reference preconditions the
 * amplitude step with the orbital-energy denominator only,
 * r = g / D (baserout/diavc.f:31-35, contract/dia4op_ev.f, called from
 * solve/optc_diis_sbsp_add.f:125-140).  Here a 4 -> 128 -> 128 -> 1 network
 * (SiLU) scales that step per amplitude, r = s * g / D with s = 1 + y(x);
 * gecco_hip_diis_set_scale (gecco_hip_solver.h) feeds s into the DIIS step.
 *
 * Parameters: GECCO_HIP_MLP_P float32 values in device memory, in this
 * order: W1[128][4], b1[128], W2[128][128] (row = output unit), b2[128],
 * w3[128], b3.  The network is
 *   z1 = bf(W1) x + b1,  a = bf(silu(z1)),  z2 = bf(W2) a + b2,
 *   y  = bf(w3) . silu(z2) + b3,           bf(): round to bfloat16.
 * Features of amplitude i, from the denominator D, g = D t + rest, the
 * amplitude t and the iteration index it (u = g / D', D' as the solver's
 * step):
 *   x = [D/4 - 1, asinh(u / rms(u)), asinh(t / rms(t)) (0 if rms(t) = 0), min(it, 32)/32]
 * Training pairs (x, u~, dt~) carry u / rms(u) and the accepted update
 * dt / rms(u) of a plain solver step; loss = mean((1 + y) u~ + dt~)^2.
 *
 * All calls queue work on `stream` (hipStream_t, may be 0) and return 0 or
 * an error code (message: gecco_hip_last_error). */
#ifndef GECCO_HIP_MLP_H
#define GECCO_HIP_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GECCO_HIP_MLP_P 17281
#define GECCO_HIP_MLP_GRAD_GROUPS 256
#define GECCO_HIP_MLP_EINVAL 1
#define GECCO_HIP_MLP_EDEVICE 3
/* device work array sizes: mlp_scale needs GECCO_HIP_MLP_SCALE_WORK
 * doubles, mlp_grad GECCO_HIP_MLP_GRAD_WORK floats */
#define GECCO_HIP_MLP_SCALE_WORK (2 + 2 * 512)
#define GECCO_HIP_MLP_GRAD_WORK (GECCO_HIP_MLP_GRAD_GROUPS * GECCO_HIP_MLP_P + 2 * GECCO_HIP_MLP_GRAD_GROUPS)

/* y[i] for the feature rows x[i][0..3] (bfloat16 weights, BF16 MFMA). */
int gecco_hip_mlp_infer(const float *params, const float *x, int64_t n, float *y, void *stream);

/* scale[i] = 1 + y(features of amplitude i) for an amplitude vector
 * (diag, rest, t of length n, iteration index it); work: device doubles,
 * GECCO_HIP_MLP_SCALE_WORK of them (work[0], work[1] = rms(u), rms(t)). */
int gecco_hip_mlp_scale(const float *params, const double *diag, const double *rest, const double *t,
                        int64_t n, double it, double *scale, double *work, void *stream);

/* Gradient of the loss over n training pairs (x[n][4], ut[n], dt[n]) into
 * grad[GECCO_HIP_MLP_P] (float32, straight through bf()) and, if loss is not
 * NULL, the loss into loss[0] (device double); work: device floats,
 * GECCO_HIP_MLP_GRAD_WORK of them.  Partial sums per workgroup are added in
 * a fixed order; inside a workgroup the biases, W1 and w3 accumulate with
 * LDS atomics, so repeated calls can differ in the last bits. */
int gecco_hip_mlp_grad(const float *params, const float *x, const float *ut, const float *dt, int64_t n,
                       float *grad, double *loss, float *work, void *stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8; step counts from 1) on
 * params with moments m, v (device float32, GECCO_HIP_MLP_P each). */
int gecco_hip_mlp_adam(float *params, float *m, float *v, const float *grad, int64_t step, double lr,
                       void *stream);

#ifdef __cplusplus
}
#endif

#endif
