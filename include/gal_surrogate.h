/*
 * gal_surrogate.h -- C ABI of the learned node-interval surrogate.
 *
 * A small MLP (32 -> 256 -> 256 -> 256 -> 31, SiLU between layers) stands in
 * for whole quiet node-interval steps of the node ODE system of gal_ode.h:
 * the node state at t0 and the interval length in, the state increment out,
 * plus one extra output that predicts the log of its own error.  Weights and
 * activations are bf16 on the MFMA units with fp32 accumulation; forward,
 * backward and Adam run in the library's own kernels.  The FP64 batched
 * integrator of gal_ode.h is the fallback for every step whose predicted
 * error or mass/metal budget check fails, and it produces the training data.
 *
 * Reference counterpart: none.  The reference evolves every node interval
 * with the integrator (mergerTreeNodeEvolverStandard%evolve,
 * source/merger_trees/node_evolver/standard.F90:385-755); the surrogate is a
 * front end to that same path, gal_surrogate_evolve_d has the argument list
 * and per-node status codes of gal_ode_batch_evolve_d, and every step it does
 * not accept is evolved by the integrator.
 *
 * Encoding (one node, t0 -> t1, dt = t1 - t0):
 *   unit U = f_b * M_halo(t0); per-variable scale s_v = 1e-4 U (masses),
 *   1e-6 U (metal masses), 1e-4 U r_vir V_vir (angular momenta), 1 (black
 *   hole mass in Msun, black hole spin); u_v = asinh(y_v / s_v).
 *   Inputs 0..21: u_v of the 22 dynamic variables (state order, see
 *   GAL_SURR_INPUT_VARS); 22 ln dt; 23 log10 M_halo - 11; 24 t0/10;
 *   25 Mdot_halo dt / M_halo; 26 ln spin parameter; 27 r_outer / r_vir;
 *   28 V_vir/200; 29 reionization suppression flag; 30 ln(dt / t_dyn); 31 0.
 *   Targets 0..29: u_v(t1) - u_v(t0) with the t0 scales (the radiated
 *   energy uses asinh(dE / s_E)), minus a first-order increment from one
 *   RHS evaluation f = f(t0, y0): asinh((y + f dt)/s) - asinh(y/s) where
 *   f >= 0 (linear growth), asinh(y exp(f dt / y)/s) - asinh(y/s) where
 *   f < 0 < y (exponential decay on the time scale y/|f|), f dt /
 *   sqrt(s^2 + y^2) where f < 0 and y <= 0, asinh(f dt / s_E) for the
 *   energy: the network learns the residual of a first-order step.
 *   Column 30 is the error head; a nonzero column 31
 *   marks a training sample whose targets were clipped (its error counts
 *   as at least 1 in the head's loss).
 *   Variables with zero rate are carried over, the spin angular momentum is
 *   advanced analytically (its rate is the node's constant spin rate).
 * Normalisation array norm (device fp32, 4 x 32):
 *   [input mean | input std | target centre m | target scale sigma].  Inputs
 *   are standardised, (x - mean) / std.  A target residual r is presented to
 *   the network as asinh((r - m) / sigma), and a network output o decodes to
 *   the residual m + sigma sinh(o) (o clipped to +-20).
 */
#ifndef GAL_SURROGATE_H
#define GAL_SURROGATE_H

#include <stdint.h>
#include "gal_ode.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GAL_MLP_INPUTS   32
#define GAL_MLP_HIDDEN   256
#define GAL_MLP_OUTPUTS  31   /* 30 standardised increments + 1 log-error head */
#define GAL_MLP_OUT_LD   32   /* row stride of outputs and targets (fp32)       */
#define GAL_MLP_HEAD     30   /* column of the log-error head                    */
#define GAL_SURR_N_INPUT_VARS 22

typedef struct gal_mlp gal_mlp;

/* Acceptance rule of a surrogate step. */
typedef struct gal_surrogate_gate {
  double log_error_max;   /* accept only if the head predicts ln(err) <= this    */
  double mass_tolerance;  /* |baryon budget residual| / (M_b0 + accreted)        */
  double metal_tolerance; /* |metal budget residual| / (Z0 + yield * dM_formed)  */
  double bh_slack;        /* extra negative slack per unit black-hole growth     */
  double negative_tolerance; /* a decoded non-negative quantity in [-tol s, 0)
                                (s its asinh scale) is set to 0, as the
                                reference's post-step fixes do; below it the
                                step is rejected                               */
} gal_surrogate_gate;

/* ---- the network ------------------------------------------------------- */

/* Number of fp32 parameters: layer by layer, W (fan_in x fan_out, row-major)
   then b (fan_out). */
int64_t gal_mlp_param_count(void);

/* Create a network for training batches of up to `batch` rows on the current
   device.  Parameters start at zero: set them with gal_mlp_set_params. */
int gal_mlp_create(int batch, gal_mlp** out);
int gal_mlp_destroy(gal_mlp* m);

/* Host fp32 parameter vector in/out (n = gal_mlp_param_count()).  Setting
   the parameters also resets the Adam moments and step counter. */
int gal_mlp_set_params(gal_mlp* m, const float* params, int64_t n);
int gal_mlp_get_params(gal_mlp* m, float* params, int64_t n);

/* Adam with a cosine learning-rate decay from lr to lr_min over total_steps. */
int gal_mlp_set_adam(gal_mlp* m, double lr, double lr_min, double beta1, double beta2, double eps,
                     int64_t total_steps);

/* Loss per row: sum_v w_v r_v^2 / sum_v w_v + head_weight (o_30 - ln(e + 1e-7))^2
   with r_v = o_v - target_v for v < 30 (targets are asinh((r - m)/sigma) of
   the residuals) and e = max over head_vars of
   target_sigma_v |sinh(o_v) - sinh(target_v)| (the step's error in asinh
   units of the state; o_v clipped to +-20), raised to at least 1 when target
   column 31 is nonzero.
   var_weight, target_sigma, head_vars: host arrays of 30. */
int gal_mlp_set_loss(gal_mlp* m, const float* var_weight, const float* target_sigma, const int* head_vars,
                     double head_weight);

/* Inference: x (device bf16, n x 32) -> out (device fp32, n x 32; column 31
   is not written).  One fused kernel runs all four layers for 128 rows per
   workgroup, activations kept in LDS. */
int gal_mlp_forward_d(gal_mlp* m, int64_t n, const uint16_t* x, float* out, void* stream);

/* The same forward as four per-layer GEMM launches through global memory
   (the training step's forward); bit-identical to gal_mlp_forward_d. */
int gal_mlp_forward_layers_d(gal_mlp* m, int64_t n, const uint16_t* x, float* out, void* stream);

/* One forward + backward pass on n <= batch rows without an update:
   grad (device fp32, param_count) and loss (device fp32, 1; mean over rows). */
int gal_mlp_gradient_d(gal_mlp* m, int n, const uint16_t* x, const float* target, float* grad, float* loss,
                       void* stream);

/* Training data (device pointers, caller-owned, must outlive training):
   x (n x 32 bf16), target (n x 32 fp32), perm (n int32: row order). */
int gal_mlp_set_dataset(gal_mlp* m, int64_t n, const uint16_t* x, const float* target, const int* perm);

/* n_steps Adam steps on consecutive batches of the permuted dataset,
   starting from its first row.  use_graph != 0 replays one captured
   hipGraph per step.  loss_trace: device fp32[n_steps] or NULL. */
int gal_mlp_train_d(gal_mlp* m, int n_steps, int use_graph, float* loss_trace, void* stream);

/* ---- node encoding and the surrogate evolve ---------------------------- */

/* Encode nodes (SoA, leading dimension ld) for the network.  norm NULL gives
   raw (unstandardised) values.  Outputs, each optional: x (bf16 n x 32),
   x_raw (fp32 n x 32), target (fp32 n x 32; needs y1). */
int gal_surrogate_encode_d(gal_ode_batch* h, int n, int ld, const double* y0, const double* y1,
                           const double* node_par, const double* t0, const double* t1, const float* norm,
                           uint16_t* x, float* x_raw, float* target, void* stream);

/* Decode network outputs and apply the gate.  Accepted nodes: y (in place),
   scale, status = GAL_STATUS_SUCCESS, accepted = 1.  Rejected nodes: y
   untouched, accepted = 0, index appended to reject_list (count in
   *reject_count, device int, incremented). */
int gal_surrogate_apply_d(gal_ode_batch* h, const gal_surrogate_gate* gate, int n, int ld, double* y,
                          const double* node_par, const double* t0, const double* t1, const float* mlp_out,
                          const float* norm, double* scale, int* status, int* accepted, int* reject_list,
                          int* reject_count, void* stream);

/* The surrogate evolve: encode -> network -> gate for all n nodes, then the
   FP64 integrator (gal_ode_batch_evolve_d semantics) on the rejected ones.
   accepted (device int n) reports the route of each node; *n_rejected
   (host) the number evolved by the integrator.  With n_rejected NULL the
   call only enqueues work on stream (the integrator reads the rejected
   count from device memory); otherwise it waits for the stream to read it.
   One network handle serves one stream at a time (it holds the inference
   buffers and the rejected list). */
int gal_surrogate_evolve_d(gal_ode_batch* h, gal_mlp* m, const float* norm, const gal_surrogate_gate* gate,
                           int n, int ld, double* y, const double* node_par, const double* t_start,
                           const double* t_end, double* h_step, double* scale, int* status, int* accepted,
                           int* n_rejected, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GAL_SURROGATE_H */
