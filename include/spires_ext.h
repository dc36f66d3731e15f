/*
 * spires_ext.h -- extensions of the spires C ABI for the MI355X engine.
 *
 * These entry points have no counterpart in the reference's include/spires.h.
 * They exist for callers that need what the reference can only do through
 * private structs (spoken_digit_recording.c reaches into them for rasters and
 * weights) or cannot do at all: reservoirs built from caller-supplied
 * weights, whole-run recording of states / spike rasters without per-step
 * API round trips, device-resident inputs for timing, a ridge readout with a
 * washout, and the choice between the FP64 exact and the FP32 fast step.
 */
#ifndef SPIRES_EXT_H
#define SPIRES_EXT_H

#include "spires.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Create from an explicit recurrent matrix in CSR form (row i lists the
 * presynaptic inputs of neuron i, columns ascending), input weights
 * W_in [N x Din] and optional W_out [Dout x N] (NULL = zeros).  No spectral
 * rescaling and no rand() draws happen, except the per-edge tau draw of the
 * PSC-heterogeneous synapse.  Zero-valued entries are dropped. */
spires_status spires_ext_create_csr(const spires_reservoir_config *cfg, size_t nnz,
                                    const int64_t *row_ptr, const int32_t *col_idx,
                                    const double *values, const double *W_in,
                                    const double *W_out, spires_reservoir **out_r);

/* The weights spires_reservoir_create would build for cfg (same rand()
 * draws: W_in, topology, synapse taus, spectral rescale, W_out), without
 * creating a reservoir or touching the GPU.  Outputs are malloc'd; the caller
 * free()s them.  Recurrent matrix as CSR with ascending columns. */
spires_status spires_ext_build_weights(const spires_reservoir_config *cfg, size_t *nnz_out,
                                       int64_t **row_ptr, int32_t **col_idx, double **values,
                                       double **W_in, double **W_out);

/* Largest |eigenvalue| estimate of a CSR matrix by the reference's power
 * iteration (sparse.c:93-127: start from ones, max-norm, |d lambda| < 1e-6 or
 * 1000 iterations), as used to rescale W to the requested spectral radius. */
double spires_ext_spectral_radius(size_t n, const int64_t *row_ptr, const int32_t *col_idx,
                                  const double *values);

/* Reservoir construction on the GPU (replaces the host rand() draws and the
 * dense N x N scratch of init_weights, reservoir.c:345-545, and the host
 * power iteration + rescale, sparse.c:89-127 / reservoir.c:548-557, for the
 * random topology).  Edge i->j (i != j) with probability p, magnitude
 * U(0,1], positive with probability ei -- generate_weight's distribution --
 * hashed from (seed, i, j) with SplitMix64: a different random stream from
 * the reference's rand(), so the matrix agrees with the reference's in
 * distribution, not entry by entry.  W is scaled by rho / lambda
 * (lambda > 1e-9) with lambda from the reference power iteration, run on the
 * device; *lambda (may be NULL) gets the unscaled estimate.  n <= 19968.
 * Outputs are malloc'd; the caller free()s them. */
spires_status spires_ext_random_er(size_t n, double p, double ei, uint64_t seed, double rho,
                                   size_t *nnz, int64_t **row_ptr, int32_t **col_idx,
                                   double **values, double *lambda);
/* spires_ext_random_er for any connectivity type (replaces the small-world
 * and scale-free branches of init_weights, reservoir.c:399-536, and the
 * rescale): the reference's sizes (ring degree K = round(conn (n-1)), even,
 * K/2 forward neighbours; m = round(conn (n-1)) edges per new node over a
 * clique of max(m, 2)) and its steps in law -- each ring edge rewired with
 * probability 0.1 to a uniform target that is neither i nor an existing edge;
 * preferential attachment proportional to deg + 1 with the uniform fallback
 * after 50 m attempts, each undirected edge oriented by a fair coin -- from
 * SplitMix64 draws instead of rand().  Then the reference power iteration and
 * the rescale to rho on the device.  n <= 19968; outputs malloc'd. */
spires_status spires_ext_build_topology(spires_connectivity_type type, size_t n, double connectivity,
                                        double ei, uint64_t seed, double rho, size_t *nnz,
                                        int64_t **row_ptr, int32_t **col_idx, double **values,
                                        double *lambda);
/* spires_reservoir_create (spires.h) with W built by
 * spires_ext_build_topology(connectivity_type, N, connectivity, ei_ratio,
 * seed, spectral_radius) and W_in / W_out from SplitMix64 streams of the same
 * seed (generate_weight's distribution; ei_ratio for W_in, 0.5 for W_out).
 * Synapse time constants still come from rand(), as in
 * spires_reservoir_create. */
spires_status spires_ext_create_device(const spires_reservoir_config *cfg, uint64_t seed,
                                       spires_reservoir **out);
/* spires_ext_spectral_radius on the device (n <= 19968): the same operations
 * in the same order, so *lambda is bit-identical; *iters (may be NULL) gets
 * the number of sparse products run. */
spires_status spires_ext_spectral_radius_device(size_t n, const int64_t *row_ptr,
                                                const int32_t *col_idx, const double *values,
                                                double *lambda, int *iters);
/* spires_ext_spectral_radius_device without any allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  d_row_ptr [n+1], d_col_idx and d_values are a valid CSR
 * (monotone row pointers, columns in [0, n)) in device memory of the current
 * device; *d_lambda (double) and *d_iters (int) are device memory too, written
 * in stream order on `stream` (a hipStream_t; NULL = the null stream); the
 * call returns without waiting.  Same operations in the same order as
 * spires_ext_spectral_radius_device, so lambda is bit-identical.  The work
 * buffers are one fixed array of 328,856 bytes per device, held until the
 * process ends: n above 19968 is rejected with SPIRES_ERR_INVALID_ARG, and
 * calls that overlap on one device race on that array. */
spires_status spires_ext_spectral_radius_device_noalloc(size_t n, const int64_t *d_row_ptr,
                                                        const int32_t *d_col_idx, const double *d_values,
                                                        double *d_lambda, int *d_iters, void *stream);

/* Run T macro steps over u [T x Din] and copy back any of: states [T x N]
 * (V after each step), raster [T x N] (u8 spikes), outputs [T x Dout]. */
spires_status spires_ext_run_record(spires_reservoir *r, const double *u, size_t T,
                                    double *states, uint8_t *raster, double *outputs);

/* Device-resident inputs: upload once, then run slices asynchronously on the
 * handle's stream; records stay in HBM until read. */
#define SPIRES_EXT_REC_STATE  1u
#define SPIRES_EXT_REC_RASTER 2u
#define SPIRES_EXT_REC_OUTPUT 4u
#define SPIRES_EXT_REC_APPEND 8u   /* record macro step t0+k at row t0+k, keeping rows [0, t0) */
spires_status spires_ext_upload_inputs(spires_reservoir *r, const double *u, size_t T);
spires_status spires_ext_reserve_records(spires_reservoir *r, size_t rows, unsigned flags);
spires_status spires_ext_run_resident(spires_reservoir *r, size_t t0, size_t T, unsigned flags);
spires_status spires_ext_read_records(spires_reservoir *r, size_t row0, size_t rows,
                                      double *states, uint8_t *raster, double *outputs);

/* Ridge readout with a washout: reset, run u [T x Din], drop the first
 * `washout` states, fit W_out on targets y [(T - washout) x Dout].
 * info (optional, 2 doubles): [0] = 1 if the LU fallback was used, plus 2 if
 * fast mode formed the FP64 Gram because its split-BF16 Gram was not positive
 * definite; [1] = relative residual of the normal equations after refinement. */
spires_status spires_ext_train_ridge_washout(spires_reservoir *r, const double *u,
                                             const double *y, size_t T, size_t washout,
                                             double lambda, double *info);

/* Ridge readout fitted on recorded states rows [row0, row1) of the last
 * resident run (no re-run), against Y[(row1-row0) x Dout]; info as in
 * spires_ext_train_ridge_washout. */
spires_status spires_ext_fit_ridge_records(spires_reservoir *r, size_t row0, size_t row1, const double *Y,
                                           double lambda, double *info);
/* spires_ext_fit_ridge_records for `count` reservoirs at once (one lambda and
 * one Y[(row1-row0) x Dout] each); exact-mode reservoirs of equal N and Dout
 * are factorised together.  info[2r] is 0 (Cholesky), 1 (LU) or -1 (singular
 * system, that W_out unchanged); info[2r+1] the relative residual.  No
 * counterpart in the reference. */
/* Readout outputs of recorded state rows [row0, row0+rows) under the current
 * W_out, into out[rows x Dout] (host).  No counterpart in the reference. */
spires_status spires_ext_predict_records(spires_reservoir *r, size_t row0, size_t rows, double *out);
spires_status spires_ext_fit_ridge_records_batch(spires_reservoir *const *rs, int count, size_t row0,
                                                 size_t row1, const double *const *Y, const double *lambda,
                                                 double *info);

/* Batched runs: `count` reservoirs advance over macro steps [t0, t0+T) of
 * their own uploaded input series in one device launch, each into its own
 * records (as spires_ext_run_resident).  They must be exact-mode FLIF-GL
 * reservoirs with the same neurons per workgroup (spires_ext_kernel_plan)
 * that fit the product-wave kernel together (spires_ext_batch_workgroups returns 0 otherwise).  No
 * counterpart in the reference, whose AGILE evaluates trials one by one. */
spires_status spires_ext_run_resident_batch(spires_reservoir *const *rs, int count, size_t t0, size_t T,
                                            unsigned flags);
int           spires_ext_batch_workgroups(spires_reservoir *const *rs, int count);
/* Time blocks under a HIP graph: one graph holds the launches that advance
 * `count` exact-mode FLIF-GL reservoirs (product-wave kernel) by K macro
 * steps; every spires_ext_timeblock_run replays it `blocks` times.  Step
 * counters live on the device, so each replay continues where the last one
 * stopped.  Inputs are read, and records (flags SPIRES_EXT_REC_STATE /
 * _RASTER) written, at row = macro step since reset; reserve the records
 * first.  No counterpart in the reference. */
#define SPIRES_EXT_TB_ONE_LAUNCH 16u   /* all reservoirs in one persistent launch: one wave per
                                          64 neurons, the history and CSR streamed from HBM;
                                          reservoirs of at most 16384 neurons whose slices are
                                          resident together */
typedef struct spires_timeblock spires_timeblock;
spires_status spires_ext_timeblock_create(spires_reservoir *const *rs, int count, size_t K, unsigned flags,
                                          spires_timeblock **out);
spires_status spires_ext_timeblock_run(spires_timeblock *tb, size_t blocks);
int           spires_ext_timeblock_launches(const spires_timeblock *tb);
void          spires_ext_timeblock_destroy(spires_timeblock *tb);

/* The exact-mode launch plan of a reservoir, info[7] (spires_dev_plan_info
 * layout: kernel, neurons per workgroup, workgroups, row-pull kind, weights
 * in LDS, LDS bytes, neurons per workgroup when batched).  No counterpart in
 * the reference. */
spires_status spires_ext_kernel_plan(spires_reservoir *r, int *info);
/* Fault injection for callers' error-path tests: the next persistent run of
 * r behaves as if a spike-exchange wait had timed out (the kernel gives up at
 * its first unfinished wait, every wave exits) and the timeout is reported as
 * SPIRES_ERR_INTERNAL by the call that waits for that run; the flag is then
 * cleared, and a reset gives a reservoir that runs normally.  No counterpart
 * in the reference. */
spires_status spires_ext_inject_exchange_timeout(spires_reservoir *r);
/* Measured floor of a persistent step: microseconds per step of the spike
 * exchange alone among `workgroups` workgroups (at most one per compute unit),
 * timed over `steps` steps with no other work.  No counterpart in the
 * reference. */
spires_status spires_ext_exchange_floor(int workgroups, size_t steps, double *us_per_step);

/* Step arithmetic. */
#define SPIRES_EXT_MODE_EXACT 0   /* FP64, bit-exact with the reference CPU path (default) */
#define SPIRES_EXT_MODE_FAST  1   /* FP32 time-blocked GL history on matrix cores          */
spires_status spires_ext_set_mode(spires_reservoir *r, int mode, int block_k);
/* Exact-mode kernel path: 0 automatic (one persistent launch per run when the
 * model supports it), 1 one launch per micro-step, 2 persistent only (FLIF-GL:
 * chain wave fed by product waves when it fits), 3 persistent with a
 * single-wave history sum streamed from HBM, 4 persistent with a single-wave
 * history sum over an LDS-resident history.  All paths give identical results. */
spires_status spires_ext_set_kernel_path(spires_reservoir *r, int path);

spires_status spires_ext_set_state(spires_reservoir *r, const double *V);
spires_status spires_ext_sync(spires_reservoir *r);
void         *spires_ext_stream(spires_reservoir *r);
size_t        spires_ext_nnz(const spires_reservoir *r);
spires_status spires_ext_get_csr(const spires_reservoir *r, int64_t *row_ptr,
                                 int32_t *col_idx, double *values);
spires_status spires_ext_get_input_weights(const spires_reservoir *r, double *W_in);
const char   *spires_ext_last_error(void);
/* Releases the engine's process-wide GPU resources (the per-device rocBLAS
 * handle and readout workspace); call after every reservoir is destroyed,
 * before the process exits.  Later calls re-create them on demand.  No
 * counterpart in the reference, whose library holds no GPU state. */
void          spires_ext_shutdown(void);

/* One evaluation of spires_optimize's objective, exactly as the optimizer
 * runs it (eval_adapter + eval_mse, spires_api.c:298-438; the reference keeps
 * them static): theta[8] = {alpha, spectral radius, input gain,
 * connectivity, E/I ratio, log10 ridge lambda, N (snapped to 200 / 400 /
 * 800 / 1200), topology (snapped to 0 / 1 / 2)}.  For each of
 * budget->num_seeds seeds a reservoir is created with
 * spires_reservoir_create (consuming rand()), its ridge readout is trained on
 * the first data_fraction of the series and its MSE measured on the rest
 * after a reset.  out[5] = {metric mean, metric std (metric = 1 / (1 + MSE)),
 * cost proxy, feasible (1 / 0), score = mean - lambda_var std - lambda_cost
 * cost (-1e6 when infeasible)}.  Writes alpha into base->neuron_params[4],
 * as spires_optimize does. */
spires_status spires_ext_optimize_evaluate(const spires_reservoir_config *base,
                                           const struct spires_opt_budget *budget,
                                           const struct spires_opt_score *score, const double *theta,
                                           const double *input_series, const double *target_series,
                                           size_t series_length, double *out);

/* AGILE (Levy exploration + SPSA refinement, the search behind
 * spires_optimize) on a caller-supplied loss.  opts[11] = {mu, eta_min,
 * eta_max, patience0, patience_decay, refine_step, refine_max_iters,
 * refine_no_improve, spsa_eps, reflect_at_bounds, seed}.  Returns 0. */
int spires_ext_agile_optimize(const double *theta0, const double *lo, const double *hi, int dim,
                              double (*loss)(const double *theta, int dim, void *ctx), void *ctx,
                              const double *opts, double *theta_best);

/* AGILE surrogate: FP32 MLP in_dim -> hidden -> hidden -> 1 (tanh hidden
 * layers, linear output) with Adam, forward and backward as HIP kernels on
 * the device.  No counterpart in the reference: it predicts a trial's
 * validation score from (hyperparameters, fidelity) so that the search can
 * prune candidates between fidelity rungs.  Parameters, in order:
 * W1[hidden x in_dim], b1[hidden], W2[hidden x hidden], b2[hidden],
 * W3[hidden], b3.  X is [B x in_dim] row-major, targets T and outputs y [B].
 * The loss is mean((y - T)^2).  in_dim <= 16, hidden <= 64. */
typedef struct spires_mlp spires_mlp;
spires_status spires_ext_mlp_create(int in_dim, int hidden, unsigned long long seed, spires_mlp **out);
void          spires_ext_mlp_destroy(spires_mlp *n);
int           spires_ext_mlp_num_params(const spires_mlp *n);
spires_status spires_ext_mlp_get_params(spires_mlp *n, float *params);
/* Sets the parameters and clears the Adam state. */
spires_status spires_ext_mlp_set_params(spires_mlp *n, const float *params);
spires_status spires_ext_mlp_predict(spires_mlp *n, const float *X, int B, float *y);
/* Gradient of the loss over the batch (grad may be NULL). */
spires_status spires_ext_mlp_gradient(spires_mlp *n, const float *X, const float *T, int B, float *grad,
                                      float *loss);
/* `epochs` full-batch Adam steps (beta1 0.9, beta2 0.999, eps 1e-8); loss of
 * the last gradient evaluation. */
spires_status spires_ext_mlp_train(spires_mlp *n, const float *X, const float *T, int B, int epochs, float lr,
                                   float *loss);
const char   *spires_ext_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
