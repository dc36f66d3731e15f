/*
 * spires_dev.h -- thin C-ABI device layer between the C++ host library and
 * the gfx950 HIP kernels.
 *
 * The host (spires_amd/csrc/host) never includes HIP headers; it hands this
 * layer plain host pointers and sizes.  Everything behind it lives in device
 * memory (HBM) of the current HIP device and runs on one HIP stream per
 * reservoir.  Functions return 0 on success, SPIRES_DEV_EALLOC on an
 * allocation failure and SPIRES_DEV_EINTERNAL otherwise;
 * spires_dev_last_error() gives the message.
 *
 * This layer has no counterpart in the reference (which is CPU-only); it
 * replaces the bodies of step_reservoir / compute_output /
 * train_output_* in src/reservoir.c:113-192, 244-280, 573-767.
 */
#ifndef SPIRES_DEV_H
#define SPIRES_DEV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPIRES_DEV_EALLOC    1
#define SPIRES_DEV_EINTERNAL 2
#define SPIRES_DEV_EINVAL    3
#define SPIRES_DEV_ESINGULAR 4   /* ridge system singular: W_out left unchanged */

typedef struct spires_dev spires_dev;

/* Per-model scalars.  Meaning by neuron model (index into np[]):
 *  LIF_DISCRETE   0 V_0, 1 V_th, 2 leak, 3 bias
 *  LIF_BIO        0 V_0, 1 V_th, 2 tau, 3 bias
 *  FLIF_GL        0 V_th, 1 V_reset, 2 V_rest, 3 tau_m, 4 pow(dt,alpha), 5 bias, 6 t_ref
 *  FLIF_CAPUTO    0 C_m, 1 g_l, 2 V_l, 3 V_th, 4 V_0, 5 kr, 6 t_ref, 7 initial t_prev
 *  FLIF_DIFFUSIVE 0 V_th, 1 V_reset, 2 V_rest, 3 g_euler, 4 bias, 5 alpha,
 *                 6 t_ref, 7 t_min, 8 t_max, 9 pow(alpha,2), 10 q,
 *                 11 pow(t_min,1-q), 12 pow(t_max,1-q), 13 1/(1-q)          */
typedef struct {
    int32_t n, din, dout, micro;
    int32_t neuron_type, synapse_type, backend;
    double  gain, dt;
    double  np[16];
    int32_t L;                 /* GL ring length / Caputo mem_len          */
    const double *coef;        /* GL c[0..L-1] / Caputo weights [L-2]      */
    /* recurrent weights: CSR (row i = inputs of neuron i) or dense n*n   */
    int64_t nnz;
    const int64_t *row_ptr;
    const int32_t *col;
    const double  *val;
    const double  *dense;      /* row-major n*n (dense backend)            */
    /* synapse tables */
    double  syn_decay;         /* PSC-homogeneous exp(-dt/tau_syn)         */
    const double *edge_decay;  /* PSC-heterogeneous exp(-dt/tau_k) per edge (CSR order or n*n) */
    int32_t frac_L;            /* fractional synapse memory length         */
    const double *frac_coef;   /* [frac_L]                                  */
    double  frac_dt_alpha, frac_tau;
    int32_t me_n;              /* multi-exponential component count        */
    const double *me_decay;    /* [me_n] exp(-dt/tau_c)                     */
    const double *me_c;        /* [me_n] mixing weights                     */
    const double *W_in;        /* [n x din]                                 */
    const double *W_out;       /* [dout x n]                                */
    const uint64_t *rng_state; /* diffusive: initial per-neuron RNG state [n] */
} spires_dev_desc;

/* Execution modes of the reservoir step. */
#define SPIRES_DEV_MODE_EXACT 0   /* FP64, bit-exact with the CPU reference order */
#define SPIRES_DEV_MODE_FAST  1   /* FP32 time-blocked history contraction        */

int  spires_dev_create(const spires_dev_desc *d, spires_dev **out);
void spires_dev_destroy(spires_dev *g);
/* Re-initialise neuron state (scalars and tables may change; synapse state
 * is kept, as in the reference's reset). */
int  spires_dev_reset_neurons(spires_dev *g, const spires_dev_desc *d);

/* One macro step with input u[din] (NULL = zeros). Asynchronous. */
int  spires_dev_step(spires_dev *g, const double *u);
/* Resident input series: upload u[T x din] once. */
int  spires_dev_upload_inputs(spires_dev *g, const double *u, size_t T);
/* Run steps t0..t0+T-1 of the resident series.  flags: */
#define SPIRES_DEV_REC_STATE  1u   /* write V after each step into the state record */
#define SPIRES_DEV_REC_RASTER 2u   /* write spikes (u8) into the raster record        */
#define SPIRES_DEV_REC_OUTPUT 4u   /* write y = W_out V into the output record        */
#define SPIRES_DEV_REC_APPEND 8u   /* record step t0+k at row t0+k (keeping rows below t0),
                                      not at row k                                      */
/* Record buffers are (re)allocated to hold T rows starting at row 0. */
int  spires_dev_run(spires_dev *g, size_t t0, size_t T, unsigned flags);
/* spires_dev_run for `count` exact-mode FLIF-GL reservoirs at once: one
 * cooperative launch in which each reservoir keeps its own workgroups,
 * exchange words, inputs and records.  They must all fit the product-wave
 * kernel with the same neurons per workgroup, and together use no more
 * workgroups than the device has compute units.  Each reservoir runs macro
 * steps [t0, t0+T) of its own uploaded input series. */
int  spires_dev_run_batch(spires_dev *const *gs, int count, size_t t0, size_t T, unsigned flags);
/* Workgroups a batch would use (0 if the reservoirs cannot run batched). */
int  spires_dev_batch_workgroups(spires_dev *const *gs, int count);
/* Time blocks replayed from a HIP graph: the launches that advance `count`
 * exact-mode FLIF-GL reservoirs (product-wave kernel) by K macro steps,
 * packed per neurons-per-workgroup class into launches of at most one
 * workgroup per CU, captured once with device-side step clocks; each replay
 * runs the next K steps of every reservoir.  Inputs are read at row
 * msteps + k; with REC_STATE / REC_RASTER, records are written at row
 * msteps + k (reserve them beforehand; reallocating inputs or records
 * invalidates the time block). */
#define SPIRES_DEV_TB_ONE_LAUNCH 16u   /* time block: every reservoir in one launch of 64-neuron
                                          one-wave slices (k_mux_gl) instead of packed
                                          product-wave launches */
typedef struct spires_dev_timeblock spires_dev_timeblock;
int  spires_dev_timeblock_create(spires_dev *const *gs, int count, long long K, unsigned flags,
                                 spires_dev_timeblock **out);
int  spires_dev_timeblock_run(spires_dev_timeblock *tb, long long blocks);
int  spires_dev_timeblock_launches(const spires_dev_timeblock *tb);
void spires_dev_timeblock_destroy(spires_dev_timeblock *tb);
/* Exact-mode launch plan of a run of this reservoir alone: info[0] kernel
 * (0 per-step launches, 1 single-wave persistent, 2 product-wave persistent,
 * 3 product-wave persistent with a window of product chunks in LDS),
 * [1] neurons per workgroup, [2] workgroups, [3] row pull (0 edge list,
 * 1 column bitmaps), [4] recurrent weights held in LDS, [5] LDS bytes per
 * workgroup; [6] neurons per workgroup in a batched launch (0: cannot batch). */
int  spires_dev_plan_info(spires_dev *g, int *info);
/* Raise the exchange abort flag, as a timed-out spike-exchange wait would:
 * the next persistent launch gives up at its first unfinished wait and the
 * launch reports the timeout (fault injection for error-path tests). */
int  spires_dev_inject_exchange_timeout(spires_dev *g);
/* Time per step of the spike exchange alone among G workgroups (one per
 * compute unit, G <= 512), over T steps of the persistent kernels' publish /
 * gather protocol with no other work: the floor the exchange puts under a
 * persistent step. */
int  spires_dev_exchange_floor(int G, long long T, double *us_per_step);
/* Pre-size the record buffers (so a later run allocates nothing). */
int  spires_dev_reserve_records(spires_dev *g, size_t rows, unsigned flags);
int  spires_dev_read_records(spires_dev *g, size_t row0, size_t rows, double *states,
                             uint8_t *raster, double *outputs);

int  spires_dev_read_state(spires_dev *g, double *V);
int  spires_dev_set_state(spires_dev *g, const double *V);
int  spires_dev_read_spikes(spires_dev *g, double *S);
int  spires_dev_compute_output(spires_dev *g, double *y);
int  spires_dev_get_readout(spires_dev *g, double *W_out);
int  spires_dev_set_readout(spires_dev *g, const double *W_out);

/* Ridge on the recorded states rows [washout, T) against targets
 * Y[(T - washout) x dout] (host): W_out = (X'X + lambda I)^-1 X'Y by FP64
 * Cholesky with iterative refinement (LU if not positive definite).
 * info_out (optional, 2 doubles): [0] 1 if LU was used, [1] final relative
 * residual of the normal equations. */
int  spires_dev_ridge(spires_dev *g, size_t washout, const double *Y, double lambda,
                      double *info_out);
/* As spires_dev_ridge on the recorded rows [row0, row1) with Y[(row1-row0) x dout]. */
int  spires_dev_ridge_rows(spires_dev *g, size_t row0, size_t row1, const double *Y, double lambda,
                           double *info_out);
/* spires_dev_ridge_rows for `count` exact-mode reservoirs of the same N and
 * Dout over the same row range, with one lambda each: the factorisations and
 * solves run batched.  info_out[2r] is 0 (Cholesky), 1 (LU) or -1 (singular:
 * that W_out is left unchanged), info_out[2r+1] the relative residual. */
int  spires_dev_ridge_rows_batch(spires_dev *const *gs, int count, size_t row0, size_t row1,
                                 const double *const *Y, const double *lambda, double *info_out);
/* Readout outputs Y[rows x Dout] = X W_out' of recorded state rows
 * [row0, row0+rows), computed on the device and copied to host `out`. */
int  spires_dev_predict_records(spires_dev *g, size_t row0, size_t rows, double *out);
/* Recursive least squares over T resident input steps with targets Y[T x dout]. */
int  spires_dev_rls(spires_dev *g, const double *Y, size_t T, double delta, double lambda);
/* Delta rule on the current state. */
int  spires_dev_online(spires_dev *g, const double *target, double lr);

int  spires_dev_set_mode(spires_dev *g, int mode, int block_k);
/* Kernel path of exact-mode runs: 0 automatic (persistent kernel when the
 * model supports it), 1 one launch per micro-step, 2 persistent only,
 * 3 persistent with a single-wave history sum streamed from HBM,
 * 4 persistent with a single-wave history sum over LDS. */
int  spires_dev_set_path(spires_dev *g, int path);
int  spires_dev_sync(spires_dev *g);
void *spires_dev_stream(spires_dev *g);
size_t spires_dev_steps(const spires_dev *g);
const char *spires_dev_last_error(void);
/* Release what the layer holds beyond any reservoir: the per-device rocBLAS
 * handle and readout workspace.  Call with no reservoir work in flight (the
 * next readout call re-creates them). */
void spires_dev_shutdown(void);

/* coarse_grain_reservoir (reservoir.c:782-967) on the current device: the
 * reference's greedy merge passes over the CSR matrix (rp, ci, val) with
 * threshold thr >= 0, in the reference's order and arithmetic.  W_in [n x din],
 * W_out [dout x n] and V [n] are merged in place (dead neurons zeroed);
 * alive[n] marks the survivors; the surviving neurons' matrix comes back as a
 * malloc'd CSR over *n_out neurons (ascending columns, zeros dropped). */
int  spires_dev_coarse_grain(int n, const int64_t *row_ptr, const int32_t *col_idx, const double *values,
                             double thr, int din, double *W_in, int dout, double *W_out, double *V,
                             int *alive, int *n_out, int64_t **row_ptr_out, int32_t **col_idx_out,
                             double **values_out, size_t *nnz_out, int *passes);
/* Reservoir construction on the current device.  Erdos-Renyi W (edge i->j,
 * i != j, with probability p; magnitude U(0,1], positive with probability ei;
 * SplitMix64 of (seed, i, j), so one seed gives one matrix on any device)
 * generated in CSR with ascending columns, its spectral radius estimated by
 * the reference power iteration and W scaled by rho / lambda when
 * lambda > 1e-9.  n <= 19968.  Outputs are malloc'd (the caller free()s
 * them); *lambda gets the unscaled estimate. */
int  spires_dev_random_er(int n, double p, double ei, unsigned long long seed, double rho, int64_t **row_ptr,
                          int32_t **col_idx, double **values, size_t *nnz, double *lambda);
/* spires_dev_random_er for any of the three topologies (type: 0 random, 1
 * small world, 2 scale free; reservoir.c:384-536), with the reference's sizes
 * from conn (ring degree K, edges per new node m) and its construction steps
 * in law, from SplitMix64 draws: the rewired ring of the small world, the
 * preferential attachment (deg + 1) with the seed clique and the uniform
 * fallback, then a fair-coin orientation, for the scale-free graph.  Then the
 * same power iteration and rescale.  n <= 19968. */
int  spires_dev_build_topology(int type, int n, double conn, double ei, unsigned long long seed, double rho,
                               int64_t **row_ptr, int32_t **col_idx, double **values, size_t *nnz,
                               double *lambda);
/* The reference power iteration (sparse.c:93-127) on the device for a host
 * CSR matrix; *lambda is bit-identical to the CPU loop's, *iters its count. */
int  spires_dev_spectral_radius(int n, const int64_t *row_ptr, const int32_t *col_idx,
                                const double *values, double *lambda, int *iters);
/* spires_dev_spectral_radius for a CSR in device memory, allocating nothing:
 * the work buffers are a fixed per-device array (n <= 19968), *lambda and
 * *iters are device pointers written in stream order on `stream` (a
 * hipStream_t, NULL = null stream), and the call does not wait.  Calls that
 * overlap on one device race on the shared array. */
int  spires_dev_spectral_radius_noalloc(int n, const int64_t *row_ptr, const int32_t *col_idx,
                                        const double *values, double *lambda, int *iters, void *stream);

#ifdef __cplusplus
}
#endif
#endif
