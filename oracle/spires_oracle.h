/*
 * spires_oracle.h -- CPU restatement of the spires reservoir path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the CPU
 * baseline: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, the product library (spires_amd) never links or calls it.
 *
 * Every function restates the reference algorithm (txmastin/spires, files
 * cited per function in spires_oracle.c) in plain C99, compiled with
 * -O2 -ffp-contract=off and no -march, so that each floating-point operation
 * is one IEEE double operation in the reference's order.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - neuron updates, GL coefficients, PSC-heterogeneous synapse and AGILE
 *     are checked bit-for-bit against the reference sources compiled from
 *     /root/reference (oracle/_ref, built by oracle/Makefile.ref);
 *   - CSR row dot (brute force), PSC-homogeneous, multi-exponential and
 *     fractional synapses (closed forms, the Mittag-Leffler spectral
 *     density, a scipy triangular solve of the GL system), RLS (weighted
 *     ridge by numpy), delta rule (geometric error contraction), power
 *     iteration (known spectra, eigvalsh), LU and ridge solves (numpy) and
 *     the construction's rand() draw order (an independent restatement of
 *     reservoir.c:307-545) are checked in tests/test_oracle_pins.py;
 *   - BLAS-order-dependent results (dense row dot, multi-input W_in*u with a
 *     non power-of-two gain, W_out*V, dgemm Gram, dgesv) are restated in
 *     sequential order; parity with the reference's BLAS build is at
 *     tolerance only ("parity unpinned" bit-for-bit).
 */
#ifndef SPIRES_ORACLE_H
#define SPIRES_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/spires.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_res or_res;

/* ---- construction ---- */
or_res *or_create(const spires_reservoir_config *cfg);
/* Build from explicit recurrent CSR (row i = inputs of neuron i), W_in and
 * optional W_out (NULL = zeros).  No rand() is consumed except by the
 * PSC-heterogeneous per-edge tau draw (as the reference's synapse build). */
or_res *or_create_csr(const spires_reservoir_config *cfg, size_t nnz,
                      const int64_t *row_ptr, const int32_t *col_idx,
                      const double *values, const double *W_in,
                      const double *W_out);
void    or_free(or_res *r);
void    or_reset(or_res *r);

/* ---- stepping ---- */
void or_step(or_res *r, const double *u);
void or_compute_output(const or_res *r, double *y);
void or_read_state(const or_res *r, double *V);
void or_read_spikes(const or_res *r, double *S);
/* T macro steps; any output pointer may be NULL.
 * states: [T x N] V after each step; raster: [T x N] spike after each step;
 * y: [T x Dout] readout after each step. */
void or_run_record(or_res *r, const double *u, size_t T, double *states,
                   uint8_t *raster, double *y);

/* ---- readout training ---- */
void or_train_ridge(or_res *r, const double *u, const double *y, size_t T,
                    double lambda);
/* Ridge on a given state matrix X [T x N] and targets Y [T x Dout];
 * writes W [Dout x N].  Gram in sequential order, LU with partial pivoting
 * per output.  Returns 0 or -1 on a singular pivot. */
int  or_ridge_solve(const double *X, const double *Y, size_t T, size_t N,
                    size_t Dout, double lambda, double *W);
void or_train_rls(or_res *r, const double *u, const double *y, size_t T,
                  double delta, double lambda);
void or_train_online(or_res *r, const double *target, double lr);
or_res *or_coarse_grain(const or_res *r, double thr);

/* ---- accessors ---- */
size_t or_num_neurons(const or_res *r);
size_t or_nnz(const or_res *r);
void   or_get_csr(const or_res *r, int64_t *row_ptr, int32_t *col_idx, double *values);
void   or_get_dense(const or_res *r, double *W);
void   or_get_W_in(const or_res *r, double *W_in);
void   or_get_W_out(const or_res *r, double *W_out);
void   or_set_W_out(or_res *r, const double *W_out);

/* ---- pieces exposed for pinning ---- */
void   or_gl_coeffs(double alpha, int n, double *c);
double or_csr_spectral_radius(size_t n, const int64_t *row_ptr,
                              const int32_t *col_idx, const double *values);
/* Topology + weight draw exactly as the reference (glibc rand()).
 * W_dense [n x n] and W_in [n x din] are outputs. */
int    or_init_weights_dense(size_t n, size_t din, int conn_type,
                             double connectivity, double ei_ratio,
                             double *W_dense, double *W_in);
int    or_lu_solve(size_t n, const double *A, const double *b, double *x);
/* Diffusive-neuron seed for neuron i (index replaces the reference's heap
 * address; see spires_oracle.c). */
uint64_t or_df_seed(size_t i, const double *p);

/* ---- probes (tests only) ----
 * or_probe_synapse: feeds spikes [T x N] to r's synapse_prepare one step at a
 * time and copies the vector the row dot then reads (trace / current) into
 * eff_out [T x N] (not meaningful for PSC-heterogeneous, whose traces are per
 * edge).  or_probe_csr_rows: every row's csr_row_dot over x.
 * or_probe_multiexp: the multi-exponential synapse's time constants and
 * mixing weights (returns their count). */
void or_probe_synapse(or_res *r, const double *spikes, size_t T, double *eff_out);
void or_probe_csr_rows(size_t n, const int64_t *rp, const int32_t *ci, const double *val,
                       const double *x, double *out);
int  or_probe_multiexp(const or_res *r, double *tau, double *c);

/* ---- AGILE (reference src/agile.c) ---- */
typedef double (*or_loss_fn)(const double *theta, int dim, void *ctx);
struct or_agile_opts {
    double mu, eta_min, eta_max, patience0, patience_decay;
    double refine_step; int refine_max_iters, refine_no_improve;
    double spsa_eps; int reflect_at_bounds; uint64_t seed; int verbose;
};
struct or_agile_report {
    double best_loss; int best_iter, total_evals, total_iters, exited_refinement;
};
int or_agile_optimize(const double *theta0, const double *lo, const double *hi,
                      int dim, or_loss_fn f, void *ctx,
                      const struct or_agile_opts *opts, double *theta_best,
                      struct or_agile_report *rep);

#ifdef __cplusplus
}
#endif
#endif
