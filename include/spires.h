/*
 * spires.h -- public C ABI of the MI355X reservoir engine.
 *
 * Drop-in replacement for the opaque-handle API of spires
 * (reference: include/spires.h:1-273).  Every type, enumerator value,
 * struct field order and function signature below is identical to the
 * reference so that existing callers recompile (or relink) unchanged.
 * The implementation behind it runs the reservoir step on one AMD
 * Instinct MI355X (gfx950) through hand-written HIP kernels; the host
 * side (construction, parameter handling, optimizer) is C++.
 *
 * Ownership and error rules follow the reference:
 *   - config param arrays are caller-owned and the POINTER is retained
 *     (re-read by spires_reservoir_reset), so they must outlive the handle;
 *   - functions returning double* hand back malloc() memory, caller free()s;
 *   - NULL handles / arguments give SPIRES_ERR_INVALID_ARG (or NULL / 0);
 *   - handles are not thread-safe; one handle = one HIP stream.
 */
#ifndef SPIRES_H
#define SPIRES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (reference include/spires.h:25-30). */
typedef enum {
    SPIRES_OK = 0,
    SPIRES_ERR_INVALID_ARG,
    SPIRES_ERR_ALLOC,
    SPIRES_ERR_INTERNAL
} spires_status;

/* Opaque reservoir handle (reference include/spires.h:33). */
typedef struct spires_reservoir spires_reservoir;

/* Recurrent topology generator (reference include/spires.h:39-43). */
typedef enum {
    SPIRES_CONN_RANDOM = 0,      /* Erdos-Renyi directed, p = connectivity   */
    SPIRES_CONN_SMALL_WORLD,     /* Watts-Strogatz directed ring, p_rw = 0.1 */
    SPIRES_CONN_SCALE_FREE       /* Barabasi-Albert, randomly oriented       */
} spires_connectivity_type;

/* Neuron model (reference include/spires.h:45-51).
 * neuron_params layouts:
 *   LIF_DISCRETE   [V_0, V_th, leak_rate, bias]
 *   LIF_BIO        [V_0, V_th, tau, bias]
 *   FLIF_CAPUTO    [C_m, g_l, V_l, V_th, V_0, V_peak, alpha, t_ref, T_mem]
 *   FLIF_GL        [V_th, V_reset, V_rest, tau_m, alpha, T_mem, bias, t_ref]
 *   FLIF_DIFFUSIVE [V_th, V_reset, V_rest, tau_m, alpha, T_mem_or_tref, bias] */
typedef enum {
    SPIRES_NEURON_LIF_DISCRETE = 0,
    SPIRES_NEURON_LIF_BIO,
    SPIRES_NEURON_FLIF_CAPUTO,
    SPIRES_NEURON_FLIF_GL,
    SPIRES_NEURON_FLIF_DIFFUSIVE
} spires_neuron_type;

/* Synapse model of the recurrent connections (reference include/spires.h:90-96).
 * synapse_params layouts:
 *   SIMPLE               unused
 *   PSC_HOMOGENEOUS      [tau_syn]
 *   PSC_HETEROGENEOUS    [tau_min, tau_max]   (per-edge tau, log-uniform)
 *   FRACTIONAL           [alpha, tau_syn, T_mem]
 *   FRACTIONAL_MULTIEXP  [alpha, tau_syn, tau_min, tau_max, n_components] */
typedef enum {
    SPIRES_SYNAPSE_SIMPLE = 0,
    SPIRES_SYNAPSE_PSC_HOMOGENEOUS,
    SPIRES_SYNAPSE_PSC_HETEROGENEOUS,
    SPIRES_SYNAPSE_FRACTIONAL,
    SPIRES_SYNAPSE_FRACTIONAL_MULTIEXP
} spires_synapse_type;

/* Storage of the recurrent matrix (reference include/spires.h:102-105). */
typedef enum {
    SPIRES_SYNAPSE_SPARSE = 0,   /* CSR row-pull */
    SPIRES_SYNAPSE_DENSE         /* dense row dot */
} spires_synapse_backend;

/* Creation-time configuration (reference include/spires.h:111-126). */
typedef struct {
    size_t num_neurons;
    size_t num_inputs;
    size_t num_outputs;
    double spectral_radius;
    double ei_ratio;
    double input_strength;
    double connectivity;
    double dt;
    spires_connectivity_type connectivity_type;
    spires_neuron_type       neuron_type;
    double *neuron_params;                    /* caller owns; pointer retained */
    spires_synapse_type      synapse_type;
    spires_synapse_backend   synapse_backend;
    double *synapse_params;                   /* caller owns; pointer retained */
} spires_reservoir_config;

/* Lifecycle (reference include/spires.h:131-134). */
spires_status spires_reservoir_create(const spires_reservoir_config *cfg,
                                      spires_reservoir **out_r);
void          spires_reservoir_destroy(spires_reservoir *r);
spires_status spires_reservoir_reset(spires_reservoir *r);

/* One macro step (1/dt micro-steps) with input u_t[num_inputs]; NULL = zeros
 * (reference include/spires.h:142). */
spires_status spires_step(spires_reservoir *r, const double *u_t);

/* Run over input_series[series_length x num_inputs]; returns a malloc'd
 * [series_length x num_outputs] array of readouts y_t = W_out * V_t
 * (reference include/spires.h:155). */
double *spires_run(spires_reservoir *r, const double *input_series, size_t series_length);

/* Readout training (reference include/spires.h:161-173). */
spires_status spires_train_online(spires_reservoir *r,
                                  const double *target_vec, double lr);
spires_status spires_train_ridge(spires_reservoir *r,
                                 const double *input_series,
                                 const double *target_series,
                                 size_t series_length, double lambda);
spires_status spires_train_rls(spires_reservoir *r,
                               const double *input_series,
                               const double *target_series,
                               size_t series_length,
                               double delta, double lambda);

/* Pairwise merge of strongly coupled neurons into a new reservoir
 * (reference include/spires.h:175-177). */
spires_status spires_coarse_grain(const spires_reservoir *r,
                                  double weight_threshold,
                                  spires_reservoir **out_r);

/* State access (reference include/spires.h:185-218). */
double       *spires_copy_reservoir_state(spires_reservoir *r);
spires_status spires_read_reservoir_state(spires_reservoir *r, double *buffer);
spires_status spires_read_spike_state(spires_reservoir *r, double *buffer);
double       *spires_copy_weights(const spires_reservoir *r);
spires_status spires_read_weights(const spires_reservoir *r, double *buffer);
spires_status spires_compute_output(spires_reservoir *r, double *out);
double       *spires_copy_readout(const spires_reservoir *r);
spires_status spires_read_readout(const spires_reservoir *r, double *buffer);

/* Introspection (reference include/spires.h:226-228). */
size_t spires_num_neurons(const spires_reservoir *r);
size_t spires_num_inputs(const spires_reservoir *r);
size_t spires_num_outputs(const spires_reservoir *r);

/* Hyperparameter search (reference include/spires.h:235-267). */
struct spires_opt_budget {
    double data_fraction;     /* 0..1 */
    int    num_seeds;
    double time_limit_sec;    /* 0 = no limit */
    int    _reserved_i0;
    double _reserved_d0;
};

enum spires_metric_kind { SPIRES_METRIC_AUROC = 0, SPIRES_METRIC_AUPRC = 1 };

struct spires_opt_score {
    double lambda_var;        /* penalty on std(metric) */
    double lambda_cost;       /* penalty on cost proxy */
    int    metric;            /* enum spires_metric_kind */
    int    _reserved_i0;
};

struct spires_opt_result {
    spires_reservoir_config best_config;
    double                  best_log10_ridge;
    double                  best_score;   /* mean - lambda_var*std - lambda_cost*cost */
    double                  metric_mean;
    double                  metric_std;
};

int spires_optimize(const spires_reservoir_config *base_config,
                    const struct spires_opt_budget *budgets,
                    int num_budgets,
                    const struct spires_opt_score *score,
                    struct spires_opt_result *out,
                    const double *input_series,
                    const double *target_series,
                    size_t series_length);

#ifdef __cplusplus
} /* extern "C" */
#endif
#endif /* SPIRES_H */
