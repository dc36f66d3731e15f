/*
 * ref_harness.c -- drives the reference's own neuron, PSC-heterogeneous
 * synapse and AGILE sources (compiled from /root/reference/src by
 * oracle/Makefile.ref into oracle/_ref/libspires_ref.so).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle restatement.
 * The reference's reservoir.c, sparse.c, synapses/simple.c and math_utils.c
 * include cblas.h / lapacke.h, which this image does not provide, so they are
 * not built; the oracle's versions of those parts are pinned by closed forms,
 * brute force and library routines instead.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "neuron.h"
#include "synapses/psc_heterogeneous.h"
#include "agile.h"
#include "spires_opt_agile.h"

void flif_diffusive_set_seed(struct flif_diffusive_neuron *n, unsigned long long seed);

/* Run n independent reference neurons of one type for T steps.
 * inputs [T x n]; V_out / S_out [T x n] (post-update V and spike).
 * seeds (diffusive only, may be NULL): per-neuron flif_diffusive_set_seed. */
int ref_neurons_run(int type, double *params, double dt, size_t n, size_t T,
                    const double *inputs, double *V_out, double *S_out,
                    const uint64_t *seeds, double *gl_coeffs_out, int ncoeffs)
{
    void **nr = malloc(n * sizeof(void *));
    void *shared = NULL;
    if (!nr) return -1;
    for (size_t i = 0; i < n; i++) {
        nr[i] = init_neuron((enum neuron_type)type, params, dt, &shared);
        if (type == FLIF_DIFFUSIVE && seeds)
            flif_diffusive_set_seed((struct flif_diffusive_neuron *)nr[i], seeds[i]);
    }
    if (type == FLIF_GL && gl_coeffs_out && ncoeffs > 0)
        memcpy(gl_coeffs_out, ((struct flif_gl_neuron *)nr[0])->coeffs, (size_t)ncoeffs * sizeof(double));
    for (size_t t = 0; t < T; t++)
        for (size_t i = 0; i < n; i++) {
            update_neuron(nr[i], (enum neuron_type)type, inputs[t * n + i], dt);
            V_out[t * n + i] = get_neuron_state(nr[i], (enum neuron_type)type);
            S_out[t * n + i] = get_neuron_spike(nr[i], (enum neuron_type)type);
        }
    for (size_t i = 0; i < n; i++) free_neuron(nr[i], (enum neuron_type)type);
    free(shared);
    free(nr);
    return 0;
}

/* Reference PSC-heterogeneous synapse: build from dense (consumes rand()),
 * optional rescale factor, then per step prepare(spikes_t) and row_dot for
 * every row.  spikes [T x n], rec_out [T x n]. */
int ref_psc_het_run(const double *dense, size_t n, const double *params, int sparse,
                    double scale, double dt, size_t T, const double *spikes,
                    double *rec_out, double *rho_out)
{
    struct psc_heterogeneous_synapse_data *d = sparse
        ? synapse_psc_heterogeneous_build_sparse(dense, n, params)
        : synapse_psc_heterogeneous_build_dense(dense, n, params);
    if (!d) return -1;
    if (rho_out) *rho_out = synapse_psc_heterogeneous_spectral_radius(d);
    if (scale != 1.0) synapse_psc_heterogeneous_scale(d, scale);
    for (size_t t = 0; t < T; t++) {
        synapse_psc_heterogeneous_prepare(d, &spikes[t * n], dt);
        for (size_t i = 0; i < n; i++) rec_out[t * n + i] = synapse_psc_heterogeneous_row_dot(d, i, NULL);
    }
    synapse_psc_heterogeneous_free(d);
    return 0;
}

/* Reference AGILE with a caller loss; gradient callback NULL (SPSA). */
int ref_agile(const double *theta0, const double *lo, const double *hi, int dim,
              agile_loss_fn f, void *ctx, const struct agile_options *opts,
              double *best, struct agile_report *rep)
{
    return agile_optimize(theta0, lo, hi, dim, f, NULL, ctx, opts, best, rep);
}

/* Reference multi-budget wrapper with a caller evaluator. */
int ref_opt_agile(spires_eval_fn eval, void *ctx, const struct spires_budget *b,
                  int nb, const struct spires_score_opts *s, struct spires_result *out)
{
    return spires_optimize_with_agile(eval, ctx, b, nb, s, out);
}
