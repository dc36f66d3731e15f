// AGILE hyperparameter search behind spires_optimize.
//
// agile_run restates AGILE (reference src/agile.c:8-351): Levy-flight
// exploration along the negative SPSA gradient with a PCG32 stream, patience
// that decays on improvement, then SPSA gradient-descent refinement from the
// best point.  The spires wrapper (src/spires_opt_agile.c:14-213) and the
// evaluator (src/spires_api.c:272-498) keep the reference's 8-dimensional
// search space, snapping, budgets and scoring; each evaluation creates a
// reservoir, trains the ridge readout and measures the hold-out MSE through
// this library's GPU path.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "host.hpp"
#include "spires_ext.h"

namespace shost {
namespace {

uint32_t pcg32(uint64_t* s) {
    uint64_t o = *s;
    *s = o * 6364136223846793005ULL + 0xda3e39cb94b95bdbULL;
    uint32_t xs = (uint32_t)(((o >> 18u) ^ o) >> 27u);
    uint32_t rot = (uint32_t)(o >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
double pcg_u01(uint64_t* s) { return pcg32(s) * (1.0 / 4294967296.0); }
double clampd(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }

struct AgileOpts {
    double mu, eta_min, eta_max, patience0, patience_decay, refine_step;
    int refine_max_iters, refine_no_improve;
    double spsa_eps;
    int reflect;
    uint64_t seed;
};

typedef double (*loss_fn)(const double* th, int d, void* ctx);

void project(double* x, const double* lo, const double* hi, int d, int refl) {
    for (int i = 0; i < d; ++i) {
        if (!refl) { x[i] = clampd(x[i], lo[i], hi[i]); continue; }
        if (x[i] < lo[i]) {
            const double over = lo[i] - x[i], span = hi[i] - lo[i];
            x[i] = lo[i] + (over - std::floor(over / span) * span);
        } else if (x[i] > hi[i]) {
            const double over = x[i] - hi[i], span = hi[i] - lo[i];
            x[i] = hi[i] - (over - std::floor(over / span) * span);
        }
    }
}

void spsa(const double* th, const double* lo, const double* hi, int d, double eps, loss_fn f, void* ctx,
          uint64_t* st, double* g) {
    double dl[64] = {0}, tp[64] = {0}, tm[64] = {0};
    for (int i = 0; i < d; ++i) {
        dl[i] = (pcg32(st) & 1) ? 1.0 : -1.0;
        tp[i] = clampd(th[i] + eps * dl[i], lo[i], hi[i]);
        tm[i] = clampd(th[i] - eps * dl[i], lo[i], hi[i]);
    }
    const double fp = f(tp, d, ctx), fm = f(tm, d, ctx);
    const double sc = (fp - fm) / (2.0 * eps);
    for (int i = 0; i < d; ++i) g[i] = sc * dl[i];
}

int agile_run(const double* th0, const double* lo, const double* hi, int d, loss_fn f, void* ctx,
              AgileOpts o, double* best) {
    if (d > 64) return -1;
    if (o.refine_step <= 0.0) o.refine_step = o.eta_min;
    double th[64], cand[64], g[64], dv[64];
    uint64_t rng = o.seed ? o.seed : 0xA57B17E5ULL;
    std::memcpy(th, th0, d * sizeof(double));
    project(th, lo, hi, d, o.reflect);
    double loss = f(th, d, ctx), bl = loss;
    std::memcpy(best, th, d * sizeof(double));
    double pat = o.patience0;
    int since = 0, iters = 0;
    bool refine = false;
    for (;;) {
        iters++;
        spsa(th, lo, hi, d, o.spsa_eps, f, ctx, &rng, g);
        double gn = 0.0;
        for (int i = 0; i < d; ++i) gn += g[i] * g[i];
        gn = std::sqrt(gn);
        if (gn > 1e-18) {
            for (int i = 0; i < d; ++i) dv[i] = (-1.0 / gn) * g[i];
        } else {
            double s = 0.0;
            for (int i = 0; i < d; ++i) { const double z = pcg_u01(&rng) * 2.0 - 1.0; dv[i] = z; s += z * z; }
            const double nn = std::sqrt(s) + 1e-18;
            for (int i = 0; i < d; ++i) dv[i] /= nn;
        }
        const double mu = o.mu <= 1.0 ? 1.000001 : o.mu;
        const double uu = pcg_u01(&rng);
        const double c1 = std::pow(o.eta_min, 1.0 - mu), c2 = std::pow(o.eta_max, 1.0 - mu);
        const double eta = std::pow(c1 + (c2 - c1) * uu, 1.0 / (1.0 - mu));
        std::memcpy(cand, th, d * sizeof(double));
        for (int i = 0; i < d; ++i) cand[i] += eta * dv[i];
        project(cand, lo, hi, d, o.reflect);
        const double cl = f(cand, d, ctx);
        if (cl < loss) {
            std::memcpy(th, cand, d * sizeof(double));
            loss = cl;
            since = 0;
            pat *= o.patience_decay;
            if (pat < 1.0) pat = 1.0;
            if (loss < bl) { bl = loss; std::memcpy(best, th, d * sizeof(double)); }
        } else {
            since++;
        }
        if (since >= (int)std::ceil(pat)) { refine = true; break; }
        if (iters > 1000000) break;
    }
    if (refine) {
        std::memcpy(th, best, d * sizeof(double));
        loss = bl;
        int ni = 0;
        for (int k = 0; k < o.refine_max_iters; ++k) {
            spsa(th, lo, hi, d, o.spsa_eps, f, ctx, &rng, g);
            for (int i = 0; i < d; ++i) th[i] += -o.refine_step * g[i];
            project(th, lo, hi, d, o.reflect);
            const double nl = f(th, d, ctx);
            if (nl < loss - 1e-12) {
                loss = nl;
                if (loss < bl) { bl = loss; std::memcpy(best, th, d * sizeof(double)); }
                ni = 0;
            } else {
                ni++;
            }
            if (ni >= o.refine_no_improve) break;
        }
    }
    return 0;
}

size_t snap_N(double x) {   // spires_api.c:272-281
    static const size_t ch[] = {200, 400, 800, 1200};
    size_t best = ch[0];
    double bd = std::fabs(x - (double)ch[0]);
    for (size_t i = 1; i < 4; ++i) {
        const double dd = std::fabs(x - (double)ch[i]);
        if (dd < bd) { bd = dd; best = ch[i]; }
    }
    return best;
}
int snap_topo(double t) {
    int k = (int)llround(t);
    if (k < 0) k = 0;
    if (k > 2) k = 2;
    return k;
}

struct OptCtx {
    const spires_reservoir_config* base;
    const double* x;
    const double* y;
    size_t n;
    const spires_opt_budget* budget;
    const spires_opt_score* score;
};

struct EvalOut { double mean, stdv, cost; int feasible; };

// eval_adapter + eval_mse, spires_api.c:298-438.
EvalOut evaluate(const double* th, OptCtx* c) {
    EvalOut out{0.0, 1.0, 1.0, 0};
    spires_reservoir_config cfg = *c->base;
    cfg.num_neurons = snap_N(th[6]);
    cfg.spectral_radius = th[1];
    cfg.input_strength = th[2];
    cfg.connectivity = th[3];
    cfg.ei_ratio = th[4];
    cfg.connectivity_type = (spires_connectivity_type)snap_topo(th[7]);
    double local[8] = {0};
    if (!cfg.neuron_params) cfg.neuron_params = local;
    cfg.neuron_params[4] = th[0];   // written into the caller's array, as the reference does
    const size_t din = cfg.num_inputs ? cfg.num_inputs : 1;
    const size_t dout = cfg.num_outputs ? cfg.num_outputs : 1;
    const size_t n_train = (size_t)((double)c->n * std::fmax(0.0, std::fmin(1.0, c->budget->data_fraction)));
    if (n_train < 8 || c->n <= n_train) return out;
    const size_t n_valid = c->n - n_train;
    const int seeds = c->budget->num_seeds > 0 ? c->budget->num_seeds : 1;
    double sum = 0.0, sum2 = 0.0;
    int ok = 0;
    std::vector<double> yhat(n_valid * dout);
    for (int s = 0; s < seeds; ++s) {
        spires_reservoir* r = nullptr;
        if (spires_reservoir_create(&cfg, &r) != SPIRES_OK || !r) continue;
        const double lambda = std::pow(10.0, th[5]);
        if (spires_train_ridge(r, c->x, c->y, n_train, lambda) != SPIRES_OK) { spires_reservoir_destroy(r); continue; }
        if (spires_reservoir_reset(r) != SPIRES_OK) { spires_reservoir_destroy(r); continue; }
        double mse = INFINITY;
        if (spires_ext_run_record(r, c->x + n_train * din, n_valid, nullptr, nullptr, yhat.data()) == SPIRES_OK) {
            double se = 0.0;
            for (size_t t = 0; t < n_valid; ++t)
                for (size_t o = 0; o < dout; ++o) {
                    const double e = c->y[(n_train + t) * dout + o] - yhat[t * dout + o];
                    se += e * e;
                }
            mse = se / (double)(n_valid * dout);
        }
        spires_reservoir_destroy(r);
        if (!std::isfinite(mse) || mse < 0.0) continue;
        const double metric = 1.0 / (1.0 + mse);
        sum += metric;
        sum2 += metric * metric;
        ok++;
    }
    if (ok == 0) return out;
    const double mean = sum / (double)ok;
    const double var = std::fmax(0.0, (sum2 / (double)ok) - mean * mean);
    out.mean = mean;
    out.stdv = std::sqrt(var);
    const double sparsity = cfg.connectivity > 1e-8 ? cfg.connectivity : 1e-8;
    out.cost = (double)cfg.num_neurons * (1.0 / sparsity) * 1e-6;
    out.feasible = 1;
    return out;
}

double score_of(const EvalOut& e, const spires_opt_score* s) {
    return e.mean - s->lambda_var * e.stdv - s->lambda_cost * e.cost;
}

double loss_fn_spires(const double* th, int, void* ctx) {   // spires_opt_agile.c:102-123
    OptCtx* c = (OptCtx*)ctx;
    double t[8];
    std::memcpy(t, th, sizeof(t));
    t[6] = (double)snap_N(t[6]);
    t[7] = (double)snap_topo(t[7]);
    const EvalOut e = evaluate(t, c);
    if (!e.feasible) return 1e6;
    return -score_of(e, c->score);
}

}  // namespace
}  // namespace shost

using namespace shost;

/* AGILE on a caller loss (same algorithm that drives spires_optimize).
 * opts: mu, eta_min, eta_max, patience0, patience_decay, refine_step,
 * refine_max_iters, refine_no_improve, spsa_eps, reflect_at_bounds, seed. */
extern "C" int spires_ext_agile_optimize(const double* theta0, const double* lo, const double* hi, int dim,
                                         double (*loss)(const double*, int, void*), void* ctx,
                                         const double* opts, double* theta_best) {
    if (!theta0 || !lo || !hi || !loss || !opts || !theta_best || dim <= 0 || dim > 64) return -1;
    AgileOpts o{opts[0], opts[1], opts[2], opts[3], opts[4], opts[5], (int)opts[6], (int)opts[7], opts[8],
                (int)opts[9], (uint64_t)opts[10]};
    return agile_run(theta0, lo, hi, dim, loss, ctx, o, theta_best);
}

extern "C" spires_status spires_ext_optimize_evaluate(const spires_reservoir_config* base,
                                                      const struct spires_opt_budget* budget,
                                                      const struct spires_opt_score* score, const double* theta,
                                                      const double* input_series, const double* target_series,
                                                      size_t series_length, double* out) {
    if (!base || !budget || !score || !theta || !input_series || !target_series || !out) return SPIRES_ERR_INVALID_ARG;
    OptCtx ctx{base, input_series, target_series, series_length, budget, score};
    double t[8];
    std::memcpy(t, theta, sizeof(t));
    t[6] = (double)snap_N(t[6]);
    t[7] = (double)snap_topo(t[7]);
    const EvalOut e = evaluate(t, &ctx);
    out[0] = e.mean;
    out[1] = e.stdv;
    out[2] = e.cost;
    out[3] = e.feasible ? 1.0 : 0.0;
    out[4] = e.feasible ? score_of(e, score) : -1e6;
    return SPIRES_OK;
}

extern "C" int spires_optimize(const spires_reservoir_config* base, const struct spires_opt_budget* budgets,
                               int num_budgets, const struct spires_opt_score* score,
                               struct spires_opt_result* out, const double* input_series,
                               const double* target_series, size_t series_length) {
    if (!base || !budgets || num_budgets <= 0 || !score || !out) return -1;
    // spires_opt_agile.c:60-83, 146-159
    const double lo[8] = {0.05, 0.75, 0.20, 0.08, 0.70, -5.0, 200.0, 0.0};
    const double hi[8] = {1.0, 1.05, 1.50, 0.35, 1.10, -1.0, 1200.0, 2.0};
    double th0[8] = {0.55, 0.90, 1.00, 0.20, 0.80, -2.0, 400.0, 0.0};
    double best[8];
    AgileOpts o{1.3, 1e-3, 0.3, 100.0, 0.9, 1e-2, 600, 60, 1e-3, 1, 0xA91EULL};
    OptCtx ctx{base, input_series, target_series, series_length, nullptr, score};
    for (int b = 0; b < num_budgets; ++b) {
        ctx.budget = &budgets[b];
        if (agile_run(th0, lo, hi, 8, loss_fn_spires, &ctx, o, best) != 0) return -2;
        std::memcpy(th0, best, sizeof(th0));
    }
    double pb[8];
    std::memcpy(pb, best, sizeof(pb));
    pb[6] = (double)snap_N(pb[6]);
    pb[7] = (double)snap_topo(pb[7]);
    ctx.budget = &budgets[num_budgets - 1];
    const EvalOut e = evaluate(pb, &ctx);
    out->best_config = *base;
    out->best_config.num_neurons = snap_N(pb[6]);
    out->best_config.spectral_radius = pb[1];
    out->best_config.input_strength = pb[2];
    out->best_config.connectivity = pb[3];
    out->best_config.ei_ratio = pb[4];
    out->best_config.connectivity_type = (spires_connectivity_type)snap_topo(pb[7]);
    out->best_log10_ridge = pb[5];
    out->metric_mean = e.mean;
    out->metric_std = e.stdv;
    out->best_score = score_of(e, score);
    return 0;
}
