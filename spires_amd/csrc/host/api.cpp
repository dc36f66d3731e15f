// The spires C ABI (include/spires.h, include/spires_ext.h) on top of the
// device layer (include/spires_dev.h).
//
// Host responsibilities: argument validation and status codes
// (spires_api.c:22-261), reservoir construction with glibc rand()
// (reservoir.c:18-78, 345-571), per-model constant tables computed with the
// same libm calls as the reference (so pow/exp/tgamma values are identical),
// and the coarse-graining transform (reservoir.c:782-967).  Stepping,
// readouts and state live on the GPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host.hpp"
#include "spires_ext.h"

namespace shost {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const std::string& last_error() { return g_err; }

namespace {

int mem_len_gl(const double* p, double dt) {
    int L = (p[5] > 0 && dt > 0) ? (int)(p[5] / dt) : 2000;   // flif_gl.c:37-38
    if (L > kMaxMemLen) L = kMaxMemLen;
    return L;
}

int mem_len_caputo(const double* p, double dt) {
    int L = (p[8] > 0) ? (int)(p[8] / dt) : kMaxMemLen;        // flif_caputo.c:27-28
    if (L > kMaxMemLen) L = kMaxMemLen;
    return L;
}

uint64_t sm64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97f4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Diffusive neuron RNG state.  The reference seeds from the neuron's heap
// address (flif_diffusive.c:118-126), which is not reproducible; the neuron
// index takes its place, mixed with the parameters the same way, and the
// state is then set as flif_diffusive_set_seed does (flif_diffusive.c:131-138).
uint64_t diffusive_state(size_t i, const double* p) {
    uint64_t seed = 0x9E3779B97f4A7C15ULL ^ ((uint64_t)(i + 1) * 0xD1B54A32D192ED03ULL);
    for (int k = 0; k < 7; ++k) {
        uint64_t b;
        std::memcpy(&b, &p[k], 8);
        seed ^= b + 0x9E3779B97f4A7C15ULL;
        seed = sm64(&seed);
    }
    if (seed == 0ULL) seed = 0xD1B54A32D192ED03ULL;
    (void)sm64(&seed);
    (void)sm64(&seed);
    return seed;
}

// Tables that must outlive a spires_dev_desc (its pointers refer into these).
struct DescStore {
    std::vector<double> coef, edge_decay, frac_coef, me_decay, me_c, W_out_zero;
    std::vector<uint64_t> rng;
};

// Neuron scalars / tables from the (retained) parameter array.  Returns
// SPIRES_ERR_INVALID_ARG for parameter sets the reference cannot run
// (zero-length memories that it would divide by or index out of bounds).
spires_status neuron_desc(spires_reservoir* r, spires_dev_desc& d, DescStore& st, bool first) {
    const double* p = r->cfg.neuron_params;
    const double dt = r->cfg.dt;
    if (!p) { set_error("neuron_params is NULL"); return SPIRES_ERR_INVALID_ARG; }
    std::memset(d.np, 0, sizeof(d.np));
    d.L = 0;
    d.coef = nullptr;
    switch (r->cfg.neuron_type) {
    case SPIRES_NEURON_LIF_DISCRETE:   // lif_discrete.c:12-15
    case SPIRES_NEURON_LIF_BIO:        // lif_bio.c:6-9
        for (int k = 0; k < 4; ++k) d.np[k] = p[k];
        break;
    case SPIRES_NEURON_FLIF_GL: {      // flif_gl.c:7-48, neuron.c:36-53
        const int L = mem_len_gl(p, dt);
        if (L < 1) { set_error("FLIF_GL memory length T_mem/dt must be >= 1"); return SPIRES_ERR_INVALID_ARG; }
        if (first) {
            // coefficient table built once at creation from alpha then (neuron.c:39-51)
            const int nc = (p[5] > 0 && dt > 0) ? (int)(p[5] / dt) : 2000;
            r->gl_ncoef = nc;
            r->gl_coef.assign((size_t)nc, 0.0);
            r->gl_coef[0] = 1.0;
            for (int k = 1; k < nc; ++k)
                r->gl_coef[k] = r->gl_coef[k - 1] * (1.0 - (p[4] + 1.0) / (float)k);
        }
        if (L > r->gl_ncoef) { set_error("FLIF_GL T_mem grew beyond the coefficient table made at creation"); return SPIRES_ERR_INVALID_ARG; }
        d.np[0] = p[0]; d.np[1] = p[1]; d.np[2] = p[2]; d.np[3] = p[3];
        d.np[4] = std::pow(dt, p[4]);
        d.np[5] = p[6]; d.np[6] = p[7];
        d.L = L;
        d.coef = r->gl_coef.data();
        break;
    }
    case SPIRES_NEURON_FLIF_CAPUTO: {  // flif_caputo.c:6-41
        const int L = mem_len_caputo(p, dt);
        if (L < 2) { set_error("FLIF_CAPUTO memory length T_mem/dt must be >= 2"); return SPIRES_ERR_INVALID_ARG; }
        const int m = L - 2;
        st.coef.assign((size_t)(m > 0 ? m : 1), 0.0);
        const double a = p[6];
        for (int i = 0; i < m; ++i) {
            double x = i + 2;
            st.coef[i] = std::pow(x, 1.0 - a) - std::pow(x - 1.0, 1.0 - a);
        }
        d.np[0] = p[0]; d.np[1] = p[1]; d.np[2] = p[2]; d.np[3] = p[3]; d.np[4] = p[4];
        d.np[5] = std::pow(dt, a) * std::tgamma(2.0 - a);
        d.np[6] = p[7];
        d.np[7] = 2 * p[7];
        d.L = L;
        d.coef = st.coef.data();
        break;
    }
    case SPIRES_NEURON_FLIF_DIFFUSIVE: {   // flif_diffusive.c:68-129
        double tau = p[3], alpha = p[4];
        const double P6 = p[5];
        if (tau <= 0.0) tau = 1.0;
        if (alpha < 0.0) alpha = 0.0;
        if (alpha > 1.0) alpha = 1.0;
        double tmin, tmax, tref;
        if (alpha < 1.0) {
            tmin = 1.0 * dt;
            tmax = (P6 > 0.0) ? P6 : 100.0 * tau;
            if (tmax <= tmin) tmax = tmin * 2.0;
            tref = dt;
        } else {
            tref = (P6 > 0.0) ? P6 : 0.0;
            tmin = 1.0 * dt;
            tmax = tref;
        }
        const double q = 6.0 * alpha, e = 1.0 - q;
        d.np[0] = p[0]; d.np[1] = p[1]; d.np[2] = p[2];
        d.np[3] = dt / tau; d.np[4] = p[6]; d.np[5] = alpha; d.np[6] = tref;
        d.np[7] = tmin; d.np[8] = tmax;
        d.np[9] = std::pow(alpha, 2.0); d.np[10] = q;
        d.np[11] = std::pow(tmin, e); d.np[12] = std::pow(tmax, e); d.np[13] = 1.0 / e;
        const size_t n = r->cfg.num_neurons;
        st.rng.resize(n);
        for (size_t i = 0; i < n; ++i) st.rng[i] = diffusive_state(i, p);
        d.rng_state = st.rng.data();
        break;
    }
    default:
        set_error("unknown neuron type");
        return SPIRES_ERR_INVALID_ARG;
    }
    return SPIRES_OK;
}

spires_status build_desc(spires_reservoir* r, spires_dev_desc& d, DescStore& st,
                         const double* W_out) {
    std::memset(&d, 0, sizeof(d));
    const auto& c = r->cfg;
    d.n = (int32_t)c.num_neurons;
    d.din = (int32_t)c.num_inputs;
    d.dout = (int32_t)c.num_outputs;
    d.micro = r->micro;
    d.neuron_type = c.neuron_type;
    d.synapse_type = c.synapse_type;
    d.backend = c.synapse_backend;
    d.gain = c.input_strength;
    d.dt = c.dt;
    spires_status s = neuron_desc(r, d, st, true);
    if (s != SPIRES_OK) return s;
    d.W_in = r->W_in.data();
    d.W_out = W_out;
    if (c.synapse_backend == SPIRES_SYNAPSE_SPARSE) {
        d.nnz = (int64_t)r->W.val.size();
        d.row_ptr = r->W.rp.data();
        d.col = r->W.ci.data();
        d.val = r->W.val.data();
    } else {
        d.dense = r->dense.data();
    }
    const double* sp = c.synapse_params;
    const double dt = c.dt;
    switch (c.synapse_type) {
    case SPIRES_SYNAPSE_SIMPLE: break;
    case SPIRES_SYNAPSE_PSC_HOMOGENEOUS: {   // psc_homogeneous.c:25, 80
        const double tau = sp ? sp[0] : 1.0;
        d.syn_decay = std::exp(-dt / tau);
        break;
    }
    case SPIRES_SYNAPSE_PSC_HETEROGENEOUS: { // psc_heterogeneous.c:219
        st.edge_decay.resize(r->het_tau.size());
        for (size_t k = 0; k < r->het_tau.size(); ++k) st.edge_decay[k] = std::exp(-dt / r->het_tau[k]);
        d.edge_decay = st.edge_decay.data();
        break;
    }
    case SPIRES_SYNAPSE_FRACTIONAL: {        // fractional.c:24-45, 103-120
        const double alpha = sp ? sp[0] : 1.0, tau = sp ? sp[1] : 1.0, Tm = sp ? sp[2] : 10.0;
        int L = (Tm > 0 && dt > 0) ? (int)(Tm / dt) : 2000;
        if (L > kMaxMemLen) L = kMaxMemLen;
        if (L < 1) L = 1;
        st.frac_coef.assign((size_t)L, 0.0);
        st.frac_coef[0] = 1.0;
        for (int k = 1; k < L; ++k) st.frac_coef[k] = st.frac_coef[k - 1] * (1.0 - (alpha + 1.0) / (double)k);
        d.frac_L = L;
        d.frac_coef = st.frac_coef.data();
        d.frac_dt_alpha = std::pow(dt, alpha);
        d.frac_tau = tau;
        break;
    }
    case SPIRES_SYNAPSE_FRACTIONAL_MULTIEXP: {   // fractional_multiexp.c:18-90
        const double PI = 3.14159265358979323846;
        const double alpha = sp ? sp[0] : 0.5, tsyn = sp ? sp[1] : 1.0;
        const double tmin = sp ? sp[2] : 1.0, tmax = sp ? sp[3] : 20.0;
        int Nc = sp ? (int)(sp[4] + 0.5) : 8;
        if (Nc < 1) Nc = 1;
        std::vector<double> tau((size_t)Nc), cc((size_t)Nc);
        const double sa = std::sin(alpha * PI);
        if (std::fabs(sa) < 1e-9) {
            Nc = 1; tau[0] = tsyn; cc[0] = 1.0;
        } else {
            const double l0 = std::log(tmin), l1 = std::log(tmax);
            const double dl = (Nc > 1) ? (l1 - l0) / (Nc - 1) : 0.0;
            const double ca = std::cos(alpha * PI);
            std::vector<double> raw((size_t)Nc);
            double sum = 0.0;
            for (int i = 0; i < Nc; ++i) {
                const double ti = std::exp(l0 + i * dl);
                const double x = tsyn / ti;
                const double K = (1.0 / PI) * std::pow(x, alpha - 1.0) * sa /
                                 (std::pow(x, 2.0 * alpha) + 2.0 * std::pow(x, alpha) * ca + 1.0);
                double w = K * x * dl;
                if (w < 0.0) w = 0.0;
                raw[i] = w * ti; tau[i] = ti; sum += raw[i];
            }
            for (int i = 0; i < Nc; ++i) cc[i] = (sum > 0.0) ? raw[i] / sum : 1.0 / Nc;
        }
        if (Nc > 64) { set_error("FRACTIONAL_MULTIEXP supports at most 64 components"); return SPIRES_ERR_INVALID_ARG; }
        st.me_decay.resize((size_t)Nc);
        st.me_c.resize((size_t)Nc);
        for (int i = 0; i < Nc; ++i) { st.me_decay[i] = std::exp(-dt / tau[i]); st.me_c[i] = cc[i]; }
        d.me_n = Nc;
        d.me_decay = st.me_decay.data();
        d.me_c = st.me_c.data();
        break;
    }
    default:
        set_error("unknown synapse type");
        return SPIRES_ERR_INVALID_ARG;
    }
    return SPIRES_OK;
}

spires_status dev_status(int rc) {
    if (rc == 0) return SPIRES_OK;
    set_error(spires_dev_last_error());
    if (rc == SPIRES_DEV_EALLOC) return SPIRES_ERR_ALLOC;
    if (rc == SPIRES_DEV_EINVAL) return SPIRES_ERR_INVALID_ARG;
    return SPIRES_ERR_INTERNAL;
}

spires_status validate_cfg(const spires_reservoir_config* cfg) {
    if (!cfg) return SPIRES_ERR_INVALID_ARG;
    // spires_api.c:29-38
    if (!(cfg->dt > 0.0) || cfg->dt > 1.0) { set_error("dt must be in (0, 1.0]"); return SPIRES_ERR_INVALID_ARG; }
    const double steps_d = 1.0 / cfg->dt;
    const int steps = (int)llround(steps_d);
    if (steps < 1 || std::fabs(steps_d - (double)steps) > 1e-12) {
        set_error("dt must evenly divide 1.0");
        return SPIRES_ERR_INVALID_ARG;
    }
    if (cfg->num_neurons < 1 || cfg->num_neurons > (size_t)INT32_MAX / 2) { set_error("num_neurons out of range"); return SPIRES_ERR_INVALID_ARG; }
    if ((int)cfg->neuron_type < 0 || (int)cfg->neuron_type > 4) { set_error("unknown neuron type"); return SPIRES_ERR_INVALID_ARG; }
    if ((int)cfg->synapse_type < 0 || (int)cfg->synapse_type > 4) { set_error("unknown synapse type"); return SPIRES_ERR_INVALID_ARG; }
    if ((int)cfg->synapse_backend < 0 || (int)cfg->synapse_backend > 1) { set_error("unknown synapse backend"); return SPIRES_ERR_INVALID_ARG; }
    if (!cfg->neuron_params) { set_error("neuron_params is NULL"); return SPIRES_ERR_INVALID_ARG; }
    return SPIRES_OK;
}

// Synapse build: dense copy for the dense backend, PSC-heterogeneous per-edge
// tau draws (psc_heterogeneous.c:33-86: log-uniform, CSR order or all n*n).
void synapse_build(spires_reservoir* r) {
    const size_t n = r->cfg.num_neurons;
    if (r->cfg.synapse_backend == SPIRES_SYNAPSE_DENSE) {
        r->dense.assign(n * n, 0.0);
        for (size_t i = 0; i < n; ++i)
            for (int64_t k = r->W.rp[i]; k < r->W.rp[i + 1]; ++k) r->dense[i * n + r->W.ci[k]] = r->W.val[k];
    }
    if (r->cfg.synapse_type == SPIRES_SYNAPSE_PSC_HETEROGENEOUS) {
        const double* sp = r->cfg.synapse_params;
        const double tmin = sp ? sp[0] : 1.0, tmax = sp ? sp[1] : 1.0;
        const size_t m = r->cfg.synapse_backend == SPIRES_SYNAPSE_SPARSE ? r->W.val.size() : n * n;
        r->het_tau.resize(m);
        const double lmin = std::log(tmin), lmax = std::log(tmax);
        for (size_t k = 0; k < m; ++k) r->het_tau[k] = std::exp(lmin + urand01() * (lmax - lmin));
    }
}

// rescale_weights, reservoir.c:548-557.
void rescale(spires_reservoir* r) {
    const size_t n = r->cfg.num_neurons;
    const bool sparse = r->cfg.synapse_backend == SPIRES_SYNAPSE_SPARSE;
    const double lam = sparse ? csr_spectral_radius(r->W) : dense_spectral_radius(r->dense, n);
    if (lam > 1e-9) {
        const double f = r->cfg.spectral_radius / lam;
        for (double& v : r->W.val) v *= f;
        if (!sparse) for (double& v : r->dense) v *= f;
    }
}

spires_status finish_create(spires_reservoir* r, const double* W_out, spires_reservoir** out_r) {
    spires_dev_desc d;
    DescStore st;
    spires_status s = build_desc(r, d, st, W_out);
    if (s != SPIRES_OK) { delete r; return s; }
    int rc = spires_dev_create(&d, &r->dev);
    if (rc) { s = dev_status(rc); delete r; return s == SPIRES_ERR_ALLOC ? SPIRES_ERR_ALLOC : SPIRES_ERR_INTERNAL; }
    *out_r = r;
    return SPIRES_OK;
}

}  // namespace
}  // namespace shost

using namespace shost;

extern "C" {

spires_status spires_reservoir_create(const spires_reservoir_config* cfg, spires_reservoir** out_r) {
    if (!cfg || !out_r) return SPIRES_ERR_INVALID_ARG;
    spires_status s = validate_cfg(cfg);
    if (s != SPIRES_OK) return s;
    spires_reservoir* r = new (std::nothrow) spires_reservoir();
    if (!r) return SPIRES_ERR_ALLOC;
    r->cfg = *cfg;
    r->micro = (int)llround(1.0 / cfg->dt);
    const size_t n = cfg->num_neurons;
    // init_reservoir: weights, rescale, random readout (reservoir.c:68-78)
    if (generate_topology(n, cfg->num_inputs, cfg->connectivity_type, cfg->connectivity, cfg->ei_ratio,
                          r->W_in, r->W) != 0) {
        delete r;
        set_error("unknown connectivity type");
        return SPIRES_ERR_INTERNAL;
    }
    synapse_build(r);
    rescale(r);
    std::vector<double> W_out(n * cfg->num_outputs);
    for (size_t k = 0; k < W_out.size(); ++k) W_out[k] = gen_weight(0.5);   // reservoir.c:559-571
    return finish_create(r, W_out.data(), out_r);
}

spires_status spires_ext_create_csr(const spires_reservoir_config* cfg, size_t nnz, const int64_t* row_ptr,
                                    const int32_t* col_idx, const double* values, const double* W_in,
                                    const double* W_out, spires_reservoir** out_r) {
    if (!cfg || !out_r || !row_ptr || (nnz && (!col_idx || !values)) || (cfg->num_inputs && !W_in))
        return SPIRES_ERR_INVALID_ARG;
    spires_status s = validate_cfg(cfg);
    if (s != SPIRES_OK) return s;
    const size_t n = cfg->num_neurons;
    if (row_ptr[0] != 0 || (size_t)row_ptr[n] != nnz) { set_error("row_ptr inconsistent with nnz"); return SPIRES_ERR_INVALID_ARG; }
    for (size_t i = 0; i < n; ++i) {
        if (row_ptr[i + 1] < row_ptr[i]) { set_error("row_ptr not monotone"); return SPIRES_ERR_INVALID_ARG; }
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
            if (col_idx[k] < 0 || (size_t)col_idx[k] >= n) { set_error("column index out of range"); return SPIRES_ERR_INVALID_ARG; }
            if (k > row_ptr[i] && col_idx[k] <= col_idx[k - 1]) { set_error("columns must be strictly ascending"); return SPIRES_ERR_INVALID_ARG; }
        }
    }
    spires_reservoir* r = new (std::nothrow) spires_reservoir();
    if (!r) return SPIRES_ERR_ALLOC;
    r->cfg = *cfg;
    r->micro = (int)llround(1.0 / cfg->dt);
    r->W_in.assign(W_in, W_in + n * cfg->num_inputs);
    r->W.n = n;
    r->W.rp.assign(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
            if (values[k] != 0.0) { r->W.ci.push_back(col_idx[k]); r->W.val.push_back(values[k]); }
        r->W.rp[i + 1] = (int64_t)r->W.ci.size();
    }
    synapse_build(r);
    return finish_create(r, W_out, out_r);
}

void spires_reservoir_destroy(spires_reservoir* r) {
    if (!r) return;
    spires_dev_destroy(r->dev);
    delete r;
}

// reset_reservoir (reservoir.c:770-780): neurons re-read from the retained
// params pointer; the FLIF-GL coefficient table of creation is kept.
spires_status spires_reservoir_reset(spires_reservoir* r) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    spires_dev_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n = (int32_t)r->cfg.num_neurons;
    d.dt = r->cfg.dt;
    d.gain = r->cfg.input_strength;
    DescStore st;
    spires_status s = neuron_desc(r, d, st, false);
    if (s != SPIRES_OK) return s;
    return dev_status(spires_dev_reset_neurons(r->dev, &d));
}

spires_status spires_step(spires_reservoir* r, const double* u_t) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_step(r->dev, u_t));
}

double* spires_run(spires_reservoir* r, const double* input_series, size_t T) {
    if (!r || !r->dev || !input_series) return nullptr;
    const size_t D = r->cfg.num_outputs;
    const size_t cnt = T * D;
    double* out = (double*)std::malloc((cnt > 0 ? cnt : 1) * sizeof(double));
    if (!out) return nullptr;
    if (T == 0) return out;
    if (spires_dev_upload_inputs(r->dev, input_series, T) ||
        spires_dev_run(r->dev, 0, T, SPIRES_DEV_REC_OUTPUT) ||
        spires_dev_read_records(r->dev, 0, T, nullptr, nullptr, out)) {
        set_error(spires_dev_last_error());
        std::free(out);
        return nullptr;
    }
    return out;
}

spires_status spires_train_online(spires_reservoir* r, const double* target_vec, double lr) {
    if (!r || !r->dev || !target_vec) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_online(r->dev, target_vec, lr));
}

namespace {
// Reset, run u, drop `washout` states, fit W_out.  Returns the device layer's
// code; a failed reset returns -1 with its status in *reset_status.
int ridge_washout_dev(spires_reservoir* r, const double* u, const double* y, size_t T, size_t washout,
                      double lambda, double* info, spires_status* reset_status) {
    *reset_status = spires_reservoir_reset(r);
    if (*reset_status != SPIRES_OK) return -1;
    if (T <= washout) {
        // empty design matrix: A = lambda I, b = 0 -> W_out = 0 (reference LU of lambda*I)
        if (lambda != 0.0) {
            std::vector<double> z(r->cfg.num_neurons * r->cfg.num_outputs, 0.0);
            if (!z.empty()) return spires_dev_set_readout(r->dev, z.data());
        }
        return 0;
    }
    int rc = spires_dev_upload_inputs(r->dev, u, T);
    if (!rc) rc = spires_dev_run(r->dev, 0, T, SPIRES_DEV_REC_STATE);
    if (!rc) rc = spires_dev_ridge(r->dev, washout, y, lambda, info);
    return rc;
}
}  // namespace

spires_status spires_ext_train_ridge_washout(spires_reservoir* r, const double* u, const double* y, size_t T,
                                             size_t washout, double lambda, double* info) {
    if (!r || !r->dev || !u || !y) return SPIRES_ERR_INVALID_ARG;
    spires_status s = SPIRES_OK;
    const int rc = ridge_washout_dev(r, u, y, T, washout, lambda, info, &s);
    return rc < 0 ? s : dev_status(rc);
}

spires_status spires_train_ridge(spires_reservoir* r, const double* input_series, const double* target_series,
                                 size_t series_length, double lambda) {
    if (!r || !r->dev || !input_series || !target_series) return SPIRES_ERR_INVALID_ARG;
    spires_status s = SPIRES_OK;
    const int rc = ridge_washout_dev(r, input_series, target_series, series_length, 0, lambda, nullptr, &s);
    if (rc < 0) return s;
    // The reference reports success when an output's solve fails (reservoir.c:663-667) and
    // keeps that W_out; run, HIP and exchange-timeout failures still reach the caller.
    if (rc == SPIRES_DEV_ESINGULAR) { set_error(spires_dev_last_error()); return SPIRES_OK; }
    return dev_status(rc);
}

spires_status spires_train_rls(spires_reservoir* r, const double* input_series, const double* target_series,
                               size_t series_length, double delta, double lambda) {
    if (!r || !r->dev || !input_series || !target_series) return SPIRES_ERR_INVALID_ARG;
    spires_status s = spires_reservoir_reset(r);
    if (s != SPIRES_OK) return s;
    int rc = spires_dev_upload_inputs(r->dev, input_series, series_length);
    if (!rc) rc = spires_dev_rls(r->dev, target_series, series_length, delta, lambda);
    return dev_status(rc);
}

// coarse_grain_reservoir, reservoir.c:782-967: the merge passes run on the
// device (spires_dev_coarse_grain, same order and arithmetic); the new
// reservoir is not rescaled and keeps the merged weights.
spires_status spires_coarse_grain(const spires_reservoir* rc_, double thr, spires_reservoir** out_r) {
    if (!rc_ || !rc_->dev || !out_r) return SPIRES_ERR_INVALID_ARG;
    spires_reservoir* r = const_cast<spires_reservoir*>(rc_);
    const size_t n = r->cfg.num_neurons, din = r->cfg.num_inputs, dout = r->cfg.num_outputs;
    if (n > (size_t)INT32_MAX) return SPIRES_ERR_INVALID_ARG;
    std::vector<double> Wi = r->W_in, Wo(n * dout), V(n);
    if (dout && spires_dev_get_readout(r->dev, Wo.data())) return dev_status(SPIRES_DEV_EINTERNAL);
    if (spires_dev_read_state(r->dev, V.data())) return dev_status(SPIRES_DEV_EINTERNAL);
    std::vector<int> alive(n, 1);
    int ns = 0;
    int64_t* rp = nullptr;
    int32_t* ci = nullptr;
    double* val = nullptr;
    size_t nnz = 0;
    spires_status s = dev_status(spires_dev_coarse_grain((int)n, r->W.rp.data(), r->W.ci.data(), r->W.val.data(), thr,
                                                         (int)din, Wi.data(), (int)dout, Wo.data(), V.data(),
                                                         alive.data(), &ns, &rp, &ci, &val, &nnz, nullptr));
    if (s != SPIRES_OK) return s;
    std::vector<double> Win((size_t)ns * din), Won((size_t)ns * dout), Vn(ns);
    size_t a = 0;
    for (size_t i = 0; i < n; ++i) {
        if (!alive[i]) continue;
        for (size_t k = 0; k < din; ++k) Win[a * din + k] = Wi[i * din + k];
        for (size_t o = 0; o < dout; ++o) Won[o * ns + a] = Wo[o * n + i];
        Vn[a] = V[i];
        ++a;
    }
    spires_reservoir_config c = r->cfg;
    c.num_neurons = (size_t)ns;
    spires_reservoir* q = nullptr;
    s = spires_ext_create_csr(&c, nnz, rp, ci, val, Win.data(), Won.data(), &q);
    std::free(rp);
    std::free(ci);
    std::free(val);
    if (s != SPIRES_OK) return s == SPIRES_ERR_ALLOC ? SPIRES_ERR_ALLOC : SPIRES_ERR_INTERNAL;
    if (spires_dev_set_state(q->dev, Vn.data())) { spires_reservoir_destroy(q); return SPIRES_ERR_INTERNAL; }
    *out_r = q;
    return SPIRES_OK;
}

double* spires_copy_reservoir_state(spires_reservoir* r) {
    if (!r || !r->dev) return nullptr;
    double* v = (double*)std::malloc(r->cfg.num_neurons * sizeof(double));
    if (!v) return nullptr;
    if (spires_dev_read_state(r->dev, v)) { std::free(v); return nullptr; }
    return v;
}

spires_status spires_read_reservoir_state(spires_reservoir* r, double* buffer) {
    if (!r || !r->dev || !buffer) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_read_state(r->dev, buffer));
}

spires_status spires_read_spike_state(spires_reservoir* r, double* buffer) {
    if (!r || !r->dev || !buffer) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_read_spikes(r->dev, buffer));
}

spires_status spires_read_weights(const spires_reservoir* r, double* buffer) {
    if (!r || !buffer) return SPIRES_ERR_INVALID_ARG;
    const size_t n = r->cfg.num_neurons;
    std::memset(buffer, 0, n * n * sizeof(double));
    for (size_t i = 0; i < n; ++i)
        for (int64_t k = r->W.rp[i]; k < r->W.rp[i + 1]; ++k) buffer[i * n + r->W.ci[k]] = r->W.val[k];
    return SPIRES_OK;
}

double* spires_copy_weights(const spires_reservoir* r) {
    if (!r) return nullptr;
    const size_t n = r->cfg.num_neurons;
    double* w = (double*)std::malloc(n * n * sizeof(double));
    if (!w) return nullptr;
    spires_read_weights(r, w);
    return w;
}

spires_status spires_compute_output(spires_reservoir* r, double* out) {
    if (!r || !r->dev || !out) return SPIRES_ERR_INVALID_ARG;
    return spires_dev_compute_output(r->dev, out) ? SPIRES_ERR_INTERNAL : SPIRES_OK;
}

spires_status spires_read_readout(const spires_reservoir* r, double* buffer) {
    if (!r || !r->dev || !buffer) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_get_readout(r->dev, buffer));
}

double* spires_copy_readout(const spires_reservoir* r) {
    if (!r || !r->dev) return nullptr;
    const size_t cnt = r->cfg.num_neurons * r->cfg.num_outputs;
    double* w = (double*)std::malloc((cnt ? cnt : 1) * sizeof(double));
    if (!w) return nullptr;
    if (cnt && spires_dev_get_readout(r->dev, w)) { std::free(w); return nullptr; }
    return w;
}

size_t spires_num_neurons(const spires_reservoir* r) { return r ? r->cfg.num_neurons : 0; }
size_t spires_num_inputs(const spires_reservoir* r) { return r ? r->cfg.num_inputs : 0; }
size_t spires_num_outputs(const spires_reservoir* r) { return r ? r->cfg.num_outputs : 0; }

// ---- extensions
spires_status spires_ext_run_record(spires_reservoir* r, const double* u, size_t T, double* states,
                                    uint8_t* raster, double* outputs) {
    if (!r || !r->dev || (!u && r->cfg.num_inputs && T)) return SPIRES_ERR_INVALID_ARG;
    if (T == 0) return SPIRES_OK;
    unsigned f = (states ? SPIRES_DEV_REC_STATE : 0u) | (raster ? SPIRES_DEV_REC_RASTER : 0u) |
                 (outputs ? SPIRES_DEV_REC_OUTPUT : 0u);
    int rc = spires_dev_upload_inputs(r->dev, u, T);
    if (!rc) rc = spires_dev_run(r->dev, 0, T, f);
    if (!rc) rc = spires_dev_read_records(r->dev, 0, T, states, raster, outputs);
    return dev_status(rc);
}

spires_status spires_ext_upload_inputs(spires_reservoir* r, const double* u, size_t T) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_upload_inputs(r->dev, u, T));
}

spires_status spires_ext_reserve_records(spires_reservoir* r, size_t rows, unsigned flags) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_reserve_records(r->dev, rows, flags));
}

spires_status spires_ext_run_resident(spires_reservoir* r, size_t t0, size_t T, unsigned flags) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_run(r->dev, t0, T, flags));
}

spires_status spires_ext_fit_ridge_records(spires_reservoir* r, size_t row0, size_t row1, const double* Y,
                                           double lambda, double* info) {
    if (!r || !r->dev || !Y) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_ridge_rows(r->dev, row0, row1, Y, lambda, info));
}

spires_status spires_ext_predict_records(spires_reservoir* r, size_t row0, size_t rows, double* out) {
    if (!r || !r->dev || (!out && rows)) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_predict_records(r->dev, row0, rows, out));
}

spires_status spires_ext_fit_ridge_records_batch(spires_reservoir* const* rs, int count, size_t row0, size_t row1,
                                                 const double* const* Y, const double* lambda, double* info) {
    if (!rs || count < 1 || !Y || !lambda) return SPIRES_ERR_INVALID_ARG;
    std::vector<spires_dev*> devs(count);
    for (int i = 0; i < count; ++i) {
        if (!rs[i] || !rs[i]->dev || !Y[i]) return SPIRES_ERR_INVALID_ARG;
        devs[i] = rs[i]->dev;
    }
    return dev_status(spires_dev_ridge_rows_batch(devs.data(), count, row0, row1, Y, lambda, info));
}

spires_status spires_ext_run_resident_batch(spires_reservoir* const* rs, int count, size_t t0, size_t T,
                                            unsigned flags) {
    if (!rs || count < 1) return SPIRES_ERR_INVALID_ARG;
    std::vector<spires_dev*> devs(count);
    for (int i = 0; i < count; ++i) {
        if (!rs[i] || !rs[i]->dev) return SPIRES_ERR_INVALID_ARG;
        devs[i] = rs[i]->dev;
    }
    return dev_status(spires_dev_run_batch(devs.data(), count, t0, T, flags));
}

struct spires_timeblock {
    spires_dev_timeblock* tb = nullptr;
};

spires_status spires_ext_timeblock_create(spires_reservoir* const* rs, int count, size_t K, unsigned flags,
                                          spires_timeblock** out) {
    if (!rs || count < 1 || K < 1 || !out) return SPIRES_ERR_INVALID_ARG;
    *out = nullptr;
    std::vector<spires_dev*> devs(count);
    for (int i = 0; i < count; ++i) {
        if (!rs[i] || !rs[i]->dev) return SPIRES_ERR_INVALID_ARG;
        devs[i] = rs[i]->dev;
    }
    spires_timeblock* t = new (std::nothrow) spires_timeblock();
    if (!t) return SPIRES_ERR_ALLOC;
    spires_status s = dev_status(spires_dev_timeblock_create(devs.data(), count, (long long)K, flags, &t->tb));
    if (s != SPIRES_OK) { delete t; return s; }
    *out = t;
    return SPIRES_OK;
}

spires_status spires_ext_timeblock_run(spires_timeblock* tb, size_t blocks) {
    if (!tb || !tb->tb) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_timeblock_run(tb->tb, (long long)blocks));
}

int spires_ext_timeblock_launches(const spires_timeblock* tb) { return tb ? spires_dev_timeblock_launches(tb->tb) : 0; }

void spires_ext_timeblock_destroy(spires_timeblock* tb) {
    if (!tb) return;
    spires_dev_timeblock_destroy(tb->tb);
    delete tb;
}

spires_status spires_ext_kernel_plan(spires_reservoir* r, int* info) {
    if (!r || !r->dev || !info) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_plan_info(r->dev, info));
}

int spires_ext_batch_workgroups(spires_reservoir* const* rs, int count) {
    if (!rs || count < 1) return 0;
    std::vector<spires_dev*> devs(count);
    for (int i = 0; i < count; ++i) {
        if (!rs[i] || !rs[i]->dev) return 0;
        devs[i] = rs[i]->dev;
    }
    return spires_dev_batch_workgroups(devs.data(), count);
}

spires_status spires_ext_read_records(spires_reservoir* r, size_t row0, size_t rows, double* states,
                                      uint8_t* raster, double* outputs) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_read_records(r->dev, row0, rows, states, raster, outputs));
}

spires_status spires_ext_set_mode(spires_reservoir* r, int mode, int block_k) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    int rc = spires_dev_set_mode(r->dev, mode, block_k);
    if (rc == 0) { r->mode = mode; r->block_k = block_k; }
    return dev_status(rc);
}

spires_status spires_ext_set_kernel_path(spires_reservoir* r, int path) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_set_path(r->dev, path));
}

spires_status spires_ext_set_state(spires_reservoir* r, const double* V) {
    if (!r || !r->dev || !V) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_set_state(r->dev, V));
}

spires_status spires_ext_sync(spires_reservoir* r) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_sync(r->dev));
}

void* spires_ext_stream(spires_reservoir* r) { return (r && r->dev) ? spires_dev_stream(r->dev) : nullptr; }

size_t spires_ext_nnz(const spires_reservoir* r) { return r ? r->W.val.size() : 0; }

spires_status spires_ext_get_csr(const spires_reservoir* r, int64_t* rp, int32_t* ci, double* val) {
    if (!r || !rp) return SPIRES_ERR_INVALID_ARG;
    std::memcpy(rp, r->W.rp.data(), r->W.rp.size() * sizeof(int64_t));
    if (!r->W.val.empty()) {
        if (!ci || !val) return SPIRES_ERR_INVALID_ARG;
        std::memcpy(ci, r->W.ci.data(), r->W.ci.size() * sizeof(int32_t));
        std::memcpy(val, r->W.val.data(), r->W.val.size() * sizeof(double));
    }
    return SPIRES_OK;
}

spires_status spires_ext_get_input_weights(const spires_reservoir* r, double* W_in) {
    if (!r || !W_in) return SPIRES_ERR_INVALID_ARG;
    std::memcpy(W_in, r->W_in.data(), r->W_in.size() * sizeof(double));
    return SPIRES_OK;
}

spires_status spires_ext_inject_exchange_timeout(spires_reservoir* r) {
    if (!r || !r->dev) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_inject_exchange_timeout(r->dev));
}

spires_status spires_ext_exchange_floor(int workgroups, size_t steps, double* us_per_step) {
    if (workgroups < 1 || steps < 2 || !us_per_step) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_exchange_floor(workgroups, (long long)steps, us_per_step));
}

const char* spires_ext_last_error(void) { return last_error().c_str(); }

void spires_ext_shutdown(void) { spires_dev_shutdown(); }

double spires_ext_spectral_radius(size_t n, const int64_t* row_ptr, const int32_t* col_idx, const double* values) {
    if (!row_ptr || n == 0) return 0.0;
    Csr W;
    W.n = n;
    W.rp.assign(row_ptr, row_ptr + n + 1);
    W.ci.assign(col_idx, col_idx + row_ptr[n]);
    W.val.assign(values, values + row_ptr[n]);
    return csr_spectral_radius(W);
}

spires_status spires_ext_random_er(size_t n, double p, double ei, uint64_t seed, double rho, size_t* nnz,
                                   int64_t** row_ptr, int32_t** col_idx, double** values, double* lambda) {
    if (n == 0 || n > (size_t)INT32_MAX || !nnz || !row_ptr || !col_idx || !values) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_random_er((int)n, p, ei, (unsigned long long)seed, rho, row_ptr, col_idx, values,
                                           nnz, lambda));
}

// init_weights' topology branches (reservoir.c:384-536) and rescale_weights
// (reservoir.c:548-557) on the device, for any connectivity type.
spires_status spires_ext_build_topology(spires_connectivity_type type, size_t n, double connectivity, double ei,
                                        uint64_t seed, double rho, size_t* nnz, int64_t** row_ptr, int32_t** col_idx,
                                        double** values, double* lambda) {
    if (n == 0 || n > (size_t)INT32_MAX || !nnz || !row_ptr || !col_idx || !values) return SPIRES_ERR_INVALID_ARG;
    if (type != SPIRES_CONN_RANDOM && type != SPIRES_CONN_SMALL_WORLD && type != SPIRES_CONN_SCALE_FREE)
        return SPIRES_ERR_INVALID_ARG;
    const int t = type == SPIRES_CONN_RANDOM ? 0 : type == SPIRES_CONN_SMALL_WORLD ? 1 : 2;
    return dev_status(spires_dev_build_topology(t, (int)n, connectivity, ei, (unsigned long long)seed, rho, row_ptr,
                                                col_idx, values, nnz, lambda));
}

// init_reservoir (reservoir.c:68-78) with the recurrent matrix built and
// rescaled on the device.  W_in and W_out keep generate_weight's distribution
// (magnitude U(0,1], sign by ei_ratio; 0.5 for W_out, reservoir.c:559-571),
// drawn from SplitMix64 streams of the same seed.
spires_status spires_ext_create_device(const spires_reservoir_config* cfg, uint64_t seed, spires_reservoir** out_r) {
    if (!cfg || !out_r) return SPIRES_ERR_INVALID_ARG;
    spires_status s = validate_cfg(cfg);
    if (s != SPIRES_OK) return s;
    const size_t n = cfg->num_neurons;
    size_t nnz = 0;
    int64_t* rp = nullptr;
    int32_t* ci = nullptr;
    double* val = nullptr;
    s = spires_ext_build_topology(cfg->connectivity_type, n, cfg->connectivity, cfg->ei_ratio, seed,
                                  cfg->spectral_radius, &nnz, &rp, &ci, &val, nullptr);
    if (s != SPIRES_OK) return s;
    std::vector<double> W_in(n * cfg->num_inputs), W_out(n * cfg->num_outputs);
    uint64_t z = seed ^ 0xD1B54A32D192ED03ULL;
    auto unit = [&]() {   // SplitMix64, 53 bits in [0, 1)
        z += 0x9E3779B97F4A7C15ULL;
        uint64_t h = z;
        h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ULL;
        h = (h ^ (h >> 27)) * 0x94D049BB133111EBULL;
        return (double)((h ^ (h >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto weight = [&](double ei) {
        const double mag = 1.0 - unit();   // (0, 1]
        return unit() < ei ? mag : -mag;
    };
    for (double& w : W_in) w = weight(cfg->ei_ratio);
    for (double& w : W_out) w = weight(0.5);
    s = spires_ext_create_csr(cfg, nnz, rp, ci, val, W_in.data(), W_out.data(), out_r);
    std::free(rp);
    std::free(ci);
    std::free(val);
    return s;
}

spires_status spires_ext_spectral_radius_device(size_t n, const int64_t* row_ptr, const int32_t* col_idx,
                                                const double* values, double* lambda, int* iters) {
    if (n == 0 || n > (size_t)INT32_MAX || !row_ptr || !lambda) return SPIRES_ERR_INVALID_ARG;
    return dev_status(spires_dev_spectral_radius((int)n, row_ptr, col_idx, values, lambda, iters));
}

spires_status spires_ext_spectral_radius_device_noalloc(size_t n, const int64_t* d_row_ptr,
                                                        const int32_t* d_col_idx, const double* d_values,
                                                        double* d_lambda, int* d_iters, void* stream) {
    if (n == 0 || !d_row_ptr || !d_col_idx || !d_values || !d_lambda || !d_iters) return SPIRES_ERR_INVALID_ARG;
    if (n > (size_t)INT32_MAX) { set_error("spectral_radius_noalloc: n too large"); return SPIRES_ERR_INVALID_ARG; }
    return dev_status(spires_dev_spectral_radius_noalloc((int)n, d_row_ptr, d_col_idx, d_values, d_lambda, d_iters,
                                                         stream));
}

spires_status spires_ext_build_weights(const spires_reservoir_config* cfg, size_t* nnz_out, int64_t** row_ptr,
                                       int32_t** col_idx, double** values, double** W_in, double** W_out) {
    if (!cfg || !nnz_out || !row_ptr || !col_idx || !values || !W_in || !W_out) return SPIRES_ERR_INVALID_ARG;
    spires_status s = validate_cfg(cfg);
    if (s != SPIRES_OK) return s;
    spires_reservoir r;
    r.cfg = *cfg;
    const size_t n = cfg->num_neurons;
    if (generate_topology(n, cfg->num_inputs, cfg->connectivity_type, cfg->connectivity, cfg->ei_ratio, r.W_in,
                          r.W) != 0) {
        set_error("unknown connectivity type");
        return SPIRES_ERR_INVALID_ARG;
    }
    synapse_build(&r);
    rescale(&r);
    const size_t nnz = r.W.val.size(), nout = n * cfg->num_outputs;
    *row_ptr = (int64_t*)std::malloc((n + 1) * sizeof(int64_t));
    *col_idx = (int32_t*)std::malloc((nnz ? nnz : 1) * sizeof(int32_t));
    *values = (double*)std::malloc((nnz ? nnz : 1) * sizeof(double));
    *W_in = (double*)std::malloc((r.W_in.size() ? r.W_in.size() : 1) * sizeof(double));
    *W_out = (double*)std::malloc((nout ? nout : 1) * sizeof(double));
    if (!*row_ptr || !*col_idx || !*values || !*W_in || !*W_out) {
        std::free(*row_ptr); std::free(*col_idx); std::free(*values); std::free(*W_in); std::free(*W_out);
        return SPIRES_ERR_ALLOC;
    }
    std::memcpy(*row_ptr, r.W.rp.data(), (n + 1) * sizeof(int64_t));
    if (nnz) {
        std::memcpy(*col_idx, r.W.ci.data(), nnz * sizeof(int32_t));
        std::memcpy(*values, r.W.val.data(), nnz * sizeof(double));
    }
    if (!r.W_in.empty()) std::memcpy(*W_in, r.W_in.data(), r.W_in.size() * sizeof(double));
    for (size_t k = 0; k < nout; ++k) (*W_out)[k] = gen_weight(0.5);   // reservoir.c:559-571
    *nnz_out = nnz;
    return SPIRES_OK;
}

}  // extern "C"
