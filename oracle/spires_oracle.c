/*
 * spires_oracle.c -- CPU restatement of the spires reservoir path (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle + CPU baseline).  Never linked into
 * the product library.  Build: oracle/Makefile (-O2 -ffp-contract=off -fopenmp,
 * no -march), so every double operation is a single IEEE operation in the
 * order written, which is the reference's order.
 *
 * Reference files restated (txmastin/spires):
 *   construction  src/reservoir.c:18-78, 307-571
 *   stepping      src/reservoir.c:113-242
 *   readout       src/reservoir.c:244-280, 573-767; src/math_utils.c:138-241
 *   reset/coarse  src/reservoir.c:770-967
 *   neurons       src/neuron.c:14-62, src/neurons/{flif_gl,lif_discrete,lif_bio,flif_caputo,flif_diffusive}.c
 *   synapses      src/synapse.c, src/synapses/{simple,psc_*,fractional*}.c, src/sparse.c:7-127
 *   optimizer     src/agile.c:8-351
 */
#include "spires_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX_MEM 20000   /* MAX_MEM_LEN, src/neurons/flif_gl.h:4 */

struct or_res {
    size_t N, Din, Dout;
    double rho, ei, gain, conn, dt;
    int ctype, ntype, stype, sback;
    double *np, *sp;             /* retained caller pointers (reservoir.c:42,45) */
    int micro;                   /* llround(1/dt), reservoir.c:127 */
    double *W_in, *W_out;
    /* recurrent weights: CSR (sparse backend) or dense row-major */
    size_t nnz; int64_t *rp; int32_t *ci; double *val;
    double *Wd;
    /* synapse state */
    double psc_tau; double *psc_tr;
    double *het_tau, *het_tr;
    double fr_alpha, fr_tau, fr_Tmem; int fr_L; double *fr_c, *fr_I, *fr_hist; long fr_step;
    int me_N; double *me_tau, *me_c, *me_tr, *me_comb;
    /* neuron state (structure of arrays) */
    double *V, *S;
    /* FLIF-GL */
    int gl_L, gl_ncoef; double *gl_c, *gl_ring, *gl_tprev; long gl_step;
    double gl_Vth, gl_Vreset, gl_Vrest, gl_tau, gl_alpha, gl_bias, gl_tref, gl_pdt;
    /* FLIF-Caputo */
    int cp_L; double *cp_c, *cp_dm, *cp_tprev; int *cp_head;
    double cp_Cm, cp_gl, cp_Vl, cp_Vth, cp_V0, cp_alpha, cp_tref, cp_kr;
    /* LIF */
    double lif_V0, lif_Vth, lif_leak, lif_tau, lif_bias;
    /* FLIF-diffusive */
    long *df_until; uint64_t *df_rng; long df_step;
    double df_Vth, df_Vreset, df_Vrest, df_tau, df_alpha, df_bias, df_g, df_tmin, df_tmax, df_tref;
    /* scratch */
    double *ext, *last, *eff;
};

/* ------------------------------------------------------------------ rand */
/* urand01 / generate_weight: reservoir.c:308-317, 330-333 */
static double u01r(void) { return (double)rand() / (double)RAND_MAX; }
static double gen_weight(double ei)
{
    double mag = (double)rand() / RAND_MAX;
    if (((double)rand() / RAND_MAX) < ei) return mag;
    return -mag;
}

/* GL derivative coefficients, neuron.c:14-21 (note the (float)k). */
void or_gl_coeffs(double alpha, int n, double *c)
{
    if (n <= 0) return;
    c[0] = 1.0;
    for (int k = 1; k < n; k++)
        c[k] = c[k - 1] * (1.0 - (alpha + 1.0) / (float)k);
}

/* ------------------------------------------------------------- topology */
/* init_weights, reservoir.c:345-545, on a dense n x n buffer. */
int or_init_weights_dense(size_t n, size_t din, int ctype, double conn,
                          double ei, double *W, double *W_in)
{
    memset(W, 0, n * n * sizeof(double));
    for (size_t k = 0; k < n * din; k++) W_in[k] = gen_weight(ei);
    if (ctype == SPIRES_CONN_RANDOM) {
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < n; j++) {
                if (i == j) continue;
                if (u01r() < conn) W[i * n + j] = gen_weight(ei);
            }
    } else if (ctype == SPIRES_CONN_SMALL_WORLD) {
        if (n < 3) return 0;
        int K = (int)(conn * (double)(n - 1) + 0.5);
        if (K < 2) K = 2;
        if (K % 2) K++;
        if (K >= (int)n) K = (int)n - 1;
        const double p = 0.1;
        int half = K / 2;
        for (size_t i = 0; i < n; i++)
            for (int s = 1; s <= half; s++) {
                size_t j = (i + (size_t)s) % n;
                if (i == j) continue;
                W[i * n + j] = gen_weight(ei);
            }
        for (size_t i = 0; i < n; i++)
            for (int s = 1; s <= half; s++) {
                size_t jo = (i + (size_t)s) % n;
                if (u01r() < p) {
                    W[i * n + jo] = 0.0;
                    size_t jn; int tries = 0;
                    for (;;) {
                        jn = (size_t)(u01r() * (double)n);
                        if (jn >= n) jn = n - 1;
                        if (++tries > 10 * (int)n) break;
                        if (!(jn == i || W[i * n + jn] != 0.0)) break;
                    }
                    if (jn != i) W[i * n + jn] = gen_weight(ei);
                }
            }
    } else if (ctype == SPIRES_CONN_SCALE_FREE) {
        if (n < 2) return 0;
        int m = (int)(conn * (double)(n - 1) + 0.5);
        if (m < 1) m = 1;
        if (m >= (int)n) m = (int)n - 1;
        int m0 = m > 2 ? m : 2;
        if (m0 >= (int)n) m0 = (int)n - 1;
        const double A = 1.0;
        unsigned char *adj = calloc(n * n, 1);
        int *deg = calloc(n, sizeof(int));
        if (!adj || !deg) { free(adj); free(deg); return -1; }
        for (int a = 0; a < m0; a++)
            for (int b = a + 1; b < m0; b++) {
                adj[(size_t)a * n + b] = adj[(size_t)b * n + a] = 1;
                deg[a]++; deg[b]++;
            }
        for (size_t v = (size_t)m0; v < n; v++) {
            int added = 0, guard = 0;
            while (added < m && guard < 50 * m) {
                double total = 0.0;
                for (size_t u = 0; u < v; u++) total += (double)deg[u] + A;
                if (total <= 0.0) total = (double)v;
                double r = u01r() * total, acc = 0.0;
                size_t pick = 0;
                for (size_t u = 0; u < v; u++) {
                    acc += (double)deg[u] + A;
                    if (r <= acc) { pick = u; break; }
                }
                if (pick != v && !adj[v * n + pick]) {
                    adj[v * n + pick] = adj[pick * n + v] = 1;
                    deg[v]++; deg[pick]++; added++;
                }
                guard++;
            }
            while (added < m) {
                size_t pick = (size_t)(u01r() * (double)v);
                if (pick >= v) pick = v - 1;
                if (pick != v && !adj[v * n + pick]) {
                    adj[v * n + pick] = adj[pick * n + v] = 1;
                    deg[v]++; deg[pick]++; added++;
                }
            }
        }
        for (size_t i = 0; i < n; i++)
            for (size_t j = i + 1; j < n; j++) {
                if (!adj[i * n + j]) continue;
                if (u01r() < 0.5) W[i * n + j] = gen_weight(ei);
                else              W[j * n + i] = gen_weight(ei);
            }
        free(adj); free(deg);
    }
    return 0;
}

/* ------------------------------------------------------------ CSR utils */
/* csr_build_from_dense, sparse.c:7-42 (row-major scan, ascending columns). */
static void csr_from_dense(or_res *r, const double *W)
{
    size_t n = r->N, nnz = 0;
    for (size_t k = 0; k < n * n; k++) if (W[k] != 0.0) nnz++;
    r->nnz = nnz;
    r->rp = malloc((n + 1) * sizeof(int64_t));
    r->ci = malloc((nnz ? nnz : 1) * sizeof(int32_t));
    r->val = malloc((nnz ? nnz : 1) * sizeof(double));
    size_t idx = 0;
    for (size_t i = 0; i < n; i++) {
        r->rp[i] = (int64_t)idx;
        for (size_t j = 0; j < n; j++)
            if (W[i * n + j] != 0.0) { r->ci[idx] = (int32_t)j; r->val[idx] = W[i * n + j]; idx++; }
    }
    r->rp[n] = (int64_t)idx;
}

/* csr_row_dot, sparse.c:68-76: sequential sum from +0.0 in CSR order. */
static double csr_dot(const int64_t *rp, const int32_t *ci, const double *val,
                      size_t row, const double *x)
{
    double s = 0.0;
    for (int64_t k = rp[row]; k < rp[row + 1]; k++) s += val[k] * x[ci[k]];
    return s;
}

static double dense_dot(const double *row, const double *x, size_t n)
{
    double s = 0.0;
    for (size_t j = 0; j < n; j++) s += row[j] * x[j];
    return s;
}

/* Power iteration, sparse.c:93-127 (also psc_heterogeneous.c:256-288 and,
 * in sequential order, math_utils.c:30-71).  mv(i) gives row i of W*x. */
typedef double (*row_fn)(const or_res *r, size_t i, const double *x);
static double power_iter(const or_res *r, size_t n, row_fn mv)
{
    double *x = malloc(n * sizeof(double)), *y = malloc(n * sizeof(double));
    for (size_t i = 0; i < n; i++) x[i] = 1.0;
    double lo = 0.0, ln = 0.0;
    for (unsigned it = 0; it < 1000; it++) {
        for (size_t i = 0; i < n; i++) y[i] = mv(r, i, x);
        ln = 0.0;
        for (size_t i = 0; i < n; i++) { double a = fabs(y[i]); if (a > ln) ln = a; }
        for (size_t i = 0; i < n; i++) x[i] = y[i] / ln;
        if (fabs(ln - lo) < 1e-6) break;
        lo = ln;
    }
    free(x); free(y);
    return ln;
}
static double mv_csr(const or_res *r, size_t i, const double *x) { return csr_dot(r->rp, r->ci, r->val, i, x); }
static double mv_dense(const or_res *r, size_t i, const double *x) { return dense_dot(&r->Wd[i * r->N], x, r->N); }

double or_csr_spectral_radius(size_t n, const int64_t *rp, const int32_t *ci, const double *val)
{
    or_res tmp; memset(&tmp, 0, sizeof(tmp));
    tmp.N = n; tmp.rp = (int64_t *)rp; tmp.ci = (int32_t *)ci; tmp.val = (double *)val;
    return power_iter(&tmp, n, mv_csr);
}

/* ------------------------------------------------------------ synapses */
/* synapse_build_from_dense, synapse.c:3-42 plus the per-type builders. */
static void synapse_build(or_res *r, const double *W)
{
    size_t n = r->N;
    const double *sp = r->sp;
    if (r->sback == SPIRES_SYNAPSE_SPARSE) csr_from_dense(r, W);
    else { r->Wd = malloc(n * n * sizeof(double)); memcpy(r->Wd, W, n * n * sizeof(double)); }
    switch (r->stype) {
    case SPIRES_SYNAPSE_PSC_HOMOGENEOUS:               /* psc_homogeneous.c:12-28 */
        r->psc_tau = sp ? sp[0] : 1.0;
        r->psc_tr = calloc(n, sizeof(double));
        break;
    case SPIRES_SYNAPSE_PSC_HETEROGENEOUS: {           /* psc_heterogeneous.c:33-86 */
        double tmin = sp ? sp[0] : 1.0, tmax = sp ? sp[1] : 1.0;
        size_t m = (r->sback == SPIRES_SYNAPSE_SPARSE) ? r->nnz : n * n;
        r->het_tau = malloc((m ? m : 1) * sizeof(double));
        r->het_tr = calloc(m ? m : 1, sizeof(double));
        double lmin = log(tmin), lmax = log(tmax);
        for (size_t k = 0; k < m; k++) r->het_tau[k] = exp(lmin + u01r() * (lmax - lmin));
        break;
    }
    case SPIRES_SYNAPSE_FRACTIONAL:                    /* fractional.c:24-45 */
        r->fr_alpha = sp ? sp[0] : 1.0;
        r->fr_tau = sp ? sp[1] : 1.0;
        r->fr_Tmem = sp ? sp[2] : 10.0;
        r->fr_L = 0; r->fr_step = 0;
        r->fr_I = calloc(n, sizeof(double));
        break;
    case SPIRES_SYNAPSE_FRACTIONAL_MULTIEXP: {         /* fractional_multiexp.c:18-90 */
        const double PI = 3.14159265358979323846;
        double alpha = sp ? sp[0] : 0.5, tsyn = sp ? sp[1] : 1.0;
        double tmin = sp ? sp[2] : 1.0, tmax = sp ? sp[3] : 20.0;
        int Nc = sp ? (int)(sp[4] + 0.5) : 8;
        if (Nc < 1) Nc = 1;
        r->me_N = Nc;
        r->me_tau = malloc(Nc * sizeof(double)); r->me_c = malloc(Nc * sizeof(double));
        double sa = sin(alpha * PI);
        if (fabs(sa) < 1e-9) {
            r->me_N = 1; r->me_tau[0] = tsyn; r->me_c[0] = 1.0;
        } else {
            double l0 = log(tmin), l1 = log(tmax);
            double dl = (Nc > 1) ? (l1 - l0) / (Nc - 1) : 0.0;
            double ca = cos(alpha * PI), sum = 0.0;
            double *raw = malloc(Nc * sizeof(double));
            for (int i = 0; i < Nc; i++) {
                double ti = exp(l0 + i * dl);
                double x = tsyn / ti;
                double K = (1.0 / PI) * pow(x, alpha - 1.0) * sa /
                           (pow(x, 2.0 * alpha) + 2.0 * pow(x, alpha) * ca + 1.0);
                double w = K * x * dl;
                if (w < 0.0) w = 0.0;
                raw[i] = w * ti; r->me_tau[i] = ti; sum += raw[i];
            }
            for (int i = 0; i < Nc; i++) r->me_c[i] = (sum > 0.0) ? raw[i] / sum : 1.0 / Nc;
            free(raw);
        }
        r->me_tr = calloc((size_t)r->me_N * n, sizeof(double));
        r->me_comb = calloc(n, sizeof(double));
        break;
    }
    default: break;
    }
}

/* rescale_weights, reservoir.c:548-557 with synapse_spectral_radius /
 * synapse_scale dispatch (synapse.c:107-143). */
static void rescale(or_res *r)
{
    size_t n = r->N;
    double lam = power_iter(r, n, r->sback == SPIRES_SYNAPSE_SPARSE ? mv_csr : mv_dense);
    if (lam > 1e-9) {
        double f = r->rho / lam;
        if (r->sback == SPIRES_SYNAPSE_SPARSE) for (size_t k = 0; k < r->nnz; k++) r->val[k] *= f;
        else for (size_t k = 0; k < n * n; k++) r->Wd[k] *= f;
    }
}

/* synapse_prepare + row dot, synapse.c:90-105,145-160 and the per-type files. */
static void synapse_prepare(or_res *r, const double *spk)
{
    size_t n = r->N;
    switch (r->stype) {
    case SPIRES_SYNAPSE_SIMPLE:
        memcpy(r->eff, spk, n * sizeof(double));
        break;
    case SPIRES_SYNAPSE_PSC_HOMOGENEOUS: {             /* psc_homogeneous.c:71-84 */
        double d = exp(-r->dt / r->psc_tau);
        for (size_t i = 0; i < n; i++) r->psc_tr[i] = r->psc_tr[i] * d + (1.0 - d) * spk[i];
        memcpy(r->eff, r->psc_tr, n * sizeof(double));
        break;
    }
    case SPIRES_SYNAPSE_PSC_HETEROGENEOUS:             /* psc_heterogeneous.c:206-229 */
        if (r->sback == SPIRES_SYNAPSE_SPARSE) {
            #pragma omp parallel for if (n >= 1024)
            for (size_t i = 0; i < n; i++)
                for (int64_t k = r->rp[i]; k < r->rp[i + 1]; k++) {
                    double d = exp(-r->dt / r->het_tau[k]);
                    r->het_tr[k] = r->het_tr[k] * d + (1.0 - d) * spk[r->ci[k]];
                }
        } else {
            #pragma omp parallel for if (n >= 1024)
            for (size_t i = 0; i < n; i++)
                for (size_t j = 0; j < n; j++) {
                    size_t k = i * n + j;
                    double d = exp(-r->dt / r->het_tau[k]);
                    r->het_tr[k] = r->het_tr[k] * d + (1.0 - d) * spk[j];
                }
        }
        break;
    case SPIRES_SYNAPSE_FRACTIONAL: {                  /* fractional.c:103-142 */
        if (r->fr_L == 0) {
            int L = (r->fr_Tmem > 0 && r->dt > 0) ? (int)(r->fr_Tmem / r->dt) : 2000;
            if (L > OR_MAX_MEM) L = OR_MAX_MEM;
            if (L < 1) L = 1;
            r->fr_L = L;
            r->fr_c = malloc((size_t)L * sizeof(double));
            r->fr_c[0] = 1.0;
            for (int k = 1; k < L; k++)
                r->fr_c[k] = r->fr_c[k - 1] * (1.0 - (r->fr_alpha + 1.0) / (double)k);
            r->fr_hist = calloc(n * (size_t)L, sizeof(double));
        }
        int L = r->fr_L;
        double da = pow(r->dt, r->fr_alpha);
        long st = r->fr_step;
        int head = (int)(st % L);
        int lim = (st < L) ? (int)st : L - 1;
        #pragma omp parallel for if (n >= 256)
        for (size_t i = 0; i < n; i++) {
            double *h = &r->fr_hist[i * (size_t)L];
            double Ip = h[(head - 1 + L) % L];
            double H = 0.0;
            for (int k = 1; k <= lim; k++) H += r->fr_c[k] * h[(head - k + L) % L];
            double rhs = (-Ip + spk[i]) / r->fr_tau;
            double In = da * rhs - H;
            h[head] = In;
            r->fr_I[i] = In;
        }
        r->fr_step++;
        memcpy(r->eff, r->fr_I, n * sizeof(double));
        break;
    }
    case SPIRES_SYNAPSE_FRACTIONAL_MULTIEXP: {         /* fractional_multiexp.c:136-157 */
        for (int c = 0; c < r->me_N; c++) {
            double d = exp(-r->dt / r->me_tau[c]);
            double *t = &r->me_tr[(size_t)c * n];
            for (size_t i = 0; i < n; i++) t[i] = t[i] * d + (1.0 - d) * spk[i];
        }
        for (size_t i = 0; i < n; i++) {
            double s = 0.0;
            for (int c = 0; c < r->me_N; c++) s += r->me_c[c] * r->me_tr[(size_t)c * n + i];
            r->me_comb[i] = s;
        }
        memcpy(r->eff, r->me_comb, n * sizeof(double));
        break;
    }
    }
}

static double synapse_row(const or_res *r, size_t i)
{
    size_t n = r->N;
    if (r->stype == SPIRES_SYNAPSE_PSC_HETEROGENEOUS) {   /* psc_heterogeneous.c:124-139 */
        double s = 0.0;
        if (r->sback == SPIRES_SYNAPSE_SPARSE) {
            for (int64_t k = r->rp[i]; k < r->rp[i + 1]; k++) s += r->val[k] * r->het_tr[k];
        } else {
            for (size_t j = 0; j < n; j++) s += r->Wd[i * n + j] * r->het_tr[i * n + j];
        }
        return s;
    }
    if (r->sback == SPIRES_SYNAPSE_SPARSE) return csr_dot(r->rp, r->ci, r->val, i, r->eff);
    return dense_dot(&r->Wd[i * n], r->eff, n);
}

/* --------------------------------------------------------------- neurons */
static uint64_t sm64(uint64_t *x)   /* SplitMix64, flif_diffusive.c:8-13 */
{
    uint64_t z = (*x += 0x9E3779B97f4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double sm_u01(uint64_t *s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Per-neuron seed of the diffusive model.  The reference seeds from the
 * neuron's heap address (flif_diffusive.c:118-126), which is not
 * reproducible; here the address is replaced by the neuron index so runs
 * are repeatable.  The state is then set exactly as the reference's own
 * flif_diffusive_set_seed does (flif_diffusive.c:131-138): zero maps to a
 * fixed constant and the generator is advanced twice.  The same formula is
 * used by the HIP kernel's host set-up. */
uint64_t or_df_seed(size_t i, const double *p)
{
    uint64_t seed = 0x9E3779B97f4A7C15ULL ^ ((uint64_t)(i + 1) * 0xD1B54A32D192ED03ULL);
    for (int k = 0; k < 7; k++) {
        uint64_t b; memcpy(&b, &p[k], 8);
        seed ^= b + 0x9E3779B97f4A7C15ULL;
        seed = sm64(&seed);
    }
    return seed;
}
static uint64_t df_state_from_seed(uint64_t seed)
{
    if (seed == 0ULL) seed = 0xD1B54A32D192ED03ULL;
    uint64_t st = seed;
    (void)sm64(&st);
    (void)sm64(&st);
    return st;
}

/* init_neuron for every neuron (neuron.c:23-62 and the neurons/ init_ functions). */
static void neurons_init(or_res *r, int first)
{
    size_t n = r->N;
    const double *p = r->np;
    double dt = r->dt;
    switch (r->ntype) {
    case SPIRES_NEURON_LIF_DISCRETE:                   /* lif_discrete.c:5-20 */
        r->lif_V0 = p[0]; r->lif_Vth = p[1]; r->lif_leak = p[2]; r->lif_bias = p[3];
        for (size_t i = 0; i < n; i++) { r->V[i] = r->lif_V0; r->S[i] = 0.0; }
        break;
    case SPIRES_NEURON_LIF_BIO:                        /* lif_bio.c:4-14 */
        r->lif_V0 = p[0]; r->lif_Vth = p[1]; r->lif_tau = p[2]; r->lif_bias = p[3];
        for (size_t i = 0; i < n; i++) { r->V[i] = r->lif_V0; r->S[i] = 0.0; }
        break;
    case SPIRES_NEURON_FLIF_GL: {                      /* flif_gl.c:7-48, neuron.c:36-53 */
        if (first) {
            int ml = (p[5] > 0 && dt > 0) ? (int)(p[5] / dt) : 2000;
            r->gl_ncoef = ml;
            r->gl_c = malloc((size_t)(ml > 0 ? ml : 1) * sizeof(double));
            or_gl_coeffs(p[4], ml, r->gl_c);
        }
        r->gl_Vth = p[0]; r->gl_Vreset = p[1]; r->gl_Vrest = p[2]; r->gl_tau = p[3];
        r->gl_alpha = p[4]; r->gl_bias = p[6]; r->gl_tref = p[7];
        int L = (p[5] > 0 && dt > 0) ? (int)(p[5] / dt) : 2000;
        if (L > OR_MAX_MEM) L = OR_MAX_MEM;
        if (first || L != r->gl_L) {
            free(r->gl_ring);
            r->gl_ring = malloc(n * (size_t)(L > 0 ? L : 1) * sizeof(double));
        }
        r->gl_L = L;
        r->gl_pdt = pow(dt, r->gl_alpha);
        r->gl_step = 0;
        for (size_t i = 0; i < n; i++) {
            r->V[i] = r->gl_Vrest; r->S[i] = 0.0; r->gl_tprev[i] = r->gl_tref + dt;
            for (int k = 0; k < L; k++) r->gl_ring[i * (size_t)L + k] = r->gl_Vrest;
        }
        break;
    }
    case SPIRES_NEURON_FLIF_CAPUTO: {                  /* flif_caputo.c:6-41 */
        r->cp_Cm = p[0]; r->cp_gl = p[1]; r->cp_Vl = p[2]; r->cp_Vth = p[3];
        r->cp_V0 = p[4]; r->cp_alpha = p[6]; r->cp_tref = p[7];
        double Tm = p[8];
        int L = (Tm > 0) ? (int)(Tm / dt) : OR_MAX_MEM;
        if (L > OR_MAX_MEM) L = OR_MAX_MEM;
        int m = L - 2 > 0 ? L - 2 : 0;
        if (first || L != r->cp_L) {
            free(r->cp_c); free(r->cp_dm);
            r->cp_c = malloc((size_t)(m ? m : 1) * sizeof(double));
            r->cp_dm = malloc(n * (size_t)(m ? m : 1) * sizeof(double));
        }
        r->cp_L = L;
        for (int i = 0; i < m; i++) {
            double x = i + 2;
            r->cp_c[i] = pow(x, 1.0 - r->cp_alpha) - pow(x - 1.0, 1.0 - r->cp_alpha);
        }
        r->cp_kr = pow(dt, r->cp_alpha) * tgamma(2.0 - r->cp_alpha);
        for (size_t i = 0; i < n; i++) {
            r->V[i] = r->cp_V0; r->S[i] = 0.0; r->cp_tprev[i] = 2 * r->cp_tref; r->cp_head[i] = 0;
            for (int k = 0; k < m; k++) r->cp_dm[i * (size_t)(m ? m : 1) + k] = 0.0;
        }
        break;
    }
    case SPIRES_NEURON_FLIF_DIFFUSIVE: {               /* flif_diffusive.c:68-129 */
        r->df_Vth = p[0]; r->df_Vreset = p[1]; r->df_Vrest = p[2]; r->df_tau = p[3];
        r->df_alpha = p[4]; double P6 = p[5]; r->df_bias = p[6];
        if (r->df_tau <= 0.0) r->df_tau = 1.0;
        if (r->df_alpha < 0.0) r->df_alpha = 0.0;
        if (r->df_alpha > 1.0) r->df_alpha = 1.0;
        r->df_g = dt / r->df_tau;
        if (r->df_alpha < 1.0) {
            r->df_tmin = 1.0 * dt;
            r->df_tmax = (P6 > 0.0) ? P6 : 100.0 * r->df_tau;
            if (r->df_tmax <= r->df_tmin) r->df_tmax = r->df_tmin * 2.0;
            r->df_tref = dt;
        } else {
            r->df_tref = (P6 > 0.0) ? P6 : 0.0;
            r->df_tmin = 1.0 * dt;
            r->df_tmax = r->df_tref;
        }
        r->df_step = 0;
        for (size_t i = 0; i < n; i++) {
            r->V[i] = r->df_Vrest; r->S[i] = 0.0; r->df_until[i] = -1; r->df_rng[i] = df_state_from_seed(or_df_seed(i, p));
        }
        break;
    }
    }
}

/* Truncated power-law refractory draw, flif_diffusive.c:21-63. */
static long df_refractory_steps(const or_res *r, uint64_t *rng)
{
    double dt = r->dt;
    if (r->df_alpha < 1.0) {
        double a = r->df_alpha;
        double m = pow(a, 2.0), q = 6.0 * a, W;
        double u0 = sm_u01(rng);
        if (u0 < m) W = r->df_tref;
        else {
            double u1 = sm_u01(rng);
            double tmin = r->df_tmin, tmax = r->df_tmax;
            if (tmax <= tmin) W = tmin;
            else if (q <= 1e-12) W = tmin + u1 * (tmax - tmin);
            else if (fabs(q - 1.0) < 1e-12) W = tmin * pow(tmax / tmin, u1);
            else {
                double e = 1.0 - q;
                double A = pow(tmin, e), B = pow(tmax, e);
                W = pow((1.0 - u1) * A + u1 * B, 1.0 / e);
            }
        }
        long s = (W <= 0.0) ? 0 : (long)ceil(W / dt);
        if (s < 0) s = 0;
        if (s < 1) s = 1;
        return s;
    }
    double W = r->df_tref;
    long s = (W <= 0.0) ? 0 : (long)ceil(W / dt);
    return s < 0 ? 0 : s;
}

/* update_* for neuron i with total input x (update_ functions under src/neurons/). */
static void neuron_update(or_res *r, size_t i, double x)
{
    double dt = r->dt;
    switch (r->ntype) {
    case SPIRES_NEURON_LIF_DISCRETE: {                 /* lif_discrete.c:22-34 */
        if (r->S[i] == 1.0) r->S[i] = 0.0;
        r->V[i] = (1.0 - r->lif_leak) * r->V[i] + x + r->lif_bias;
        if (r->V[i] > r->lif_Vth) { r->V[i] = r->lif_V0; r->S[i] = 1.0; }
        break;
    }
    case SPIRES_NEURON_LIF_BIO: {                      /* lif_bio.c:16-29 */
        if (r->S[i] == 1.0) r->S[i] = 0.0;
        double dV = (-r->V[i] + x + r->lif_bias) / r->lif_tau;
        r->V[i] += dV * dt;
        if (r->V[i] >= r->lif_Vth) { r->V[i] = r->lif_V0; r->S[i] = 1.0; }
        break;
    }
    case SPIRES_NEURON_FLIF_GL: {                      /* flif_gl.c:50-100 */
        int L = r->gl_L;
        double *h = &r->gl_ring[i * (size_t)L];
        long st = r->gl_step;
        if (r->S[i] == 1.0) r->S[i] = 0.0;
        int head = (int)(st % L);
        if (r->gl_tprev[i] > r->gl_tref) {
            double Vp = h[(head - 1 + L) % L];
            double H = 0.0;
            int lim = (st < L) ? (int)st : L - 1;
            for (int k = 1; k <= lim; k++) H += r->gl_c[k] * h[(head - k + L) % L];
            double rhs = (-(Vp - r->gl_Vrest) / r->gl_tau) + x + r->gl_bias;
            r->V[i] = r->gl_pdt * rhs - H;
            if (r->V[i] >= r->gl_Vth) { r->V[i] = r->gl_Vreset; r->S[i] = 1.0; r->gl_tprev[i] = 0.0; }
        } else {
            r->V[i] = r->gl_Vreset;
        }
        h[head] = r->V[i];
        r->gl_tprev[i] += dt;
        break;
    }
    case SPIRES_NEURON_FLIF_CAPUTO: {                  /* flif_caputo.c:43-71 */
        if (r->S[i] == 1.0) r->S[i] = 0.0;
        int m = r->cp_L - 2;
        if (r->cp_tprev[i] > r->cp_tref) {
            double dV = (-r->cp_gl * (r->V[i] - r->cp_Vl) + x) / r->cp_Cm;
            double mk = r->cp_kr * dV;
            double mem = 0.0;
            if (m > 0) {
                /* shift register stored as a ring: slot of delta[j] is (head - j) mod m */
                double *d = &r->cp_dm[i * (size_t)m];
                int hd = r->cp_head[i] + 1; if (hd == m) hd = 0;
                r->cp_head[i] = hd;
                d[hd] = r->V[i] - r->cp_V0;
                for (int j = 0; j < m; j++) { int s = hd - j; if (s < 0) s += m; mem += r->cp_c[j] * d[s]; }
            }
            r->V[i] += mk - mem;
            if (r->V[i] >= r->cp_Vth) { r->S[i] = 1.0; r->V[i] = r->cp_V0; r->cp_tprev[i] = 0.0; }
        } else {
            r->V[i] = r->cp_V0;
        }
        r->cp_tprev[i] += dt;
        break;
    }
    case SPIRES_NEURON_FLIF_DIFFUSIVE: {               /* flif_diffusive.c:140-175 */
        if (r->S[i] == 1.0) r->S[i] = 0.0;
        double Vp = r->V[i];
        double rhs = -(Vp - r->df_Vrest) + x + r->df_bias;
        r->V[i] = Vp + r->df_g * rhs;
        if (r->df_step >= r->df_until[i] && r->V[i] >= r->df_Vth) {
            r->V[i] = r->df_Vreset; r->S[i] = 1.0;
            r->df_until[i] = r->df_step + df_refractory_steps(r, &r->df_rng[i]);
        }
        break;
    }
    }
}

/* ----------------------------------------------------------- lifecycle */
static or_res *alloc_common(const spires_reservoir_config *cfg)
{
    if (!cfg || cfg->dt <= 0.0 || cfg->dt > 1.0) return NULL;
    or_res *r = calloc(1, sizeof(*r));
    r->N = cfg->num_neurons; r->Din = cfg->num_inputs; r->Dout = cfg->num_outputs;
    r->rho = cfg->spectral_radius; r->ei = cfg->ei_ratio; r->gain = cfg->input_strength;
    r->conn = cfg->connectivity; r->dt = cfg->dt;
    r->ctype = cfg->connectivity_type; r->ntype = cfg->neuron_type;
    r->stype = cfg->synapse_type; r->sback = cfg->synapse_backend;
    r->np = cfg->neuron_params; r->sp = cfg->synapse_params;
    r->micro = (int)llround(1.0 / r->dt);
    size_t n = r->N;
    r->V = calloc(n, sizeof(double)); r->S = calloc(n, sizeof(double));
    r->gl_tprev = calloc(n, sizeof(double)); r->cp_tprev = calloc(n, sizeof(double));
    r->cp_head = calloc(n, sizeof(int));
    r->df_until = calloc(n, sizeof(long)); r->df_rng = calloc(n, sizeof(uint64_t));
    r->ext = calloc(n, sizeof(double)); r->last = calloc(n, sizeof(double));
    r->eff = calloc(n, sizeof(double));
    r->W_in = calloc(n * r->Din + 1, sizeof(double));
    r->W_out = calloc(n * r->Dout + 1, sizeof(double));
    neurons_init(r, 1);
    return r;
}

/* spires_reservoir_create -> create_reservoir + init_reservoir
 * (spires_api.c:22-75, reservoir.c:18-78). */
or_res *or_create(const spires_reservoir_config *cfg)
{
    or_res *r = alloc_common(cfg);
    if (!r) return NULL;
    size_t n = r->N;
    double *W = malloc(n * n * sizeof(double) + 1);
    or_init_weights_dense(n, r->Din, r->ctype, r->conn, r->ei, W, r->W_in);
    synapse_build(r, W);
    free(W);
    rescale(r);
    for (size_t k = 0; k < n * r->Dout; k++) r->W_out[k] = gen_weight(0.5);  /* reservoir.c:559-571 */
    return r;
}

or_res *or_create_csr(const spires_reservoir_config *cfg, size_t nnz,
                      const int64_t *rp, const int32_t *ci, const double *val,
                      const double *W_in, const double *W_out)
{
    or_res *r = alloc_common(cfg);
    if (!r) return NULL;
    size_t n = r->N;
    double *W = calloc(n * n + 1, sizeof(double));
    for (size_t i = 0; i < n; i++)
        for (int64_t k = rp[i]; k < rp[i + 1]; k++) W[i * n + ci[k]] = val[k];
    (void)nnz;
    synapse_build(r, W);
    free(W);
    memcpy(r->W_in, W_in, n * r->Din * sizeof(double));
    if (W_out) memcpy(r->W_out, W_out, n * r->Dout * sizeof(double));
    return r;
}

void or_free(or_res *r)
{
    if (!r) return;
    free(r->W_in); free(r->W_out); free(r->rp); free(r->ci); free(r->val); free(r->Wd);
    free(r->psc_tr); free(r->het_tau); free(r->het_tr);
    free(r->fr_c); free(r->fr_I); free(r->fr_hist);
    free(r->me_tau); free(r->me_c); free(r->me_tr); free(r->me_comb);
    free(r->V); free(r->S); free(r->gl_c); free(r->gl_ring); free(r->gl_tprev);
    free(r->cp_c); free(r->cp_dm); free(r->cp_tprev); free(r->cp_head);
    free(r->df_until); free(r->df_rng);
    free(r->ext); free(r->last); free(r->eff);
    free(r);
}

/* reset_reservoir, reservoir.c:770-780: neurons only, from the retained
 * params pointer; the FLIF-GL coefficient table made at creation is kept;
 * synapse state is not touched. */
void or_reset(or_res *r) { if (r) neurons_init(r, 0); }

/* ------------------------------------------------------------- stepping */
/* step_reservoir, reservoir.c:113-165.  u == NULL is treated as zeros. */
void or_step(or_res *r, const double *u)
{
    size_t n = r->N, din = r->Din;
    for (size_t i = 0; i < n; i++) {
        double s = 0.0;
        if (u) for (size_t j = 0; j < din; j++) s += r->W_in[i * din + j] * u[j];
        r->ext[i] = r->gain * s;
    }
    memcpy(r->last, r->S, n * sizeof(double));
    for (int t = 0; t < r->micro; t++) {
        synapse_prepare(r, r->last);
        #pragma omp parallel for if (n >= 256)
        for (size_t i = 0; i < n; i++) {
            double rec = synapse_row(r, i);
            double tot = r->ext[i] + rec;
            neuron_update(r, i, tot);
        }
        if (r->ntype == SPIRES_NEURON_FLIF_GL) r->gl_step++;
        if (r->ntype == SPIRES_NEURON_FLIF_DIFFUSIVE) r->df_step++;
        memcpy(r->last, r->S, n * sizeof(double));
    }
}

/* compute_output, reservoir.c:169-192 (dgemv restated sequentially). */
void or_compute_output(const or_res *r, double *y)
{
    for (size_t o = 0; o < r->Dout; o++) y[o] = dense_dot(&r->W_out[o * r->N], r->V, r->N);
}
void or_read_state(const or_res *r, double *V) { memcpy(V, r->V, r->N * sizeof(double)); }
void or_read_spikes(const or_res *r, double *S) { memcpy(S, r->S, r->N * sizeof(double)); }

void or_run_record(or_res *r, const double *u, size_t T, double *states,
                   uint8_t *raster, double *y)
{
    size_t n = r->N;
    for (size_t t = 0; t < T; t++) {
        or_step(r, u ? &u[t * r->Din] : NULL);
        if (states) memcpy(&states[t * n], r->V, n * sizeof(double));
        if (raster) for (size_t i = 0; i < n; i++) raster[t * n + i] = (uint8_t)(r->S[i] != 0.0);
        if (y) or_compute_output(r, &y[t * r->Dout]);
    }
}

/* --------------------------------------------------------------- readout */
/* LU with partial pivoting and forward/back substitution: the algorithm of
 * LAPACK dgesv used at math_utils.c:187-241 (unblocked form). */
int or_lu_solve(size_t n, const double *A0, const double *b0, double *x)
{
    double *A = malloc(n * n * sizeof(double) + 1), *b = malloc(n * sizeof(double) + 1);
    memcpy(A, A0, n * n * sizeof(double)); memcpy(b, b0, n * sizeof(double));
    int rc = 0;
    for (size_t j = 0; j < n; j++) {
        size_t p = j; double best = fabs(A[j * n + j]);
        for (size_t i = j + 1; i < n; i++) if (fabs(A[i * n + j]) > best) { best = fabs(A[i * n + j]); p = i; }
        if (A[p * n + j] == 0.0) { rc = -1; break; }
        if (p != j) {
            for (size_t k = 0; k < n; k++) { double t = A[j * n + k]; A[j * n + k] = A[p * n + k]; A[p * n + k] = t; }
            double t = b[j]; b[j] = b[p]; b[p] = t;
        }
        double piv = A[j * n + j];
        for (size_t i = j + 1; i < n; i++) {
            double l = A[i * n + j] / piv;
            A[i * n + j] = l;
            if (l != 0.0) for (size_t k = j + 1; k < n; k++) A[i * n + k] -= l * A[j * n + k];
        }
    }
    if (rc == 0) {
        for (size_t i = 0; i < n; i++) { double s = b[i]; for (size_t k = 0; k < i; k++) s -= A[i * n + k] * b[k]; b[i] = s; }
        for (size_t ii = n; ii-- > 0;) {
            double s = b[ii];
            for (size_t k = ii + 1; k < n; k++) s -= A[ii * n + k] * x[k];
            x[ii] = s / A[ii * n + ii];
        }
    }
    free(A); free(b);
    return rc;
}

/* Gram + per-output solve of train_output_ridge_regression, reservoir.c:601-673. */
int or_ridge_solve(const double *X, const double *Y, size_t T, size_t N,
                   size_t Dout, double lambda, double *W)
{
    double *A = calloc(N * N + 1, sizeof(double));
    #pragma omp parallel for schedule(dynamic, 8)
    for (size_t i = 0; i < N; i++)
        for (size_t j = i; j < N; j++) {
            double s = 0.0;
            for (size_t t = 0; t < T; t++) s += X[t * N + i] * X[t * N + j];
            A[i * N + j] = s; A[j * N + i] = s;
        }
    for (size_t i = 0; i < N; i++) A[i * N + i] += lambda;
    int rc = 0;
    double *b = malloc(N * sizeof(double) + 1);
    for (size_t o = 0; o < Dout; o++) {
        for (size_t i = 0; i < N; i++) {
            double s = 0.0;
            for (size_t t = 0; t < T; t++) s += X[t * N + i] * Y[t * Dout + o];
            b[i] = s;
        }
        if (or_lu_solve(N, A, b, &W[o * N]) != 0) rc = -1;   /* row kept on failure */
    }
    free(b); free(A);
    return rc;
}

void or_train_ridge(or_res *r, const double *u, const double *y, size_t T, double lambda)
{
    size_t n = r->N;
    double *X = malloc(T * n * sizeof(double) + 1);
    or_reset(r);
    or_run_record(r, u, T, X, NULL, NULL);
    double *W = malloc(r->Dout * n * sizeof(double) + 1);
    memcpy(W, r->W_out, r->Dout * n * sizeof(double));
    or_ridge_solve(X, y, T, n, r->Dout, lambda, W);
    memcpy(r->W_out, W, r->Dout * n * sizeof(double));
    free(W); free(X);
}

/* train_output_rls, reservoir.c:681-767 (BLAS calls restated sequentially). */
void or_train_rls(or_res *r, const double *u, const double *y, size_t T, double delta, double lambda)
{
    size_t n = r->N, D = r->Dout;
    double *P = calloc(n * n, sizeof(double)), *x = malloc(n * sizeof(double));
    double *k = malloc(n * sizeof(double)), *e = malloc((D ? D : 1) * sizeof(double));
    double id = 1.0 / delta;
    for (size_t i = 0; i < n; i++) P[i * n + i] = id;
    memset(r->W_out, 0, D * n * sizeof(double));
    or_reset(r);
    for (size_t t = 0; t < T; t++) {
        or_step(r, &u[t * r->Din]);
        memcpy(x, r->V, n * sizeof(double));
        for (size_t i = 0; i < n; i++) k[i] = dense_dot(&P[i * n], x, n);
        double den = lambda + dense_dot(x, k, n);
        double sc = 1.0 / den;
        for (size_t i = 0; i < n; i++) k[i] = sc * k[i];
        for (size_t o = 0; o < D; o++) e[o] = y[t * D + o] - dense_dot(&r->W_out[o * n], x, n);
        for (size_t o = 0; o < D; o++) for (size_t j = 0; j < n; j++) r->W_out[o * n + j] += (1.0 * e[o]) * k[j];
        for (size_t i = 0; i < n; i++) for (size_t j = 0; j < n; j++) P[i * n + j] += (-den * k[i]) * k[j];
        if (lambda != 1.0) { double il = 1.0 / lambda; for (size_t q = 0; q < n * n; q++) P[q] = il * P[q]; }
    }
    free(P); free(x); free(k); free(e);
}

/* train_output_iteratively, reservoir.c:244-280. */
void or_train_online(or_res *r, const double *target, double lr)
{
    size_t n = r->N;
    double *st = malloc(n * sizeof(double)), *pred = malloc((r->Dout ? r->Dout : 1) * sizeof(double));
    memcpy(st, r->V, n * sizeof(double));
    or_compute_output(r, pred);
    for (size_t o = 0; o < r->Dout; o++) {
        double a = lr * (target[o] - pred[o]);
        for (size_t j = 0; j < n; j++) r->W_out[o * n + j] += a * st[j];
    }
    free(st); free(pred);
}

/* coarse_grain_reservoir, reservoir.c:782-967. */
or_res *or_coarse_grain(const or_res *r, double thr)
{
    size_t n = r->N, din = r->Din, dout = r->Dout;
    double *W = malloc(n * n * sizeof(double) + 1), *Wi = malloc(n * din * sizeof(double) + 1);
    double *Wo = malloc(n * dout * sizeof(double) + 1), *V = malloc(n * sizeof(double) + 1);
    int *alive = malloc(n * sizeof(int) + 1);
    or_get_dense(r, W);
    memcpy(Wi, r->W_in, n * din * sizeof(double));
    memcpy(Wo, r->W_out, n * dout * sizeof(double));
    for (size_t i = 0; i < n; i++) { V[i] = r->V[i]; alive[i] = 1; }
    int merged = 1;
    while (merged) {
        merged = 0;
        for (size_t i = 0; i < n; i++)
            for (size_t j = i + 1; j < n; j++) {
                if (W[i * n + j] <= thr && W[j * n + i] <= thr) continue;
                merged = 1;
                V[i] = (V[i] + V[j]) / 2.0; V[j] = 0.0; alive[j] = 0;
                for (size_t k = 0; k < n; k++) {
                    if (k == i || k == j) continue;
                    double a = W[i * n + k], b = W[j * n + k];
                    double nw = (a * b == 0.0) ? a + b : (a + b) / 2.0;
                    W[i * n + k] = nw; W[k * n + i] = nw;
                }
                for (size_t k = 0; k < din; k++) {
                    double a = Wi[i * din + k], b = Wi[j * din + k];
                    Wi[i * din + k] = (a * b == 0.0) ? a + b : (a + b) / 2.0;
                    Wi[j * din + k] = 0.0;
                }
                for (size_t o = 0; o < dout; o++) { Wo[o * n + i] += Wo[o * n + j]; Wo[o * n + j] = 0.0; }
                for (size_t k = 0; k < n; k++) { W[j * n + k] = 0.0; W[k * n + j] = 0.0; }
            }
    }
    size_t ns = 0;
    for (size_t i = 0; i < n; i++) ns += alive[i] ? 1 : 0;
    size_t *map = malloc(n * sizeof(size_t) + 1), idx = 0;
    for (size_t i = 0; i < n; i++) map[i] = alive[i] ? idx++ : (size_t)-1;
    double *Wn = calloc(ns * ns + 1, sizeof(double)), *Win = calloc(ns * din + 1, sizeof(double));
    double *Won = calloc(ns * dout + 1, sizeof(double));
    for (size_t i = 0; i < n; i++) {
        if (!alive[i]) continue;
        size_t a = map[i];
        for (size_t j = 0; j < n; j++) if (alive[j]) Wn[a * ns + map[j]] = W[i * n + j];
        for (size_t k = 0; k < din; k++) Win[a * din + k] = Wi[i * din + k];
        for (size_t o = 0; o < dout; o++) Won[o * ns + a] = Wo[o * n + i];
    }
    spires_reservoir_config c;
    memset(&c, 0, sizeof(c));
    c.num_neurons = ns; c.num_inputs = din; c.num_outputs = dout;
    c.spectral_radius = r->rho; c.ei_ratio = r->ei; c.input_strength = r->gain;
    c.connectivity = r->conn; c.dt = r->dt;
    c.connectivity_type = (spires_connectivity_type)r->ctype;
    c.neuron_type = (spires_neuron_type)r->ntype; c.neuron_params = r->np;
    c.synapse_type = (spires_synapse_type)r->stype;
    c.synapse_backend = (spires_synapse_backend)r->sback; c.synapse_params = r->sp;
    or_res *q = alloc_common(&c);
    if (q) {
        synapse_build(q, Wn);
        memcpy(q->W_in, Win, ns * din * sizeof(double));
        memcpy(q->W_out, Won, ns * dout * sizeof(double));
        idx = 0;
        for (size_t i = 0; i < n; i++) if (alive[i]) q->V[idx++] = V[i];
    }
    free(W); free(Wi); free(Wo); free(V); free(alive); free(map); free(Wn); free(Win); free(Won);
    return q;
}

/* ------------------------------------------------------------- accessors */
/* ---- probes for pinning (tests only) ---- */
void or_probe_synapse(or_res *r, const double *spikes, size_t T, double *eff_out)
{
    size_t n = r->N;
    for (size_t t = 0; t < T; t++) {
        synapse_prepare(r, &spikes[t * n]);
        memcpy(&eff_out[t * n], r->eff, n * sizeof(double));
    }
}

void or_probe_csr_rows(size_t n, const int64_t *rp, const int32_t *ci, const double *val,
                       const double *x, double *out)
{
    for (size_t i = 0; i < n; i++) out[i] = csr_dot(rp, ci, val, i, x);
}

int or_probe_multiexp(const or_res *r, double *tau, double *c)
{
    for (int k = 0; k < r->me_N; k++) { tau[k] = r->me_tau[k]; c[k] = r->me_c[k]; }
    return r->me_N;
}

size_t or_num_neurons(const or_res *r) { return r ? r->N : 0; }
size_t or_nnz(const or_res *r) { return r ? r->nnz : 0; }
void or_get_csr(const or_res *r, int64_t *rp, int32_t *ci, double *val)
{
    memcpy(rp, r->rp, (r->N + 1) * sizeof(int64_t));
    memcpy(ci, r->ci, r->nnz * sizeof(int32_t));
    memcpy(val, r->val, r->nnz * sizeof(double));
}
void or_get_dense(const or_res *r, double *W)
{
    size_t n = r->N;
    if (r->sback == SPIRES_SYNAPSE_SPARSE) {
        memset(W, 0, n * n * sizeof(double));
        for (size_t i = 0; i < n; i++)
            for (int64_t k = r->rp[i]; k < r->rp[i + 1]; k++) W[i * n + r->ci[k]] = r->val[k];
    } else memcpy(W, r->Wd, n * n * sizeof(double));
}
void or_get_W_in(const or_res *r, double *W) { memcpy(W, r->W_in, r->N * r->Din * sizeof(double)); }
void or_get_W_out(const or_res *r, double *W) { memcpy(W, r->W_out, r->N * r->Dout * sizeof(double)); }
void or_set_W_out(or_res *r, const double *W) { memcpy(r->W_out, W, r->N * r->Dout * sizeof(double)); }

/* ----------------------------------------------------------------- AGILE */
/* PCG32 (agile.c:10-23). */
static uint32_t pcg32(uint64_t *s)
{
    uint64_t o = *s;
    *s = o * 6364136223846793005ULL + 0xda3e39cb94b95bdbULL;
    uint32_t xs = (uint32_t)(((o >> 18u) ^ o) >> 27u);
    uint32_t rot = (uint32_t)(o >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
static double pcg_u01(uint64_t *s) { return pcg32(s) * (1.0 / 4294967296.0); }
static double clampd(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }

/* reflect_inplace / project, agile.c:32-49,146-156. */
static void project(double *x, const double *lo, const double *hi, int d, int refl)
{
    for (int i = 0; i < d; i++) {
        if (!refl) { x[i] = clampd(x[i], lo[i], hi[i]); continue; }
        if (x[i] < lo[i]) {
            double over = lo[i] - x[i], span = hi[i] - lo[i];
            x[i] = lo[i] + (over - floor(over / span) * span);
        } else if (x[i] > hi[i]) {
            double over = x[i] - hi[i], span = hi[i] - lo[i];
            x[i] = hi[i] - (over - floor(over / span) * span);
        }
    }
}

/* spsa_grad, agile.c:114-142. */
static void spsa(const double *th, const double *lo, const double *hi, int d, double eps,
                 or_loss_fn f, void *ctx, uint64_t *st, double *g)
{
    double dl[64], tp[64], tm[64];
    for (int i = 0; i < d; i++) {
        dl[i] = (pcg32(st) & 1) ? 1.0 : -1.0;
        tp[i] = clampd(th[i] + eps * dl[i], lo[i], hi[i]);
        tm[i] = clampd(th[i] - eps * dl[i], lo[i], hi[i]);
    }
    double fp = f(tp, d, ctx), fm = f(tm, d, ctx);
    double sc = (fp - fm) / (2.0 * eps);
    for (int i = 0; i < d; i++) g[i] = sc * dl[i];
}

/* agile_optimize, agile.c:159-351 (gradient callback path omitted: the
 * spires wrapper always forces SPSA, spires_opt_agile.c:128-135). */
int or_agile_optimize(const double *th0, const double *lo, const double *hi, int d,
                      or_loss_fn f, void *ctx, const struct or_agile_opts *oi,
                      double *best, struct or_agile_report *rep)
{
    if (d > 64) return -1;
    struct or_agile_opts o = { 1.3, 1e-3, 5e-1, 100.0, 0.9, 0.0, 200, 30, 1e-3, 1, 0xC0FFEEULL, 0 };
    if (oi) o = *oi;
    if (o.refine_step <= 0.0) o.refine_step = o.eta_min;
    double th[64], cand[64], g[64], dv[64];
    uint64_t rng = o.seed ? o.seed : 0xA57B17E5ULL;
    memcpy(th, th0, d * sizeof(double));
    project(th, lo, hi, d, o.reflect_at_bounds);
    double loss = f(th, d, ctx), bl = loss;
    memcpy(best, th, d * sizeof(double));
    int evals = 1, iters = 0, bit = 0, refine = 0;
    double pat = o.patience0;
    int since = 0;
    for (;;) {
        iters++;
        spsa(th, lo, hi, d, o.spsa_eps, f, ctx, &rng, g);
        evals += 2;
        double gn = 0.0;
        for (int i = 0; i < d; i++) gn += g[i] * g[i];
        gn = sqrt(gn);
        if (gn > 1e-18) { for (int i = 0; i < d; i++) dv[i] = (-1.0 / gn) * g[i]; }
        else {
            double s = 0.0;
            for (int i = 0; i < d; i++) { double z = pcg_u01(&rng) * 2.0 - 1.0; dv[i] = z; s += z * z; }
            double nn = sqrt(s) + 1e-18;
            for (int i = 0; i < d; i++) dv[i] /= nn;
        }
        double mu = o.mu <= 1.0 ? 1.000001 : o.mu;
        double uu = pcg_u01(&rng);
        double c1 = pow(o.eta_min, 1.0 - mu), c2 = pow(o.eta_max, 1.0 - mu);
        double eta = pow(c1 + (c2 - c1) * uu, 1.0 / (1.0 - mu));
        memcpy(cand, th, d * sizeof(double));
        for (int i = 0; i < d; i++) cand[i] += eta * dv[i];
        project(cand, lo, hi, d, o.reflect_at_bounds);
        double cl = f(cand, d, ctx);
        evals++;
        if (cl < loss) {
            memcpy(th, cand, d * sizeof(double));
            loss = cl; since = 0; pat *= o.patience_decay;
            if (pat < 1.0) pat = 1.0;
            if (loss < bl) { bl = loss; bit = iters; memcpy(best, th, d * sizeof(double)); }
        } else since++;
        if (since >= (int)ceil(pat)) { refine = 1; break; }
        if (iters > 1000000) break;
    }
    if (refine) {
        memcpy(th, best, d * sizeof(double));
        loss = bl;
        int ni = 0;
        for (int k = 0; k < o.refine_max_iters; ++k) {
            iters++;
            spsa(th, lo, hi, d, o.spsa_eps, f, ctx, &rng, g);
            evals += 2;
            for (int i = 0; i < d; i++) th[i] += -o.refine_step * g[i];
            project(th, lo, hi, d, o.reflect_at_bounds);
            double nl = f(th, d, ctx);
            evals++;
            if (nl < loss - 1e-12) {
                loss = nl;
                if (loss < bl) { bl = loss; bit = iters; memcpy(best, th, d * sizeof(double)); }
                ni = 0;
            } else ni++;
            if (ni >= o.refine_no_improve) break;
        }
    }
    if (rep) { rep->best_loss = bl; rep->best_iter = bit; rep->total_evals = evals; rep->total_iters = iters; rep->exited_refinement = refine; }
    return 0;
}
