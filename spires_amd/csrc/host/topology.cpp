// Recurrent topology and weight generation on the host.
//
// Restates init_weights (reference src/reservoir.c:345-545) with sparse
// per-row storage instead of the reference's dense N x N scratch buffer, while
// drawing glibc rand() values in exactly the same order, so that a caller who
// seeds with srand() gets the reference's W bit for bit:
//   W_in: N*Din generate_weight draws (magnitude, then sign)
//   RANDOM: one draw per ordered pair i != j, plus generate_weight per edge
//   SMALL_WORLD: K/2 forward ring edges per node, then a rewiring pass
//   SCALE_FREE: clique seed + preferential attachment on (deg + 1), then a
//               random orientation draw per undirected edge.
// The preferential-attachment roulette uses a Fenwick tree over the integer
// weights deg+1.  Every partial sum of the reference's double accumulator is an
// exact integer (< 2^53), so "first u with r <= prefix(u)" picks the same node.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <vector>

#include "host.hpp"

namespace shost {

double urand01() { return (double)rand() / (double)RAND_MAX; }

double gen_weight(double ei) {
    double mag = (double)rand() / RAND_MAX;
    if (((double)rand() / RAND_MAX) < ei) return mag;
    return -mag;
}

namespace {

// Fenwick tree over int64 weights with "smallest index whose prefix >= target".
struct Fenwick {
    std::vector<int64_t> t;
    int n = 0;
    explicit Fenwick(int n_) : t(n_ + 1, 0), n(n_) {}
    void add(int i, int64_t v) { for (++i; i <= n; i += i & -i) t[i] += v; }
    int64_t total() const {
        int64_t s = 0;
        for (int i = n; i > 0; i -= i & -i) s += t[i];
        return s;
    }
    // smallest idx with prefix(idx) >= target (target >= 1); n if none
    int lower_bound(int64_t target) const {
        int pos = 0;
        int step = 1;
        while (step * 2 <= n) step *= 2;
        int64_t acc = 0;
        for (; step > 0; step >>= 1) {
            if (pos + step <= n && acc + t[pos + step] < target) {
                pos += step;
                acc += t[pos];
            }
        }
        return pos;  // 0-based index
    }
};

void rows_to_csr(const std::vector<std::map<int32_t, double>>& rows, Csr& W) {
    const size_t n = rows.size();
    W.n = n;
    W.rp.assign(n + 1, 0);
    W.ci.clear();
    W.val.clear();
    for (size_t i = 0; i < n; ++i) {
        for (const auto& kv : rows[i])
            if (kv.second != 0.0) { W.ci.push_back(kv.first); W.val.push_back(kv.second); }
        W.rp[i + 1] = (int64_t)W.ci.size();
    }
}

}  // namespace

int generate_topology(size_t n, size_t din, int ctype, double conn, double ei,
                      std::vector<double>& W_in, Csr& W) {
    W_in.resize(n * din);
    for (size_t k = 0; k < n * din; ++k) W_in[k] = gen_weight(ei);
    std::vector<std::map<int32_t, double>> rows(n);
    if (ctype == SPIRES_CONN_RANDOM) {
        W.n = n;
        W.rp.assign(n + 1, 0);
        W.ci.clear();
        W.val.clear();
        for (size_t i = 0; i < n; ++i) {
            for (size_t j = 0; j < n; ++j) {
                if (i == j) continue;
                if (urand01() < conn) {
                    double w = gen_weight(ei);
                    if (w != 0.0) { W.ci.push_back((int32_t)j); W.val.push_back(w); }
                }
            }
            W.rp[i + 1] = (int64_t)W.ci.size();
        }
        return 0;
    }
    if (ctype == SPIRES_CONN_SMALL_WORLD) {
        if (n >= 3) {
            int K = (int)(conn * (double)(n - 1) + 0.5);
            if (K < 2) K = 2;
            if (K % 2) K++;
            if (K >= (int)n) K = (int)n - 1;
            const double p = 0.1;
            const int half = K / 2;
            for (size_t i = 0; i < n; ++i)
                for (int s = 1; s <= half; ++s) {
                    size_t j = (i + (size_t)s) % n;
                    if (i == j) continue;
                    rows[i][(int32_t)j] = gen_weight(ei);
                }
            auto has_edge = [&](size_t i, size_t j) {
                auto it = rows[i].find((int32_t)j);
                return it != rows[i].end() && it->second != 0.0;
            };
            for (size_t i = 0; i < n; ++i)
                for (int s = 1; s <= half; ++s) {
                    size_t jo = (i + (size_t)s) % n;
                    if (urand01() < p) {
                        rows[i][(int32_t)jo] = 0.0;
                        size_t jn;
                        int tries = 0;
                        for (;;) {
                            jn = (size_t)(urand01() * (double)n);
                            if (jn >= n) jn = n - 1;
                            if (++tries > 10 * (int)n) break;
                            if (!(jn == i || has_edge(i, jn))) break;
                        }
                        if (jn != i) rows[i][(int32_t)jn] = gen_weight(ei);
                    }
                }
        }
        rows_to_csr(rows, W);
        return 0;
    }
    if (ctype == SPIRES_CONN_SCALE_FREE) {
        if (n >= 2) {
            int m = (int)(conn * (double)(n - 1) + 0.5);
            if (m < 1) m = 1;
            if (m >= (int)n) m = (int)n - 1;
            int m0 = m > 2 ? m : 2;
            if (m0 >= (int)n) m0 = (int)n - 1;
            std::vector<std::vector<int32_t>> adj(n);   // undirected neighbour lists
            std::vector<int64_t> deg(n, 0);
            auto adjacent = [&](size_t v, size_t u) {
                const auto& a = adj[v].size() <= adj[u].size() ? adj[v] : adj[u];
                const size_t other = adj[v].size() <= adj[u].size() ? u : v;
                for (int32_t x : a) if ((size_t)x == other) return true;
                return false;
            };
            auto link = [&](size_t v, size_t u) {
                adj[v].push_back((int32_t)u);
                adj[u].push_back((int32_t)v);
                deg[v]++; deg[u]++;
            };
            for (int a = 0; a < m0; ++a)
                for (int b = a + 1; b < m0; ++b) link((size_t)a, (size_t)b);
            Fenwick fw((int)n);
            for (int a = 0; a < m0; ++a) fw.add(a, deg[a] + 1);
            for (size_t v = (size_t)m0; v < n; ++v) {
                int added = 0, guard = 0;
                while (added < m && guard < 50 * m) {
                    const int64_t itotal = fw.total();           // sum over u < v of deg[u] + 1
                    double total = (double)itotal;
                    if (total <= 0.0) total = (double)v;
                    const double r = urand01() * total;
                    size_t pick = 0;
                    int64_t target = (int64_t)std::ceil(r);
                    if (target < 1) target = 1;
                    const int idx = fw.lower_bound(target);
                    if (idx < (int)v) pick = (size_t)idx;
                    if (pick != v && !adjacent(v, pick)) {
                        link(v, pick);
                        fw.add((int)pick, 1);
                        added++;
                    }
                    guard++;
                }
                while (added < m) {
                    size_t pick = (size_t)(urand01() * (double)v);
                    if (pick >= v) pick = v - 1;
                    if (pick != v && !adjacent(v, pick)) {
                        link(v, pick);
                        fw.add((int)pick, 1);
                        added++;
                    }
                }
                fw.add((int)v, deg[v] + 1);
            }
            for (size_t i = 0; i < n; ++i) {
                std::vector<int32_t> hi;
                for (int32_t j : adj[i]) if ((size_t)j > i) hi.push_back(j);
                std::sort(hi.begin(), hi.end());
                for (int32_t j : hi) {
                    if (urand01() < 0.5) rows[i][j] = gen_weight(ei);
                    else rows[(size_t)j][(int32_t)i] = gen_weight(ei);
                }
            }
        }
        rows_to_csr(rows, W);
        return 0;
    }
    return -1;
}

double csr_spectral_radius(const Csr& W) {
    const size_t n = W.n;
    std::vector<double> x(n, 1.0), y(n);
    double lo = 0.0, ln = 0.0;
    for (unsigned it = 0; it < 1000; ++it) {
        for (size_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t k = W.rp[i]; k < W.rp[i + 1]; ++k) s += W.val[k] * x[W.ci[k]];
            y[i] = s;
        }
        ln = 0.0;
        for (size_t i = 0; i < n; ++i) { double a = std::fabs(y[i]); if (a > ln) ln = a; }
        for (size_t i = 0; i < n; ++i) x[i] = y[i] / ln;
        if (std::fabs(ln - lo) < 1e-6) break;
        lo = ln;
    }
    return ln;
}

double dense_spectral_radius(const std::vector<double>& W, size_t n) {
    std::vector<double> x(n, 1.0), y(n);
    double lo = 0.0, ln = 0.0;
    for (unsigned it = 0; it < 1000; ++it) {
        for (size_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (size_t j = 0; j < n; ++j) s += W[i * n + j] * x[j];
            y[i] = s;
        }
        ln = 0.0;
        for (size_t i = 0; i < n; ++i) { double a = std::fabs(y[i]); if (a > ln) ln = a; }
        for (size_t i = 0; i < n; ++i) x[i] = y[i] / ln;
        if (std::fabs(ln - lo) < 1e-6) break;
        lo = ln;
    }
    return ln;
}

}  // namespace shost
