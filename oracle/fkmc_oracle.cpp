// fkmc_oracle.cpp -- CPU restatement of the fk_mc per-move diagonalisation path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is compared
// against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  The product library (fk_mc_amd) never links or calls it.
//
// It restates, in plain C++17 with no third-party libraries, the algorithms of the
// reference files cited at each function.  Random numbers come from the real
// libstdc++ std::mt19937 and std::*_distribution objects, exactly as the reference
// draws them, so an accept/reject trace produced here is the reference's trace for
// the same seed (same toolchain).
//
// The dense eigensolver is of the same algorithm class as Eigen's
// SelfAdjointEigenSolver (used at src/configuration.cpp:213): scaling by the largest
// coefficient, Householder tridiagonalisation, implicit symmetric QR steps with the
// Wilkinson shift, eigenvalues sorted ascending.  It is pinned against LAPACK
// (numpy eigvalsh) and closed forms in tests/test_oracle.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace orc {

typedef std::mt19937 rng_t;

// ---------------------------------------------------------------------------
// Lattice: hopping matrix (src/lattice/hypercubic.cpp:31-51 index order,
// :115-131 nearest neighbours, :137-155 triangular).  Dense row-major n x n.
// ---------------------------------------------------------------------------
struct lattice {
    int kind = 0;  // 0 hypercubic, 1 triangular
    int D = 2, L = 4;
    double t = 1.0, tp = 0.0;
    int n = 0;
    std::vector<double> hop;               // dense n*n
    std::vector<std::vector<int>> nbr;     // neighbour list in neighbor_index order

    void pos(int index, int* p) const {
        for (int i = D - 1; i >= 0; --i) { p[i] = index % L; index /= L; }
    }
    int index(const int* p) const {
        int out = 0, mult = 1;
        for (int i = D - 1; i >= 0; --i) { out += p[i] * mult; mult *= L; }
        return out;
    }
};

static void build_lattice(lattice& l) {
    if (l.L < 3) throw std::invalid_argument("lattice size L must be >= 3");
    if (l.kind == 1 && l.D != 2) throw std::invalid_argument("triangular lattice is 2-D");
    l.n = 1;
    for (int i = 0; i < l.D; ++i) l.n *= l.L;
    l.hop.assign((size_t)l.n * l.n, 0.0);
    l.nbr.assign(l.n, {});
    int p[3], q[3];
    for (int i = 0; i < l.n; ++i) {
        l.pos(i, p);
        for (int d = 0; d < l.D; ++d) {
            std::memcpy(q, p, sizeof(p));
            q[d] = p[d] > 0 ? p[d] - 1 : l.L - 1;
            int jl = l.index(q);
            std::memcpy(q, p, sizeof(p));
            q[d] = p[d] < l.L - 1 ? p[d] + 1 : 0;
            int jr = l.index(q);
            l.hop[(size_t)i * l.n + jl] += -l.t;
            l.hop[(size_t)i * l.n + jr] += -l.t;
            l.nbr[i].push_back(jl);
            l.nbr[i].push_back(jr);
        }
        if (l.kind == 1) {
            int ql[3], qr[3];
            for (int d = 0; d < 2; ++d) {
                ql[d] = p[d] > 0 ? p[d] - 1 : l.L - 1;
                qr[d] = p[d] < l.L - 1 ? p[d] + 1 : 0;
            }
            int jl = l.index(ql), jr = l.index(qr);
            l.hop[(size_t)i * l.n + jl] += -l.tp;
            l.hop[(size_t)i * l.n + jr] += -l.tp;
            l.nbr[i].push_back(jl);
            l.nbr[i].push_back(jr);
        }
    }
}

// ---------------------------------------------------------------------------
// Dense symmetric eigensolver (Eigen SelfAdjointEigenSolver algorithm class).
// ---------------------------------------------------------------------------

// Householder reduction of symmetric A (row-major, full storage) to tridiagonal
// T = Q^T A Q.  d[n], e[n-1].  If Q != nullptr it receives Q (row-major).
static void tridiagonalize(int n, std::vector<double>& A, double* d, double* e, double* Q) {
    std::vector<double> v(n), p(n), w(n);
    std::vector<double> taus(n > 1 ? n - 1 : 1, 0.0);
    std::vector<double> vs(Q ? (size_t)n * n : 0, 0.0);  // reflector k stored in row k
    for (int k = 0; k + 2 < n; ++k) {
        int m = n - k - 1;
        double alpha = A[(size_t)(k + 1) * n + k];
        double xnorm2 = 0.0;
        for (int i = k + 2; i < n; ++i) xnorm2 += A[(size_t)i * n + k] * A[(size_t)i * n + k];
        double tau = 0.0, beta = alpha;
        if (xnorm2 != 0.0) {
            beta = -std::copysign(std::sqrt(alpha * alpha + xnorm2), alpha);
            tau = (beta - alpha) / beta;
            double scal = 1.0 / (alpha - beta);
            v[k + 1] = 1.0;
            for (int i = k + 2; i < n; ++i) v[i] = A[(size_t)i * n + k] * scal;
        }
        e[k] = beta;
        taus[k] = tau;
        if (tau != 0.0) {
            // p = tau * A22 v
            for (int i = k + 1; i < n; ++i) {
                double s = 0.0;
                for (int j = k + 1; j < n; ++j) s += A[(size_t)i * n + j] * v[j];
                p[i] = tau * s;
            }
            double pv = 0.0;
            for (int i = k + 1; i < n; ++i) pv += p[i] * v[i];
            double c = 0.5 * tau * pv;
            for (int i = k + 1; i < n; ++i) w[i] = p[i] - c * v[i];
            for (int i = k + 1; i < n; ++i)
                for (int j = k + 1; j < n; ++j)
                    A[(size_t)i * n + j] -= v[i] * w[j] + w[i] * v[j];
            if (Q) {
                for (int i = k + 1; i < n; ++i) vs[(size_t)k * n + i] = v[i];
            }
        }
        (void)m;
    }
    for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i];
    if (n >= 2) e[n - 2] = A[(size_t)(n - 1) * n + (n - 2)];
    if (Q) {
        // Q = H_0 H_1 ... H_{n-3}; accumulate backwards starting from identity.
        std::fill(Q, Q + (size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
        for (int k = n - 3; k >= 0; --k) {
            double tau = taus[k];
            if (tau == 0.0) continue;
            const double* vk = &vs[(size_t)k * n];
            for (int j = k + 1; j < n; ++j) {
                double s = 0.0;
                for (int i = k + 1; i < n; ++i) s += vk[i] * Q[(size_t)i * n + j];
                s *= tau;
                for (int i = k + 1; i < n; ++i) Q[(size_t)i * n + j] -= s * vk[i];
            }
        }
    }
}

// Implicit symmetric QR with the Wilkinson shift on (d, e); rotations are
// accumulated into the columns of Z (row-major n x n) when Z != nullptr.
static void tridiagonal_qr(int n, double* d, double* e, double* Z) {
    const double eps = std::numeric_limits<double>::epsilon();
    const double tiny = std::numeric_limits<double>::min();
    int end = n - 1, start = 0;
    int iter = 0;
    const int maxit = 30 * n;
    while (end > 0) {
        for (int i = start; i < end; ++i) {
            if (std::abs(e[i]) <= eps * (std::abs(d[i]) + std::abs(d[i + 1])) || std::abs(e[i]) <= tiny)
                e[i] = 0.0;
        }
        while (end > 0 && e[end - 1] == 0.0) --end;
        if (end <= 0) break;
        if (++iter > maxit) throw std::runtime_error("tridiagonal QR did not converge");
        start = end - 1;
        while (start > 0 && e[start - 1] != 0.0) --start;
        // Wilkinson shift from the trailing 2x2 block.
        double td = (d[end - 1] - d[end]) * 0.5;
        double e2 = e[end - 1] * e[end - 1];
        double mu = d[end];
        if (td == 0.0) {
            mu -= std::abs(e[end - 1]);
        } else if (e2 != 0.0) {
            double h = std::hypot(td, e[end - 1]);
            mu -= e2 / (td + (td > 0 ? h : -h));
        }
        // Bulge chase.  R_k = [[c,-s],[s,c]] acts on columns k,k+1 (T <- R^T T R);
        // (x, z) is the pair the rotation annihilates: the first column of
        // T - mu I at k == start, afterwards (e[k-1], bulge at (k+1,k-1)).
        double x = d[start] - mu;
        double z = e[start];
        for (int k = start; k < end; ++k) {
            double r = std::hypot(x, z);
            double c = 1.0, s = 0.0;
            if (r != 0.0) { c = x / r; s = z / r; }
            if (k > start) e[k - 1] = r;
            double a = d[k], b = e[k], dd = d[k + 1];
            d[k] = c * c * a + 2.0 * c * s * b + s * s * dd;
            d[k + 1] = s * s * a - 2.0 * c * s * b + c * c * dd;
            e[k] = c * s * (dd - a) + (c * c - s * s) * b;
            if (k + 1 < end) {
                z = s * e[k + 1];
                e[k + 1] = c * e[k + 1];
            }
            x = e[k];
            if (Z) {
                for (int i = 0; i < n; ++i) {
                    double zk = Z[(size_t)i * n + k], zk1 = Z[(size_t)i * n + k + 1];
                    Z[(size_t)i * n + k] = c * zk + s * zk1;
                    Z[(size_t)i * n + k + 1] = -s * zk + c * zk1;
                }
            }
        }
    }
}

// Full eigensolve: evals ascending; evecs (row-major, column k = eigenvector k) if wanted.
static void eigh(int n, const double* H, double* evals, double* evecs) {
    if (n == 0) return;
    std::vector<double> A(H, H + (size_t)n * n);
    double scale = 0.0;
    for (double a : A) scale = std::max(scale, std::abs(a));
    if (scale == 0.0) scale = 1.0;
    for (double& a : A) a /= scale;
    std::vector<double> d(n), e(std::max(n - 1, 1)), Q;
    if (evecs) Q.resize((size_t)n * n);
    if (n == 1) {
        d[0] = A[0];
        if (evecs) Q[0] = 1.0;
    } else {
        tridiagonalize(n, A, d.data(), e.data(), evecs ? Q.data() : nullptr);
        tridiagonal_qr(n, d.data(), e.data(), evecs ? Q.data() : nullptr);
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int k = 0; k < n; ++k) evals[k] = d[order[k]] * scale;
    if (evecs)
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) evecs[(size_t)i * n + k] = Q[(size_t)i * n + order[k]];
}

// ---------------------------------------------------------------------------
// Omega / logZ (src/configuration.cpp:316-336): shifted log-sum, exactly as
// written there, plus the cached exp and Fermi arrays.
// ---------------------------------------------------------------------------
static double logz_reference(int n, const double* ev, double beta, double* cexp, double* cfermi) {
    double e0 = ev[0];
    double logw0 = beta * e0;
    double weight0 = std::exp(logw0);
    double logz = 0.0;
    for (int i = 0; i < n; ++i) {
        double e = ev[i];
        double w = std::exp(-beta * (e - e0));
        double ce = std::exp(beta * e);
        if (cexp) cexp[i] = ce;
        if (cfermi) cfermi[i] = 1.0 / (1.0 + ce);
        logz += std::log(weight0 + w) - logw0;
    }
    return logz;
}

// Overflow-free form log(1+e^{-x}) summed with Neumaier compensation.
static double logz_stable(int n, const double* ev, double beta) {
    double s = 0.0, c = 0.0;
    for (int i = 0; i < n; ++i) {
        double x = beta * ev[i];
        double term = x > 0 ? std::log1p(std::exp(-x)) : -x + std::log1p(std::exp(x));
        double t = s + term;
        if (std::abs(s) >= std::abs(term)) c += (s - t) + term; else c += (term - t) + s;
        s = t;
    }
    return s + c;
}

// ---------------------------------------------------------------------------
// Configuration (include/fk_mc/configuration.hpp:52-87, src/configuration.cpp)
// ---------------------------------------------------------------------------
struct config {
    const lattice* lat;
    double beta, U, mu_c, mu_f;
    std::vector<double> W;
    std::vector<int> f;
    // ed cache
    int status = 0;  // 0 empty, 1 spectrum, 2 full
    std::vector<double> spectrum, cexp, cfermi, evecs;
    double logZ = 0.0;
    int use_stable_logz = 0;

    int nf() const { return std::accumulate(f.begin(), f.end(), 0); }

    // src/configuration.cpp:47-56
    void randomize_f(rng_t& rnd, size_t nfs = 0) {
        std::uniform_int_distribution<> distr(0, lat->n - 1);
        if (!nfs) nfs = distr(rnd);
        std::fill(f.begin(), f.end(), 0);
        for (size_t i = 0; i < nfs; ++i) {
            size_t ind = distr(rnd);
            while (f[ind] == 1) ind = distr(rnd);
            f[ind] = 1;
        }
    }
    // src/configuration.cpp:59-77 (1-D only)
    double ff_energy() const {
        if (lat->D != 1) return 0.0;
        double e = 0.0;
        int V = lat->n;
        for (int i = 0; i < V; ++i) {
            if (!f[i]) continue;
            for (int l = 0; l < (int)W.size(); ++l) {
                int left = (i - l + V) % V, right = (i + l) % V;
                e += W[l] * f[left];
                e += W[l] * f[right] * (l > 0);
            }
        }
        return e;
    }
    // src/configuration.cpp:79-91
    void hamiltonian(std::vector<double>& H) const {
        int n = lat->n;
        H = lat->hop;
        for (int i = 0; i < n; ++i) H[(size_t)i * n + i] += -mu_c + U * f[i];
    }
    void calc_hamiltonian() { status = 0; }
    // src/configuration.cpp:208-246
    void calc_ed(bool evecs_wanted) {
        if ((status == 1 && !evecs_wanted) || (status == 2 && evecs_wanted)) return;
        int n = lat->n;
        std::vector<double> H;
        hamiltonian(H);
        spectrum.resize(n);
        if (evecs_wanted) evecs.resize((size_t)n * n);
        eigh(n, H.data(), spectrum.data(), evecs_wanted ? evecs.data() : nullptr);
        status = evecs_wanted ? 2 : 1;
        cexp.resize(n);
        cfermi.resize(n);
        logZ = logz_reference(n, spectrum.data(), beta, cexp.data(), cfermi.data());
        if (use_stable_logz) logZ = logz_stable(n, spectrum.data(), beta);
    }
};

// ---------------------------------------------------------------------------
// Moves (src/moves.cpp) plus the nearest-neighbour exchange (no reference
// counterpart; symmetric proposal over the hopping graph's neighbour list).
// Each attempt records what the trace needs.
// ---------------------------------------------------------------------------
enum { MOVE_ADDREMOVE = 0, MOVE_FLIP = 1, MOVE_HOP = 2, MOVE_RANDOMIZE = 3 };

// Test hooks.  g_mutation plants a deliberate mistake in the nearest-neighbour
// exchange so that the tests can show their checks catch it (0: none; 1: the sign
// of dlogZ in its weight flipped; 2: the sign of the f-f term flipped; 3: a
// neighbour slot whose occupation equals site i's is replaced by the next slot
// with a different occupation, an f-dependent proposal with no Hastings factor).
// g_chain_W is the f-f interaction W given to the chains orc_run_chain builds.
static int g_mutation = 0;
static std::vector<double> g_chain_W;
// g_flip_ff: the global exchange weighs -beta dE_ff as the Chebyshev move_flip does
// (src/moves_chebyshev.cpp:22); 0 (default) is the ED move_flip (src/moves.cpp:19-20).
static int g_flip_ff = 0;

struct attempt_rec {
    int site_a = -1, site_b = -1;
    int solved = 0;
    double logz_new = 0.0, logw = -INFINITY, w = 0.0;
};

struct chain {
    config cur, prop;
    rng_t rng;
    std::vector<int> moves;
    std::vector<double> probs;
    std::discrete_distribution<> move_distrib;
    std::uniform_real_distribution<> metropolis = std::uniform_real_distribution<>(0, 1);
    long naccept = 0;

    chain(const lattice* l, double beta, double U, double mu_c, double mu_f, uint32_t seed)
        : cur{l, beta, U, mu_c, mu_f, {}, std::vector<int>(l->n, 0)},
          prop{l, beta, U, mu_c, mu_f, {}, std::vector<int>(l->n, 0)}, rng(seed) {}

    void add_move(int kind, double p) {
        moves.push_back(kind);
        probs.push_back(p);
        move_distrib = std::discrete_distribution<>(probs.begin(), probs.end());
    }

    // src/moves.cpp:52-67
    double attempt_addremove(attempt_rec& r) {
        std::uniform_int_distribution<> distr(0, cur.lat->n - 1);
        prop.f = cur.f;
        size_t to = distr(rng);
        prop.f[to] = 1 - cur.f[to];
        r.site_a = (int)to;
        cur.calc_ed(false);
        prop.calc_hamiltonian();
        prop.calc_ed(false);
        r.solved = 1;
        r.logz_new = prop.logZ;
        double exp_beta_mu_f = std::exp(cur.beta * cur.mu_f);
        double ff_diff = prop.ff_energy() - cur.ff_energy();
        double ratio = std::exp(prop.logZ - cur.logZ);
        double out = (prop.f[to] ? ratio * exp_beta_mu_f : ratio / exp_beta_mu_f) * std::exp(-cur.beta * ff_diff);
        r.logw = (prop.logZ - cur.logZ) + (prop.f[to] ? 1 : -1) * cur.beta * cur.mu_f - cur.beta * ff_diff;
        return out;
    }
    // src/moves.cpp:5-21
    double attempt_flip(attempt_rec& r) {
        std::uniform_int_distribution<> distr(0, cur.lat->n - 1);
        int nf = cur.nf();
        if (nf == 0 || nf == cur.lat->n) return 0;
        prop.f = cur.f;
        size_t from = distr(rng);
        while (prop.f[from] == 0) from = distr(rng);
        size_t to = distr(rng);
        while (prop.f[to] == 1) to = distr(rng);
        cur.calc_ed(false);
        prop.f[from] = 0;
        prop.f[to] = 1;
        r.site_a = (int)from;
        r.site_b = (int)to;
        prop.calc_hamiltonian();
        prop.calc_ed(false);
        r.solved = 1;
        r.logz_new = prop.logZ;
        const double ff_diff = g_flip_ff ? prop.ff_energy() - cur.ff_energy() : 0.0;
        r.logw = prop.logZ - cur.logZ - cur.beta * ff_diff;
        return std::exp(prop.logZ - cur.logZ - cur.beta * ff_diff);
    }
    // Nearest-neighbour exchange: site i ~ U{0..V-1}, neighbour slot ~ U{0..z_i-1};
    // the pair is swapped when the occupations differ, otherwise it is a null move
    // with weight 1 (no eigensolve).
    double attempt_hop(attempt_rec& r) {
        int n = cur.lat->n;
        std::uniform_int_distribution<> distr(0, n - 1);
        int i = distr(rng);
        const std::vector<int>& nb = cur.lat->nbr[i];
        std::uniform_int_distribution<> dn(0, (int)nb.size() - 1);
        int slot = dn(rng);
        int j = nb[slot];
        if (g_mutation == 3) {
            for (int k = 1; k < (int)nb.size() && cur.f[j] == cur.f[i]; ++k) j = nb[(slot + k) % nb.size()];
        }
        r.site_a = i;
        r.site_b = j;
        prop.f = cur.f;
        if (cur.f[i] == cur.f[j]) { r.logw = 0.0; return 1.0; }
        std::swap(prop.f[i], prop.f[j]);
        cur.calc_ed(false);
        prop.calc_hamiltonian();
        prop.calc_ed(false);
        r.solved = 1;
        r.logz_new = prop.logZ;
        double ff_diff = prop.ff_energy() - cur.ff_energy();
        double dlz = prop.logZ - cur.logZ;
        if (g_mutation == 1) dlz = -dlz;
        if (g_mutation == 2) ff_diff = -ff_diff;
        r.logw = dlz - cur.beta * ff_diff;
        return std::exp(dlz) * std::exp(-cur.beta * ff_diff);
    }
    // src/moves.cpp:35-49 (including its clamp quirk)
    double attempt_randomize(attempt_rec& r) {
        prop.f = cur.f;
        prop.randomize_f(rng);
        prop.calc_hamiltonian();
        cur.calc_ed(false);
        prop.calc_ed(false);
        r.solved = 1;
        r.logz_new = prop.logZ;
        double log_ratio = prop.logZ - cur.logZ;
        double ff_diff = prop.ff_energy() - cur.ff_energy();
        double beta = cur.beta;
        double dn = (double)(prop.nf() - cur.nf());
        r.logw = log_ratio + beta * cur.mu_f * dn - ff_diff;
        if (beta * cur.mu_f * dn - ff_diff > 2.7182818 - log_ratio) return 1;
        else if (beta * cur.mu_f * dn - ff_diff + log_ratio < 0) return 0;
        return std::exp(log_ratio) * std::exp(beta * (cur.mu_f * dn - ff_diff));
    }

    // src/mc_metropolis.cpp:34-52 -- one move.  Returns accepted flag.
    int step(int forced, int* kind_out, attempt_rec& r, double* u_out) {
        int mi = move_distrib(rng);
        int kind = moves[mi];
        double w = 0.0;
        switch (kind) {
            case MOVE_ADDREMOVE: w = attempt_addremove(r); break;
            case MOVE_FLIP: w = attempt_flip(r); break;
            case MOVE_HOP: w = attempt_hop(r); break;
            case MOVE_RANDOMIZE: w = attempt_randomize(r); break;
        }
        r.w = w;
        double u = metropolis(rng);
        int acc = std::abs(w) > u;
        if (forced >= 0) acc = forced;
        if (acc) {
            cur.f = prop.f;
            if (r.solved) {
                cur.spectrum = prop.spectrum;
                cur.cexp = prop.cexp;
                cur.cfermi = prop.cfermi;
                cur.logZ = prop.logZ;
                cur.status = prop.status == 2 ? 1 : prop.status;
            }
            ++naccept;
        }
        *kind_out = kind;
        *u_out = u;
        return acc;
    }
};

}  // namespace orc

using namespace orc;

// ===========================================================================
// C ABI for the Python test harness (ctypes).  All functions return 0 on
// success and a negative number on invalid input.
// ===========================================================================
extern "C" {

int orc_lattice_hop(int kind, int D, int L, double t, double tp, double* hop_out, int* nbr_out, int* z_out) {
    try {
        lattice l;
        l.kind = kind; l.D = D; l.L = L; l.t = t; l.tp = tp;
        build_lattice(l);
        if (hop_out) std::memcpy(hop_out, l.hop.data(), sizeof(double) * l.hop.size());
        if (nbr_out && z_out) {
            int zmax = 0;
            for (auto& v : l.nbr) zmax = std::max<int>(zmax, (int)v.size());
            for (int i = 0; i < l.n; ++i) {
                z_out[i] = (int)l.nbr[i].size();
                for (int k = 0; k < (int)l.nbr[i].size(); ++k) nbr_out[(size_t)i * zmax + k] = l.nbr[i][k];
            }
        }
        return l.n;
    } catch (...) { return -1; }
}

int orc_eigh(int n, const double* H, double* evals, double* evecs) {
    try { eigh(n, H, evals, evecs); return 0; } catch (...) { return -1; }
}

int orc_tridiagonalize(int n, const double* H, double* d, double* e) {
    std::vector<double> A(H, H + (size_t)n * n);
    tridiagonalize(n, A, d, e, nullptr);
    return 0;
}

double orc_logz_reference(int n, const double* ev, double beta, double* cexp, double* cfermi) {
    return logz_reference(n, ev, beta, cexp, cfermi);
}

double orc_logz_stable(int n, const double* ev, double beta) { return logz_stable(n, ev, beta); }

// Hamiltonian + eigenvalues + logZ of one configuration (calc_hamiltonian + calc_ed).
int orc_calc_ed(int kind, int D, int L, double t, double tp, const int* f, double beta, double U, double mu_c,
                int want_vectors, double* evals, double* evecs, double* logZ, double* cexp, double* cfermi) {
    try {
        lattice l;
        l.kind = kind; l.D = D; l.L = L; l.t = t; l.tp = tp;
        build_lattice(l);
        config c{&l, beta, U, mu_c, 0.0, {}, std::vector<int>(f, f + l.n)};
        c.calc_hamiltonian();
        c.calc_ed(want_vectors != 0);
        std::memcpy(evals, c.spectrum.data(), sizeof(double) * l.n);
        if (want_vectors && evecs) std::memcpy(evecs, c.evecs.data(), sizeof(double) * l.n * l.n);
        if (logZ) *logZ = c.logZ;
        if (cexp) std::memcpy(cexp, c.cexp.data(), sizeof(double) * l.n);
        if (cfermi) std::memcpy(cfermi, c.cfermi.data(), sizeof(double) * l.n);
        return 0;
    } catch (...) { return -1; }
}

// randomize_f with a fresh mt19937(seed) (src/configuration.cpp:47-56).
int orc_randomize_f(int n, uint32_t seed, int nf, int* f_out) {
    lattice l; l.n = n;
    config c{&l, 1, 0, 0, 0, {}, std::vector<int>(n, 0)};
    rng_t r(seed);
    c.randomize_f(r, (size_t)nf);
    std::memcpy(f_out, c.f.data(), sizeof(int) * n);
    return 0;
}

double orc_ff_energy(int D, int n, const int* f, int nW, const double* W) {
    lattice l; l.D = D; l.n = n;
    config c{&l, 1, 0, 0, 0, std::vector<double>(W, W + nW), std::vector<int>(f, f + n)};
    return c.ff_energy();
}

// Runs one Metropolis chain of nmoves single moves (sweep_len = 1 per update
// call does not change the draw order) and records the per-move trace.
// move_kinds/move_probs: registered moves in add_move order.
// forced (nullable): per move, -1 = decide normally, 0/1 = force the decision
// (used to follow another implementation's choice at a counted near-tie).
// Trace arrays (each nmoves long, nullable): kind, site_a, site_b, u, logw,
// logz_new, accepted, nf_after, logz_cur_after.
int orc_run_chain(int kind, int D, int L, double t, double tp,
                  double beta, double U, double mu_c, double mu_f, uint32_t seed,
                  const int* f_init, int nmoves_kinds, const int* move_kinds, const double* move_probs,
                  int nmoves, const int* forced, int stable_logz,
                  int* tr_kind, int* tr_a, int* tr_b, double* tr_u, double* tr_logw, double* tr_logz,
                  int* tr_acc, int* tr_nf, double* tr_logz_cur, int* f_final, double* spectrum_final) {
    try {
        lattice l;
        l.kind = kind; l.D = D; l.L = L; l.t = t; l.tp = tp;
        build_lattice(l);
        chain ch(&l, beta, U, mu_c, mu_f, seed);
        ch.cur.use_stable_logz = ch.prop.use_stable_logz = stable_logz;
        ch.cur.f.assign(f_init, f_init + l.n);
        ch.cur.W = ch.prop.W = g_chain_W;
        for (int m = 0; m < nmoves_kinds; ++m) ch.add_move(move_kinds[m], move_probs[m]);
        ch.cur.calc_hamiltonian();
        ch.cur.calc_ed(false);
        for (int s = 0; s < nmoves; ++s) {
            attempt_rec r;
            int k = 0;
            double u = 0;
            int acc = ch.step(forced ? forced[s] : -1, &k, r, &u);
            if (tr_kind) tr_kind[s] = k;
            if (tr_a) tr_a[s] = r.site_a;
            if (tr_b) tr_b[s] = r.site_b;
            if (tr_u) tr_u[s] = u;
            if (tr_logw) tr_logw[s] = r.logw;
            if (tr_logz) tr_logz[s] = r.logz_new;
            if (tr_acc) tr_acc[s] = acc;
            if (tr_nf) tr_nf[s] = ch.cur.nf();
            if (tr_logz_cur) { ch.cur.calc_ed(false); tr_logz_cur[s] = ch.cur.logZ; }
        }
        if (f_final) std::memcpy(f_final, ch.cur.f.data(), sizeof(int) * l.n);
        if (spectrum_final) { ch.cur.calc_ed(false); std::memcpy(spectrum_final, ch.cur.spectrum.data(), sizeof(double) * l.n); }
        return (int)ch.naccept;
    } catch (...) { return -1; }
}

void orc_set_mutation(int m) { g_mutation = m; }
void orc_set_flip_ff(int on) { g_flip_ff = on; }
void orc_set_chain_ff(int nW, const double* W) { g_chain_W.assign(W, W + nW); }

// Raw libstdc++ draws, for pinning the device port of the generator and
// distributions: mode 0 = mt19937 words, 1 = uniform_int(0, hi), 2 =
// uniform_real(0,1), 3 = discrete_distribution over probs[0..nprobs).
int orc_rng_draws(uint32_t seed, int mode, int hi, int nprobs, const double* probs, int count, double* out) {
    rng_t r(seed);
    std::uniform_int_distribution<> ui(0, hi);
    std::uniform_real_distribution<> ur(0, 1);
    std::discrete_distribution<> dd;
    if (mode == 3) dd = std::discrete_distribution<>(probs, probs + nprobs);
    for (int i = 0; i < count; ++i) {
        switch (mode) {
            case 0: out[i] = (double)r(); break;
            case 1: out[i] = ui(r); break;
            case 2: out[i] = ur(r); break;
            case 3: out[i] = dd(r); break;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Measures (src/measures/energy.cpp:6-26, include/fk_mc/measures/ipr.hpp:39-56,
// include/fk_mc/measures/stiffness.hpp:69-187).
// ---------------------------------------------------------------------------
int orc_measure_energy(int n, const double* spectrum, const double* cexp, double mu_f, int nf, double e_ff,
                       double* e_out, double* d2e_out, double* ec_out) {
    double se = 0.0, sd2 = 0.0;
    // Eigen's .sum() reduction order is unspecified; a plain left-to-right sum is used.
    for (int i = 0; i < n; ++i) {
        se += spectrum[i] / (1.0 + cexp[i]);
        sd2 += spectrum[i] * spectrum[i] / (1.0 + 0.5 * (cexp[i] + 1.0 / cexp[i]));
    }
    *ec_out = se;
    *e_out = se - mu_f * nf + e_ff;
    *d2e_out = sd2 / 2.0;
    return 0;
}

// evecs row-major, column k = eigenvector k.  out[k] = ||v_k||_4 / ||v_k||_2^2.
int orc_measure_ipr(int n, const double* evecs, double* out) {
    for (int k = 0; k < n; ++k) {
        double s2 = 0.0, s4 = 0.0;
        for (int i = 0; i < n; ++i) {
            double v = evecs[(size_t)i * n + k];
            s2 += v * v;
            s4 += v * v * v * v;
        }
        out[k] = std::sqrt(std::sqrt(s4)) / s2;
    }
    return 0;
}

// Stiffness of a 2-D or 3-D hypercubic lattice along direction 0.
// Returns (T+V)/N in *stiffness; cond_out[w] = sum of Lorentzian resonant terms at wgrid[w].
int orc_measure_stiffness(int D, int L, int n, const double* evecs, const double* evals, const double* fermi,
                          int nw, const double* wgrid, double offset, double* stiffness, double* cond_out,
                          double* T_out, double* V_out) {
    // T_x and J_x: hops along direction 0 (the slowest index).
    std::vector<int> left(n), right(n);
    int stride = 1;
    for (int d = 1; d < D; ++d) stride *= L;
    for (int i = 0; i < n; ++i) {
        int x = i / stride, rest = i % stride;
        left[i] = ((x - 1 + L) % L) * stride + rest;
        right[i] = ((x + 1) % L) * stride + rest;
    }
    const double t_ = 1.0;
    // mT = V^T Tm V diagonal, mJ = V^T Jm V (full).  Tm(i,left)=-t, Tm(i,right)=-t; Jm(i,left)=-1, Jm(i,right)=+1.
    std::vector<double> JV((size_t)n * n), TV((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            double vl = evecs[(size_t)left[i] * n + k], vr = evecs[(size_t)right[i] * n + k];
            // duplicates (L<3) would be summed by setFromTriplets
            TV[(size_t)i * n + k] = -t_ * vl - t_ * vr;
            JV[(size_t)i * n + k] = -vl + vr;
        }
    double T = 0.0;
    for (int k = 0; k < n; ++k) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += evecs[(size_t)i * n + k] * TV[(size_t)i * n + k];
        T += s * fermi[k];
    }
    T *= -M_PI;
    std::vector<double> mJ((size_t)n * n);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += evecs[(size_t)i * n + a] * JV[(size_t)i * n + b];
            mJ[(size_t)a * n + b] = s;
        }
    double V = 0.0;
    std::vector<double> cond(nw, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            bool neq = std::abs(evals[i] - evals[j]) > 1e-12;
            if (!neq) continue;
            double sigma = M_PI * (fermi[j] - fermi[i]) * mJ[(size_t)j * n + i] * mJ[(size_t)i * n + j];
            double Vv = 2.0 * sigma / (evals[i] - evals[j]);
            if (std::abs(sigma) > 1e-13) {
                V += Vv;
                for (int w = 0; w < nw; ++w) {
                    double e1 = evals[j] - evals[i], e2 = evals[i] - evals[j];
                    double x1 = wgrid[w] + e1, x2 = wgrid[w] + e2;
                    cond[w] += offset / M_PI / (x1 * x1 + offset * offset) * sigma;
                    cond[w] += offset / M_PI / (x2 * x2 + offset * offset) * (-sigma);
                }
            }
        }
    *stiffness = (V + T) / n;
    if (T_out) *T_out = T;
    if (V_out) *V_out = V;
    for (int w = 0; w < nw; ++w) cond_out[w] = cond[w];
    return 0;
}

// ---------------------------------------------------------------------------
// Chebyshev evaluator (include/fk_mc/chebyshev.hpp:15-63) and exact-trace
// Chebyshev logZ (src/configuration.cpp:94-206), dense matrices.
// ---------------------------------------------------------------------------
struct cheb_eval {
    int M, G;
    std::vector<double> angle, x, T;  // T[k*G+i]
    cheb_eval(int max_moment, int grid) : G(grid) {
        M = max_moment + max_moment % 2;
        angle.resize(G); x.resize(G); T.resize((size_t)M * G);
        for (int i = 0; i < G; ++i) {
            angle[i] = G > 1 ? (double)i / (double)(G - 1) : 0.0;
            x[i] = -std::cos(M_PI * angle[i]);
            for (int k = 0; k < M; ++k) T[(size_t)k * G + i] = std::cos(k * std::acos(x[i]));
        }
    }
    double moment(const std::vector<double>& vals, int order) const {
        double s = 0.0;
        for (int i = 0; i + 1 < G; ++i)
            s += (vals[i + 1] * T[(size_t)order * G + i + 1] + vals[i] * T[(size_t)order * G + i]) * (angle[i + 1] - angle[i]);
        return s * 0.5;
    }
};

// fn: 0 -> 1, 1 -> x^2, 2 -> sin(x)
double orc_cheb_moment_f(int max_moment, int grid, int fn, int order) {
    cheb_eval c(max_moment, grid);
    std::vector<double> v(grid);
    for (int i = 0; i < grid; ++i) v[i] = fn == 0 ? 1.0 : fn == 1 ? c.x[i] * c.x[i] : std::sin(c.x[i]);
    return c.moment(v, order);
}

// g_m = moment_f(w -> n*log(1+exp(-beta*(a*w+b))), m) for m < M.
int orc_cheb_logz_coeffs(int max_moment, int grid, int n, double beta, double a, double b, double* g_out) {
    cheb_eval c(max_moment, grid);
    std::vector<double> v(grid);
    for (int i = 0; i < grid; ++i) v[i] = n * std::log(1.0 + std::exp(-beta * (a * c.x[i] + b)));
    for (int m = 0; m < c.M; ++m) g_out[m] = c.moment(v, m);
    return c.M;
}

// Exact-trace moments mu_m = Tr T_m(X)/n, X = (H-b)/a with a,b from the exact
// extreme eigenvalues; doubling identities as in src/configuration.cpp:137-194
// (T_{2m-1} = 2 T_m T_{m-1} - X, T_{2m} = 2 T_m^2 - 1, traced).
int orc_calc_chebyshev(int kind, int D, int L, double t, double tp, const int* f, double beta, double U, double mu_c,
                       int max_moment, int grid, double* moments_out, double* logZ_out, double* ab_out) {
    try {
        lattice l;
        l.kind = kind; l.D = D; l.L = L; l.t = t; l.tp = tp;
        build_lattice(l);
        int n = l.n;
        config c{&l, beta, U, mu_c, 0.0, {}, std::vector<int>(f, f + n)};
        std::vector<double> H;
        c.hamiltonian(H);
        std::vector<double> ev(n);
        eigh(n, H.data(), ev.data(), nullptr);
        double e_min = ev[0], e_max = ev[n - 1];
        double a = (e_max - e_min) / 2.0, b = (e_max + e_min) / 2.0;
        std::vector<double> X(H);
        for (int i = 0; i < n; ++i) X[(size_t)i * n + i] -= b;
        for (double& v : X) v /= a;
        cheb_eval ce(max_moment, grid);
        int M = ce.M;
        std::vector<double> mom(M, 0.0);
        std::vector<char> is_set(M, 0);
        std::vector<double> c0((size_t)n * n, 0.0), c1(X), tmp((size_t)n * n);
        for (int i = 0; i < n; ++i) c0[(size_t)i * n + i] = 1.0;
        auto trace = [&](const std::vector<double>& A) { double s = 0; for (int i = 0; i < n; ++i) s += A[(size_t)i * n + i]; return s; };
        auto trace_prod = [&](const std::vector<double>& A, const std::vector<double>& B) {
            double s = 0; for (int i = 0; i < n; ++i) for (int k = 0; k < n; ++k) s += A[(size_t)i * n + k] * B[(size_t)k * n + i]; return s; };
        double trX = trace(X);
        mom[0] = 1.0; is_set[0] = 1;
        mom[1] = trace(c1) / n; is_set[1] = 1;
        int m = 2;
        for (; m <= M / 2; ++m) {
            // tmp = 2 X c1 - c0
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < n; ++k) s += X[(size_t)i * n + k] * c1[(size_t)k * n + j];
                    tmp[(size_t)i * n + j] = 2.0 * s - c0[(size_t)i * n + j];
                }
            std::swap(c0, c1);
            std::swap(c1, tmp);
            if (!is_set[m]) { mom[m] = trace(c1) / n; is_set[m] = 1; }
            int k = 2 * m - 1;
            if (k < M && k >= M / 2) {
                mom[k] = (trace_prod(c0, c1) * 2.0 - trX) / n;
                is_set[k] = 1;
                if (k != M - 1) {
                    ++k;
                    mom[k] = trace_prod(c1, c1) / n * 2.0 - 1.0;
                    is_set[k] = 1;
                }
            }
        }
        std::vector<double> v(grid);
        for (int i = 0; i < grid; ++i) v[i] = n * std::log(1.0 + std::exp(-beta * (a * ce.x[i] + b)));
        double s = ce.moment(v, 0);
        for (int mm = 1; mm < M; ++mm) s += 2.0 * ce.moment(v, mm) * mom[mm];
        for (int mm = 0; mm < M; ++mm) moments_out[mm] = mom[mm];
        *logZ_out = s;
        ab_out[0] = a; ab_out[1] = b; ab_out[2] = e_min; ab_out[3] = e_max;
        return M;
    } catch (...) { return -1; }
}

// CPU baseline: nthreads independent addremove chains of nmoves each (one chain
// per std::thread), each move one full FP64 eigensolve.  Returns moves done.
long orc_bench_chains(int kind, int D, int L, double t, double tp, double beta, double U, double mu,
                      uint32_t seed, int nthreads, int nmoves) {
    lattice l;
    l.kind = kind; l.D = D; l.L = L; l.t = t; l.tp = tp;
    build_lattice(l);
    std::vector<std::thread> th;
    std::vector<long> done(nthreads, 0);
    for (int k = 0; k < nthreads; ++k)
        th.emplace_back([&, k]() {
            chain ch(&l, beta, U, mu, mu, seed + k);
            std::bernoulli_distribution bd(0.5);
            rng_t init(seed + 1000003u * (k + 1));
            for (int i = 0; i < l.n; ++i) ch.cur.f[i] = bd(init);
            ch.add_move(MOVE_ADDREMOVE, 1.0);
            ch.cur.calc_ed(false);
            for (int s = 0; s < nmoves; ++s) {
                attempt_rec r; int kd; double u;
                ch.step(-1, &kd, r, &u);
            }
            done[k] = nmoves;
        });
    for (auto& x : th) x.join();
    long tot = 0;
    for (long d : done) tot += d;
    return tot;
}

// Chains with their own (beta, U, mu_c = mu_f = mu, starting f, seed) and move mix, one
// per thread, nmoves moves each: the CPU figure of a batched workload.  Returns the total
// number of moves; *nsolves_out (nullable) the number of eigensolves they needed.
long orc_bench_chain_set(int kind, int D, int L, double t, double tp, int nchains, const double* beta,
                         const double* U, const double* mu, const int* f0, const uint32_t* seeds, int nkinds,
                         const int* kinds, const double* probs, int nmoves, long* nsolves_out) {
    lattice l;
    l.kind = kind; l.D = D; l.L = L; l.t = t; l.tp = tp;
    try { build_lattice(l); } catch (...) { return -1; }
    std::vector<std::thread> th;
    std::vector<long> done(nchains, 0), solves(nchains, 0);
    for (int k = 0; k < nchains; ++k)
        th.emplace_back([&, k]() {
            chain ch(&l, beta[k], U[k], mu[k], mu[k], seeds[k]);
            ch.cur.f.assign(f0 + (size_t)k * l.n, f0 + (size_t)(k + 1) * l.n);
            for (int m = 0; m < nkinds; ++m) ch.add_move(kinds[m], probs[m]);
            ch.cur.calc_ed(false);
            for (int s = 0; s < nmoves; ++s) {
                attempt_rec r; int kd; double u;
                ch.step(-1, &kd, r, &u);
                solves[k] += r.solved;
            }
            done[k] = nmoves;
        });
    for (auto& x : th) x.join();
    long tot = 0, ns = 0;
    for (int k = 0; k < nchains; ++k) { tot += done[k]; ns += solves[k]; }
    if (nsolves_out) *nsolves_out = ns;
    return tot;
}

}  // extern "C"
