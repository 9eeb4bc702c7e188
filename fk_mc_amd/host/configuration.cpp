// configuration.cpp -- see fkmc/configuration.hpp.
#include "fkmc/configuration.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <string>

namespace fk {

bool config_params::operator==(const config_params& rhs) const {
    const double tol = std::numeric_limits<double>::epsilon();
    return std::abs(beta - rhs.beta) < tol && std::abs(U - rhs.U) < tol && std::abs(mu_c - rhs.mu_c) < tol &&
           std::abs(mu_f - rhs.mu_f) < tol;
}

configuration_t::configuration_t(const abstract_lattice& lattice, double beta, double U, double mu_c, double mu_f,
                                 std::vector<double> W)
    : lattice_(lattice), params_{beta, U, mu_c, mu_f, W}, f_config_(lattice.volume(), 0) {
    hamilt_.n = (int)lattice.msize();
    hamilt_.row_ptr.assign(hamilt_.n + 1, 0);
}

configuration_t& configuration_t::operator=(const configuration_t& rhs) {
    f_config_ = rhs.f_config_;
    hamilt_ = rhs.hamilt_;
    ed_data_ = rhs.ed_data_;
    cheb_data_ = rhs.cheb_data_;
    if (!(params_ == rhs.params_)) throw std::logic_error("Mismatched parameters in config assignment");
    return *this;
}

size_t configuration_t::get_nf() const {
    return (size_t)std::accumulate(f_config_.begin(), f_config_.begin() + lattice_.volume(), 0);
}

// src/configuration.cpp:47-56: nf sites drawn by rejection; nf = 0 draws nf first.
void configuration_t::randomize_f(random_generator& rnd, size_t nf) {
    std::uniform_int_distribution<> distr(0, (int)lattice_.volume() - 1);
    if (!nf) nf = distr(rnd);
    f_config_.setZero();
    for (size_t i = 0; i < nf; ++i) {
        size_t ind = distr(rnd);
        while (f_config_(ind) == 1) ind = distr(rnd);
        f_config_(ind) = 1;
    }
}

// src/configuration.cpp:59-77 (1-D lattices only)
double configuration_t::calc_ff_energy() const {
    if (lattice_.ndim() != 1) return 0;
    double e = 0;
    const int V = (int)lattice_.volume();
    for (int i = 0; i < V; ++i) {
        if (!f_config_(i)) continue;
        for (int l = 0; l < (int)params_.W.size(); ++l) {
            const int left = (i - l + V) % V, right = (i + l) % V;
            e += params_.W[l] * f_config_(left);
            e += params_.W[l] * f_config_(right) * (l > 0);
        }
    }
    return e;
}

// src/configuration.cpp:79-91: H = hopping + diag(-mu_c + U n_f); resets the caches.
// (The device assembles H on chip from the same f; this host copy serves callers of hamilt_.)
const sparse_m& configuration_t::calc_hamiltonian() {
    reset_cache();
    const sparse_m& hop = lattice_.hopping_m();
    const size_t norbs = lattice_.norbs();
    sparse_m h;
    h.n = hop.n;
    h.row_ptr.assign(1, 0);
    for (int i = 0; i < hop.n; ++i) {
        const double dg = -params_.mu_c + params_.U * f_config_((size_t)i / norbs);
        bool placed = false;
        for (int p = hop.row_ptr[i]; p < hop.row_ptr[i + 1]; ++p) {
            const int j = hop.col[p];
            if (!placed && j > i) { h.col.push_back(i); h.val.push_back(dg); placed = true; }
            h.col.push_back(j);
            h.val.push_back(hop.val[p] + (j == i ? dg : 0.0));
            if (j == i) placed = true;
        }
        if (!placed) { h.col.push_back(i); h.val.push_back(dg); }
        h.row_ptr.push_back((int)h.col.size());
    }
    hamilt_ = std::move(h);
    return hamilt_;
}

static void check(fkgpu_ctx* ctx, int r, const char* what) {
    if (r != FKGPU_OK) throw std::logic_error(std::string(what) + ": " + fkgpu_last_error(ctx));
}

// src/configuration.cpp:208-246 with the eigensolve and the Omega sum on the GPU.
void configuration_t::calc_ed(bool calc_evecs) {
    if ((ed_data_.status == ed_cache::spectrum && !calc_evecs) || (ed_data_.status == ed_cache::full && calc_evecs))
        return;
    fkgpu_ctx* ctx = lattice_.device();
    const size_t n = lattice_.msize();
    ed_data_.cached_spectrum.resize(n);
    ed_data_.cached_exp.resize(n);
    ed_data_.cached_fermi.resize(n);
    if (calc_evecs) ed_data_.cached_evecs.resize(n * n);
    const int32_t* f = f_config_.data();
    check(ctx,
          fkgpu_eig_logz(ctx, 1, f, &params_.beta, &params_.U, &params_.mu_c, calc_evecs ? 1 : 0,
                         ed_data_.cached_spectrum.data(), calc_evecs ? ed_data_.cached_evecs.data() : nullptr,
                         &ed_data_.logZ, ed_data_.cached_exp.data(), ed_data_.cached_fermi.data()),
          "calc_ed");
    ed_data_.status = calc_evecs ? ed_cache::full : ed_cache::spectrum;
}

// src/configuration.cpp:94-206 (exact trace) on the GPU.
void configuration_t::calc_chebyshev(const chebyshev::chebyshev_eval& cheb) {
    if ((int)cheb_data_.status >= (int)chebyshev_cache::logz) return;
    fkgpu_ctx* ctx = lattice_.device();
    cheb_data_.moments.assign(cheb.cheb_size(), 0.0);
    double ab[4];
    check(ctx,
          fkgpu_cheb_logz(ctx, f_config_.data(), params_.beta, params_.U, params_.mu_c, cheb.cheb_size(),
                          cheb.grid_size(), cheb_data_.moments.data(), &cheb_data_.logZ, ab),
          "calc_chebyshev");
    cheb_data_.a = ab[0];
    cheb_data_.b = ab[1];
    cheb_data_.e_min = ab[2];
    cheb_data_.e_max = ab[3];
    cheb_data_.status = chebyshev_cache::logz;
}

}  // namespace fk
