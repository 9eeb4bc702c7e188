// measures.hpp -- Monte Carlo measures with the reference's interface
// (include/fk_mc/measures/*.hpp, src/measures/*.cpp): accumulate(double sign) is
// called once per measurement by mc_metropolis::measure().  The eigenvector-based
// measures (IPR, stiffness) run their O(N^3) work on the GPU from the current
// f-configuration; energy and the spectra read the ed_cache filled by calc_ed.
// collect_results() is the single-process reduction (one chain per process).
#pragma once
#include <string>
#include <vector>

#include "configuration.hpp"

namespace fk {

struct measure_energy {  // src/measures/energy.cpp:6-26
    typedef configuration_t::real_array_t real_array_t;
    double beta;
    configuration_t& config;
    int _Z = 0;
    double _average_energy = 0.0;
    double _average_d2energy = 0.0;
    std::vector<double>& _energies;
    std::vector<double>& _d2energies;
    std::vector<double>& _c_energies;
    measure_energy(double beta, configuration_t& in, std::vector<double>& energies, std::vector<double>& d2energies,
                   std::vector<double>& c_energies)
        : beta(beta), config(in), _energies(energies), _d2energies(d2energies), _c_energies(c_energies) {}
    void accumulate(double sign);
    void collect_results();
    double average_energy() const { return _Z ? _average_energy / _Z : 0.0; }
    double average_d2energy() const { return _Z ? _average_d2energy / _Z : 0.0; }
};

struct measure_spectrum {  // src/measures/spectrum.cpp:13-21 (running mean)
    configuration_t& config;
    int _Z = 0;
    std::vector<double>& _average_spectrum;
    measure_spectrum(configuration_t& in, std::vector<double>& average_spectrum);
    void accumulate(double sign);
    void collect_results() {}
};

struct measure_spectrum_history {  // src/measures/spectrum_history.cpp:13-19
    configuration_t& config;
    int _Z = 0;
    std::vector<std::vector<double>>& _spectrum_history;
    measure_spectrum_history(configuration_t& in, std::vector<std::vector<double>>& spectrum_history);
    void accumulate(double sign);
    void collect_results() {}
};

struct measure_focc {  // src/measures/focc_history.cpp
    const configuration_t& config;
    int _Z = 0;
    std::vector<std::vector<double>>& focc_;
    measure_focc(const configuration_t& in, std::vector<std::vector<double>>& focc) : config(in), focc_(focc) {
        focc_.resize(config.lattice_.volume());
    }
    void accumulate(double sign);
    void collect_results() {}
};

// Eigenvectors per measurement, row-major msize x msize (column k = eigenvector k);
// src/measures/eigenfunctions.cpp:12-18.
struct measure_eigenfunctions {
    configuration_t& config;
    int _Z = 0;
    std::vector<std::vector<double>>& eigenfunctions_;
    measure_eigenfunctions(configuration_t& in, std::vector<std::vector<double>>& eigenfunctions)
        : config(in), eigenfunctions_(eigenfunctions) {}
    void accumulate(double sign);
    void collect_results() {}
};

// IPR per eigenvector, ||v_k||_4 / ||v_k||_2^2 (include/fk_mc/measures/ipr.hpp:39-56).
template <typename lattice_t>
struct measure_ipr {
    const lattice_t& lattice_;
    configuration_t& config_;
    int _Z = 0;
    std::vector<std::vector<double>>& ipr_vals_;
    measure_ipr(configuration_t& in, const lattice_t& lattice, std::vector<std::vector<double>>& ipr_vals)
        : lattice_(lattice), config_(in), ipr_vals_(ipr_vals) {
        ipr_vals_.resize(config_.lattice_.volume());
    }
    void accumulate(double sign);
    void collect_results() {}
};

// Superfluid stiffness and conductivity along direction 0 of a hypercubic lattice
// (include/fk_mc/measures/stiffness.hpp:69-187).
template <typename lattice_t>
struct measure_stiffness {
    measure_stiffness(configuration_t& in, const lattice_t& lattice, std::vector<double>& stiffness_vals,
                      std::vector<std::vector<double>>& cond_history, std::vector<double> wgrid, double offset)
        : lattice_(lattice), config_(in), stiffness_vals_(stiffness_vals), cond_history_(cond_history),
          wgrid_(std::move(wgrid)), offset_(offset) {
        cond_history_.resize(wgrid_.size());
    }
    void accumulate(double sign);
    void collect_results() {}

protected:
    const lattice_t& lattice_;
    configuration_t& config_;
    std::vector<double>& stiffness_vals_;
    std::vector<std::vector<double>>& cond_history_;
    std::vector<double> wgrid_;
    double offset_;
};

// DOS of a set of spectra (prog/data_save.hxx:264-345): out[s][w] =
// (1/N) sum_k (eta/pi) / ((wgrid[w] - e_k)^2 + eta^2), on the lattice's device.
std::vector<std::vector<double>> density_of_states(const abstract_lattice& lattice,
                                                   const std::vector<std::vector<double>>& spectra,
                                                   const std::vector<double>& wgrid, double eta);

void check_status(fkgpu_ctx* ctx, int r, const char* what);  // FKGPU_* -> std::logic_error

template <typename lattice_t>
void measure_ipr<lattice_t>::accumulate(double) {
    const size_t n = config_.lattice_.msize();
    std::vector<double> evals(n), ipr(n);
    fkgpu_ctx* ctx = config_.lattice_.device();
    const double beta = config_.params_.beta, U = config_.params_.U, mu = config_.params_.mu_c;
    check_status(ctx, fkgpu_measure_ipr(ctx, 1, config_.f_config_.data(), &beta, &U, &mu, evals.data(), ipr.data()),
                 "measure_ipr");
    for (size_t i = 0; i < n && i < ipr_vals_.size(); ++i) ipr_vals_[i].push_back(ipr[i]);
    _Z++;
}

template <typename lattice_t>
void measure_stiffness<lattice_t>::accumulate(double) {
    fkgpu_ctx* ctx = config_.lattice_.device();
    const double beta = config_.params_.beta, U = config_.params_.U, mu = config_.params_.mu_c;
    double s = 0.0;
    std::vector<double> cond(wgrid_.size());
    check_status(ctx,
                 fkgpu_measure_stiffness(ctx, 1, config_.f_config_.data(), &beta, &U, &mu, (int)wgrid_.size(),
                                         wgrid_.data(), offset_, &s, cond.data()),
                 "measure_stiffness");
    stiffness_vals_.push_back(s);
    for (size_t w = 0; w < wgrid_.size(); ++w) cond_history_[w].push_back(cond[w]);
}

}  // namespace fk
