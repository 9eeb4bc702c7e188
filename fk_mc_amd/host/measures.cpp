// measures.cpp -- see fkmc/measures.hpp.
#include "fkmc/measures.hpp"

#include <iostream>
#include <string>

namespace fk {

void check_status(fkgpu_ctx* ctx, int r, const char* what) {
    if (r != FKGPU_OK) throw std::logic_error(std::string(what) + ": " + fkgpu_last_error(ctx));
}

void measure_energy::accumulate(double) {
    config.calc_ed(false);
    const auto& spectrum = config.ed_data_.cached_spectrum;
    const auto& exp_e = config.ed_data_.cached_exp;
    _Z++;
    double e_c = 0.0, d2e = 0.0;
    for (size_t i = 0; i < spectrum.size(); ++i) {
        e_c += spectrum[i] / (1.0 + exp_e[i]);
        d2e += spectrum[i] * spectrum[i] / (1.0 + 0.5 * (exp_e[i] + 1.0 / exp_e[i]));
    }
    const double e = e_c - config.params_.mu_f * (double)config.get_nf() + config.calc_ff_energy();
    d2e /= 2.0;
    _average_energy += e;
    _average_d2energy += d2e;
    _energies.push_back(e);
    _d2energies.push_back(d2e);
    _c_energies.push_back(e_c);
}

void measure_energy::collect_results() {
    std::cout << "Total energy: " << average_energy() << std::endl;
    std::cout << "Total d2energy: " << average_d2energy() << std::endl;
}

measure_spectrum::measure_spectrum(configuration_t& in, std::vector<double>& average_spectrum)
    : config(in), _average_spectrum(average_spectrum) {
    _average_spectrum.resize(config.lattice_.msize(), 0.0);
}

void measure_spectrum::accumulate(double) {
    config.calc_ed(false);
    const auto& s = config.ed_data_.cached_spectrum;
    for (size_t i = 0; i < _average_spectrum.size(); ++i)
        _average_spectrum[i] = (_average_spectrum[i] * _Z + s[i]) / (_Z + 1);
    _Z++;
}

measure_spectrum_history::measure_spectrum_history(configuration_t& in,
                                                   std::vector<std::vector<double>>& spectrum_history)
    : config(in), _spectrum_history(spectrum_history) {
    _spectrum_history.resize(config.lattice_.msize());
}

void measure_spectrum_history::accumulate(double) {
    config.calc_ed(false);
    const auto& s = config.ed_data_.cached_spectrum;
    for (size_t i = 0; i < s.size(); ++i) _spectrum_history[i].push_back(s[i]);
    _Z++;
}

void measure_focc::accumulate(double) {
    for (size_t i = 0; i < config.f_config_.size(); ++i) focc_[i].push_back(config.f_config_(i));
    _Z++;
}

void measure_eigenfunctions::accumulate(double) {
    config.calc_ed(true);
    eigenfunctions_.push_back(config.ed_data_.cached_evecs);
    _Z++;
}

std::vector<std::vector<double>> density_of_states(const abstract_lattice& lattice,
                                                   const std::vector<std::vector<double>>& spectra,
                                                   const std::vector<double>& wgrid, double eta) {
    const size_t n = lattice.msize(), ns = spectra.size(), nw = wgrid.size();
    std::vector<std::vector<double>> out(ns, std::vector<double>(nw));
    if (!ns || !nw) return out;
    std::vector<double> flat(ns * n), dos(ns * nw);
    for (size_t s = 0; s < ns; ++s) {
        if (spectra[s].size() != n) throw std::logic_error("density_of_states: spectrum size != msize");
        std::copy(spectra[s].begin(), spectra[s].end(), flat.begin() + s * n);
    }
    fkgpu_ctx* ctx = lattice.device();
    check_status(ctx, fkgpu_dos(ctx, (int)ns, flat.data(), (int)nw, wgrid.data(), eta, dos.data()), "dos");
    for (size_t s = 0; s < ns; ++s) std::copy(dos.begin() + s * nw, dos.begin() + (s + 1) * nw, out[s].begin());
    return out;
}

}  // namespace fk
