// configuration.hpp -- configuration_t with the reference's interface
// (include/fk_mc/configuration.hpp:14-87).  calc_ed and calc_chebyshev run on the
// GPU through the fkgpu C ABI; the caches, their status rules and the error
// convention (std::logic_error) are those of src/configuration.cpp.
#pragma once
#include <random>
#include <stdexcept>
#include <vector>

#include "chebyshev.hpp"
#include "lattice.hpp"

namespace fk {

typedef std::mt19937 random_generator;

// Eigen::ArrayXi / ArrayXd stand-ins: element access by () and [].
template <class T>
struct array_t : std::vector<T> {
    using std::vector<T>::vector;
    T& operator()(size_t i) { return (*this)[i]; }
    const T& operator()(size_t i) const { return (*this)[i]; }
    void setZero() { std::fill(this->begin(), this->end(), T(0)); }
    T sum() const { T s = 0; for (const T& x : *this) s += x; return s; }
};

struct config_params {
    double beta, U, mu_c, mu_f;
    std::vector<double> W;  // f-f interaction (1-D lattices)
    bool operator==(const config_params& rhs) const;
};

struct ed_cache {
    enum status_eval { empty, spectrum, full };
    typedef array_t<double> real_array_t;
    status_eval status = empty;
    real_array_t cached_spectrum, cached_exp, cached_fermi;
    std::vector<double> cached_evecs;  // row-major msize x msize, column k = eigenvector k
    double logZ = 0.0;
    double evec(size_t i, size_t k) const { return cached_evecs[i * cached_spectrum.size() + k]; }
};

struct chebyshev_cache {
    enum status_eval { empty, logz };
    status_eval status = empty;
    double e_max = 0, e_min = 0, a = 0, b = 0;
    std::vector<double> moments;
    double logZ = 0.0;
};

struct configuration_t {
    typedef fk::sparse_m sparse_m;
    typedef ed_cache::real_array_t real_array_t;
    typedef array_t<int> int_array_t;

    configuration_t(const abstract_lattice& lattice, double beta, double U, double mu_c, double mu_f,
                    std::vector<double> W = {});
    configuration_t(const configuration_t& rhs) = default;
    configuration_t& operator=(const configuration_t& rhs);  // throws on mismatched parameters

    size_t get_nf() const;
    void randomize_f(random_generator& rnd, size_t nf = 0);
    const sparse_m& calc_hamiltonian();
    void reset_cache() { ed_data_.status = ed_cache::empty; cheb_data_.status = chebyshev_cache::empty; }
    void calc_ed(bool calc_evecs = false);
    void calc_chebyshev(const chebyshev::chebyshev_eval& cheb);
    double calc_ff_energy() const;

    const config_params& params() const { return params_; }
    const ed_cache& ed_data() const { return ed_data_; }
    const chebyshev_cache& cheb_data() const { return cheb_data_; }

    const abstract_lattice& lattice_;
    const config_params params_;
    int_array_t f_config_;
    sparse_m hamilt_;
    ed_cache ed_data_;
    chebyshev_cache cheb_data_;
};

}  // namespace fk
