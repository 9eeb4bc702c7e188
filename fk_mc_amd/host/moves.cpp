// moves.cpp -- see fkmc/moves.hpp.  The draw order of each attempt() is that of
// src/moves.cpp so that a chain consumes the generator exactly as the reference.
#include "fkmc/moves.hpp"

#include <stdexcept>

namespace fk {

move_flip::mc_weight_type move_flip::attempt() {
    std::uniform_int_distribution<> distr(0, (int)config.lattice_.volume() - 1);
    last_a = last_b = -1;
    if (config.get_nf() == 0 || config.get_nf() == config.lattice_.msize()) return 0;
    new_config = config;
    size_t from = distr(RND);
    while (new_config.f_config_(from) == 0) from = distr(RND);
    size_t to = distr(RND);
    while (new_config.f_config_(to) == 1) to = distr(RND);
    config.calc_ed(false);
    new_config.f_config_(from) = 0;
    new_config.f_config_(to) = 1;
    last_a = (int)from;
    last_b = (int)to;
    new_config.calc_hamiltonian();
    new_config.calc_ed(false);
    return std::exp(new_config.ed_data_.logZ - config.ed_data_.logZ);
}

move_flip::mc_weight_type move_randomize::attempt() {
    new_config = config;
    new_config.randomize_f(RND);
    new_config.calc_hamiltonian();
    config.calc_ed(false);
    new_config.calc_ed(false);
    const double log_ratio = new_config.ed_data_.logZ - config.ed_data_.logZ;
    const double ff_diff = new_config.calc_ff_energy() - config.calc_ff_energy();
    const double dn = (double)new_config.get_nf() - (double)config.get_nf();
    if (beta * config.params_.mu_f * dn - ff_diff > 2.7182818 - log_ratio) return 1;
    if (beta * config.params_.mu_f * dn - ff_diff + log_ratio < 0) return 0;
    return std::exp(log_ratio) * std::exp(beta * (config.params_.mu_f * dn - ff_diff));
}

move_flip::mc_weight_type move_addremove::attempt() {
    std::uniform_int_distribution<> distr(0, (int)config.lattice_.volume() - 1);
    new_config = config;
    const size_t to = distr(RND);
    new_config.f_config_(to) = 1 - config.f_config_(to);
    last_a = (int)to;
    last_b = -1;
    config.calc_ed(false);
    new_config.calc_hamiltonian();
    new_config.calc_ed(false);
    const double ff_diff = new_config.calc_ff_energy() - config.calc_ff_energy();
    const double ratio = std::exp(new_config.ed_data_.logZ - config.ed_data_.logZ);
    return (new_config.f_config_(to) ? ratio * exp_beta_mu_f : ratio / exp_beta_mu_f) * std::exp(-beta * ff_diff);
}

move_flip::mc_weight_type move_hop::attempt() {
    const int n = (int)config.lattice_.volume();
    std::uniform_int_distribution<> distr(0, n - 1);
    const int i = distr(RND);
    const std::vector<size_t> nb = config.lattice_.hop_neighbors(i);
    if (nb.empty()) throw std::logic_error("nearest-neighbour exchange from an isolated site");
    std::uniform_int_distribution<> dn(0, (int)nb.size() - 1);
    const int j = (int)nb[dn(RND)];
    last_a = i;
    last_b = j;
    new_config = config;
    if (config.f_config_(i) == config.f_config_(j)) return 1.0;
    std::swap(new_config.f_config_(i), new_config.f_config_(j));
    config.calc_ed(false);
    new_config.calc_hamiltonian();
    new_config.calc_ed(false);
    const double ff_diff = new_config.calc_ff_energy() - config.calc_ff_energy();
    return std::exp(new_config.ed_data_.logZ - config.ed_data_.logZ) * std::exp(-beta * ff_diff);
}

namespace chebyshev {

move_flip::mc_weight_type move_flip::attempt() {
    config.calc_chebyshev(cheb_);
    last_a = last_b = -1;
    if (config.get_nf() == 0 || config.get_nf() == config.lattice_.msize()) return 0;
    std::uniform_int_distribution<> distr(0, (int)config.lattice_.msize() - 1);
    new_config = config;
    size_t from = distr(RND);
    while (new_config.f_config_(from) == 0) from = distr(RND);
    size_t to = distr(RND);
    while (new_config.f_config_(to) == 1) to = distr(RND);
    new_config.f_config_(from) = 0;
    new_config.f_config_(to) = 1;
    last_a = (int)from;
    last_b = (int)to;
    new_config.calc_hamiltonian();
    new_config.calc_chebyshev(cheb_);
    const double ff_diff = new_config.calc_ff_energy() - config.calc_ff_energy();
    return std::exp(new_config.cheb_data_.logZ - config.cheb_data_.logZ - beta * ff_diff);
}

move_flip::mc_weight_type move_randomize::attempt() {
    config.calc_chebyshev(cheb_);
    new_config = config;
    new_config.randomize_f(RND);
    new_config.calc_hamiltonian();
    new_config.calc_chebyshev(cheb_);
    const double log_ratio = new_config.cheb_data_.logZ - config.cheb_data_.logZ;
    const double ff_diff = new_config.calc_ff_energy() - config.calc_ff_energy();
    const double dn = (double)new_config.get_nf() - (double)config.get_nf();
    if (beta * config.params_.mu_f * dn - ff_diff > 2.7182818 - log_ratio) return 1;
    if (beta * config.params_.mu_f * dn - ff_diff + log_ratio < 0) return 0;
    return std::exp(log_ratio) * std::exp(beta * (config.params_.mu_f * dn - ff_diff));
}

move_flip::mc_weight_type move_addremove::attempt() {
    std::uniform_int_distribution<> distr(0, (int)config.lattice_.msize() - 1);
    config.calc_chebyshev(cheb_);
    new_config = config;
    const size_t to = distr(RND);
    new_config.f_config_(to) = 1 - config.f_config_(to);
    last_a = (int)to;
    last_b = -1;
    new_config.calc_hamiltonian();
    new_config.calc_chebyshev(cheb_);
    const double ff_diff = new_config.calc_ff_energy() - config.calc_ff_energy();
    const double ratio = std::exp(new_config.cheb_data_.logZ - config.cheb_data_.logZ);
    return (new_config.f_config_(to) ? ratio * exp_beta_mu_f : ratio / exp_beta_mu_f) * std::exp(-beta * ff_diff);
}

}  // namespace chebyshev
}  // namespace fk
