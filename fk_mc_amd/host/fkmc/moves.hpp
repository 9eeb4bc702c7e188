// moves.hpp -- Metropolis moves with the reference's interface
// (include/fk_mc/moves.hpp, src/moves.cpp; include/fk_mc/moves_chebyshev.hpp,
// src/moves_chebyshev.cpp): attempt() returns the weight, accept() copies the
// proposal into the shared configuration, reject() does nothing.  Each move also
// records the sites it touched (last_a, last_b) for tracing.
#pragma once
#include <cmath>
#include <random>

#include "configuration.hpp"

namespace fk {

struct move_flip {  // global exchange of one f-electron (src/moves.cpp:5-21)
    typedef double mc_weight_type;
    double beta;
    configuration_t& config;
    configuration_t new_config;
    random_generator& RND;
    int last_a = -1, last_b = -1;
    move_flip(double beta, configuration_t& current_config, random_generator& RND_)
        : beta(beta), config(current_config), new_config(current_config), RND(RND_) {}
    mc_weight_type attempt();
    mc_weight_type accept() { config = new_config; return 1.0; }
    void reject() {}
};

struct move_randomize : move_flip {  // src/moves.cpp:35-49 (with its clamp of the weight)
    move_randomize(double beta, configuration_t& c, random_generator& r) : move_flip(beta, c, r) {}
    mc_weight_type attempt();
};

struct move_addremove : move_flip {  // single-site f flip, src/moves.cpp:52-67
    double exp_beta_mu_f;
    move_addremove(double beta, configuration_t& c, random_generator& r)
        : move_flip(beta, c, r), exp_beta_mu_f(std::exp(beta * c.params_.mu_f)) {}
    mc_weight_type attempt();
};

// Nearest-neighbour exchange (no reference counterpart): site i uniform, neighbour
// slot uniform over hop_neighbors(i); occupations swapped when they differ,
// otherwise a null move of weight 1.  Symmetric proposal, so w = e^{dlogZ - beta dE_ff}.
struct move_hop : move_flip {
    move_hop(double beta, configuration_t& c, random_generator& r) : move_flip(beta, c, r) {}
    mc_weight_type attempt();
};

namespace chebyshev {

// Chebyshev-weighted moves (src/moves_chebyshev.cpp:6-72): logZ from calc_chebyshev.
struct move_flip {
    typedef double mc_weight_type;
    double beta;
    configuration_t& config;
    configuration_t new_config;
    const chebyshev_eval& cheb_;
    random_generator& RND;
    int last_a = -1, last_b = -1;
    move_flip(double beta, configuration_t& current_config, const chebyshev_eval& cheb, random_generator& RND_)
        : beta(beta), config(current_config), new_config(current_config), cheb_(cheb), RND(RND_) {}
    mc_weight_type attempt();
    mc_weight_type accept() { config = new_config; return 1.0; }
    void reject() {}
};
struct move_randomize : move_flip {
    move_randomize(double beta, configuration_t& c, const chebyshev_eval& ch, random_generator& r)
        : move_flip(beta, c, ch, r) {}
    mc_weight_type attempt();
};
struct move_addremove : move_flip {
    double exp_beta_mu_f;
    move_addremove(double beta, configuration_t& c, const chebyshev_eval& ch, random_generator& r)
        : move_flip(beta, c, ch, r), exp_beta_mu_f(std::exp(beta * c.params_.mu_f)) {}
    mc_weight_type attempt();
};

}  // namespace chebyshev
}  // namespace fk
