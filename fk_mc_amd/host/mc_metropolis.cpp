// mc_metropolis.cpp -- see fkmc/mc_metropolis.hpp.
#include "fkmc/mc_metropolis.hpp"

#include <cmath>
#include <stdexcept>

namespace fk {

void mc_metropolis::update() {
    if (moves_.empty()) throw std::logic_error("No registered moves");
    for (long m = 0; m < sweep_len_; ++m) {
        const int move_index = move_distrib_(random);
        mc_weight_t weight = moves_[move_index].attempt();
        const double u = metropolis_distrib_(random);
        const bool acc = std::abs(weight) > u;
        if (acc) {
            weight *= moves_[move_index].accept();
            ++naccept_;
            phase_ *= (weight > 0) - (weight < 0);
        } else {
            moves_[move_index].reject();
        }
        if (tracer_) {
            move_record r{move_index, -1, -1, acc ? 1 : 0, weight, u};
            moves_[move_index].sites_(r.site_a, r.site_b);
            tracer_(r);
        }
    }
    ++sweep_count_;
}

void mc_metropolis::measure() {
    if (measure_count_ >= thermalization_sweeps_)
        for (auto& m : measures_) m.second.accumulate(phase_);
    ++measure_count_;
}

void mc_metropolis::run() {
    while (fraction_completed() < 1.0) {
        update();
        measure();
    }
}

}  // namespace fk
