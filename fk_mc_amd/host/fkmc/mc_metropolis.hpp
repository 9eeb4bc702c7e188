// mc_metropolis.hpp -- the Metropolis engine with the reference's interface
// (include/fk_mc/mc_metropolis.hpp:17-193, src/mc_metropolis.cpp): type-erased
// moves and measures, move choice by std::discrete_distribution over the
// registered probabilities, accept iff |w| > U[0,1), one std::mt19937 seeded
// SEED + rank.  The ALPS mcbase scheduling is replaced by run(): nsweeps +
// ntherm_sweeps calls of update() (sweep_len moves) and measure().
#pragma once
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

namespace fk {

typedef double mc_weight_t;

struct move_wrap {
    template <class M, class = typename std::enable_if<!std::is_same<typename std::decay<M>::type, move_wrap>::value>::type>
    move_wrap(M&& in) {
        typedef typename std::decay<M>::type m_type;
        auto m = std::make_shared<m_type>(std::forward<M>(in));
        ptr_ = m;
        attempt_ = [m]() { return m->attempt(); };
        accept_ = [m]() { return m->accept(); };
        reject_ = [m]() { m->reject(); };
        sites_ = [m](int& a, int& b) { a = m->last_a; b = m->last_b; };
    }
    mc_weight_t attempt() { return attempt_(); }
    mc_weight_t accept() { return accept_(); }
    void reject() { reject_(); }
    std::shared_ptr<void> ptr_;
    std::function<mc_weight_t()> attempt_, accept_;
    std::function<void()> reject_;
    std::function<void(int&, int&)> sites_;
};

struct measure_wrap {
    template <class M, class = typename std::enable_if<!std::is_same<typename std::decay<M>::type, measure_wrap>::value>::type>
    measure_wrap(M&& in) {
        typedef typename std::decay<M>::type m_type;
        auto m = std::make_shared<m_type>(std::forward<M>(in));
        ptr_ = m;
        accumulate_ = [m](mc_weight_t s) { m->accumulate(s); };
    }
    void accumulate(mc_weight_t phase) { accumulate_(phase); }
    template <class M> M& cast() { return *static_cast<M*>(ptr_.get()); }
    std::shared_ptr<void> ptr_;
    std::function<void(mc_weight_t)> accumulate_;
};

struct mc_metropolis {
    typedef std::mt19937 random_generator;
    // One record per move when a tracer is installed (tests and diagnostics).
    struct move_record {
        int move_index, site_a, site_b, accepted;
        double weight, u;
    };

    // Parameter defaults of src/mc_metropolis.cpp:11-19.
    explicit mc_metropolis(long seed, int rank = 0, long nsweeps = 1024, long sweep_len = 16, long ntherm_sweeps = 1)
        : random((unsigned long)(seed + rank)), rank_(rank), measure_sweeps_(nsweeps), sweep_len_(sweep_len),
          thermalization_sweeps_(ntherm_sweeps) {}

    template <class Move_t>
    bool add_move(Move_t&& move, std::string name, double move_prob = 1.0) {
        moves_.emplace_back(std::forward<Move_t>(move));
        move_names_.push_back(name);
        move_probs_.push_back(move_prob);
        move_distrib_ = std::discrete_distribution<>(move_probs_.begin(), move_probs_.end());
        return true;
    }
    template <class Measure_t>
    bool add_measure(Measure_t&& measure, std::string name) {
        measures_.emplace(name, measure_wrap(std::forward<Measure_t>(measure)));
        return true;
    }

    void update();          // src/mc_metropolis.cpp:34-52
    void measure();         // src/mc_metropolis.cpp:54-61
    void run();             // thermalisation + measurement sweeps
    double fraction_completed() const {
        return double(sweep_count_) / double(measure_sweeps_ + thermalization_sweeps_);
    }
    double acceptance_rate() const { return double(naccept_) / double(sweep_count_ * sweep_len_); }
    long naccepted() const { return naccept_; }
    random_generator& rng() { return random; }
    void set_tracer(std::function<void(const move_record&)> t) { tracer_ = std::move(t); }
    template <class M> M& extract_measurement(const std::string& name) { return measures_.at(name).cast<M>(); }

protected:
    std::vector<move_wrap> moves_;
    std::vector<std::string> move_names_;
    std::vector<double> move_probs_;
    std::map<std::string, measure_wrap> measures_;  // alphabetical order of measure() as in the reference
    random_generator random;
    int rank_;
    long measure_sweeps_, sweep_len_, thermalization_sweeps_;
    long sweep_count_ = 0, measure_count_ = 0, naccept_ = 0;
    std::discrete_distribution<> move_distrib_;
    std::uniform_real_distribution<> metropolis_distrib_ = std::uniform_real_distribution<>(0, 1);
    mc_weight_t phase_ = 1.0;
    std::function<void(const move_record&)> tracer_;
};

}  // namespace fk
