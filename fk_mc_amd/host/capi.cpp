// capi.cpp -- include/fkmc_host.h: the reference program's set-up
// (include/fk_mc/fk_mc.hxx:40-140) behind C entry points.
#include "fkmc_host.h"

#include <cstring>
#include <memory>
#include <string>

#include "binning.hpp"
#include "fkmc/mc_metropolis.hpp"
#include "fkmc/measures.hpp"
#include "fkmc/moves.hpp"

namespace {

thread_local std::string g_err;

struct model_lattice {
    std::unique_ptr<fk::abstract_lattice> lat;
};

model_lattice make_lattice(const fkmc_model& m) {
    model_lattice out;
    if (m.lattice_kind == FKGPU_LATTICE_TRIANGULAR) {
        if (m.D != 2) throw std::logic_error("triangular lattice needs D = 2");
        auto l = std::make_unique<fk::hypercubic_lattice<2>>(m.L);
        fk::fill_triangular(*l, m.t, m.tp);
        out.lat = std::move(l);
    } else if (m.lattice_kind == FKGPU_LATTICE_HYPERCUBIC) {
        if (m.D == 1) { auto l = std::make_unique<fk::hypercubic_lattice<1>>(m.L); fk::fill_nearest_neighbors(*l, m.t); out.lat = std::move(l); }
        else if (m.D == 2) { auto l = std::make_unique<fk::hypercubic_lattice<2>>(m.L); fk::fill_nearest_neighbors(*l, m.t); out.lat = std::move(l); }
        else if (m.D == 3) { auto l = std::make_unique<fk::hypercubic_lattice<3>>(m.L); fk::fill_nearest_neighbors(*l, m.t); out.lat = std::move(l); }
        else throw std::logic_error("hypercubic lattice needs D in 1..3");
    } else {
        throw std::logic_error("unknown lattice kind");
    }
    if (m.L < 3) throw std::logic_error("L >= 3 required");
    return out;
}

std::vector<double> model_W(const fkmc_model& m) {
    return m.nW > 0 && m.W ? std::vector<double>(m.W, m.W + m.nW) : std::vector<double>();
}

// add_move calls of fk_mc.hxx:60-80 for the requested kinds.
void register_moves(fk::mc_metropolis& mc, const fkmc_model& m, fk::configuration_t& config,
                    const fk::chebyshev::chebyshev_eval* cheb, int nkinds, const int* kinds, const double* probs) {
    for (int k = 0; k < nkinds; ++k) {
        const double p = probs ? probs[k] : 1.0;
        auto& rng = mc.rng();
        if (cheb) {
            switch (kinds[k]) {
                case FKGPU_MOVE_ADDREMOVE: mc.add_move(fk::chebyshev::move_addremove(m.beta, config, *cheb, rng), "addremove", p); break;
                case FKGPU_MOVE_FLIP: mc.add_move(fk::chebyshev::move_flip(m.beta, config, *cheb, rng), "flip", p); break;
                case FKGPU_MOVE_RANDOMIZE: mc.add_move(fk::chebyshev::move_randomize(m.beta, config, *cheb, rng), "reshuffle", p); break;
                default: throw std::logic_error("move kind not available with Chebyshev weights");
            }
        } else {
            switch (kinds[k]) {
                case FKGPU_MOVE_ADDREMOVE: mc.add_move(fk::move_addremove(m.beta, config, rng), "addremove", p); break;
                case FKGPU_MOVE_FLIP: mc.add_move(fk::move_flip(m.beta, config, rng), "flip", p); break;
                case FKGPU_MOVE_HOP: mc.add_move(fk::move_hop(m.beta, config, rng), "hop", p); break;
                case FKGPU_MOVE_RANDOMIZE: mc.add_move(fk::move_randomize(m.beta, config, rng), "reshuffle", p); break;
                default: throw std::logic_error("unknown move kind");
            }
        }
    }
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return FKGPU_EINVAL;
    }
}

}  // namespace

extern "C" {

const char* fkmc_host_last_error(void) { return g_err.c_str(); }

int fkmc_run_chain(const fkmc_model* m, uint32_t seed, const int32_t* f_init, int nkinds, const int* move_kinds,
                   const double* move_probs, int nmoves, int* tr_kind, int* tr_a, int* tr_b, int* tr_acc,
                   double* tr_w, double* tr_u, double* tr_logz, int32_t* f_final, double* spectrum_final) {
    return guarded([&]() -> int {
        if (!m || !f_init || nkinds <= 0 || !move_kinds || nmoves < 0) throw std::logic_error("invalid arguments");
        model_lattice ml = make_lattice(*m);
        const size_t n = ml.lat->msize();
        fk::configuration_t config(*ml.lat, m->beta, m->U, m->mu_c, m->mu_f, model_W(*m));
        std::memcpy(config.f_config_.data(), f_init, sizeof(int32_t) * n);
        std::unique_ptr<fk::chebyshev::chebyshev_eval> cheb;
        if (m->cheb_size > 0) cheb = std::make_unique<fk::chebyshev::chebyshev_eval>(m->cheb_size, m->cheb_grid);
        fk::mc_metropolis mc((long)seed, 0, nmoves, 1, 0);
        register_moves(mc, *m, config, cheb.get(), nkinds, move_kinds, move_probs);
        config.calc_hamiltonian();
        long s = 0;
        mc.set_tracer([&](const fk::mc_metropolis::move_record& r) {
            if (tr_kind) tr_kind[s] = r.move_index;
            if (tr_a) tr_a[s] = r.site_a;
            if (tr_b) tr_b[s] = r.site_b;
            if (tr_acc) tr_acc[s] = r.accepted;
            if (tr_w) tr_w[s] = r.weight;
            if (tr_u) tr_u[s] = r.u;
            if (tr_logz) {
                if (cheb) { config.calc_chebyshev(*cheb); tr_logz[s] = config.cheb_data_.logZ; }
                else { config.calc_ed(false); tr_logz[s] = config.ed_data_.logZ; }
            }
            ++s;
        });
        for (int i = 0; i < nmoves; ++i) mc.update();
        if (f_final) std::memcpy(f_final, config.f_config_.data(), sizeof(int32_t) * n);
        if (spectrum_final) {
            config.calc_ed(false);
            std::memcpy(spectrum_final, config.ed_data_.cached_spectrum.data(), sizeof(double) * n);
        }
        return (int)mc.naccepted();
    });
}

int fkmc_run_mc(const fkmc_model* m, uint32_t seed, const int32_t* f_init, int nkinds, const int* move_kinds,
                const double* move_probs, int nsweeps, int sweep_len, int ntherm, double* energies,
                double* d2energies, double* c_energies, double* avg_spectrum, double* focc, double* acceptance) {
    return guarded([&]() -> int {
        if (!m || nkinds <= 0 || !move_kinds || nsweeps <= 0 || sweep_len <= 0 || ntherm < 0)
            throw std::logic_error("invalid arguments");
        model_lattice ml = make_lattice(*m);
        const size_t n = ml.lat->msize();
        fk::configuration_t config(*ml.lat, m->beta, m->U, m->mu_c, m->mu_f, model_W(*m));
        std::unique_ptr<fk::chebyshev::chebyshev_eval> cheb;
        if (m->cheb_size > 0) cheb = std::make_unique<fk::chebyshev::chebyshev_eval>(m->cheb_size, m->cheb_grid);
        fk::mc_metropolis mc((long)seed, 0, nsweeps, sweep_len, ntherm);
        // fk_mc.hxx:50-53: random initial f-configuration, then H
        if (f_init) std::memcpy(config.f_config_.data(), f_init, sizeof(int32_t) * n);
        else config.randomize_f(mc.rng(), (size_t)(n / 2));
        config.calc_hamiltonian();
        register_moves(mc, *m, config, cheb.get(), nkinds, move_kinds, move_probs);
        std::vector<double> e, d2e, ce, spec;
        std::vector<std::vector<double>> fh;
        mc.add_measure(fk::measure_energy(m->beta, config, e, d2e, ce), "energy");
        mc.add_measure(fk::measure_spectrum(config, spec), "spectrum");
        mc.add_measure(fk::measure_focc(config, fh), "focc");
        mc.run();
        for (int i = 0; i < nsweeps; ++i) {
            if (energies) energies[i] = e.at(i);
            if (d2energies) d2energies[i] = d2e.at(i);
            if (c_energies) c_energies[i] = ce.at(i);
            if (focc)
                for (size_t j = 0; j < n; ++j) focc[(size_t)i * n + j] = fh[j].at(i);
        }
        if (avg_spectrum) std::memcpy(avg_spectrum, spec.data(), sizeof(double) * n);
        if (acceptance) *acceptance = mc.acceptance_rate();
        return 0;
    });
}

int fkmc_measure_config(const fkmc_model* m, const int32_t* f, double* energy, double* spectrum, double* ipr,
                        double* evecs, int nw, const double* wgrid, double offset, double* stiffness, double* cond,
                        double eta, double* dos) {
    return guarded([&]() -> int {
        if (!m || !f || nw < 0 || (nw > 0 && !wgrid)) throw std::logic_error("invalid arguments");
        model_lattice ml = make_lattice(*m);
        const size_t n = ml.lat->msize();
        fk::configuration_t config(*ml.lat, m->beta, m->U, m->mu_c, m->mu_f, model_W(*m));
        std::memcpy(config.f_config_.data(), f, sizeof(int32_t) * n);
        config.calc_hamiltonian();
        if (energy) {
            std::vector<double> e, d2e, ce;
            fk::measure_energy me(m->beta, config, e, d2e, ce);
            me.accumulate(1.0);
            energy[0] = e[0];
            energy[1] = d2e[0];
            energy[2] = ce[0];
        }
        if (spectrum) {
            std::vector<double> s;
            fk::measure_spectrum ms(config, s);
            ms.accumulate(1.0);
            std::memcpy(spectrum, s.data(), sizeof(double) * n);
        }
        if (ipr) {
            std::vector<std::vector<double>> iv;
            fk::measure_ipr<fk::abstract_lattice> mi(config, *ml.lat, iv);
            mi.accumulate(1.0);
            for (size_t k = 0; k < n; ++k) ipr[k] = iv[k].at(0);
        }
        if (evecs) {
            std::vector<std::vector<double>> ev;
            fk::measure_eigenfunctions me(config, ev);
            me.accumulate(1.0);
            std::memcpy(evecs, ev.at(0).data(), sizeof(double) * n * n);
        }
        if (stiffness) {
            if (m->lattice_kind != FKGPU_LATTICE_HYPERCUBIC) throw std::logic_error("stiffness needs a hypercubic lattice");
            std::vector<double> sv;
            std::vector<std::vector<double>> ch;
            std::vector<double> wg(wgrid, wgrid + nw);
            fk::measure_stiffness<fk::abstract_lattice> ms(config, *ml.lat, sv, ch, wg, offset);
            ms.accumulate(1.0);
            stiffness[0] = sv.at(0);
            if (cond)
                for (int w = 0; w < nw; ++w) cond[w] = ch[w].at(0);
        }
        if (dos && nw > 0) {
            config.calc_ed(false);
            std::vector<std::vector<double>> sp{std::vector<double>(config.ed_data_.cached_spectrum.begin(),
                                                                    config.ed_data_.cached_spectrum.end())};
            auto d = fk::density_of_states(*ml.lat, sp, std::vector<double>(wgrid, wgrid + nw), eta);
            std::memcpy(dos, d[0].data(), sizeof(double) * nw);
        }
        return 0;
    });
}

int fkmc_binning(int n, const double* x, int depth, double* table) {
    if (n < 1 || !x || depth < 0 || depth > 30 || !table) {
        g_err = "invalid binning arguments";
        return FKGPU_EINVAL;
    }
    namespace b = fkmc::binning;
    const auto st = b::accumulate(std::vector<double>(x, x + n), depth);
    const auto cl = b::cor_length(st);
    for (size_t i = 0; i < st.size(); ++i) {
        double* r = table + 5 * i;
        r[0] = st[i].size, r[1] = st[i].mean, r[2] = st[i].disp, r[3] = st[i].sqerror, r[4] = cl[i];
    }
    return b::estimate_bin(st);
}

int fkmc_max_bin(int nmeasures) { return nmeasures < 1 ? FKGPU_EINVAL : fkmc::binning::max_bin((size_t)nmeasures); }

int fkmc_jackknife_cv(int n, const double* e, const double* d2e, double beta, int volume, int depth, double* table) {
    if (n < 2 || !e || !d2e || volume < 1 || depth < 0 || depth > 30 || !table) {
        g_err = "invalid jackknife arguments";
        return FKGPU_EINVAL;
    }
    namespace b = fkmc::binning;
    std::vector<double> E(e, e + n), E2(n), D(d2e, d2e + n);
    for (int i = 0; i < n; ++i) E2[i] = e[i] * e[i];
    const auto F = [beta, volume](const std::vector<double>& m) {
        return beta * beta * (m[1] - m[2] - m[0] * m[0]) / volume;
    };
    const auto st = b::accumulate_jackknife(F, {E, E2, D}, depth);
    for (size_t i = 0; i < st.size(); ++i) {
        double* r = table + 4 * i;
        r[0] = st[i].size, r[1] = st[i].mean, r[2] = st[i].disp, r[3] = st[i].sqerror;
    }
    return b::estimate_bin(st);
}

}  // extern "C"
