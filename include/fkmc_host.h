/* fkmc_host.h -- C entry points of the C++ host layer (libfkmc_host.so).
 *
 * The host layer keeps the reference's class interfaces (configuration_t,
 * move_*, mc_metropolis, measure_*) and routes calc_ed / calc_chebyshev and the
 * eigenvector measures to libfkgpu.so.  These functions drive that layer as the
 * reference's own program would (include/fk_mc/fk_mc.hxx:40-140: lattice,
 * configuration, add_move, add_measure, run) so that a caller without C++ --
 * and the tests -- can run the plain per-move path end to end.
 *
 * All functions return >= 0 on success and a negative FKGPU_E* code on failure;
 * fkmc_host_last_error() then holds the message (thread-local).
 */
#ifndef FKMC_HOST_H
#define FKMC_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model and lattice of one chain (the parameters fk_mc.hxx reads from its ini file). */
typedef struct fkmc_model {
    int lattice_kind; /* FKGPU_LATTICE_HYPERCUBIC (D = 1..3) or FKGPU_LATTICE_TRIANGULAR (D = 2) */
    int D, L;
    double t, tp;
    double beta, U, mu_c, mu_f;
    int nW;           /* f-f interaction W[0..nW) (1-D lattices; calc_ff_energy) */
    const double* W;
    int cheb_size;    /* > 0: Chebyshev moves (moves_chebyshev.cpp) with this many moments */
    int cheb_grid;    /* quadrature grid of chebyshev_eval */
} fkmc_model;

const char* fkmc_host_last_error(void);

/* One Metropolis chain of nmoves single moves (mc_metropolis with sweep_len 1),
 * seeded `seed` (the reference seeds SEED + rank), from f_init (msize int32).
 * move_kinds: FKGPU_MOVE_* in add_move order, with probabilities move_probs.
 * Per-move trace arrays (nmoves each, nullable): kind index, sites, accepted,
 * weight returned by attempt(), the uniform u, and logZ of the current
 * configuration after the move.  f_final / spectrum_final: msize (nullable;
 * the spectrum is that of calc_ed, also for Chebyshev chains).
 * Returns the number of accepted moves. */
int fkmc_run_chain(const fkmc_model* m, uint32_t seed, const int32_t* f_init, int nkinds, const int* move_kinds,
                   const double* move_probs, int nmoves, int* tr_kind, int* tr_a, int* tr_b, int* tr_acc,
                   double* tr_w, double* tr_u, double* tr_logz, int32_t* f_final, double* spectrum_final);

/* A full run as the reference program drives it: nsweeps measured sweeps after
 * ntherm thermalisation sweeps of sweep_len moves, measures energy, spectrum
 * (running mean) and the f-occupation history.  Outputs: energies / d2energies /
 * c_energies (nsweeps each), avg_spectrum (msize), focc (nsweeps x msize,
 * nullable), *acceptance = accepted / attempted.  f_init NULL: randomize_f with
 * msize/2 f-electrons from the chain's generator (fk_mc.hxx:46, Nf_start). */
int fkmc_run_mc(const fkmc_model* m, uint32_t seed, const int32_t* f_init, int nkinds, const int* move_kinds,
                const double* move_probs, int nsweeps, int sweep_len, int ntherm, double* energies,
                double* d2energies, double* c_energies, double* avg_spectrum, double* focc, double* acceptance);

/* Every measure once on one configuration f: energy[3] = {E, d2E, E_c}
 * (measure_energy), spectrum (msize), ipr (msize, measure_ipr), evecs (msize^2
 * row-major, measure_eigenfunctions; nullable), stiffness[1] and cond[nw]
 * (measure_stiffness; hypercubic lattices, skipped when stiffness is NULL), and
 * dos[nw] of the spectrum with width eta (nullable). */
int fkmc_measure_config(const fkmc_model* m, const int32_t* f, double* energy, double* spectrum, double* ipr,
                        double* evecs, int nw, const double* wgrid, double offset, double* stiffness, double* cond,
                        double eta, double* dos);

/* Binning analysis of a series, in the order given (the reference bins its histories from
 * the last measurement backwards): table[(depth + 1) x 5] = {size, mean, variance, error,
 * correlation length} per level (include/fk_mc/binning.hpp:21-171).  Returns the level
 * estimate_bin picks (prog/data_save.hpp:107-122). */
int fkmc_binning(int n, const double* x, int depth, double* table);
/* max_bin_ of data_saver for nmeasures measurements (prog/data_save.hpp:90). */
int fkmc_max_bin(int nmeasures);
/* Jackknife of the specific heat beta^2 (<E^2> - <d2E> - <E>^2) / volume
 * (prog/data_save.hxx:183-198, include/fk_mc/jackknife.hpp:40-81) on binning levels
 * 0..depth: table[(depth + 1) x 4] = {size, value, variance, error}.  Returns the level
 * estimate_bin picks. */
int fkmc_jackknife_cv(int n, const double* e, const double* d2e, double beta, int volume, int depth, double* table);

#ifdef __cplusplus
}
#endif
#endif
