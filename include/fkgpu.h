/* fkgpu.h -- C ABI of the MI355X per-move diagonalisation path of fk_mc.
 *
 * The reference has no FFI: its operator API for this path is the C++ surface of
 * fk::configuration_t (include/fk_mc/configuration.hpp:52-87) consumed by the
 * moves (src/moves.cpp), the Metropolis engine (src/mc_metropolis.cpp) and the
 * measures (include/fk_mc/measures/).  Each entry point below replaces one piece
 * of that surface; the C++ host layer in fk_mc_amd/host/ keeps the reference's
 * class interfaces and calls these functions.
 *
 * Conventions: every function returns an int status (FKGPU_OK = 0, negative on
 * error; fkgpu_last_error() holds the message).  No C++ exception crosses the
 * ABI.  Pointers named d_* are device pointers; all others are host pointers.
 * Host-pointer calls are synchronous; device-pointer calls are ordered on the
 * context's HIP stream (fkgpu_stream()).  One context per host thread.
 * Occupations f are int32 0/1, one per lattice site, site order as
 * hypercubic_lattice::index_to_pos (last coordinate fastest).
 * Matrices are row-major; eigenvectors are stored as columns (evecs[i*N + k] is
 * component i of eigenvector k), eigenvalues ascending.
 */
#ifndef FKGPU_H
#define FKGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FKGPU_OK 0
#define FKGPU_EINVAL (-1)       /* invalid argument: null pointer, size, NaN/Inf parameter */
#define FKGPU_ENOLATTICE (-2)   /* no lattice set on the context */
#define FKGPU_EUNSUPPORTED (-3) /* size or option outside what the library implements */
#define FKGPU_EHIP (-4)         /* HIP runtime error */
#define FKGPU_ENOMEM (-5)       /* device allocation failed */

#define FKGPU_LATTICE_HYPERCUBIC 0 /* fill_nearest_neighbors<D>, src/lattice/hypercubic.cpp:115-131 */
#define FKGPU_LATTICE_TRIANGULAR 1 /* fill_triangular, src/lattice/hypercubic.cpp:137-155 */

#define FKGPU_MOVE_ADDREMOVE 0 /* move_addremove, src/moves.cpp:52-67 (single-site f flip) */
#define FKGPU_MOVE_FLIP 1      /* move_flip, src/moves.cpp:5-21 (global exchange)       */
#define FKGPU_MOVE_HOP 2       /* nearest-neighbour exchange; no reference counterpart   */
#define FKGPU_MOVE_RANDOMIZE 3 /* move_randomize, src/moves.cpp:35-49                   */

typedef struct fkgpu_ctx fkgpu_ctx;
typedef struct fkgpu_chains fkgpu_chains;

/* Context on one HIP device.  Replaces nothing in the reference (which has no
 * device); owns the lattice, the stream and the workspaces. */
int fkgpu_create(int device, fkgpu_ctx** out);
void fkgpu_destroy(fkgpu_ctx* ctx);
const char* fkgpu_last_error(const fkgpu_ctx* ctx);
void* fkgpu_stream(fkgpu_ctx* ctx); /* hipStream_t */

/* Lattice hopping structure.  Replaces hypercubic_lattice<D>(L) +
 * fill_nearest_neighbors(l, t) (include/fk_mc/lattice/hypercubic.hpp:136,166;
 * src/lattice/hypercubic.cpp:115-131) and fill_triangular(l, t, tp) (:137-155):
 * H_ij = -t for nearest neighbours, -tp on the triangular diagonal. L >= 3, at most
 * 16384 sites (the dense eigensolvers and the chains need N <= 2048 and N <= 4096). */
int fkgpu_set_lattice(fkgpu_ctx* ctx, int kind, int ndim, int L, double t, double tp);
/* Arbitrary symmetric hopping matrix in CSR form (abstract_lattice::set_hopping_m,
 * include/fk_mc/lattice.hpp:39; graph lattices). */
int fkgpu_set_lattice_csr(fkgpu_ctx* ctx, int n, const int32_t* row_ptr, const int32_t* col, const double* val);
int fkgpu_msize(const fkgpu_ctx* ctx);

/* Dense eigensolver variants (no reference counterpart; the reference always uses
 * Eigen's SelfAdjointEigenSolver).  All are FP64 and give the same eigenvalues to
 * rounding; they differ in how the Householder tridiagonalisation is mapped on the GPU:
 *   FKGPU_SOLVER_AUTO       by matrix size (or the fkgpu_autotune choice)
 *   0, 1, 2                 register-resident one-stage, P x P threads holding NB x NB
 *                           blocks: (P, NB) = (16, 4) N <= 64, (16, 8) N <= 128, (32, 8) N <= 256
 *   3                       unblocked, matrix in an HBM workspace, N <= 2048
 *   4                       blocked (panel + MFMA rank-2nb trailing update), HBM workspace, N <= 1024
 *   5                       register-resident one-stage, 16 waves x 64 lanes (wave = row residue,
 *                           lane = column residue; row values by v_readlane), N <= 256
 *   6                       register-resident one-stage, one wave (lane = column), N <= 64
 *   7, 8, 9                 variant 4 in other tile shapes: 4 and 16 waves per matrix (256 and
 *                           1024 threads), and panels of at most 16 columns, N <= 1024;
 *                           8 is the default for 256 < N <= 1024
 *   10                      variant 5 with 8 waves per matrix (80 doubles a lane), N <= 256;
 *                           the default for 128 < N <= 256 */
#define FKGPU_SOLVER_AUTO (-1)
#define FKGPU_SOLVER_VARIANTS 11
int fkgpu_set_solver(fkgpu_ctx* ctx, int variant);
int fkgpu_get_solver(const fkgpu_ctx* ctx); /* the variant used for the current lattice */
/* Autotuner: times every variant that fits the lattice's N on `batch` synthetic
 * configurations (one warm-up and three timed launches each), keeps the fastest for
 * this context and returns it; ms[v] (FKGPU_SOLVER_VARIANTS entries, nullable) gets
 * the time per launch, or -1 where the variant does not fit. */
int fkgpu_autotune(fkgpu_ctx* ctx, int batch, double* ms);

/* Mixed-precision calc_ed (eigenvalues only; no reference counterpart, the reference
 * is FP64 throughout): FP32 Householder reduction with MFMA trailing updates, FP64
 * tridiagonal eigenvalues and eigenvectors, FP64 Rayleigh-Ritz refinement of the
 * eigenvalues (second-order perturbation with near-degenerate groups diagonalised
 * exactly), compensated FP64 log-sum.  Configurations the refinement cannot certify
 * are re-solved in FP64; *nfallback (nullable) counts them.  N <= 1024. */
int fkgpu_eig_logz_mixed(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                         const double* mu_c, double* evals, double* logZ, int64_t* nfallback);

/* configuration_t::calc_hamiltonian() + calc_ed(bool calc_evecs) for `batch`
 * configurations (src/configuration.cpp:79-91, 208-246).  Outputs per config:
 * evals[N] (ed_cache::cached_spectrum), logZ (ed_cache::logZ), cached_exp[N],
 * cached_fermi[N] (nullable), evecs[N*N] (ed_cache::cached_evecs, only when
 * want_vectors != 0).  Parameters are per configuration. */
int fkgpu_eig_logz(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                   const double* mu_c, int want_vectors, double* evals, double* evecs, double* logZ,
                   double* cached_exp, double* cached_fermi);
/* Same, device pointers, asynchronous on the context stream (eigenvalues only). */
int fkgpu_eig_logz_device(fkgpu_ctx* ctx, int batch, const int32_t* d_f, const double* d_beta, const double* d_U,
                          const double* d_mu_c, double* d_evals, double* d_logZ);

/* Same as fkgpu_eig_logz_device, allocating nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The HBM-workspace solvers (N > 256)
 * work in a static device array of 256 MiB per GPU (held from library load to process
 * end) instead of a hipMalloc'd buffer; a batch whose workspace exceeds it, N x N
 * doubles per configuration, is refused with FKGPU_EUNSUPPORTED and a message naming
 * the largest batch at that N.  All calls on one GPU share that array: calls that
 * overlap in time on one GPU (other streams, other contexts, other threads) race; calls
 * ordered on one stream do not. */
int fkgpu_eig_logz_device_noalloc(fkgpu_ctx* ctx, int batch, const int32_t* d_f, const double* d_beta,
                                  const double* d_U, const double* d_mu_c, double* d_evals, double* d_logZ);

/* ---- Eigenvector measures and the adjoint (eigenvectors stay on the device) ----
 * Each call runs calc_hamiltonian + calc_ed(true) for `batch` configurations and
 * then the measure, without copying the N x N eigenvector matrices to the host. */

/* measure_ipr::accumulate (include/fk_mc/measures/ipr.hpp:39-56):
 * ipr[b*N + k] = ||v_k||_4 / ||v_k||_2^2 (the reference's lpNorm<4> form). */
int fkgpu_measure_ipr(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                      const double* mu_c, double* evals, double* ipr);
/* measure_stiffness::accumulate (include/fk_mc/measures/stiffness.hpp:129-187), hypercubic
 * lattices, direction 0, T_x hopping 1 as in the reference: stiffness[b] = (T + V)/N and
 * cond[b*nw + w] = the Lorentzian conductivity sum at wgrid[w] with width `offset`. */
int fkgpu_measure_stiffness(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                            const double* mu_c, int nw, const double* wgrid, double offset, double* stiffness,
                            double* cond);
/* DOS of spectra (prog/data_save.hxx:264-345, dos0_f): dos[s*nw + w] =
 * (1/N) sum_k (eta/pi) / ((wgrid[w] - spectra[s*N + k])^2 + eta^2); N = fkgpu_msize. */
int fkgpu_dos(fkgpu_ctx* ctx, int nspectra, const double* spectra, int nw, const double* wgrid, double eta,
              double* dos);
/* ipr_f of data_saver::save_ipr (prog/data_save.hxx:490-501) per spectrum and frequency:
 * out[s*nw + w] = sum_k L_k ipr[s*N + k]^4 / sum_k L_k, L_k = eta / ((wgrid[w] - e_k)^2 + eta^2),
 * with ipr as the IPR measure records it (the 4th root of sum_i v_i^4). */
int fkgpu_ipr_w(fkgpu_ctx* ctx, int nspectra, const double* spectra, const double* ipr, int nw, const double* wgrid,
                double eta, double* out);
/* Adjoint of Omega = -logZ/beta (no reference counterpart; the closest reference data are the
 * eigenvector densities sum_k f_k V_ik^2 from ed_cache::cached_evecs / cached_fermi,
 * include/fk_mc/configuration.hpp:28-31).  dOmega/dH = rho = V f(Lambda) V^T.
 * Outputs: logZ[b]; dens[b*N + i] = rho_ii = dOmega/d eps_i (local conduction density);
 * grads[b*8 + {0..5}] = {dOmega/dU, dOmega/dt, dOmega/dtp, dOmega/dmu_c, N_c, dN_c/dmu_c};
 * rho[b*N*N] (nullable) = the full matrix V f(Lambda) V^T. */
int fkgpu_adjoint(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                  const double* mu_c, double* logZ, double* dens, double* grads, double* rho);

/* mu_c tuned to a conduction filling (no reference counterpart): Newton's method on
 * N_c(mu_c) with N_c and dN_c/dmu_c from fkgpu_adjoint (dOmega/dmu_c = -N_c), bisection
 * when a step leaves the bracket.  mu_c[batch] in: starting values, out: tuned values with
 * |N_c/N - filling| <= tol; iters[b] (nullable) = adjoint evaluations used, negative if
 * maxit was reached first. */
int fkgpu_tune_mu(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U, double* mu_c,
                  double filling, double tol, int maxit, int32_t* iters);

/* ---- Chebyshev paths ----------------------------------------------------------
 * configuration_t::calc_chebyshev (src/configuration.cpp:94-206) with
 * chebyshev_eval(cheb_size, grid_size) (include/fk_mc/chebyshev.hpp:15-63): exact
 * trace moments mu_m = Tr T_m(X)/N of X = (H - b)/a, a = (e_max - e_min)/2,
 * b = (e_max + e_min)/2 from the extreme eigenvalues (ARPACK in the reference; a device
 * Lanczos here, run until both extreme Ritz values are stable to 4 eps), two N x N
 * matrices in HBM (N <= 16384), the reference's doubling identities for the upper
 * half of the moments, and logZ = g_0 + 2 sum_m g_m mu_m with g_m the trapezoid
 * quadrature of N log(1 + e^{-beta(a w + b)}) on the Lobatto grid.  No kernel damping.
 * cheb_size is rounded up to even as in the reference.  Outputs: moments[cheb_size],
 * logZ, ab[4] = {a, b, e_min, e_max}. */
int fkgpu_cheb_logz(fkgpu_ctx* ctx, const int32_t* f, double beta, double U, double mu_c, int cheb_size,
                    int grid_size, double* moments, double* logZ, double* ab);
/* Stochastic-trace Chebyshev (kernel polynomial method; no reference counterpart, the
 * reference's path is the exact trace above): R Rademacher vectors from `seed` (the same
 * vectors for every configuration), M moments (even, <= 512), FP32 recursion with FP64
 * moment accumulation, Jackson damping when jackson != 0, spectral bounds from the
 * Gershgorin discs widened by 1%; N <= 4096, at most 8 neighbours per site.
 * Outputs logZ[b], moments[b*M + m] (nullable), ab[b*2 + {0,1}] = {a, b} (nullable). */
int fkgpu_kpm_logz(fkgpu_ctx* ctx, int batch, const int32_t* f, const double* beta, const double* U,
                   const double* mu_c, int R, int M, uint64_t seed, int jackson, double* logZ, double* moments,
                   double* ab);

/* ---- Batched persistent Metropolis chains ------------------------------------
 * Replaces mc_metropolis::update (src/mc_metropolis.cpp:34-52) with the moves
 * of src/moves.cpp for many independent chains at once; each chain owns its
 * std::mt19937 state (seeded like src/mc_metropolis.cpp:25) and consumes random
 * words in the reference's order, so a chain reproduces the reference chain of
 * the same seed and initial configuration. */
typedef struct {
    int32_t* kind;     /* move kind                                   */
    int32_t* site_a;   /* toggled / source site (-1 if none)         */
    int32_t* site_b;   /* destination site (-1 if none)              */
    double* u;         /* Metropolis uniform                         */
    double* logw;      /* log of the move weight (-inf: zero weight) */
    double* logz_new;  /* logZ of the proposal (0 if not solved)     */
    int8_t* accepted;  /* 1 if accepted                              */
} fkgpu_trace;         /* host arrays of nsteps*nchains, index [step*nchains + chain] */

int fkgpu_chains_create(fkgpu_ctx* ctx, int nchains, const int32_t* f_init, const double* beta, const double* U,
                        const double* mu_c, const double* mu_f, const uint32_t* seeds, fkgpu_chains** out);
void fkgpu_chains_destroy(fkgpu_chains* ch);
/* Chains on several lattices in one chain set (no reference counterpart: the reference runs
 * one lattice per program): chain c lives on the lattice of contexts[lattice[c]] (at most 8
 * contexts, one device, every N <= 1024).  f_init is nchains x Nmax with Nmax the largest N
 * (row c uses its first N of lattice[c] entries); fkgpu_chains_state returns f and spectrum
 * rows in the same layout, zero past each chain's N.  The FP64 dense move steps of all chains
 * run in one persistent launch (k_sweep_q), every chain with its own size and neighbour table;
 * one solver serves all of them (the register-resident wave-row solver when every N <= 256,
 * else the blocked HBM-workspace solver).  The contexts must outlive the chain set (they are
 * held until it is destroyed).  randomize_f, the rank-one, mixed-precision and Chebyshev modes
 * and the on-device measurements are not available on such a set. */
int fkgpu_chains_create_multi(int ncontexts, fkgpu_ctx* const* contexts, int nchains, const int32_t* lattice,
                              const int32_t* f_init, const double* beta, const double* U, const double* mu_c,
                              const double* mu_f, const uint32_t* seeds, fkgpu_chains** out);
/* configuration_t::randomize_f(rng, nf) (src/configuration.cpp:47-56) of every chain,
 * drawn from the chain's own generator, as fk_mc::initialize draws the starting
 * configuration from the Metropolis generator (include/fk_mc/fk_mc.hxx:46): nf_start[c]
 * f-electrons on distinct uniform sites (0: a uniform count first).  Recomputes logZ and
 * the spectrum.  On a new chain set, before any other option is set. */
int fkgpu_chains_randomize_f(fkgpu_chains* ch, const int32_t* nf_start);
/* mc_metropolis::add_move in order (include/fk_mc/mc_metropolis.hpp:176-185):
 * move kinds and their probabilities (up to 4). */
int fkgpu_chains_set_moves(fkgpu_chains* ch, int nmoves, const int32_t* kinds, const double* weights);
/* f-f interaction W of 1-D lattices (configuration_t::calc_ff_energy, src/configuration.cpp:59-77). */
int fkgpu_chains_set_ff(fkgpu_chains* ch, int nW, const double* W);
/* Precision of the chains' dense eigensolve: FKGPU_PRECISION_FP64 (default) or
 * FKGPU_PRECISION_MIXED (the fkgpu_eig_logz_mixed pipeline in three launches per move
 * step: proposal, solve, decision).  Recomputes logZ of the current configurations. */
#define FKGPU_PRECISION_FP64 0
#define FKGPU_PRECISION_MIXED 1
int fkgpu_chains_set_precision(fkgpu_chains* ch, int precision);
/* How a chain's proposal is diagonalised (no reference counterpart; the reference
 * re-diagonalises every proposal):
 *   FKGPU_UPDATE_DENSE  a full FP64 eigensolve per proposed move (default)
 *   FKGPU_UPDATE_RANK1  the eigenvectors are kept on the device; a move of one or two
 *                       sites is a rank-one (two rank-one) update of the spectrum by the
 *                       secular equation, O(N^2) per proposal, and an accepted move
 *                       updates the eigenvectors by one (two) N x N MFMA products.
 *                       Every refresh_steps move steps the eigendecomposition is
 *                       recomputed from scratch (0: never); the steps are counted on the
 *                       device from this call on, across fkgpu_sweep calls and graph
 *                       replays.  Moves: addremove, flip, hop.  N <= 1024.
 * Switching to FKGPU_UPDATE_RANK1 computes the eigendecomposition of the current
 * configurations. */
#define FKGPU_UPDATE_DENSE 0
#define FKGPU_UPDATE_RANK1 1
int fkgpu_chains_set_update(fkgpu_chains* ch, int mode, int refresh_steps);
/* Configurations the mixed-precision refinement handed to the FP64 solver so far. */
int fkgpu_chains_fallbacks(const fkgpu_chains* ch, int64_t* n);
/* Switch the chain set to the Chebyshev path (the moves of src/moves_chebyshev.cpp)
 * with the kernel-polynomial estimate of logZ instead of the exact trace: R probe
 * vectors per chain (fixed for the whole run), M moments, Jackson damping if
 * jackson != 0, spectral bounds valid for every configuration of the chain.  The chains'
 * logZ is estimated from the probes; a proposal's dlogZ then comes from the local
 * estimator below (fkgpu_chains_set_kpm_delta), each move step in four launches
 * (propose, moments, frequencies, accept).  Required for lattices above the dense
 * solver's size (N > 2048). */
int fkgpu_chains_set_kpm(fkgpu_chains* ch, int R, int M, uint64_t seed, int jackson);
/* How the chains on the Chebyshev path weigh a proposal (the dlogZ of
 * src/moves_chebyshev.cpp:6-72; the reference takes it as the difference of two exact
 * traces, src/configuration.cpp:94-206):
 *   FKGPU_KPM_DELTA_LOCAL       (default after fkgpu_chains_set_kpm) the exact dlogZ of a
 *                               move of one or two sites, to truncation error: FP64
 *                               Chebyshev vectors started at the changed sites give
 *                               G_pq(z) = <e_p|(z - H)^-1|e_q> at the Matsubara
 *                               frequencies z = i(2n+1)pi/beta, and dlogZ =
 *                               -beta Tr(Delta)/2 + sum_n log|det(1 - G_P Delta_P)|^2
 *                               with a fitted 1/w^2 + 1/w^4 tail after `frequencies`
 *                               terms.  `moments` Chebyshev moments (0: enough for a
 *                               truncation below 1e-12 at the chains' beta and spectral
 *                               scale; at most 2048), `frequencies` (0: 512).  N <= 4096;
 *                               the randomize move is refused.
 *   FKGPU_KPM_DELTA_STOCHASTIC  the difference of the two configurations' stochastic
 *                               traces with the chain's fixed probe vectors (biased by the
 *                               probe set; every move kind). */
#define FKGPU_KPM_DELTA_LOCAL 0
#define FKGPU_KPM_DELTA_STOCHASTIC 1
int fkgpu_chains_set_kpm_delta(fkgpu_chains* ch, int mode, int moments, int frequencies);
/* The estimator in use (-1 when the chain set is not on the Chebyshev path) and the local
 * estimator's moment and frequency counts (nullable outputs). */
int fkgpu_chains_kpm_delta(const fkgpu_chains* ch, int* mode, int* moments, int* frequencies);
/* Move steps of every chain fused into one launch in the FP64 dense mode (no reference
 * counterpart; default 16, 1 = one launch per step).  One workgroup per (step, chain) item, chains taken from a queue of
 * ready chains; a chain's step starts after its previous step, so every chain draws and
 * decides exactly as with one launch per step, while the solves of all chains and steps
 * fill the CUs without a launch boundary between steps. */
int fkgpu_chains_set_steps_per_launch(fkgpu_chains* ch, int steps);
/* Bound of one wait in that queue (default 30 s; no move step comes near it).  When a
 * wait exceeds it the launch ends early and the sweep call returns FKGPU_EHIP; the chain
 * set's state is then undefined (destroy it). */
int fkgpu_chains_set_queue_timeout(fkgpu_chains* ch, double seconds);
/* Run nsteps single moves on every chain.  trace may be null. */
int fkgpu_sweep(fkgpu_chains* ch, int nsteps, const fkgpu_trace* trace);
/* Replay nsteps move steps from a captured HIP graph (no trace). */
int fkgpu_sweep_graph(fkgpu_chains* ch, int nsteps);
/* Device time of the last fkgpu_sweep / fkgpu_sweep_graph / fkgpu_sweep_measure call, or of the
 * kernels of the last fkgpu_chains_green / fkgpu_chains_local_dos call (HIP events on the
 * context stream). */
int fkgpu_chains_last_ms(const fkgpu_chains* ch, double* ms);
/* Current state of every chain (nullable outputs): f[nchains*N], nf, logZ,
 * spectrum[nchains*N] (of the current configuration), accepted and solved counts. */
int fkgpu_chains_state(fkgpu_chains* ch, int32_t* f, int32_t* nf, double* logZ, double* spectrum,
                       int64_t* naccept, int64_t* nsolve);

/* ---- On-device measurements of a chain set ----------------------------------
 * The measures that mc_metropolis::measure (src/mc_metropolis.cpp:53-60) calls after
 * every sweep, taken for every chain on the device and kept there until read:
 *   measure_energy::accumulate            src/measures/energy.cpp:6-26
 *     E = sum_k e_k f(e_k) - mu_f N_f + E_ff, d2E = sum_k e_k^2 f(1 - f), E_c
 *   measure_spectrum::accumulate          src/measures/spectrum.cpp:13-21 (running mean)
 *   measure_spectrum_history::accumulate  src/measures/spectrum_history.cpp:13-19
 *   measure_focc::accumulate              src/measures/focc_history.cpp:8-13
 *   measure_ipr::accumulate               include/fk_mc/measures/ipr.hpp:39-56
 *   measure_stiffness::accumulate         include/fk_mc/measures/stiffness.hpp:129-187
 *   measure_eigenfunctions::accumulate    src/measures/eigenfunctions.cpp:12-18
 *   data_saver::save_gwr (G(w,r), typical DOS)  prog/data_save.hxx:535-708
 * The spectrum is the chain's own (the rank-one, mixed and FP64 modes keep it); chain
 * sets on the stochastic trace diagonalise their configurations at each measurement,
 * as the reference's calc_ed(false) does, and the IPR, stiffness and eigenfunction
 * records take the eigenvectors of the rank-one mode or of an FP64 decomposition
 * (N <= 1024).  A measurement changes no
 * chain state: the chains draw and decide exactly as without it.
 * The records of the first `capacity` measurements are kept (series and histories,
 * chain-major); later ones enter the means and the count only. */
#define FKGPU_MEASURE_SPECTRUM_HISTORY 1
#define FKGPU_MEASURE_FOCC_HISTORY 2
#define FKGPU_MEASURE_IPR 4
#define FKGPU_MEASURE_STIFFNESS 8 /* hypercubic lattices, D >= 2, N <= 1024 */
#define FKGPU_MEASURE_EIGENFUNCTIONS 16 /* N <= 1024; capacity x N^2 doubles per chain */
/* Occupations only (mean and, with FKGPU_MEASURE_FOCC_HISTORY, history): no eigensolve at a
 * measurement, any N; the energy series are zero, the spectrum mean and history stay empty.  The
 * reference's Chebyshev runs measure so without measure_ipr (include/fk_mc/fk_mc.hxx:115-124).
 * Not with the spectrum history or the eigenvector measures. */
#define FKGPU_MEASURE_NO_SPECTRUM 32
/* Allocate the records and clear them (measurement count 0).  Replaces the
 * observables the reference's measures are registered on (include/fk_mc/fk_mc.hxx:97-123). */
int fkgpu_chains_measure_enable(fkgpu_chains* ch, int capacity, int flags);
/* Frequencies and Lorentzian width of the conductivity that the stiffness measure
 * records (measure_stiffness's wgrid and offset, include/fk_mc/measures/stiffness.hpp:69-127;
 * default one frequency 0.0, width 0.05).  Takes effect at fkgpu_chains_measure_enable. */
int fkgpu_chains_measure_conductivity(fkgpu_chains* ch, int nw, const double* wgrid, double offset);
/* One measurement of every chain now (measure_*::accumulate). */
int fkgpu_chains_measure(fkgpu_chains* ch);
/* nsweeps x (sweep_len move steps, then one measurement), replayed from one captured
 * HIP graph: the measured part of mcbase::run with a sweep of sweep_len moves. */
int fkgpu_sweep_measure(fkgpu_chains* ch, int nsweeps, int sweep_len);
/* Measurements taken so far (*count), and the energy series of the first
 * n = min(count, capacity) of them, [nchains][n] (nullable outputs):
 * the reference's _energies, _d2energies, _c_energies. */
int fkgpu_chains_energies(fkgpu_chains* ch, int64_t* count, double* energies, double* d2energies,
                          double* c_energies);
/* Means over all measurements (nullable outputs): energy and d2energy [nchains]
 * (_average_energy / _Z, _average_d2energy / _Z), spectrum [nchains*N] (the reference's
 * running mean, same order of operations), occupation of each site [nchains*N]. */
int fkgpu_chains_means(fkgpu_chains* ch, int64_t* count, double* energy, double* d2energy, double* spectrum,
                       double* focc);
/* Histories of the first n = min(count, capacity) measurements, [nchains][n][N]
 * (nullable outputs; each must have been enabled): spectra, occupations, IPRs. */
int fkgpu_chains_histories(fkgpu_chains* ch, int64_t* count, double* spectrum, int32_t* focc, double* ipr);
/* Stiffness records (nullable outputs): stiffness [nchains][n] = (T + V)/N, cond
 * [nchains][n][nw] on the conductivity grid, and their means over all measurements
 * ([nchains], [nchains][nw]); n = min(count, capacity). */
int fkgpu_chains_stiffness(fkgpu_chains* ch, int64_t* count, double* stiffness, double* cond, double* stiffness_mean,
                           double* cond_mean);
/* Eigenvector records [nchains][n][N][N] (evecs[i*N + k] of a record is component i of
 * eigenvector k, eigenvalues ascending), n = min(count, capacity).  Signs, and bases inside
 * degenerate eigenspaces, are those of this library's solvers, not Eigen's. */
int fkgpu_chains_eigenfunctions(fkgpu_chains* ch, int64_t* count, double* evecs);

/* Lattice Green's function G(w + i offset)_{ij} = (1/n) sum_r sum_k V_r[i][k] V_r[j][k] /
 * (w + i offset - e_rk), averaged over every recorded measurement r of every chain
 * (n = nchains x min(count, capacity)), as gwr_re / gwr_im of data_saver::save_gwr
 * (prog/data_save.hxx:535-621).  Needs FKGPU_MEASURE_EIGENFUNCTIONS and
 * FKGPU_MEASURE_SPECTRUM_HISTORY; offset > 0.  g_re, g_im: N x N, row-major. */
int fkgpu_chains_green(fkgpu_chains* ch, double w, double offset, double* g_re, double* g_im);

/* Local densities of states rho_i(w) = -Im G(w + i offset)_{ii} / pi over the same records,
 * for nw frequencies: ldos[q * N + i] (the gw_diag of prog/data_save.hxx:625; the DOS of
 * save_gwr is their site mean and the typical DOS exp(mean_i log rho_i)). */
int fkgpu_chains_local_dos(fkgpu_chains* ch, int nw, const double* wgrid, double offset, double* ldos);

/* Diagnostic (host only, no device needed; no reference counterpart): draws
 * from the library's port of std::mt19937(seed) and the libstdc++
 * distributions used by the reference -- mode 0 raw words, 1
 * uniform_int(0, hi), 2 uniform_real(0, 1), 3 discrete_distribution(probs).
 * Used to pin the port against libstdc++ itself. */
int fkgpu_rng_draws(uint32_t seed, int mode, int hi, int nprobs, const double* probs, int count, double* out);

#ifdef __cplusplus
}
#endif
#endif /* FKGPU_H */
