// fk_internal.h -- kernel argument blocks and launcher prototypes shared by the
// kernels and the C ABI implementation.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fkgpu.h"

#define FK_MAX_LATTICES 8  // lattices of one chain set
#define FK_SOLVER_MIXED32 20  // internal solver kind: blocked FP32 reduction of the mixed path

// Batched calc_hamiltonian + calc_ed.  All pointers are device pointers.
struct eig_args {
    int B, N;
    const double* hop;     // N x N dense hopping template
    const int32_t* f;      // B x N occupations
    const double* beta;    // B
    const double* U;       // B
    const double* mu_c;    // B
    double* evals;         // B x N (ascending)
    double* logz;          // B
    double* cexp;          // B x N or null
    double* cfermi;        // B x N or null
    double* ws;            // grid x ws_stride workspace (HBM-resident solvers) or null
    size_t ws_stride;      // doubles per workgroup
    const int32_t* skip;   // B flags: no solve where nonzero (null: solve all)
    double* vout;          // B x N x N Householder vectors or null
    double* tauout;        // B x N or null
    double* d_out;         // B x N tridiagonal diagonal or null
    double* e_out;         // B x N tridiagonal off-diagonal or null
};

// One Metropolis move for each of B chains.
struct sweep_args {
    int B, N;
    const double* hop;
    const int32_t* nbr;    // N x zmax neighbour table (hopping-graph order)
    const int32_t* z;      // N neighbour counts
    int zmax;
    uint8_t* f;            // B x N
    int32_t* nf;           // B
    double* logz;          // B
    double* spectrum;      // B x N or null
    uint32_t* mt;          // B x 625
    const double* beta;
    const double* U;
    const double* mu_c;
    const double* mu_f;
    long long* naccept;    // B
    long long* nsolve;     // B
    double* ws;            // grid x ws_stride workspace or null
    size_t ws_stride;      // doubles per workgroup
    int ncp;               // number of registered moves
    double cp[4];          // cumulative move probabilities (libstdc++ discrete_distribution)
    int kinds[4];          // move kind of each registered move
    int nW;                // f-f interaction range (1-D lattices only)
    double W[8];
    // chains on several lattices (fkgpu_chains_create_multi): chain c lives on lattice
    // lat[c] (N = lN[.], dense template lhop[.], neighbour table lnbr[.] / lz[.] / lzmax[.]);
    // f and spectrum rows have the stride ld (the largest N).  lat == null: every chain on
    // (N, hop, nbr, z, zmax) with stride N.
    const int32_t* lat;
    int ld;
    int lN[FK_MAX_LATTICES];
    const double* lhop[FK_MAX_LATTICES];
    const int32_t* lnbr[FK_MAX_LATTICES];
    const int32_t* lz[FK_MAX_LATTICES];
    int lzmax[FK_MAX_LATTICES];
    int flip_ff;           // the global exchange's weight includes -beta dE_ff (the Chebyshev
                           // move_flip, src/moves_chebyshev.cpp:22; the ED one has none)
    int trace;             // record the per-move trace
    int step;              // trace row
    int32_t* tr_kind;
    int32_t* tr_a;
    int32_t* tr_b;
    double* tr_u;
    double* tr_logw;
    double* tr_logz;
    int8_t* tr_acc;
};

// One measurement of every chain of a chain set, recorded on the device at record
// m = count[0] (series and histories only while m < cap).
struct chain_meas_args {
    int B, N, cap;
    int nW;                        // f-f interaction range (1-D lattices only)
    double W[8];
    const uint8_t* f;              // B x N
    const int32_t* nf;             // B
    const double* spectrum;        // B x N, ascending
    const double* beta;            // B
    const double* mu_f;            // B
    const double* ipr;             // B x N of this measurement or null
    const long long* count;        // measurements taken so far
    double *e, *d2e, *ce;          // B x cap
    double *esum, *d2sum;          // B
    double* spec_mean;             // B x N running mean
    long long* focc_sum;           // B x N
    double* spec_hist;             // B x cap x N or null
    uint8_t* focc_hist;            // B x cap x N or null
    double* ipr_hist;              // B x cap x N or null
    // stiffness and conductivity of this measurement (null: not measured)
    int nw;
    const double* stiff;           // B
    const double* cond;            // B x nw
    double *stiff_hist, *stiff_sum;  // B x cap, B
    double *cond_hist, *cond_sum;    // B x cap x nw, B x nw
    // eigenvectors of this measurement (null: not recorded)
    const double* V;               // B x N x N
    double* evec_hist;             // B x cap x N x N
};
hipError_t fk_launch_chain_measure(const chain_meas_args& a, long long* count, hipStream_t st);
// fermi[b][k] = 1 / (1 + exp(beta_b ev[b][k])) (ed_cache::cached_fermi, src/configuration.cpp:238)
hipError_t fk_launch_fermi(int B, int N, const double* ev, const double* beta, double* fermi, hipStream_t st);

// One proposed move (kept between the launches of a stochastic-trace move step).
struct move_prop {
    int kind, sa, sb, solve, mode, nf_new;  // mode: 0 weighed, 1 zero weight, 2 unit weight
    double u, ffdiff;
};
hipError_t fk_launch_kpm_propose(const sweep_args& a, move_prop* props, int32_t* fprop, int32_t* skip,
                                 hipStream_t st);
hipError_t fk_launch_kpm_accept(const sweep_args& a, const move_prop* props, const int32_t* fprop,
                                const double* lz_new, hipStream_t st, const double* spec_new = nullptr);
hipError_t fk_launch_mixed_refine(int B, int N, const double* X, const int32_t* f, const double* U,
                                  const double* mu_c, const int32_t* nbr, const int32_t* z, int zmax,
                                  const double* hopv, double* Y, double* S, double* T, double* Q,
                                  const double* beta, double* evals, double* logz, int32_t* fb_skip,
                                  const int32_t* skip, unsigned long long* nflag, hipStream_t st);
int fk_mixed_cluster_max();
hipError_t fk_launch_fu_step(const sweep_args& a, const double* V, double* Q1, double* Q2, int32_t* upd,
                             double* T, double* Vout, hipStream_t st);
hipError_t fk_launch_f8to32(int n, const uint8_t* f8, int32_t* f32, hipStream_t st, const int32_t* skip0 = nullptr);
// rank-one chains: count += 1 on the device, skip[b] = (count % period != 0) for b < B
hipError_t fk_launch_fu_tick(int B, int period, long long* count, int32_t* skip, hipStream_t st);
int fk_fu_nmax();

int fk_solver_kind(int N);          // -1: unsupported size
int fk_solver_threads(int kind);
int fk_solver_needs_ws(int kind);
int fk_solver_nmax(int kind);
size_t fk_solver_ws_doubles(int kind, int N);
hipError_t fk_launch_eig(int kind, const eig_args& a, int grid, hipStream_t st);

// Static device scratch of fkgpu_eig_logz_device_noalloc (fk_noalloc.hip): 256 MiB per
// device.  fk_noalloc_scratch returns this device's address of it (resolved once per
// device with hipGetSymbolAddress) and its size in doubles.
#define FK_NOALLOC_SCRATCH_BYTES (256u << 20)
hipError_t fk_noalloc_scratch(int device, double** ptr, size_t* doubles);
hipError_t fk_launch_sweep(int kind, const sweep_args& a, int grid, hipStream_t st);
// randomize_f of every chain from its own generator (nf_start: B counts, 0 = drawn)
hipError_t fk_launch_randomize_f(int B, int N, uint32_t* mt, uint8_t* f, int32_t* nf, const int32_t* nf_start,
                                 int32_t* f32, hipStream_t st);
// Work queue of a multi-step sweep launch: take and append counters (2), a ring of
// max_steps * B slots, an error flag set when a wait exceeds wait_ticks of the wall
// clock.
struct sweep_queue {
    unsigned* ctr;
    int* ring;
    int max_steps;
    int* err;
    long long wait_ticks;
};
hipError_t fk_launch_sweep_q(int kind, const sweep_args& a, int nsub, const sweep_queue& q, hipStream_t st);

// scratch: fk_evecs_scratch_doubles(B, N) doubles (at least 3 B N^2)
hipError_t fk_launch_evecs(int B, int N, const double* d, const double* e, const double* ev, const double* vh,
                           const double* tau, double* Z, double* scratch, hipStream_t st,
                           const int32_t* skip = nullptr);
size_t fk_evecs_scratch_doubles(int B, int N);
hipError_t fk_launch_ipr(int B, int N, const double* V, double* out, hipStream_t st);
hipError_t fk_launch_stiffness(int B, int N, int stride0, int L0, double t, const double* V, const double* ev,
                               const double* fermi, int nw, const double* wgrid, double offset, double* JV,
                               double* mJ, double* tdiag, double* stiff, double* cond, hipStream_t st);
hipError_t fk_launch_dos(int S, int N, const double* spec, int nw, const double* wgrid, double eta, double* out,
                         hipStream_t st);
hipError_t fk_launch_ipr_w(int S, int N, const double* spec, const double* ipr, int nw, const double* wgrid,
                           double eta, double* out, hipStream_t st);
// Green's function / local DOS over the records of a chain set (n per chain, capacity cap):
// part buffers hold splits x N^2 (each of re, im) or splits x nw x N doubles
int fk_green_splits(int nrec, int tiles);
hipError_t fk_launch_green(int n, int cap, int nchains, int N, const double* V, const double* ev, double w,
                           double eta, int splits, double* part_re, double* part_im, double* g_re, double* g_im,
                           hipStream_t st);
hipError_t fk_launch_ldos(int n, int cap, int nchains, int N, const double* V, const double* ev, int nw,
                          const double* wgrid, double eta, int splits, double* part, double* ldos, hipStream_t st);
hipError_t fk_launch_adjoint(int B, int N, const double* V, const double* fermi, const int32_t* f,
                             const double* beta, const int32_t* nbr, const int32_t* z, int zmax, int nt_bonds,
                             double* dens, double* bond_t, double* bond_tp, double* out, double* rho_full,
                             hipStream_t st);
hipError_t fk_cheb_exact_moments(int N, const int32_t* nbr, const int32_t* z, int zmax, const double* hopv,
                                 const double* diag, double a, double b, int M, double* T0, double* T1,
                                 double* parts, double* scal, double* traces, hipStream_t st);
// hopq (host pointer, zmax values or nullptr): the hopping value of neighbour slot q when it
// is the same at every site (every site then has zmax neighbours)
// dlogZ of one- or two-site proposals of the stochastic-trace chains from local
// Chebyshev moments (M of them, FP64) and n0 Matsubara frequencies (fk_kpm_local.hip);
// lz_out = logz + dlogZ.  s2c, s4c: sums of (2n+1)^-2 and (2n+1)^-4 over n >= n0.
#define FK_KPM_LOCAL_NMAX 4096
#define FK_KPM_LOCAL_MMAX 2048
hipError_t fk_launch_kpm_local(int B, int N, int M, int n0, const uint8_t* f, const int32_t* fprop,
                               const move_prop* props, const int32_t* skip, const double* U, const double* mu_c,
                               const double* beta, const double* ab, const int32_t* nbr, const int32_t* z, int zmax,
                               const double* hopv, const double* hopq, double s2c, double s4c, double* mom,
                               const double* logz, double* lz_out, hipStream_t st);
// Lanczos steps on H = hop + diag(d) (fk_cheb.hip): work = 3N + 2 doubles of state; first != 0
// starts from a fixed vector; alpha, beta receive the tridiagonal entries of the steps run.
hipError_t fk_launch_lanczos(int N, const int32_t* nbr, const int32_t* z, int zmax, const double* hopv,
                             const double* diag, int first, int iters, double* work, double* alpha, double* beta,
                             hipStream_t st);
hipError_t fk_launch_kpm(int B, int N, int R, int M, const int32_t* f, const double* U, const double* mu_c,
                         const double* ab, const int32_t* nbr, const int32_t* z, int zmax, const double* hopv,
                         const double* hopq, uint64_t seed, int per_item_seed, const int32_t* skip, const double* coef, int jackson,
                         double* mom, double* mu_out, double* logz, hipStream_t st);
