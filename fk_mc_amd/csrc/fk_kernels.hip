// fk_kernels.hip -- HIP/CDNA4 kernels of the per-move diagonalisation path.
//
//   k_eig_logz_reg<P,NB>   calc_hamiltonian + calc_ed for a batch of f-configs
//                          (src/configuration.cpp:79-91, 208-246), N <= P*NB,
//                          matrix resident in VGPRs, one workgroup per config.
//   k_eig_logz_glob        same for larger N: matrix in an HBM workspace.
//   k_sweep_q<Solver>      several move steps of every chain in one launch (one
//                          workgroup per (step, chain) item, chains taken from a queue
//                          of ready chains; each chain's steps stay in order)
//   k_sweep<Solver>        one Metropolis move for every chain of a batch in one
//                          launch: move choice, proposal, H assembly, eigensolve,
//                          logZ, accept test, state update
//                          (src/mc_metropolis.cpp:34-52, src/moves.cpp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_blocked.h"
#include "fk_internal.h"
#include "fk_rng.h"
#include "fk_solve.h"
#include "fk_moves.h"
#include "fk_wave.h"
#include "fk_wrow.h"

namespace fkk {

using namespace fks;
using fkb::blocked_solver;

// ---------------------------------------------------------------------------
// Solver policies.  Each provides:
//   NT                      threads per workgroup
//   smem                    shared-memory struct (holds ev[], e2[] etc.)
//   solve(...) -> logZ      eigenvalues of H(f_lds) into sm.ev, returns logZ
// ---------------------------------------------------------------------------
template <int P, int NB>
struct reg_solver {
    static constexpr int NT = P * P;
    static constexpr int NP = P * NB;
    static constexpr int NMAX = NP;
    struct smem {
        tri_smem<P, NB> t;
        double ev[NP];
        double de[2 * NP];
        int cnt[NT + 2];
        double scratch[3 * (NT / WAVE) + 8];
    };
    __device__ static const double* dptr(const smem& sm) { return sm.t.d; }
    __device__ static const double* eptr(const smem& sm) { return sm.t.e; }
    __device__ static double* evptr(smem& sm) { return sm.ev; }
    __device__ __forceinline__ static double solve(smem& sm, const double* __restrict__ hop, const uint8_t* f_lds, int N,
                                   double beta, double U, double mu_c, double* ws, double* vout, double* tauout,
                                   double* cexp, double* cfermi) {
        const int tid = threadIdx.x;
        const int lane = tid & (WAVE - 1);
        const int r = (tid / WAVE) * (WAVE / P) + lane / P;
        const int c = lane % P;
        double A[NB * (NB + 1) / 2];
        assemble<P, NB>(A, hop, f_lds, U, mu_c, N, r, c);
        tridiagonalize<P, NB>(A, sm.t, N, vout, tauout);
        tridiag_eigvals<NT>(sm.t.d, sm.t.e, sm.de, sm.cnt, N, sm.ev, sm.scratch);
        return logz_block<NT>(sm.ev, N, beta, cexp, cfermi, sm.scratch);
    }
};

// Matrix in an N x N HBM workspace (ws), one workgroup per matrix.
struct glob_solver {
    static constexpr int NT = 1024;
    static constexpr int NMAX = 2048;
    struct smem {
        double xs[NMAX], vs[NMAX], ps[NMAX], wsv[NMAX];
        double d[NMAX], e[NMAX], de[2 * NMAX], ev[NMAX];
        int cnt[NT + 2];
        double scratch[3 * (NT / WAVE) + 8];
        double red[4];
    };
    __device__ static const double* dptr(const smem& sm) { return sm.d; }
    __device__ static const double* eptr(const smem& sm) { return sm.e; }
    __device__ static double* evptr(smem& sm) { return sm.ev; }
    __device__ __forceinline__ static double solve(smem& sm, const double* __restrict__ hop, const uint8_t* f_lds, int N,
                                   double beta, double U, double mu_c, double* __restrict__ W, double* vout,
                                   double* tauout, double* cexp, double* cfermi) {
        const int tid = threadIdx.x;
        const int lane = tid & (WAVE - 1), wid = tid / WAVE;
        constexpr int NW = NT / WAVE;
        const size_t NN = (size_t)N * N;
        for (size_t idx = tid; idx < NN; idx += NT) {
            int i = (int)(idx / N), j = (int)(idx % N);
            double v = hop[idx];
            if (i == j) v += -mu_c + U * (double)f_lds[i];
            W[idx] = v;
        }
        for (int i = tid; i < N; i += NT) { sm.d[i] = 0.0; sm.e[i] = 0.0; }
        __syncthreads();
        for (int k = 0; k + 2 < N; ++k) {
            for (int i = k + 1 + tid; i < N; i += NT) sm.xs[i] = W[(size_t)i * N + k];
            if (tid == 0) sm.d[k] = W[(size_t)k * N + k];
            __syncthreads();
            double alpha = sm.xs[k + 1];
            double part = 0.0;
            for (int i = k + 2 + lane; i < N; i += WAVE) part += sm.xs[i] * sm.xs[i];
            double xn2 = wave_sum(part);
            double tau = 0.0, beta_h = alpha, scal = 0.0;
            if (xn2 != 0.0) {
                beta_h = -copysign(sqrt(alpha * alpha + xn2), alpha);
                tau = (beta_h - alpha) / beta_h;
                scal = 1.0 / (alpha - beta_h);
            }
            if (tid == 0) sm.e[k] = beta_h;
            for (int i = tid; i < N; i += NT)
                sm.vs[i] = (i == k + 1) ? 1.0 : ((i > k + 1) ? sm.xs[i] * scal : 0.0);
            if (vout) {
                for (int i = tid; i < N; i += NT)
                    vout[(size_t)k * N + i] = (i == k + 1) ? 1.0 : ((i > k + 1) ? sm.xs[i] * scal : 0.0);
                if (tid == 0) tauout[k] = tau;
            }
            __syncthreads();
            if (tau != 0.0) {
                for (int i = k + 1 + wid; i < N; i += NW) {
                    const double* row = W + (size_t)i * N;
                    double s = 0.0;
                    for (int j = k + 1 + lane; j < N; j += WAVE) s = fma(row[j], sm.vs[j], s);
                    s = wave_sum(s);
                    if (lane == 0) sm.ps[i] = tau * s;
                }
                __syncthreads();
                double pvp = 0.0;
                for (int i = k + 1 + lane; i < N; i += WAVE) pvp += sm.ps[i] * sm.vs[i];
                const double coef = 0.5 * tau * wave_sum(pvp);
                for (int i = k + 1 + tid; i < N; i += NT) sm.wsv[i] = sm.ps[i] - coef * sm.vs[i];
                __syncthreads();
                const int m = N - k - 1;
                const size_t mm = (size_t)m * m;
                for (size_t idx = tid; idx < mm; idx += NT) {
                    int i = k + 1 + (int)(idx / m), j = k + 1 + (int)(idx % m);
                    double* wp = W + (size_t)i * N + j;
                    *wp = fma(-sm.vs[i], sm.wsv[j], fma(-sm.wsv[i], sm.vs[j], *wp));
                }
            }
            __syncthreads();
        }
        if (N >= 2 && tid == 0) {
            sm.d[N - 2] = W[(size_t)(N - 2) * N + (N - 2)];
            sm.e[N - 2] = W[(size_t)(N - 1) * N + (N - 2)];
            sm.d[N - 1] = W[(size_t)(N - 1) * N + (N - 1)];
        } else if (N == 1 && tid == 0) {
            sm.d[0] = W[0];
        }
        if (vout) {
            for (int i = tid; i < N; i += NT) {
                if (N >= 2) vout[(size_t)(N - 2) * N + i] = 0.0;
                vout[(size_t)(N - 1) * N + i] = 0.0;
            }
            if (tid == 0) { if (N >= 2) tauout[N - 2] = 0.0; tauout[N - 1] = 0.0; }
        }
        __syncthreads();
        tridiag_eigvals<NT>(sm.d, sm.e, sm.de, sm.cnt, N, sm.ev, sm.scratch);
        return logz_block<NT>(sm.ev, N, beta, cexp, cfermi, sm.scratch);
    }
};

// ---------------------------------------------------------------------------
// Batched calc_hamiltonian + calc_ed.
// ---------------------------------------------------------------------------
template <class S>
__global__ void __launch_bounds__(S::NT) k_eig_logz(eig_args a) {
    __shared__ typename S::smem sm;
    __shared__ uint8_t f_lds[S::NMAX];
    const int tid = threadIdx.x;
    const int N = a.N;
    // one workgroup per configuration (a grid-stride loop here costs ~25 spilled VGPRs)
    {
        const int bidx = blockIdx.x;
        if (a.skip && a.skip[bidx]) return;  // uniform over the workgroup
        for (int i = tid; i < S::NMAX; i += S::NT) f_lds[i] = (i < N) ? (uint8_t)(a.f[(size_t)bidx * N + i] != 0) : 0;
        __syncthreads();
        double* ws = a.ws ? a.ws + (size_t)blockIdx.x * a.ws_stride : nullptr;
        double* vout = a.vout ? a.vout + (size_t)bidx * N * N : nullptr;
        double* tauout = a.tauout ? a.tauout + (size_t)bidx * N : nullptr;
        double lz = S::solve(sm, a.hop, f_lds, N, a.beta[bidx], a.U[bidx], a.mu_c[bidx], ws, vout, tauout,
                             a.cexp ? a.cexp + (size_t)bidx * N : nullptr,
                             a.cfermi ? a.cfermi + (size_t)bidx * N : nullptr);
        const double* ev = S::evptr(sm);
        for (int i = tid; i < N; i += S::NT) a.evals[(size_t)bidx * N + i] = ev[i];
        if (tid == 0) a.logz[bidx] = lz;
        if (a.d_out) {
            // tridiagonal (d, e) for the eigenvector stage
            const double* dd = S::dptr(sm);
            const double* ee = S::eptr(sm);
            for (int i = tid; i < N; i += S::NT) {
                a.d_out[(size_t)bidx * N + i] = dd[i];
                a.e_out[(size_t)bidx * N + i] = (i + 1 < N) ? ee[i] : 0.0;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// One Metropolis move for every chain (one launch per move step).
// ---------------------------------------------------------------------------
template <class S>
__global__ void __launch_bounds__(S::NT) k_sweep(sweep_args a) {
    __shared__ typename S::smem sm;
    __shared__ uint8_t f_lds[S::NMAX];
    __shared__ uint32_t mt[fkrng::MT_WORDS];
    __shared__ move_prop prop;
    __shared__ int acc_s;
    // the move helpers take the argument block by reference: an LDS copy keeps the
    // compiler from materialising a private copy of it in every lane
    __shared__ sweep_args sa;
    const int tid = threadIdx.x;
    const int ch = blockIdx.x;  // one workgroup per chain
    const chain_lattice cl = chain_lattice_of(a, ch);
    const int N = cl.N;
    if (tid == 0) {
        sa = a;
        cl.apply(sa);
    }
    uint32_t* mtg = a.mt + (size_t)ch * fkrng::MT_WORDS;
    for (int i = tid; i < fkrng::MT_WORDS; i += S::NT) mt[i] = mtg[i];
    for (int i = tid; i < S::NMAX; i += S::NT) f_lds[i] = (i < N) ? a.f[(size_t)ch * cl.ld + i] : 0;
    __syncthreads();
    const double beta = a.beta[ch], U = a.U[ch], mu_c = a.mu_c[ch], mu_f = a.mu_f[ch];
    if (tid == 0) prop = propose_move(mt, f_lds, N, a.nf[ch], sa);
    __syncthreads();
    double lz_new = 0.0;
    if (prop.solve) {
        double* ws = a.ws ? a.ws + (size_t)blockIdx.x * a.ws_stride : nullptr;
        lz_new = S::solve(sm, cl.hop, f_lds, N, beta, U, mu_c, ws, nullptr, nullptr, nullptr, nullptr);
    }
    if (tid == 0) {
        double logw;
        const int acc = decide_move(prop, prop.sa >= 0 ? f_lds[prop.sa] : 0, lz_new, a.logz[ch], beta, mu_f,
                                    a.nf[ch], &logw);
        record_move(sa, ch, prop, acc, logw, lz_new);
        acc_s = acc;
    }
    __syncthreads();
    if (acc_s) {
        for (int i = tid; i < N; i += S::NT) a.f[(size_t)ch * cl.ld + i] = f_lds[i];
        if (prop.solve && a.spectrum)
            for (int i = tid; i < N; i += S::NT) a.spectrum[(size_t)ch * cl.ld + i] = S::evptr(sm)[i];
    }
    for (int i = tid; i < fkrng::MT_WORDS; i += S::NT) mtg[i] = mt[i];
}

// ---------------------------------------------------------------------------
// nsub Metropolis move steps of every chain in one launch of nsub * B workgroups,
// scheduled through a queue of ready chains: ring[0..B) holds every chain (step 0);
// the workgroup that finishes step s < nsub - 1 of chain c appends (s + 1) * B + c
// (chain and step in one word, so the slot alone tells the next workgroup both).
// Every wave's stores of the chain's state reach the L2 before the workgroup's last
// barrier, and thread 0's release writes the L2 back before it fills the slot.  A
// workgroup takes
// the next ring slot when it starts and waits until the slot is filled, so a free CU
// always runs the chain that became ready first, each chain's steps stay in order and
// consume the same random words as one launch per step, and the solves of all chains
// and steps stream through the CUs without a launch boundary between steps.  Slot
// fills are agent-scope releases (the chain's state was written by a workgroup that
// may sit on another XCD), slot reads acquires.  Slot q is filled by the B initial
// entries or by a completion; if every running workgroup waited, all filled slots would
// be completed items, which makes the fills equal B plus the completions that append,
// i.e. every chain would have finished its last step and no slot would be left to wait
// on.  Each wait is also bounded by wait_ticks of the wall clock, after which err is
// set and every later workgroup leaves at once.
// ---------------------------------------------------------------------------
template <class S>
__global__ void __launch_bounds__(S::NT) k_sweep_q(sweep_args a, int nsub, unsigned* __restrict__ qctr,
                                                   int* __restrict__ ring, int* __restrict__ err,
                                                   long long wait_ticks) {
    __shared__ typename S::smem sm;
    __shared__ uint8_t f_lds[S::NMAX];
    __shared__ uint32_t mt[fkrng::MT_WORDS];
    __shared__ move_prop prop;
    __shared__ int acc_s, chain_s, step_s;
    __shared__ __align__(16) double bmu_s[2];  // beta, mu_f for the decision after the solve (no register across it)
    __shared__ sweep_args sa;
    const int tid = threadIdx.x;
    const int B = a.B;
    if (tid == 0) {
        sa = a;
        int c = -1, st = 0;
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            const unsigned q = atomicAdd(&qctr[0], 1u);
            if (q < (unsigned)nsub * (unsigned)B) {
                const long long t0 = wall_clock64();
                int v;
                // relaxed polls, one acquire fence after the poll that sees the slot filled:
                // an agent-scope acquire invalidates the XCD's L2 (buffer_inv sc1), and a
                // workgroup polling with acquires would keep stripping the L2 under the
                // workgroups that solve on its XCD (with few chains per CU most workgroups
                // of a launch wait; a solver with an HBM workspace then ran 2-3x slower)
                while ((v = __hip_atomic_load(&ring[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < 0) {
                    if (wall_clock64() - t0 > wait_ticks) {
                        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (v >= 0) {
                    c = v % B;
                    st = v / B;
                }
            }
        }
        chain_s = c;
        step_s = st;
    }
    __syncthreads();
    // every wave acquires at agent scope after thread 0 saw the slot filled, so its loads
    // of the chain's state follow the stores of the step before (which may have run on
    // another XCD); chain and step are uniform (held in scalar registers)
    const int ch = __builtin_amdgcn_readfirstlane(chain_s);
    if (ch < 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int s = __builtin_amdgcn_readfirstlane(step_s);
    const chain_lattice cl = chain_lattice_of(a, ch);
    const int N = cl.N;
    uint32_t* mtg = a.mt + (size_t)ch * fkrng::MT_WORDS;
    for (int i = tid; i < fkrng::MT_WORDS; i += S::NT) mt[i] = mtg[i];
    for (int i = tid; i < S::NMAX; i += S::NT) f_lds[i] = (i < N) ? a.f[(size_t)ch * cl.ld + i] : 0;
    __syncthreads();
    const double beta = a.beta[ch], U = a.U[ch], mu_c = a.mu_c[ch];
    if (tid == 0) {
        sa.step = a.step + s;
        cl.apply(sa);
        prop = propose_move(mt, f_lds, N, a.nf[ch], sa);
        bmu_s[0] = beta;
        bmu_s[1] = a.mu_f[ch];
    }
    __syncthreads();
    double lz_new = 0.0;
    if (prop.solve) {
        // workspace solvers: one workspace per chain (a chain's steps never overlap)
        double* ws = a.ws ? a.ws + (size_t)ch * a.ws_stride : nullptr;
        lz_new = S::solve(sm, cl.hop, f_lds, N, beta, U, mu_c, ws, nullptr, nullptr, nullptr, nullptr);
    }
    __syncthreads();
    if (tid == 0) {
        double logw;
        const int acc = decide_move(prop, prop.sa >= 0 ? f_lds[prop.sa] : 0, lz_new, a.logz[ch], bmu_s[0], bmu_s[1],
                                    a.nf[ch], &logw);
        record_move(sa, ch, prop, acc, logw, lz_new);
        acc_s = acc;
    }
    __syncthreads();
    if (acc_s) {
        for (int i = tid; i < N; i += S::NT) a.f[(size_t)ch * cl.ld + i] = f_lds[i];
        if (prop.solve && a.spectrum)
            for (int i = tid; i < N; i += S::NT) a.spectrum[(size_t)ch * cl.ld + i] = S::evptr(sm)[i];
    }
    for (int i = tid; i < fkrng::MT_WORDS; i += S::NT) mtg[i] = mt[i];
    // every wave waits until its stores have reached the L2 (s_waitcnt vmcnt(0)) before
    // the barrier; thread 0's agent-scope release then writes the XCD's L2 back before
    // the slot is filled
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (tid == 0 && s + 1 < nsub) {
        const unsigned slot = atomicAdd(&qctr[1], 1u);
        if (slot < (unsigned)nsub * (unsigned)B)  // always true: one append per step but the last
            __hip_atomic_store(&ring[slot], (s + 1) * B + ch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        else
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Queue of a k_sweep_q launch: slots [0, B) hold the chains in order at step 0, the
// rest are empty; take counter 0, append counter B.
__global__ void k_sweep_q_reset(int B, int nslots, unsigned* __restrict__ qctr, int* __restrict__ ring) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        qctr[0] = 0u;
        qctr[1] = (unsigned)B;
    }
    if (i < nslots) ring[i] = i < B ? i : -1;
}

// ---- Metropolis moves with the stochastic-trace logZ (three launches per step) ----
constexpr int KC_NT = 64;
constexpr int KC_NMAX = 4096;

// proposals: f' (int32) for the KPM kernel, proposal record, skip flag
__global__ void __launch_bounds__(KC_NT) k_kpm_propose(sweep_args a, move_prop* __restrict__ props,
                                                       int32_t* __restrict__ fprop, int32_t* __restrict__ skip) {
    __shared__ uint8_t f_lds[KC_NMAX];
    __shared__ uint32_t mt[fkrng::MT_WORDS];
    __shared__ sweep_args sa;
    const int ch = blockIdx.x, tid = threadIdx.x, N = a.N;
    uint32_t* mtg = a.mt + (size_t)ch * fkrng::MT_WORDS;
    if (tid == 0) sa = a;
    for (int i = tid; i < fkrng::MT_WORDS; i += KC_NT) mt[i] = mtg[i];
    for (int i = tid; i < N; i += KC_NT) f_lds[i] = a.f[(size_t)ch * N + i];
    __syncthreads();
    if (tid == 0) {
        const move_prop p = propose_move(mt, f_lds, N, a.nf[ch], sa);
        props[ch] = p;
        skip[ch] = !p.solve;
    }
    __syncthreads();
    for (int i = tid; i < N; i += KC_NT) fprop[(size_t)ch * N + i] = f_lds[i];
    for (int i = tid; i < fkrng::MT_WORDS; i += KC_NT) mtg[i] = mt[i];
}

__global__ void __launch_bounds__(KC_NT) k_kpm_accept(sweep_args a, const move_prop* __restrict__ props,
                                                      const int32_t* __restrict__ fprop,
                                                      const double* __restrict__ lz_new_g,
                                                      const double* __restrict__ spec_new) {
    __shared__ int acc_s;
    __shared__ sweep_args sa;
    const int ch = blockIdx.x, tid = threadIdx.x, N = a.N;
    const move_prop p = props[ch];
    const double lz_new = p.solve ? lz_new_g[ch] : 0.0;
    if (tid == 0) {
        sa = a;
        double logw;
        const int acc = decide_move(p, p.sa >= 0 ? fprop[(size_t)ch * N + p.sa] : 0, lz_new, a.logz[ch], a.beta[ch],
                                    a.mu_f[ch], a.nf[ch], &logw);
        record_move(sa, ch, p, acc, logw, lz_new);
        acc_s = acc;
    }
    __syncthreads();
    if (acc_s) {
        for (int i = tid; i < N; i += KC_NT) a.f[(size_t)ch * N + i] = (uint8_t)fprop[(size_t)ch * N + i];
        if (spec_new && p.solve && a.spectrum)
            for (int i = tid; i < N; i += KC_NT) a.spectrum[(size_t)ch * N + i] = spec_new[(size_t)ch * N + i];
    }
}

// configuration_t::randomize_f(rnd, nf) (src/configuration.cpp:47-56) from each chain's
// own generator, as fk_mc::initialize draws the starting configuration
// (include/fk_mc/fk_mc.hxx:46): nf = nf_start, or a uniform draw when nf_start is 0,
// then rejection-sampled distinct sites.  One workgroup per chain; f32 receives the
// configuration as int32 for the eigensolve that follows.
__global__ void __launch_bounds__(KC_NT) k_randomize_f(int B, int N, uint32_t* __restrict__ mtg_all,
                                                       uint8_t* __restrict__ f, int32_t* __restrict__ nf,
                                                       const int32_t* __restrict__ nf_start,
                                                       int32_t* __restrict__ f32) {
    __shared__ uint8_t f_lds[KC_NMAX];
    __shared__ uint32_t mt[fkrng::MT_WORDS];
    const int ch = blockIdx.x, tid = threadIdx.x;
    uint32_t* mtg = mtg_all + (size_t)ch * fkrng::MT_WORDS;
    for (int i = tid; i < fkrng::MT_WORDS; i += KC_NT) mt[i] = mtg[i];
    for (int i = tid; i < N; i += KC_NT) f_lds[i] = 0;
    __syncthreads();
    if (tid == 0) {
        int n = nf_start[ch];
        if (n == 0) n = fkrng::uniform_int(mt, 0, N - 1);
        for (int k = 0; k < n; ++k) {
            int ind = fkrng::uniform_int(mt, 0, N - 1);
            while (f_lds[ind] == 1) ind = fkrng::uniform_int(mt, 0, N - 1);
            f_lds[ind] = 1;
        }
        nf[ch] = n;
    }
    __syncthreads();
    for (int i = tid; i < N; i += KC_NT) {
        f[(size_t)ch * N + i] = f_lds[i];
        f32[(size_t)ch * N + i] = f_lds[i];
    }
    for (int i = tid; i < fkrng::MT_WORDS; i += KC_NT) mtg[i] = mt[i];
}

// ---------------------------------------------------------------------------
// Host-side launchers.
// ---------------------------------------------------------------------------
template <class S>
static hipError_t launch_eig(const eig_args& a, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_eig_logz<S>, dim3(grid), dim3(S::NT), 0, st, a);
    return hipGetLastError();
}
template <class S>
static hipError_t launch_sweep(const sweep_args& a, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_sweep<S>, dim3(grid), dim3(S::NT), 0, st, a);
    return hipGetLastError();
}
template <class S>
static hipError_t launch_sweep_q(const sweep_args& a, int nsub, const sweep_queue& q, hipStream_t st) {
    const int nslots = nsub * a.B;
    hipLaunchKernelGGL(k_sweep_q_reset, dim3((nslots + 255) / 256), dim3(256), 0, st, a.B, nslots, q.ctr, q.ring);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_sweep_q<S>, dim3(nslots), dim3(S::NT), 0, st, a, nsub, q.ctr, q.ring, q.err, q.wait_ticks);
    return hipGetLastError();
}

}  // namespace fkk

using namespace fkk;

// Solver kinds: 0-2 register-resident one-stage, block-cyclic P x P grid (N <= 64 /
// 128 / 256), 3 unblocked HBM workspace (N <= 2048), 4 blocked FP64 with MFMA
// trailing updates (N <= 1024), 5 register-resident one-stage, wave-row layout
// (N <= 256), 6 register-resident one-wave (N <= 64); 7, 8, 9 the blocked FP64 solver in
// other tile shapes for the autotuner: 4 and 16 waves per matrix, and panels of at most 16
// columns; 10 the wave-row layout with 8 waves per matrix (N <= 256); 20 blocked FP32 (first
// stage of the mixed-precision path, N <= 1024).
using fkw::wrow_solver;
using fkw::wrow8_solver;
using fkq::wave_solver;
typedef blocked_solver<double, 512> blk64;
typedef blocked_solver<double, 256> blk64w4;
typedef blocked_solver<double, 1024> blk64w16;
typedef blocked_solver<double, 512, 16> blk64nb16;
typedef blocked_solver<float, 512> blk32;

int fk_solver_kind(int N) {
    if (N <= 64) return 0;
    if (N <= 128) return 1;
    if (N <= 256) return 10;
    if (N <= blk64w16::NMAX) return 8;
    if (N <= glob_solver::NMAX) return 3;
    return -1;
}
int fk_solver_nmax(int kind) {
    switch (kind) {
        case 0: return reg_solver<16, 4>::NMAX;
        case 1: return reg_solver<16, 8>::NMAX;
        case 2: return reg_solver<32, 8>::NMAX;
        case 3: return glob_solver::NMAX;
        case 4: return blk64::NMAX;
        case 5: return wrow_solver::NMAX;
        case 6: return wave_solver::NMAX;
        case 7: return blk64w4::NMAX;
        case 8: return blk64w16::NMAX;
        case 9: return blk64nb16::NMAX;
        case 10: return wrow8_solver::NMAX;
        case FK_SOLVER_MIXED32: return blk32::NMAX;
    }
    return 0;
}
int fk_solver_threads(int kind) {
    switch (kind) {
        case 0: return reg_solver<16, 4>::NT;
        case 1: return reg_solver<16, 8>::NT;
        case 2: return reg_solver<32, 8>::NT;
        case 3: return glob_solver::NT;
        case 4: return blk64::NT;
        case 5: return wrow_solver::NT;
        case 6: return wave_solver::NT;
        case 7: return blk64w4::NT;
        case 8: return blk64w16::NT;
        case 9: return blk64nb16::NT;
        case 10: return wrow8_solver::NT;
        case FK_SOLVER_MIXED32: return blk32::NT;
    }
    return 0;
}
int fk_solver_needs_ws(int kind) { return kind == 3 || kind == 4 || (kind >= 7 && kind <= 9) || kind == FK_SOLVER_MIXED32; }
size_t fk_solver_ws_doubles(int kind, int N) {
    switch (kind) {
        case 3: return (size_t)N * N;
        case 4: return blk64::ws_doubles(N);
        case 7: return blk64w4::ws_doubles(N);
        case 8: return blk64w16::ws_doubles(N);
        case 9: return blk64nb16::ws_doubles(N);
        case FK_SOLVER_MIXED32: return blk32::ws_doubles(N);
    }
    return 0;
}

hipError_t fk_launch_eig(int kind, const eig_args& a, int grid, hipStream_t st) {
    switch (kind) {
        case 0: return launch_eig<reg_solver<16, 4>>(a, grid, st);
        case 1: return launch_eig<reg_solver<16, 8>>(a, grid, st);
        case 2: return launch_eig<reg_solver<32, 8>>(a, grid, st);
        case 3: return launch_eig<glob_solver>(a, grid, st);
        case 4: return launch_eig<blk64>(a, grid, st);
        case 5: return launch_eig<wrow_solver>(a, grid, st);
        case 6: return launch_eig<wave_solver>(a, grid, st);
        case 7: return launch_eig<blk64w4>(a, grid, st);
        case 8: return launch_eig<blk64w16>(a, grid, st);
        case 9: return launch_eig<blk64nb16>(a, grid, st);
        case 10: return launch_eig<wrow8_solver>(a, grid, st);
        case FK_SOLVER_MIXED32: return launch_eig<blk32>(a, grid, st);
    }
    return hipErrorInvalidValue;
}
hipError_t fk_launch_kpm_propose(const sweep_args& a, move_prop* props, int32_t* fprop, int32_t* skip,
                                 hipStream_t st) {
    if (a.N > KC_NMAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_kpm_propose, dim3(a.B), dim3(KC_NT), 0, st, a, props, fprop, skip);
    return hipGetLastError();
}
hipError_t fk_launch_kpm_accept(const sweep_args& a, const move_prop* props, const int32_t* fprop,
                                const double* lz_new, hipStream_t st, const double* spec_new) {
    hipLaunchKernelGGL(k_kpm_accept, dim3(a.B), dim3(KC_NT), 0, st, a, props, fprop, lz_new, spec_new);
    return hipGetLastError();
}
hipError_t fk_launch_randomize_f(int B, int N, uint32_t* mt, uint8_t* f, int32_t* nf, const int32_t* nf_start,
                                 int32_t* f32, hipStream_t st) {
    if (N < 1 || N > KC_NMAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_randomize_f, dim3(B), dim3(KC_NT), 0, st, B, N, mt, f, nf, nf_start, f32);
    return hipGetLastError();
}
hipError_t fk_launch_sweep_q(int kind, const sweep_args& a, int nsub, const sweep_queue& q, hipStream_t st) {
    if (nsub < 1 || nsub > q.max_steps || (long long)nsub * a.B > 0x7fffffffLL) return hipErrorInvalidValue;
    switch (kind) {
        case 0: return launch_sweep_q<reg_solver<16, 4>>(a, nsub, q, st);
        case 1: return launch_sweep_q<reg_solver<16, 8>>(a, nsub, q, st);
        case 2: return launch_sweep_q<reg_solver<32, 8>>(a, nsub, q, st);
        case 3: return launch_sweep_q<glob_solver>(a, nsub, q, st);
        case 4: return launch_sweep_q<blk64>(a, nsub, q, st);
        case 5: return launch_sweep_q<wrow_solver>(a, nsub, q, st);
        case 6: return launch_sweep_q<wave_solver>(a, nsub, q, st);
        case 7: return launch_sweep_q<blk64w4>(a, nsub, q, st);
        case 8: return launch_sweep_q<blk64w16>(a, nsub, q, st);
        case 9: return launch_sweep_q<blk64nb16>(a, nsub, q, st);
        case 10: return launch_sweep_q<wrow8_solver>(a, nsub, q, st);
    }
    return hipErrorInvalidValue;
}

hipError_t fk_launch_sweep(int kind, const sweep_args& a, int grid, hipStream_t st) {
    switch (kind) {
        case 0: return launch_sweep<reg_solver<16, 4>>(a, grid, st);
        case 1: return launch_sweep<reg_solver<16, 8>>(a, grid, st);
        case 2: return launch_sweep<reg_solver<32, 8>>(a, grid, st);
        case 3: return launch_sweep<glob_solver>(a, grid, st);
        case 4: return launch_sweep<blk64>(a, grid, st);
        case 5: return launch_sweep<wrow_solver>(a, grid, st);
        case 6: return launch_sweep<wave_solver>(a, grid, st);
        case 7: return launch_sweep<blk64w4>(a, grid, st);
        case 8: return launch_sweep<blk64w16>(a, grid, st);
        case 9: return launch_sweep<blk64nb16>(a, grid, st);
        case 10: return launch_sweep<wrow8_solver>(a, grid, st);
    }
    return hipErrorInvalidValue;
}
