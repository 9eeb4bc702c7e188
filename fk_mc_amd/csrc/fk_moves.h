// fk_moves.h -- device code of one Metropolis move shared by the move-step kernels:
// the proposal with the reference's draw order, the weight and accept test, and
// the per-move bookkeeping (src/mc_metropolis.cpp:34-52, src/moves.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_internal.h"
#include "fk_rng.h"

namespace fkk {

__device__ inline double ff_energy_dev(const uint8_t* f, int V, int nW, const double* W) {
    // src/configuration.cpp:59-77 (called only for 1-D lattices)
    double e = 0.0;
    for (int i = 0; i < V; ++i) {
        if (!f[i]) continue;
        for (int l = 0; l < nW; ++l) {
            int left = (i - l + V) % V, right = (i + l) % V;
            e += W[l] * f[left];
            e += W[l] * f[right] * (l > 0);
        }
    }
    return e;
}

// Lattice of chain ch: its size, dense template, row stride of f / spectrum, and (apply)
// its neighbour table patched into a workgroup's copy of the argument block.
struct chain_lattice {
    int N, ld, li;
    const double* hop;
    __device__ void apply(sweep_args& s) const {
        if (li < 0) return;
        s.nbr = s.lnbr[li];
        s.z = s.lz[li];
        s.zmax = s.lzmax[li];
    }
};
__device__ inline chain_lattice chain_lattice_of(const sweep_args& a, int ch) {
    chain_lattice c;
    if (!a.lat) {
        c.N = a.N;
        c.ld = a.N;
        c.li = -1;
        c.hop = a.hop;
        return c;
    }
    c.li = __builtin_amdgcn_readfirstlane(a.lat[ch]);
    c.N = a.lN[c.li];
    c.ld = a.ld;
    c.hop = a.lhop[c.li];
    return c;
}

// Proposal of one move (thread 0 of the chain's workgroup): the draws of
// mc_metropolis::update (src/mc_metropolis.cpp:41) and of the move's attempt()
// (src/moves.cpp) in the reference's order, then the Metropolis uniform
// (src/mc_metropolis.cpp:43; the eigensolve between them draws nothing).
// f is modified in place into the proposed configuration.
__device__ __noinline__ move_prop propose_move(uint32_t* mt, uint8_t* f, int N, int nf, const sweep_args& a) {
    move_prop p;
    p.kind = a.kinds[fkrng::discrete(mt, a.cp, a.ncp)];
    p.sa = p.sb = -1;
    p.solve = 0;
    p.mode = 0;
    p.nf_new = nf;
    double ff_old = 0.0, ff_new = 0.0;
    if (a.nW > 0) ff_old = ff_energy_dev(f, N, a.nW, a.W);
    if (p.kind == FKGPU_MOVE_ADDREMOVE) {
        p.sa = fkrng::uniform_int(mt, 0, N - 1);
        f[p.sa] ^= 1;
        p.nf_new = nf + (f[p.sa] ? 1 : -1);
        p.solve = 1;
    } else if (p.kind == FKGPU_MOVE_FLIP) {
        if (nf == 0 || nf == N) {
            p.mode = 1;  // src/moves.cpp:8: weight 0, nothing drawn
        } else {
            int from = fkrng::uniform_int(mt, 0, N - 1);
            while (f[from] == 0) from = fkrng::uniform_int(mt, 0, N - 1);
            int to = fkrng::uniform_int(mt, 0, N - 1);
            while (f[to] == 1) to = fkrng::uniform_int(mt, 0, N - 1);
            f[from] = 0;
            f[to] = 1;
            p.sa = from;
            p.sb = to;
            p.solve = 1;
        }
    } else if (p.kind == FKGPU_MOVE_HOP) {
        const int i = fkrng::uniform_int(mt, 0, N - 1);
        const int j = a.nbr[(size_t)i * a.zmax + fkrng::uniform_int(mt, 0, a.z[i] - 1)];
        p.sa = i;
        p.sb = j;
        if (f[i] == f[j]) {
            p.mode = 2;  // null move, weight 1
        } else {
            const uint8_t t = f[i];
            f[i] = f[j];
            f[j] = t;
            p.solve = 1;
        }
    } else {  // FKGPU_MOVE_RANDOMIZE: randomize_f(rnd) with nf = 0 (src/configuration.cpp:47-56)
        const int nfs = fkrng::uniform_int(mt, 0, N - 1);
        for (int i = 0; i < N; ++i) f[i] = 0;
        for (int s2 = 0; s2 < nfs; ++s2) {
            int ind = fkrng::uniform_int(mt, 0, N - 1);
            while (f[ind] == 1) ind = fkrng::uniform_int(mt, 0, N - 1);
            f[ind] = 1;
        }
        p.nf_new = nfs;
        p.solve = 1;
    }
    if (a.nW > 0 && p.solve) ff_new = ff_energy_dev(f, N, a.nW, a.W);
    p.u = fkrng::canonical(mt);
    // the ED flip weighs no f-f term (src/moves.cpp:19-20), the Chebyshev one does
    // (src/moves_chebyshev.cpp:22)
    p.ffdiff = (p.kind == FKGPU_MOVE_FLIP && !a.flip_ff) ? 0.0 : ff_new - ff_old;
    return p;
}

// Weight of the proposal and the accept test |w| > u (src/mc_metropolis.cpp:43).
// f_new_sa: the proposed occupation of site sa.  Returns accept; logw via pointer.
__device__ __noinline__ int decide_move(const move_prop& p, int f_new_sa, double lz_new, double lz_old, double beta,
                                        double mu_f, int nf_old, double* logw_out) {
    double w = 0.0, logw = -INFINITY;
    if (p.mode == 1) {
        w = 0.0;
    } else if (p.mode == 2) {
        w = 1.0;
        logw = 0.0;
    } else if (p.kind == FKGPU_MOVE_ADDREMOVE) {
        // src/moves.cpp:63-66
        const double ebmf = exp(beta * mu_f);
        const double ratio = exp(lz_new - lz_old);
        w = (f_new_sa ? ratio * ebmf : ratio / ebmf) * exp(-beta * p.ffdiff);
        logw = (lz_new - lz_old) + (f_new_sa ? 1.0 : -1.0) * beta * mu_f - beta * p.ffdiff;
    } else if (p.kind == FKGPU_MOVE_FLIP) {
        // src/moves.cpp:19-20 (ffdiff 0), src/moves_chebyshev.cpp:22
        w = exp(lz_new - lz_old - beta * p.ffdiff);
        logw = lz_new - lz_old - beta * p.ffdiff;
    } else if (p.kind == FKGPU_MOVE_HOP) {
        w = exp(lz_new - lz_old) * exp(-beta * p.ffdiff);
        logw = lz_new - lz_old - beta * p.ffdiff;
    } else {
        // src/moves.cpp:43-48 including the clamp
        const double log_ratio = lz_new - lz_old;
        const double dn = (double)(p.nf_new - nf_old);
        const double g = beta * mu_f * dn - p.ffdiff;
        logw = log_ratio + g;
        if (g > 2.7182818 - log_ratio) w = 1.0;
        else if (g + log_ratio < 0) w = 0.0;
        else w = exp(log_ratio) * exp(beta * (mu_f * dn - p.ffdiff));
    }
    *logw_out = logw;
    return fabs(w) > p.u;
}

// Bookkeeping of one decided move (thread 0): counters, trace, chain scalars.
__device__ __noinline__ void record_move(const sweep_args& a, int ch, const move_prop& p, int acc, double logw,
                                         double lz_new) {
    if (a.trace) {
        const size_t t = (size_t)a.step * a.B + ch;
        a.tr_kind[t] = p.kind;
        a.tr_a[t] = p.sa;
        a.tr_b[t] = p.sb;
        a.tr_u[t] = p.u;
        a.tr_logw[t] = logw;
        a.tr_logz[t] = p.solve ? lz_new : 0.0;
        a.tr_acc[t] = acc;
    }
    if (acc) {
        a.naccept[ch] += 1;
        a.nf[ch] = p.nf_new;
        if (p.solve) a.logz[ch] = lz_new;
    }
    if (p.solve) a.nsolve[ch] += 1;
}


}  // namespace fkk
