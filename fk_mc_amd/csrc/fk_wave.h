// fk_wave.h -- one-wave register-resident Householder tridiagonalisation for N <= 64:
// one workgroup of a single wave per matrix, no barriers between waves.
//
// Lane j holds column j of H in 64 VGPR pairs (a[i] = A[i][j]).  For a Householder
// step k (Golub & Van Loan 8.3.1, the same reflector convention as the other solvers:
// v_{k+1} = 1, v_i = x_i / (alpha - beta), tau = (beta - alpha) / beta):
//   * lane k writes its column to LDS, so lane i reads x_i = A[i][k];
//   * p_j = tau sum_i A[i][j] v_i is the lane's own column against v (broadcast LDS
//     reads), w_j = p_j - (tau p.v / 2) v_j;
//   * the rank-2 update A[i][j] -= v_i w_j + w_i v_j touches the lane's own column.
// Rows i <= k take part with v_i = w_i = 0 (no effect); whole 16-row blocks below the
// active part are skipped.  A single-chain move at N = 64 is latency-bound, and a wave
// needs no workgroup barrier between its phases.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_solve.h"

namespace fkq {

using namespace fks;

struct wave_solver {
    static constexpr int NT = 64;
    static constexpr int NP = 64;
    static constexpr int NMAX = NP;
    struct smem {
        double x[NP];   // the column being reduced
        double v[NP];
        double w[NP];
        double d[NP], e[NP];
        double de[2 * NP];
        double ev[NP];
        int cnt[2 * NT + 2];
        double scratch[8];
    };
    __device__ static const double* dptr(const smem& sm) { return sm.d; }
    __device__ static const double* eptr(const smem& sm) { return sm.e; }
    __device__ static double* evptr(smem& sm) { return sm.ev; }

    // column `c` of the lane that holds it into sm.x, rows >= from (8-row groups: the groups
    // above the first needed row are not stored; a single lane's LDS stores are the costliest
    // part of a column step)
    __device__ __forceinline__ static void put_column(const double (&a)[NP], smem& sm, int c, int from) {
        if ((int)threadIdx.x == c) {
#pragma unroll
            for (int g = 0; g < NP / 8; ++g)
                if (8 * g + 7 >= from) {
#pragma unroll
                    for (int i = 8 * g; i < 8 * g + 8; ++i) sm.x[i] = a[i];
                }
        }
        __syncthreads();
    }

    __device__ __forceinline__ static void tridiagonalize(double (&a)[NP], smem& sm, int N, double* __restrict__ vout,
                                                          double* __restrict__ tauout) {
        const int j = threadIdx.x;
        for (int k = 0; k + 2 < N; ++k) {
            put_column(a, sm, k, k);  // rows >= k are read
            const double xj = sm.x[j];  // A[j][k]
            const double alpha = sm.x[k + 1];
            if (j == 0) sm.d[k] = sm.x[k];
            const double xn2 = wave_sum((j > k + 1 && j < N) ? xj * xj : 0.0);
            double tau = 0.0, beta = alpha, scal = 0.0;
            if (xn2 != 0.0) {
                beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                tau = (beta - alpha) * rcp_refined(beta);
                scal = rcp_refined(alpha - beta);
            }
            if (j == 0) sm.e[k] = beta;
            const double vj = (j == k + 1) ? 1.0 : ((j > k + 1 && j < N) ? xj * scal : 0.0);
            if (vout) {
                if (j < N) vout[(size_t)k * N + j] = vj;
                if (j == 0) tauout[k] = tau;
            }
            if (tau != 0.0) {  // uniform over the wave
                sm.v[j] = vj;
                __syncthreads();
                const int b0 = (k + 1) >> 4;  // first 16-row block with active rows
                double s = 0.0;
#pragma unroll
                for (int b = 0; b < NP / 16; ++b) {
                    if (b < b0) continue;
#pragma unroll
                    for (int ii = 0; ii < 16; ++ii) s = fma(a[16 * b + ii], sm.v[16 * b + ii], s);
                }
                const double pj = (j > k && j < N) ? tau * s : 0.0;
                const double coef = 0.5 * tau * wave_sum(pj * vj);
                const double wj = fma(-coef, vj, pj);
                sm.w[j] = wj;
                __syncthreads();
#pragma unroll
                for (int b = 0; b < NP / 16; ++b) {
                    if (b < b0) continue;
#pragma unroll
                    for (int ii = 0; ii < 16; ++ii) {
                        const int i = 16 * b + ii;
                        a[i] = fma(-sm.v[i], wj, fma(-sm.w[i], vj, a[i]));
                    }
                }
            }
            __syncthreads();
        }
        // trailing 2 x 2 (or 1 x 1) block
        if (N >= 2) {
            put_column(a, sm, N - 2, N - 2);
            if (j == 0) {
                sm.d[N - 2] = sm.x[N - 2];
                sm.e[N - 2] = sm.x[N - 1];
            }
            __syncthreads();
            put_column(a, sm, N - 1, N - 1);
            if (j == 0) sm.d[N - 1] = sm.x[N - 1];
            if (vout) {
                if (j < N) {
                    vout[(size_t)(N - 2) * N + j] = 0.0;
                    vout[(size_t)(N - 1) * N + j] = 0.0;
                }
                if (j == 0) { tauout[N - 2] = 0.0; tauout[N - 1] = 0.0; }
            }
        } else {
            put_column(a, sm, 0, 0);
            if (j == 0) sm.d[0] = sm.x[0];
            if (vout && j == 0) { vout[0] = 0.0; tauout[0] = 0.0; }
        }
        __syncthreads();
    }

    __device__ __forceinline__ static double solve(smem& sm, const double* __restrict__ hop, const uint8_t* f_lds,
                                                   int N, double beta, double U, double mu_c, double* ws,
                                                   double* vout, double* tauout, double* cexp, double* cfermi) {
        const int j = threadIdx.x;
        double a[NP];
        // the column's template entries, all loads in flight together (indices clamped,
        // out-of-range entries zeroed afterwards), then the diagonal
        const int jj = min(j, N - 1);
#pragma unroll
        for (int i = 0; i < NP; ++i) a[i] = hop[(size_t)min(i, N - 1) * N + jj];
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            double v = (i < N && j < N) ? a[i] : 0.0;
            if (i == j && i < N) v += -mu_c + U * (double)f_lds[i];
            a[i] = v;
        }
        tridiagonalize(a, sm, N, vout, tauout);
        tridiag_eigvals<NT, 2>(sm.d, sm.e, sm.de, sm.cnt, N, sm.ev, sm.scratch);
        return logz_block<NT>(sm.ev, N, beta, cexp, cfermi, sm.scratch);
    }
};

}  // namespace fkq
