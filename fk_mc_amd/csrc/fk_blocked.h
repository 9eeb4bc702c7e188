// fk_blocked.h -- blocked Householder tridiagonalisation with MFMA trailing
// updates (LAPACK sytrd/latrd organisation, lower storage convention), one
// workgroup per matrix, the matrix in a per-workgroup HBM workspace.
//
// For a panel of nb columns starting at k0 the reflectors are built one column
// at a time against the stored trailing matrix corrected by the panel's earlier
// reflectors (V, W kept in LDS); the matrix is then updated once per panel,
//     A22 <- A22 - V W^T - W V^T,
// as 16x16 MFMA tiles with K = 2 nb (f32: v_mfma_f32_16x16x4f32, f64:
// v_mfma_f64_16x16x4f64).  The symmetric matrix-vector product of each column
// reads the stored matrix by rows (A[r][i], r the summation index), which is
// coalesced because the stored matrix is kept full and symmetric.
//
// T = double: the FP64 solver for 256 < N <= 1024.  T = float: the first stage
// of the mixed-precision path (FP32 reflectors and tridiagonal, FP64 tridiagonal
// eigenvalues).  The tridiagonal (d, e) is always held in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_solve.h"

namespace fkb {

using namespace fks;

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: lane l supplies A[l%16][l/16] and B[l/16][l%16]; accumulator
// register r of lane l is D[row(l, r)][l%16] (the f32 and f64 instructions
// order the rows differently).
template <class T> struct mf;
template <> struct mf<float> {
    typedef v4f acc_t;
    __device__ __forceinline__ static acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ __forceinline__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};
template <> struct mf<double> {
    typedef v4d acc_t;
    __device__ __forceinline__ static acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ __forceinline__ static int row(int l, int r) { return (l >> 4) + 4 * r; }
};

__host__ __device__ constexpr int blk_ld(int N) { return (N + 15) & ~15; }

// NT_: threads of the workgroup (the waves per matrix); NBCAP: largest panel width.
template <class T, int NT_, int NBCAP = 32>
struct blocked_solver {
    static constexpr int NT = NT_;
    static constexpr int NW = NT / WAVE;
    static constexpr int NMAX = 1024;
    static constexpr int PANEL_BYTES = 96 * 1024;
    static constexpr int PANEL_ELEMS = PANEL_BYTES / (2 * (int)sizeof(T));  // per panel array
    static constexpr int YPART = 4 * NT;                                      // symv partial sums
    struct eig_area {
        double de[2 * NMAX];
        double ev[NMAX];
        int cnt[NT + 2];
        double scratch[3 * NW + 8];
    };
    struct smem {
        union {
            struct { T V[PANEL_ELEMS]; T W[PANEL_ELEMS]; } p;
            eig_area e;
        } u;
        T x[NMAX];          // current column, then y = A v
        T ypart[YPART];
        double d[NMAX], e[NMAX];
        T corr[64];         // W^T v and V^T v of the panel's earlier columns
        double red[NW + 2];
    };
    static_assert(sizeof(eig_area) <= 2 * PANEL_ELEMS * sizeof(T), "eigenvalue area must fit in the panel area");

    // panel width for a given N (even, so that K = 2 nb is a multiple of 4)
    __host__ __device__ static int panel_width(int N) {
        int nb = PANEL_ELEMS / blk_ld(N);
        if (nb > NBCAP) nb = NBCAP;
        return nb & ~1;
    }
    // workspace elements (double units) per workgroup
    __host__ __device__ static size_t ws_doubles(int N) {
        const size_t ld = blk_ld(N);
        return (ld * ld * sizeof(T) + sizeof(double) - 1) / sizeof(double);
    }
    __device__ static const double* dptr(const smem& sm) { return sm.d; }
    __device__ static const double* eptr(const smem& sm) { return sm.e; }
    __device__ static double* evptr(smem& sm) { return sm.u.e.ev; }

    __device__ __forceinline__ static double block_sum_d(double v, smem& sm) {
        const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
        double s = wave_sum(v);
        __syncthreads();
        if (lane == 0) sm.red[wid] = s;
        __syncthreads();
        double t = 0.0;
        for (int w = 0; w < NW; ++w) t += sm.red[w];
        return t;
    }
    __device__ __forceinline__ static T block_sum(T v, smem& sm) {
        const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
        double s = wave_sum((double)v);
        __syncthreads();
        if (lane == 0) sm.red[wid] = s;
        __syncthreads();
        double t = 0.0;
        for (int w = 0; w < NW; ++w) t += sm.red[w];
        return (T)t;
    }

    // Householder tridiagonalisation of the ld x ld workspace A (rows >= N zero).
    // Writes d, e (FP64) and, when vout != null, the reflectors (vout[c*N + i],
    // tauout[c], the layout of the eigenvector back-transform).
    __device__ static void tridiagonalize(smem& sm, T* __restrict__ A, int N, double* vout, double* tauout) {
        const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
        const int ld = blk_ld(N);
        const int nbmax = panel_width(N);
        T* __restrict__ Vp = sm.u.p.V;
        T* __restrict__ Wp = sm.u.p.W;
        const int Qp = ld / 4;        // column quads of the symmetric product
        const int G = NT / Qp;        // row groups
        const int q = tid % Qp, g = tid / Qp;
        for (int k0 = 0; k0 + 1 < N; k0 += nbmax) {
            const int jn = min(nbmax, N - 1 - k0);
            for (int j = 0; j < jn; ++j) {
                const int c = k0 + j;
                // (1) column c (= row c, the stored matrix is symmetric) corrected by the panel so far
                for (int i = c + tid; i < N; i += NT) {
                    T x = A[(size_t)c * ld + i];
                    for (int t = 0; t < j; ++t)
                        x -= Vp[t * ld + i] * Wp[t * ld + c] + Wp[t * ld + i] * Vp[t * ld + c];
                    sm.x[i] = x;
                }
                __syncthreads();
                // (2) reflector (LAPACK larfg) annihilating x[c+2 .. N-1]
                T part = 0;
                for (int i = c + 2 + tid; i < N; i += NT) part += sm.x[i] * sm.x[i];
                const T sigma = block_sum(part, sm);
                const T alpha = sm.x[c + 1];
                T tau = 0, beta_h = alpha, scal = 0;
                if (sigma != (T)0) {
                    beta_h = -copysign(sqrt(alpha * alpha + sigma), alpha);
                    tau = (beta_h - alpha) / beta_h;
                    scal = (T)1 / (alpha - beta_h);
                }
                if (tid == 0) {
                    sm.d[c] = (double)sm.x[c];
                    sm.e[c] = (double)beta_h;
                }
                double vv = 0.0;
                for (int i = tid; i < ld; i += NT) {
                    const T v = (i == c + 1) ? (T)1 : ((i > c + 1 && i < N) ? sm.x[i] * scal : (T)0);
                    Vp[j * ld + i] = v;
                    if (vout && i < N) {
                        vout[(size_t)c * N + i] = (double)v;
                        vv += (double)v * (double)v;
                    }
                }
                if (vout) {
                    // Reflectors for the FP64 back-transform.  In FP32 the factor is
                    // recomputed as 2 / (v^T v) in FP64 so that I - tau v v^T is orthogonal
                    // to FP64 rounding (the basis then only approximates the eigenvectors to
                    // FP32 accuracy, which the FP64 refinement corrects).
                    if (sizeof(T) == sizeof(double)) {
                        if (tid == 0) tauout[c] = (double)tau;
                    } else {
                        const double vvs = block_sum_d(vv, sm);
                        if (tid == 0) tauout[c] = (tau == (T)0) ? 0.0 : 2.0 / vvs;
                    }
                }
                __syncthreads();
                if (tau == (T)0) {
                    for (int i = tid; i < ld; i += NT) Wp[j * ld + i] = 0;
                    __syncthreads();
                    continue;
                }
                // (3) y = A22 v over rows r >= c+1 (partial sums per row group)
                if (g < G) {
                    const int i0 = 4 * q;
                    T acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
                    if (i0 + 3 >= c + 1) {
                        int r = c + 1 + g;
                        for (; r + 3 * G < N; r += 4 * G) {
                            T vr[4], w[4][4];
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const int rr = r + u * G;
                                vr[u] = Vp[j * ld + rr];
                                const T* row = A + (size_t)rr * ld + i0;
#pragma unroll
                                for (int s = 0; s < 4; ++s) w[u][s] = row[s];
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                acc0 += vr[u] * w[u][0];
                                acc1 += vr[u] * w[u][1];
                                acc2 += vr[u] * w[u][2];
                                acc3 += vr[u] * w[u][3];
                            }
                        }
                        for (; r < N; r += G) {
                            const T vr = Vp[j * ld + r];
                            const T* row = A + (size_t)r * ld + i0;
                            acc0 += vr * row[0];
                            acc1 += vr * row[1];
                            acc2 += vr * row[2];
                            acc3 += vr * row[3];
                        }
                    }
                    T* yp = sm.ypart + g * ld + i0;
                    yp[0] = acc0; yp[1] = acc1; yp[2] = acc2; yp[3] = acc3;
                }
                // (4a) W^T v and V^T v of the earlier panel columns
                for (int dd = wid; dd < 2 * j; dd += NW) {
                    const T* col = (dd < j) ? (Wp + dd * ld) : (Vp + (dd - j) * ld);
                    double s = 0.0;
                    for (int i = c + 1 + lane; i < N; i += WAVE) s += (double)(col[i] * Vp[j * ld + i]);
                    s = wave_sum(s);
                    if (lane == 0) sm.corr[dd] = (T)s;
                }
                __syncthreads();
                // (4b) y = sum of the partials - V (W^T v) - W (V^T v); then w = tau y
                T pw = 0;
                for (int i = c + 1 + tid; i < N; i += NT) {
                    T y = 0;
                    for (int gg = 0; gg < G; ++gg) y += sm.ypart[gg * ld + i];
                    for (int t = 0; t < j; ++t) y -= Vp[t * ld + i] * sm.corr[t] + Wp[t * ld + i] * sm.corr[j + t];
                    y *= tau;
                    sm.x[i] = y;
                    pw += y * Vp[j * ld + i];
                }
                // (5) w = y - (tau/2)(y^T v) v
                const T a2 = (T)(-0.5) * tau * block_sum(pw, sm);
                for (int i = tid; i < ld; i += NT)
                    Wp[j * ld + i] = (i >= c + 1 && i < N) ? sm.x[i] + a2 * Vp[j * ld + i] : (T)0;
                __syncthreads();
            }
            // trailing update A[i][l] -= sum_t V[i][t] W[l][t] + W[i][t] V[l][t] for the
            // 16x16 tiles covering [t0, ld)^2 (rows below k0 + jn are finished and not read again)
            const int cend = k0 + jn;
            if (cend < N) {
                const int t0 = cend & ~15;
                const int nt = (ld - t0) / 16;
                const int K4 = (2 * jn + 3) / 4;
                const int li = lane & 15, kq = lane >> 4;
                for (int tt = wid; tt < nt * nt; tt += NW) {
                    const int I0 = t0 + 16 * (tt / nt), J0 = t0 + 16 * (tt % nt);
                    typename mf<T>::acc_t acc;
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[r] = A[(size_t)(I0 + mf<T>::row(lane, r)) * ld + J0 + li];
                    for (int kk = 0; kk < K4; ++kk) {
                        const int kx = 4 * kk + kq;
                        T av = 0, bv = 0;
                        if (kx < jn) {
                            av = -Vp[kx * ld + I0 + li];
                            bv = Wp[kx * ld + J0 + li];
                        } else if (kx < 2 * jn) {
                            av = -Wp[(kx - jn) * ld + I0 + li];
                            bv = Vp[(kx - jn) * ld + J0 + li];
                        }
                        acc = mf<T>::op(av, bv, acc);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) A[(size_t)(I0 + mf<T>::row(lane, r)) * ld + J0 + li] = acc[r];
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            sm.d[N - 1] = (double)A[(size_t)(N - 1) * ld + (N - 1)];
            sm.e[N - 1] = 0.0;
            if (tauout) tauout[N - 1] = 0.0;
        }
        if (vout)
            for (int i = tid; i < N; i += NT) vout[(size_t)(N - 1) * N + i] = 0.0;
        __syncthreads();
    }

    // Assemble H = hop + diag(-mu_c + U f) into the workspace, tridiagonalise,
    // FP64 tridiagonal eigenvalues into sm.u.e.ev, return logZ.
    __device__ __forceinline__ static double solve(smem& sm, const double* __restrict__ hop, const uint8_t* f_lds,
                                                   int N, double beta, double U, double mu_c, double* ws,
                                                   double* vout, double* tauout, double* cexp, double* cfermi) {
        const int tid = threadIdx.x;
        const int ld = blk_ld(N);
        T* __restrict__ A = reinterpret_cast<T*>(ws);
        const size_t tot = (size_t)ld * ld;
        for (size_t idx = tid; idx < tot; idx += NT) {
            const int i = (int)(idx / ld), j = (int)(idx % ld);
            double v = 0.0;
            if (i < N && j < N) {
                v = hop[(size_t)i * N + j];
                if (i == j) v += -mu_c + U * (double)f_lds[i];
            }
            A[idx] = (T)v;
        }
        __syncthreads();
        tridiagonalize(sm, A, N, vout, tauout);
        tridiag_eigvals<NT>(sm.d, sm.e, sm.u.e.de, sm.u.e.cnt, N, sm.u.e.ev, sm.u.e.scratch);
        return logz_block<NT>(sm.u.e.ev, N, beta, cexp, cfermi, sm.u.e.scratch);
    }
};

}  // namespace fkb
