// fk_mixed.hip -- FP64 refinement of eigenvalues from an FP32 Householder reduction.
//
// Stage 1 (fk_kernels.hip, blocked_solver<float>): FP32 reflectors and tridiagonal,
// FP64 tridiagonal eigenvalues.  Stage 2 (fk_vectors.hip): FP64 inverse iteration
// on T and back-transform with the FP32 reflectors, each made exactly orthogonal in
// FP64 (tau = 2 / v^T v), so the basis X is orthonormal to FP64 rounding while it
// approximates the eigenvectors of H only to FP32 accuracy.
//
// This file (Ogita-Aishima style refinement, FP64 throughout):
//   k_mx_hx      Y = H X (H = hop + diag(-mu_c + U f), neighbour-table product)
//   k_mx_gemm    S = X^T Y with v_mfma_f64_16x16x4f64, 32x32 output blocks per
//                wave, lower blocks only (S is symmetric), mirrored on store
//   k_mx_refine  per matrix: S~ = D^-1/2 S D^-1/2 (D = diag X^T X); groups of
//                consecutive Ritz values closer than 1e3 max|S~_ij| (clusters) are
//                diagonalised exactly by Jacobi rotations Q; S' = Q^T S~ Q; then
//                lambda_a = S'_aa + sum_{b outside a's cluster} S'_ab^2 / (S'_aa - S'_bb)
//                (second-order perturbation, error O(|S'_ab|^3 / gap^2)).  A matrix is
//                flagged for the FP64 solver when a cluster exceeds 16 members or a
//                coupling is not small against its gap (|S'_ab| > 5e-3 |S'_aa - S'_bb|).
//                Sorted eigenvalues and the compensated log-sum as on the FP64 path.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_blocked.h"
#include "fk_internal.h"
#include "fk_solve.h"

namespace fkm {

using namespace fks;
using fkb::mf;

__global__ void k_mx_hx(int B, int N, const double* __restrict__ Xg, const int32_t* __restrict__ f,
                        const double* __restrict__ U, const double* __restrict__ mu_c,
                        const int32_t* __restrict__ nbr, const int32_t* __restrict__ z, int zmax,
                        const double* __restrict__ hopv, double* __restrict__ Yg, const int32_t* __restrict__ skip) {
    const int b = blockIdx.y;
    if (skip && skip[b]) return;
    const size_t NN = (size_t)N * N;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= NN) return;
    const int i = (int)(idx / N), k = (int)(idx % N);
    const double* X = Xg + (size_t)b * NN;
    double y = (-mu_c[b] + U[b] * (double)f[(size_t)b * N + i]) * X[idx];
    for (int s = 0; s < z[i]; ++s) y = fma(hopv[(size_t)i * zmax + s], X[(size_t)nbr[(size_t)i * zmax + s] * N + k], y);
    Yg[(size_t)b * NN + idx] = y;
}

// S = X^T Y: wave w of the workgroup computes the 32 x 32 block (I, J), I >= J, of the
// lower block triangle; the four 16 x 16 MFMA tiles share their operand loads.
constexpr int GW = 4;  // waves per workgroup
__global__ void __launch_bounds__(GW * WAVE) k_mx_gemm(int B, int N, const double* __restrict__ Xg,
                                                        const double* __restrict__ Yg, double* __restrict__ Sg,
                                                        const int32_t* __restrict__ skip) {
    const int b = blockIdx.y;
    if (skip && skip[b]) return;
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int nb = (N + 31) / 32;
    const int blk = blockIdx.x * GW + wid;
    if (blk >= nb * (nb + 1) / 2) return;
    // lower-triangle block index -> (I, J), I >= J
    int I = (int)((sqrt(8.0 * blk + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= blk) ++I;
    while (I * (I + 1) / 2 > blk) --I;
    const int J = blk - I * (I + 1) / 2;
    const size_t NN = (size_t)N * N;
    const double* X = Xg + (size_t)b * NN;
    const double* Y = Yg + (size_t)b * NN;
    double* S = Sg + (size_t)b * NN;
    const int li = lane & 15, kq = lane >> 4;
    const int ra0 = 32 * I + li, ra1 = ra0 + 16;  // S rows (columns of X)
    const int cb0 = 32 * J + li, cb1 = cb0 + 16;  // S columns (columns of Y)
    mf<double>::acc_t c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    for (int r0 = 0; r0 < N; r0 += 4) {
        const int r = r0 + kq;
        double a0 = 0, a1 = 0, y0 = 0, y1 = 0;
        if (r < N) {
            const double* xr = X + (size_t)r * N;
            const double* yr = Y + (size_t)r * N;
            if (ra0 < N) a0 = xr[ra0];
            if (ra1 < N) a1 = xr[ra1];
            if (cb0 < N) y0 = yr[cb0];
            if (cb1 < N) y1 = yr[cb1];
        }
        c00 = mf<double>::op(a0, y0, c00);
        c01 = mf<double>::op(a0, y1, c01);
        c10 = mf<double>::op(a1, y0, c10);
        c11 = mf<double>::op(a1, y1, c11);
    }
    auto store = [&](const mf<double>::acc_t& acc, int r0b, int c0b) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = r0b + mf<double>::row(lane, q), col = c0b + li;
            if (row < N && col < N) {
                S[(size_t)row * N + col] = acc[q];
                S[(size_t)col * N + row] = acc[q];  // symmetric mirror (same value on the diagonal)
            }
        }
    };
    store(c00, 32 * I, 32 * J);
    store(c01, 32 * I, 32 * J + 16);
    store(c10, 32 * I + 16, 32 * J);
    store(c11, 32 * I + 16, 32 * J + 16);
}

constexpr int RT = 512;       // threads of the refinement kernel
constexpr int RW = RT / WAVE;
constexpr int MC = 16;        // largest cluster diagonalised here
constexpr int RNMAX = 1024;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ double block_max(double v, double* red) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    v = wave_max(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double m = red[0];
    for (int w = 1; w < RW; ++w) m = fmax(m, red[w]);
    return m;
}

// Cyclic Jacobi on the m x m symmetric block a (row stride MC), rotations accumulated in q.
__device__ void wave_jacobi(double* a, double* q, int m) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int e = lane; e < m * m; e += WAVE) {
            const int r = e / m, c = e % m;
            const double v = a[r * MC + c];
            if (r < c) off += v * v;
            if (r == c) dia += v * v;
        }
        off = wave_sum(off);
        dia = wave_sum(dia);
        if (off <= 1e-34 * dia || off == 0.0) break;
        for (int p = 0; p < m - 1; ++p)
            for (int r = p + 1; r < m; ++r) {
                const double apq = a[p * MC + r];
                const double app = a[p * MC + p], aqq = a[r * MC + r];
                wave_sync();
                if (apq == 0.0) continue;
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                if (lane < m && lane != p && lane != r) {
                    const double akp = a[lane * MC + p], akq = a[lane * MC + r];
                    const double nkp = c * akp - s * akq, nkq = s * akp + c * akq;
                    a[lane * MC + p] = nkp;
                    a[p * MC + lane] = nkp;
                    a[lane * MC + r] = nkq;
                    a[r * MC + lane] = nkq;
                }
                if (lane < m) {
                    const double qkp = q[lane * MC + p], qkq = q[lane * MC + r];
                    q[lane * MC + p] = c * qkp - s * qkq;
                    q[lane * MC + r] = s * qkp + c * qkq;
                }
                if (lane == 0) {
                    a[p * MC + p] = app - t * apq;
                    a[r * MC + r] = aqq + t * apq;
                    a[p * MC + r] = 0.0;
                    a[r * MC + p] = 0.0;
                }
                wave_sync();
            }
    }
}

// One workgroup per matrix.  S is overwritten by Q^T S~ Q; T (N x N) and Qg (N x MC)
// are scratch.  fb_skip[b] = 0 asks for the FP64 solver (flagged), 1 otherwise.
__global__ void __launch_bounds__(RT) k_mx_refine(int B, int N, const double* __restrict__ Xg, double* __restrict__ Sg,
                                                  double* __restrict__ Tg, double* __restrict__ Qg,
                                                  const double* __restrict__ beta, double* __restrict__ evals,
                                                  double* __restrict__ logz, int32_t* __restrict__ fb_skip,
                                                  const int32_t* __restrict__ skip,
                                                  unsigned long long* __restrict__ nflag) {
    __shared__ double rg[RNMAX], kd[RNMAX], lam[RNMAX];
    __shared__ int cs[RNMAX], cm[RNMAX], clist[RNMAX];
    __shared__ double jac[RW][2][MC * MC];
    __shared__ double red[RW + 2];
    __shared__ double scratch[3 * RW + 8];
    __shared__ int nclus_s, flag_s;
    const int b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    if (skip && skip[b]) {
        if (tid == 0) fb_skip[b] = 1;
        return;
    }
    const size_t NN = (size_t)N * N;
    const double* X = Xg + (size_t)b * NN;
    double* S = Sg + (size_t)b * NN;
    double* T = Tg + (size_t)b * NN;
    double* Q = Qg + (size_t)b * N * MC;
    // normalisation D = diag(X^T X) and the Ritz values
    for (int i = tid; i < N; i += RT) {
        double g = 0.0;
        for (int r = 0; r < N; ++r) g = fma(X[(size_t)r * N + i], X[(size_t)r * N + i], g);
        rg[i] = 1.0 / sqrt(g);
    }
    if (tid == 0) { nclus_s = 0; flag_s = 0; }
    __syncthreads();
    for (int i = tid; i < N; i += RT) kd[i] = S[(size_t)i * N + i] * rg[i] * rg[i];
    double sm = 0.0;
    for (int j = tid; j < N; j += RT)
        for (int i = 0; i < N; ++i)
            if (i != j) sm = fmax(sm, fabs(S[(size_t)i * N + j]) * rg[i] * rg[j]);
    const double smax = block_max(sm, red);
    const double delta = fmax(1e3 * smax, 1e-300);
    // clusters: runs of consecutive Ritz values closer than delta
    if (tid == 0) {
        int start = 0, nc = 0, fl = 0;
        for (int i = 1; i <= N; ++i) {
            if (i == N || fabs(kd[i] - kd[i - 1]) >= delta) {
                const int m = i - start;
                for (int j = start; j < i; ++j) { cs[j] = start; cm[j] = m; }
                if (m > MC) fl = 1;
                else if (m >= 2) clist[nc++] = start;
                start = i;
            }
        }
        nclus_s = nc;
        flag_s = fl;
    }
    __syncthreads();
    const int nclus = nclus_s;
    // exact diagonalisation of each cluster block (one wave per cluster)
    for (int ci = wid; ci < nclus; ci += RW) {
        const int c0 = clist[ci], m = cm[c0];
        double* a = jac[wid][0];
        double* q = jac[wid][1];
        for (int e = lane; e < m * m; e += WAVE) {
            const int r = e / m, c = e % m;
            a[r * MC + c] = S[(size_t)(c0 + r) * N + c0 + c] * rg[c0 + r] * rg[c0 + c];
            q[r * MC + c] = (r == c) ? 1.0 : 0.0;
        }
        wave_sync();
        wave_jacobi(a, q, m);
        wave_sync();
        for (int e = lane; e < m * m; e += WAVE) {
            const int r = e / m, c = e % m;
            Q[(size_t)(c0 + r) * MC + c] = q[r * MC + c];
        }
    }
    __threadfence_block();
    __syncthreads();
    if (nclus > 0) {
        // T = S~ Q (cluster columns rotated), then S <- Q^T T (cluster rows rotated)
        for (size_t idx = tid; idx < NN; idx += RT) {
            const int i = (int)(idx / N), be = (int)(idx % N);
            const int c = cs[be], m = cm[be];
            double v;
            if (m == 1 || m > MC) {
                v = S[idx] * rg[i] * rg[be];
            } else {
                v = 0.0;
                for (int j = c; j < c + m; ++j) v = fma(S[(size_t)i * N + j] * rg[i] * rg[j], Q[(size_t)j * MC + (be - c)], v);
            }
            T[idx] = v;
        }
        __threadfence_block();
        __syncthreads();
        for (size_t idx = tid; idx < NN; idx += RT) {
            const int al = (int)(idx / N), be = (int)(idx % N);
            const int c = cs[al], m = cm[al];
            double v;
            if (m == 1 || m > MC) {
                v = T[idx];
            } else {
                v = 0.0;
                for (int i = c; i < c + m; ++i) v = fma(Q[(size_t)i * MC + (al - c)], T[(size_t)i * N + be], v);
            }
            S[idx] = v;
        }
        __threadfence_block();
        __syncthreads();
        for (int i = tid; i < N; i += RT) { kd[i] = S[(size_t)i * N + i]; rg[i] = 1.0; }
        __syncthreads();
    }
    // second-order correction (S' symmetric: column access, coalesced over a)
    int fl = 0;
    for (int al = tid; al < N; al += RT) {
        const double saa = kd[al];
        const int c = cs[al], m = cm[al];
        double corr = 0.0;
        for (int be = 0; be < N; ++be) {
            if (be >= c && be < c + m) continue;
            const double sab = S[(size_t)be * N + al] * rg[al] * rg[be];
            const double gap = saa - kd[be];
            if (sab == 0.0) continue;
            if (fabs(sab) > 5e-3 * fabs(gap)) { fl = 1; continue; }
            corr += sab * sab / gap;
        }
        lam[al] = saa + corr;
    }
    if (fl) atomicOr(&flag_s, 1);
    // ascending order (bitonic sort over RNMAX slots)
    for (int i = N + tid; i < RNMAX; i += RT) lam[i] = INFINITY;
    __syncthreads();
    for (int k = 2; k <= RNMAX; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < RNMAX; i += RT) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const double x = lam[i], y = lam[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { lam[i] = y; lam[ixj] = x; }
                }
            }
            __syncthreads();
        }
    const double lz = logz_block<RT>(lam, N, beta[b], nullptr, nullptr, scratch);
    for (int i = tid; i < N; i += RT) evals[(size_t)b * N + i] = lam[i];
    if (tid == 0) {
        logz[b] = lz;
        fb_skip[b] = flag_s ? 0 : 1;
        if (flag_s) atomicAdd(nflag, 1ull);
    }
}

}  // namespace fkm

hipError_t fk_launch_mixed_refine(int B, int N, const double* X, const int32_t* f, const double* U,
                                  const double* mu_c, const int32_t* nbr, const int32_t* z, int zmax,
                                  const double* hopv, double* Y, double* S, double* T, double* Q,
                                  const double* beta, double* evals, double* logz, int32_t* fb_skip,
                                  const int32_t* skip, unsigned long long* nflag, hipStream_t st) {
    if (N > fkm::RNMAX) return hipErrorInvalidValue;
    const size_t NN = (size_t)N * N;
    hipLaunchKernelGGL(fkm::k_mx_hx, dim3((unsigned)((NN + 255) / 256), B), dim3(256), 0, st, B, N, X, f, U, mu_c,
                       nbr, z, zmax, hopv, Y, skip);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    const int nb = (N + 31) / 32;
    const int nblk = nb * (nb + 1) / 2;
    hipLaunchKernelGGL(fkm::k_mx_gemm, dim3((nblk + fkm::GW - 1) / fkm::GW, B), dim3(fkm::GW * fks::WAVE), 0, st, B, N, X,
                       Y, S, skip);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    hipLaunchKernelGGL(fkm::k_mx_refine, dim3(B), dim3(fkm::RT), 0, st, B, N, X, S, T, Q, beta, evals, logz,
                       fb_skip, skip, nflag);
    return hipGetLastError();
}
int fk_mixed_cluster_max() { return fkm::MC; }
