// fk_update.hip -- Metropolis moves by rank-one spectral updates.
//
// A single-site move changes H by rho e_s e_s^T (rho = +-U), an exchange of two
// sites by two such terms.  With the chain's eigendecomposition H = V diag(d) V^T
// kept on the device, the proposal's eigenvalues are those of
// diag(d) + rho z z^T, z = V^T e_s (row s of V): the roots of the secular equation
//     1 + rho sum_k z_k^2 / (d_k - lambda) = 0,
// one in each interval between consecutive poles (Golub 1973; Bunch, Nielsen and
// Sorensen 1978).  Deflation follows LAPACK dlaed2 (|rho z_k| <= tol, and a Givens
// rotation for pairs of poles closer than the tolerance); the eigenvectors of the
// update are taken from the Gu-Eisenstat (1994) recomputed z-hat so that they are
// orthogonal to working precision.  The proposal costs O(N^2); only an accepted
// move updates V <- V Q, as an MFMA GEMM (k_fu_gemm).  A second term (exchange
// moves) uses z' = Q1^T (V^T e_s2).
//
//   k_fu_step   one workgroup per chain: proposal (reference draw order), one or
//               two rank-one updates, logZ, Metropolis decision, bookkeeping; the
//               update matrices Q1 (Q2) go to HBM for the accepted chains
//   k_fu_gemm   T = V Q1, then V = T Q2 (or V = T), 64 x 64 output blocks per
//               workgroup, K staged through LDS, v_mfma_f64_16x16x4f64
//   k_fu_f8to32 occupations as int32 for the periodic full re-decomposition
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_blocked.h"
#include "fk_internal.h"
#include "fk_moves.h"
#include "fk_solve.h"

namespace fku {

using namespace fks;
using fkk::decide_move;
using fkk::propose_move;
using fkk::record_move;

constexpr int UT = 256;
constexpr int UW = UT / WAVE;
constexpr int UNMAX = 1024;

template <int UNMAX>
struct usmem {
    double d[UNMAX];               // current eigenvalues, ascending
    double dv[UNMAX], zv[UNMAX];   // working poles and weights (sign-mapped)
    double z2[UNMAX];              // weights of the second term
    double da[UNMAX], za[UNMAX], zh[UNMAX];  // active (non-deflated) poles, weights, z-hat
    double mu[UNMAX], rn[UNMAX];   // roots relative to their origin pole, 1/|eigenvector|
    double key[UNMAX];
    double rc[UNMAX], rs[UNMAX];   // Givens rotations of the deflation
    int org[UNMAX], act[UNMAX], cidx[UNMAX], src[UNMAX], rp[UNMAX], rk[UNMAX], dl[UNMAX];
    double red[UW + 8];
    double scratch[3 * UW + 8];
    int m, nrot, nd;
    uint8_t f[UNMAX];
    uint32_t mt[fkrng::MT_WORDS];
    move_prop prop;
    int acc;
};

__device__ double block_max(double v, double* red) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    v = wave_max(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < UW; ++w) r = fmax(r, red[w]);
    __syncthreads();
    return r;
}
__device__ double block_sum(double v, double* red) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < UW; ++w) r += red[w];
    __syncthreads();
    return r;
}

// Chunked reads of two LDS arrays (indices clamped to the last valid entry, no branches), so
// that a chunk's loads are in flight together before the terms that use them.
constexpr int RC = 8;
__device__ __forceinline__ void load_chunk(const double* a, const double* b, int k0, int m, double (&ab)[RC],
                                           double (&bb)[RC]) {
#pragma unroll
    for (int q = 0; q < RC; ++q) {
        const int k = min(k0 + q, m - 1);
        ab[q] = a[k];
        bb[q] = b[k];
    }
}

// Eigen-update of diag(sm.d) + rho z z^T.  On return sm.d holds the new eigenvalues
// in ascending order and Q (N x N, row-major, HBM) the orthogonal matrix with
// diag(d) + rho z z^T = Q diag(d_new) Q^T.  All threads of the workgroup call it.
// Reciprocals are v_rcp_f64 refined by Newton steps (full double precision).
template <class SM>
__device__ void rank1(SM& sm, const double* z, double rho, int N, double* __restrict__ Q) {
    const int tid = threadIdx.x;
    const double eps = 2.220446049250313e-16;
    const bool neg = rho < 0.0;
    const double rh = fabs(rho);
    // v-space: index t <-> eigen index ix(t) = neg ? N-1-t : t, poles negated when rho < 0
    for (int t = tid; t < N; t += UT) {
        const int o = neg ? N - 1 - t : t;
        sm.dv[t] = neg ? -sm.d[o] : sm.d[o];
        sm.zv[t] = z[o];
    }
    __syncthreads();
    double dm = 0.0, zm = 0.0;
    for (int t = tid; t < N; t += UT) {
        dm = fmax(dm, fabs(sm.dv[t]));
        zm = fmax(zm, fabs(sm.zv[t]));
    }
    dm = block_max(dm, sm.red);
    zm = block_max(zm, sm.red);
    const double tol = 8.0 * eps * fmax(dm, rh * zm);
    // deflation (LAPACK dlaed2): small weights, then close poles by Givens rotations
    if (tid == 0) {
        int m = 0, nrot = 0, nd = 0, p = -1;
        for (int t = 0; t < N; ++t) {
            if (rh * fabs(sm.zv[t]) <= tol) {
                sm.cidx[t] = -1;
                sm.dl[nd++] = t;
                continue;
            }
            if (p >= 0) {
                const double tau = sqrt(sm.zv[p] * sm.zv[p] + sm.zv[t] * sm.zv[t]);  // |z| <= 1: no overflow
                const double c = sm.zv[t] / tau, s = -sm.zv[p] / tau;
                const double dt = sm.dv[t] - sm.dv[p];
                if (fabs(dt * c * s) <= tol) {
                    sm.rp[nrot] = p;
                    sm.rk[nrot] = t;
                    sm.rc[nrot] = c;
                    sm.rs[nrot] = s;
                    ++nrot;
                    const double dp = sm.dv[p] * c * c + sm.dv[t] * s * s;
                    const double dk = sm.dv[p] * s * s + sm.dv[t] * c * c;
                    sm.dv[p] = dp;
                    sm.dv[t] = dk;
                    sm.zv[p] = 0.0;
                    sm.zv[t] = tau;
                    --m;  // p was the last active entry
                    sm.cidx[p] = -1;
                    sm.dl[nd++] = p;
                }
            }
            sm.cidx[t] = m;
            sm.act[m++] = t;
            p = t;
        }
        sm.m = m;
        sm.nrot = nrot;
        sm.nd = nd;
    }
    __syncthreads();
    const int m = sm.m;
    for (int k = tid; k < m; k += UT) {
        sm.da[k] = sm.dv[sm.act[k]];
        sm.za[k] = sm.zv[sm.act[k]];
    }
    __syncthreads();
    double zz = 0.0;
    for (int k = tid; k < m; k += UT) zz += sm.za[k] * sm.za[k];
    zz = block_sum(zz, sm.red);
    const double irh = 1.0 / rh;
    // secular roots: root j lies in (da_j, da_{j+1}), the last in (da_{m-1}, da_{m-1} + rho |z|^2];
    // it is sought relative to the nearer pole o as mu = lambda - da_o, by Newton's method on
    // the pole-free function gh(mu) = -mu (1/rho + sum_{k != o} za_k^2 / (da_k - da_o - mu)) + za_o^2
    // safeguarded by the sign of the secular function (bisection when a step leaves the bracket).
    for (int j = tid; j < m; j += UT) {
        int o;
        double lo, hi;
        if (j < m - 1) {
            const double half = 0.5 * (sm.da[j + 1] - sm.da[j]);
            double g = irh;
            const double dj = sm.da[j];
            for (int k0 = 0; k0 < m; k0 += RC) {
                double db[RC], zb[RC];
                load_chunk(sm.da, sm.za, k0, m, db, zb);
#pragma unroll
                for (int q = 0; q < RC; ++q)
                    if (k0 + q < m) g += zb[q] * zb[q] * rcp_refined((db[q] - dj) - half);
            }
            if (g >= 0.0) { o = j; lo = 0.0; hi = half; }
            else { o = j + 1; lo = -half; hi = 0.0; }
        } else {
            o = m - 1;
            lo = 0.0;
            hi = rh * zz;
        }
        const double zo2 = sm.za[o] * sm.za[o];
        double x = 0.5 * (lo + hi);
        for (int it = 0; it < 64; ++it) {
            double S = 0.0, S1 = 0.0, Sa = 0.0;
            const double dao = sm.da[o];
            for (int k0 = 0; k0 < m; k0 += RC) {
                double db[RC], zb[RC];
                load_chunk(sm.da, sm.za, k0, m, db, zb);
#pragma unroll
                for (int q = 0; q < RC; ++q) {
                    const int k = k0 + q;
                    if (k >= m || k == o) continue;
                    // pole distances relative to the origin pole first (exact for near poles)
                    const double qq = zb[q] * rcp_refined((db[q] - dao) - x);
                    S += zb[q] * qq;
                    Sa += fabs(zb[q] * qq);
                    S1 += qq * qq;
                }
            }
            const double base = irh + S;
            const double gh = -x * base + zo2;
            // |gh| at the level of its rounding error: x is a root to working precision
            if (fabs(gh) <= 16.0 * eps * (fabs(x) * (irh + Sa) + zo2)) break;
            const double g = base - zo2 / x;  // secular function (x != 0 inside the bracket)
            if (g > 0.0) hi = x; else lo = x;
            const double dgh = -base - x * S1;
            double xn = (dgh != 0.0) ? x - gh / dgh : 0.5 * (lo + hi);
            if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
            const bool done = fabs(xn - x) <= 4.0 * eps * fabs(xn) || hi - lo <= 4.0 * eps * fmax(fabs(lo), fabs(hi));
            x = xn;
            if (done) break;
        }
        sm.mu[j] = x;
        sm.org[j] = o;
    }
    __syncthreads();
    // Gu-Eisenstat: zh_k^2 = (lambda_{m-1} - da_k) prod_{j<k} (lambda_j - da_k)/(da_j - da_k)
    //                        prod_{j>k} (lambda_{j-1} - da_k)/(da_j - da_k) / rho
    // (numerator and denominator multiplied in groups of 8 factors, one division per group)
    for (int k = tid; k < m; k += UT) {
        const double dk = sm.da[k];
        double pr = (sm.da[sm.org[m - 1]] - dk) + sm.mu[m - 1];
        double nu = 1.0, de = 1.0;
        int cnt = 0;
        for (int j = 0; j < m; ++j) {
            if (j == k) continue;
            const int jn = j < k ? j : j - 1;  // numerator root paired with pole j
            nu *= (sm.da[sm.org[jn]] - dk) + sm.mu[jn];
            de *= sm.da[j] - dk;
            if (++cnt == 8) { pr *= nu / de; nu = de = 1.0; cnt = 0; }
        }
        pr *= nu / de;
        sm.zh[k] = copysign(sqrt(fmax(pr * irh, 0.0)), sm.za[k]);
    }
    __syncthreads();
    for (int j = tid; j < m; j += UT) {
        const double dorg = sm.da[sm.org[j]], mj = sm.mu[j];
        double s = 0.0;
        for (int k = 0; k < m; ++k) {
            const double q = sm.zh[k] * rcp_refined((sm.da[k] - dorg) - mj);
            s += q * q;
        }
        sm.rn[j] = 1.0 / sqrt(s);
    }
    // new eigenvalues (roots and deflated poles), ascending in the original sign
    int P2 = 1;
    while (P2 < N) P2 <<= 1;
    for (int q = tid; q < P2; q += UT) {
        double v = INFINITY;
        int sr = 0x7fffffff;
        if (q < m) {
            v = sm.da[sm.org[q]] + sm.mu[q];
            sr = q;
        } else if (q < N) {
            const int t = sm.dl[q - m];
            v = sm.dv[t];
            sr = -(t + 1);
        }
        if (q < N && neg) v = -v;
        sm.key[q] = v;
        sm.src[q] = sr;
    }
    __syncthreads();
    for (int kk = 2; kk <= P2; kk <<= 1)
        for (int jj = kk >> 1; jj > 0; jj >>= 1) {
            for (int i = tid; i < P2; i += UT) {
                const int ixj = i ^ jj;
                if (ixj > i) {
                    const double a = sm.key[i], b = sm.key[ixj];
                    const bool up = (i & kk) == 0;
                    if ((a > b) == up) {
                        sm.key[i] = b;
                        sm.key[ixj] = a;
                        const int t = sm.src[i];
                        sm.src[i] = sm.src[ixj];
                        sm.src[ixj] = t;
                    }
                }
            }
            __syncthreads();
        }
    // Q'' columns in ascending order: root j -> (zh_k / (da_k - lambda_j)) |.|^-1 on the active rows,
    // deflated t -> unit vector; rows are eigen indices (ix(t))
    for (int c = tid; c < N; c += UT) {
        const int sr = sm.src[c];
        for (int i = 0; i < N; ++i) {
            const int t = neg ? N - 1 - i : i;
            double v = 0.0;
            if (sr >= 0) {
                const int k = sm.cidx[t];
                if (k >= 0) v = sm.zh[k] * rcp_refined((sm.da[k] - sm.da[sm.org[sr]]) - sm.mu[sr]) * sm.rn[sr];
            } else {
                v = (t == -sr - 1) ? 1.0 : 0.0;
            }
            Q[(size_t)i * N + c] = v;
        }
    }
    for (int c = tid; c < N; c += UT) sm.d[c] = sm.key[c];
    __threadfence_block();
    __syncthreads();
    // deflation rotations: Q <- G_1^T ... G_r^T Q'' (the last rotation first), rows (p, k)
    for (int r = sm.nrot - 1; r >= 0; --r) {
        const int p = neg ? N - 1 - sm.rp[r] : sm.rp[r];
        const int k = neg ? N - 1 - sm.rk[r] : sm.rk[r];
        const double c = sm.rc[r], s = sm.rs[r];
        for (int col = tid; col < N; col += UT) {
            const double qp = Q[(size_t)p * N + col], qk = Q[(size_t)k * N + col];
            Q[(size_t)p * N + col] = c * qp - s * qk;
            Q[(size_t)k * N + col] = s * qp + c * qk;
        }
        __threadfence_block();
        __syncthreads();
    }
}

// One move of every chain.  V: B x N x N eigenvectors (V[c][i][k], column k for d_k).
template <int NM>
__global__ void __launch_bounds__(UT) k_fu_step(sweep_args a, const double* __restrict__ Vg, double* __restrict__ Q1g,
                                                double* __restrict__ Q2g, int32_t* __restrict__ upd) {
    __shared__ usmem<NM> sm;
    __shared__ sweep_args sa;  // LDS copy for the by-reference move helpers (no private copy per lane)
    const int ch = blockIdx.x, tid = threadIdx.x, N = a.N;
    if (tid == 0) sa = a;
    const size_t NN = (size_t)N * N;
    uint32_t* mtg = a.mt + (size_t)ch * fkrng::MT_WORDS;
    for (int i = tid; i < fkrng::MT_WORDS; i += UT) sm.mt[i] = mtg[i];
    for (int i = tid; i < N; i += UT) {
        sm.f[i] = a.f[(size_t)ch * N + i];
        sm.d[i] = a.spectrum[(size_t)ch * N + i];
    }
    __syncthreads();
    const int nf_old = a.nf[ch];
    if (tid == 0) sm.prop = propose_move(sm.mt, sm.f, N, nf_old, sa);
    __syncthreads();
    const move_prop p = sm.prop;
    const double beta = a.beta[ch], U = a.U[ch];
    double lz_new = 0.0;
    int nupd = 0;
    if (p.solve && p.sa >= 0) {  // moves of at most two sites (the host rejects randomize here)
        const double* V = Vg + (size_t)ch * NN;
        // the changed sites and their diagonal changes U (f_new - f_old)
        const int s1 = p.sa, s2 = p.sb;
        const double r1 = U * (sm.f[s1] ? 1.0 : -1.0);
        const double* z1 = V + (size_t)s1 * N;
        rank1(sm, z1, r1, N, Q1g + (size_t)ch * NN);
        nupd = 1;
        if (s2 >= 0) {
            const double* Q1 = Q1g + (size_t)ch * NN;
            const double* zr = V + (size_t)s2 * N;
            // z2 = Q1^T zr, column c per thread; loads in branch-free chunks so that a chunk is
            // in flight together (rows clamped to the last one, their terms dropped)
            for (int c = tid; c < N; c += UT) {
                double s = 0.0;
                for (int i0 = 0; i0 < N; i0 += RC) {
                    double qb[RC], zb[RC];
#pragma unroll
                    for (int q = 0; q < RC; ++q) {
                        const int i = min(i0 + q, N - 1);
                        qb[q] = Q1[(size_t)i * N + c];
                        zb[q] = zr[i];
                    }
#pragma unroll
                    for (int q = 0; q < RC; ++q)
                        if (i0 + q < N) s = fma(qb[q], zb[q], s);
                }
                sm.z2[c] = s;
            }
            __syncthreads();
            const double r2 = U * (sm.f[s2] ? 1.0 : -1.0);
            rank1(sm, sm.z2, r2, N, Q2g + (size_t)ch * NN);
            nupd = 2;
        }
        lz_new = logz_block<UT>(sm.d, N, beta, nullptr, nullptr, sm.scratch);
    }
    if (tid == 0) {
        double logw;
        const int acc = decide_move(p, p.sa >= 0 ? sm.f[p.sa] : 0, lz_new, a.logz[ch], beta, a.mu_f[ch], nf_old,
                                    &logw);
        record_move(sa, ch, p, acc, logw, lz_new);
        sm.acc = acc;
        upd[ch] = (acc && p.solve) ? nupd : 0;
    }
    __syncthreads();
    if (sm.acc) {
        for (int i = tid; i < N; i += UT) a.f[(size_t)ch * N + i] = sm.f[i];
        if (p.solve)
            for (int i = tid; i < N; i += UT) a.spectrum[(size_t)ch * N + i] = sm.d[i];
    }
    for (int i = tid; i < fkrng::MT_WORDS; i += UT) mtg[i] = sm.mt[i];
}

// C = A B for the chains with upd[b] >= 1 (stage 1) or the second factor (stage 2:
// C = A B where upd[b] == 2, C = A where upd[b] == 1).  64 x 64 output block per
// workgroup, four waves of 32 x 32 (2 x 2 MFMA tiles), K in chunks of 16 through LDS.
constexpr int GB = 64, GK = 16;
__global__ void __launch_bounds__(256) k_fu_gemm(int B, int N, const int32_t* __restrict__ upd, int stage,
                                                 const double* __restrict__ Ag, const double* __restrict__ Bg,
                                                 double* __restrict__ Cg) {
    const int b = blockIdx.y;
    const int u = upd[b];
    if (u < 1) return;
    const size_t NN = (size_t)N * N;
    const double* A = Ag + (size_t)b * NN;
    const double* Bm = Bg + (size_t)b * NN;
    double* C = Cg + (size_t)b * NN;
    const int nbt = (N + GB - 1) / GB;
    const int I0 = (blockIdx.x / nbt) * GB, J0 = (blockIdx.x % nbt) * GB;
    const int tid = threadIdx.x;
    if (stage == 2 && u == 1) {
        for (int e = tid; e < GB * GB; e += 256) {
            const int i = I0 + e / GB, j = J0 + e % GB;
            if (i < N && j < N) C[(size_t)i * N + j] = A[(size_t)i * N + j];
        }
        return;
    }
    __shared__ double As[GB][GK + 1];
    __shared__ double Bs[GK][GB + 1];
    const int lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int wr = wv >> 1, wc = wv & 1;
    const int li = lane & 15, kq = lane >> 4;
    fkb::v4d c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    for (int k0 = 0; k0 < N; k0 += GK) {
        for (int e = tid; e < GB * GK; e += 256) {
            const int r = e / GK, kk = e % GK;
            As[r][kk] = (I0 + r < N && k0 + kk < N) ? A[(size_t)(I0 + r) * N + k0 + kk] : 0.0;
        }
        for (int e = tid; e < GK * GB; e += 256) {
            const int kk = e / GB, cc = e % GB;
            Bs[kk][cc] = (k0 + kk < N && J0 + cc < N) ? Bm[(size_t)(k0 + kk) * N + J0 + cc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK / 4; ++ks) {
            const int kx = 4 * ks + kq;
            const double a0 = As[32 * wr + li][kx], a1 = As[32 * wr + 16 + li][kx];
            const double b0 = Bs[kx][32 * wc + li], b1 = Bs[kx][32 * wc + 16 + li];
            c00 = fkb::mf<double>::op(a0, b0, c00);
            c01 = fkb::mf<double>::op(a0, b1, c01);
            c10 = fkb::mf<double>::op(a1, b0, c10);
            c11 = fkb::mf<double>::op(a1, b1, c11);
        }
        __syncthreads();
    }
    auto store = [&](const fkb::v4d& acc, int r0, int c0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = r0 + fkb::mf<double>::row(lane, q), col = c0 + li;
            if (row < N && col < N) C[(size_t)row * N + col] = acc[q];
        }
    };
    store(c00, I0 + 32 * wr, J0 + 32 * wc);
    store(c01, I0 + 32 * wr, J0 + 32 * wc + 16);
    store(c10, I0 + 32 * wr + 16, J0 + 32 * wc);
    store(c11, I0 + 32 * wr + 16, J0 + 32 * wc + 16);
}

// The same product for N a multiple of GB (no bounds checks): the next K chunk is loaded into
// registers (16-byte loads) while the MFMAs of the current one run from LDS, so the global
// load latency is hidden behind the matrix work instead of exposed once per chunk.
__global__ void __launch_bounds__(256) k_fu_gemm_full(int B, int N, const int32_t* __restrict__ upd, int stage,
                                                      const double* __restrict__ Ag, const double* __restrict__ Bg,
                                                      double* __restrict__ Cg) {
    const int b = blockIdx.y;
    const int u = upd[b];
    if (u < 1) return;
    const size_t NN = (size_t)N * N;
    const double* A = Ag + (size_t)b * NN;
    const double* Bm = Bg + (size_t)b * NN;
    double* C = Cg + (size_t)b * NN;
    const int nbt = N / GB;
    const int I0 = (blockIdx.x / nbt) * GB, J0 = (blockIdx.x % nbt) * GB;
    const int tid = threadIdx.x;
    if (stage == 2 && u == 1) {
        for (int e = tid; e < GB * GB / 2; e += 256) {
            const int i = I0 + e / (GB / 2), j = J0 + 2 * (e % (GB / 2));
            *reinterpret_cast<double2*>(C + (size_t)i * N + j) = *reinterpret_cast<const double2*>(A + (size_t)i * N + j);
        }
        return;
    }
    __shared__ double As[GB][GK + 1];
    __shared__ double Bs[GK][GB + 1];
    const int lane = tid & (WAVE - 1), wv = tid / WAVE;
    const int wr = wv >> 1, wc = wv & 1;
    const int li = lane & 15, kq = lane >> 4;
    // this thread's share of a chunk: two double2 of A (rows ar, ar + 32; columns 2 ac, 2 ac + 1)
    // and two of B (rows bk, bk + 8; columns 2 bc, 2 bc + 1)
    const int ar = tid / (GK / 2), ac = tid % (GK / 2);
    const int bk = tid / (GB / 2), bc = tid % (GB / 2);
    double2 ra0, ra1, rb0, rb1;
    auto fetch = [&](int k0) {
        ra0 = *reinterpret_cast<const double2*>(A + (size_t)(I0 + ar) * N + k0 + 2 * ac);
        ra1 = *reinterpret_cast<const double2*>(A + (size_t)(I0 + ar + 32) * N + k0 + 2 * ac);
        rb0 = *reinterpret_cast<const double2*>(Bm + (size_t)(k0 + bk) * N + J0 + 2 * bc);
        rb1 = *reinterpret_cast<const double2*>(Bm + (size_t)(k0 + bk + 8) * N + J0 + 2 * bc);
    };
    fkb::v4d c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
    fetch(0);
    for (int k0 = 0; k0 < N; k0 += GK) {
        As[ar][2 * ac] = ra0.x; As[ar][2 * ac + 1] = ra0.y;
        As[ar + 32][2 * ac] = ra1.x; As[ar + 32][2 * ac + 1] = ra1.y;
        Bs[bk][2 * bc] = rb0.x; Bs[bk][2 * bc + 1] = rb0.y;
        Bs[bk + 8][2 * bc] = rb1.x; Bs[bk + 8][2 * bc + 1] = rb1.y;
        __syncthreads();
        if (k0 + GK < N) fetch(k0 + GK);  // in flight during the MFMAs below
#pragma unroll
        for (int ks = 0; ks < GK / 4; ++ks) {
            const int kx = 4 * ks + kq;
            const double a0 = As[32 * wr + li][kx], a1 = As[32 * wr + 16 + li][kx];
            const double b0 = Bs[kx][32 * wc + li], b1 = Bs[kx][32 * wc + 16 + li];
            c00 = fkb::mf<double>::op(a0, b0, c00);
            c01 = fkb::mf<double>::op(a0, b1, c01);
            c10 = fkb::mf<double>::op(a1, b0, c10);
            c11 = fkb::mf<double>::op(a1, b1, c11);
        }
        __syncthreads();
    }
    auto store = [&](const fkb::v4d& acc, int r0, int c0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) C[(size_t)(r0 + fkb::mf<double>::row(lane, q)) * N + c0 + li] = acc[q];
    };
    store(c00, I0 + 32 * wr, J0 + 32 * wc);
    store(c01, I0 + 32 * wr, J0 + 32 * wc + 16);
    store(c10, I0 + 32 * wr + 16, J0 + 32 * wc);
    store(c11, I0 + 32 * wr + 16, J0 + 32 * wc + 16);
}

// skip0 (nullable): nothing to do when skip0[0] is set (the refresh is not due)
__global__ void k_fu_f8to32(int n, const uint8_t* __restrict__ f8, int32_t* __restrict__ f32,
                            const int32_t* __restrict__ skip0) {
    if (skip0 && skip0[0]) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (size_t)n) f32[i] = f8[i];
}

// Move-step counter of a rank-one chain set: count += 1; the periodic re-decomposition
// is due when count is a multiple of period; skip[b] = !due for every chain.  One block.
__global__ void __launch_bounds__(256) k_fu_tick(int B, int period, long long* __restrict__ count,
                                                 int32_t* __restrict__ skip) {
    __shared__ int due;
    if (threadIdx.x == 0) {
        const long long c = count[0] + 1;
        count[0] = c;
        due = (c % period) == 0;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < B; b += blockDim.x) skip[b] = !due;
}

}  // namespace fku

hipError_t fk_launch_fu_step(const sweep_args& a, const double* V, double* Q1, double* Q2, int32_t* upd,
                             double* T, double* Vout, hipStream_t st) {
    if (a.N > fku::UNMAX) return hipErrorInvalidValue;
    // shared memory sized to N: four chains per CU for N <= 256
    if (a.N <= 256)
        hipLaunchKernelGGL(fku::k_fu_step<256>, dim3(a.B), dim3(fku::UT), 0, st, a, V, Q1, Q2, upd);
    else
        hipLaunchKernelGGL(fku::k_fu_step<fku::UNMAX>, dim3(a.B), dim3(fku::UT), 0, st, a, V, Q1, Q2, upd);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    const int nbt = (a.N + fku::GB - 1) / fku::GB;
    auto gemm = fku::k_fu_gemm;
    if (a.N % fku::GB == 0) gemm = fku::k_fu_gemm_full;
    hipLaunchKernelGGL(gemm, dim3(nbt * nbt, a.B), dim3(256), 0, st, a.B, a.N, upd, 1, V, Q1, T);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    hipLaunchKernelGGL(gemm, dim3(nbt * nbt, a.B), dim3(256), 0, st, a.B, a.N, upd, 2, T, Q2, Vout);
    return hipGetLastError();
}
hipError_t fk_launch_fu_tick(int B, int period, long long* count, int32_t* skip, hipStream_t st) {
    if (period < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fku::k_fu_tick, dim3(1), dim3(256), 0, st, B, period, count, skip);
    return hipGetLastError();
}

hipError_t fk_launch_f8to32(int n, const uint8_t* f8, int32_t* f32, hipStream_t st, const int32_t* skip0) {
    hipLaunchKernelGGL(fku::k_fu_f8to32, dim3((n + 255) / 256), dim3(256), 0, st, n, f8, f32, skip0);
    return hipGetLastError();
}
int fk_fu_nmax() { return fku::UNMAX; }
