// fk_vectors.hip -- eigenvectors for calc_ed(true) (src/configuration.cpp:304-311,
// ed_cache::cached_evecs) on top of the eigenvalue path.
//
//   k_tri_evecs     eigenvectors of the tridiagonal T by inverse iteration, one
//                   thread per eigenvalue (LU of T - lambda with partial
//                   pivoting fused with the right-hand side, O(N) per solve);
//                   eigenvalues closer than ortol form clusters: members of
//                   tight clusters are computed in rounds and orthogonalised
//                   (modified Gram-Schmidt) against the earlier members, those
//                   of loose clusters projected off the earlier members at once.
//   k_bt_tfac, k_bt_apply
//                   V = Q Z with Q = H_0 H_1 ... H_{N-3} the Householder
//                   reflectors stored by the tridiagonaliser, applied in
//                   blocks of 16 (compact WY: H_k0 ... H_k0+15 = I - Y T Y^T,
//                   Golub & Van Loan 5.1.7 / LAPACK dlarft, forward), last block
//                   first.  A workgroup holds a panel of 16 or 32 columns of Z
//                   in MFMA accumulator registers for the whole product; per
//                   block W = Y^T Z and Z -= Y (T W) are 16x16x4 FP64 MFMAs.
// Layouts: Z and V row-major N x N per matrix (column k = vector k); the
// per-thread solve scratch is [i][j] so that a step i is one coalesced row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include "fk_internal.h"
#include "fk_solve.h"

namespace fkv {

using namespace fks;

// Deterministic start vector of inverse iteration for eigenvalue j (an LCG sequence).
struct start_vec {
    uint32_t h;
    __device__ explicit start_vec(int j) : h(2654435761u * (uint32_t)(j + 1)) {}
    __device__ double next() {
        h = h * 1664525u + 1013904223u;
        return 1.0 + 0.5 * ((double)(h >> 8) / 16777216.0 - 0.5);
    }
};

// Solve (T - lam) z = s x in place (x overwritten by z), T = (d, e), with the LU of
// T - lam by partial pivoting formed on the fly (U's three diagonals to u0/u1/u2, stride
// ld, this thread's column; zero pivots replaced by +-tiny, the usual inverse-iteration
// rule).  With gen the right-hand side is the start vector of eigenvalue j and x is not
// read.  Loads are issued in chunks ahead of their use, branch-free so that a chunk is in
// flight as a whole (the recurrences themselves are sequential).  Returns ||z||^2; *zbig gets the entry of largest magnitude.
// Chunks (CF, CB forward and backward, CV the single-vector passes -- final scaling and the
// tight-cluster Gram-Schmidt --, CL the loose-cluster Gram-Schmidt with OL + 1 loads an
// element) are template arguments: a wave holds at most 63 vector loads in flight, and the
// kernel's register budget (its workgroup size) bounds how many values a chunk may stage.
template <int CF, int CB>
__device__ double tri_solve(const double* __restrict__ d, const double* __restrict__ e, int N, double lam,
                            double tiny, double* x, double s, bool gen, int j, double* u0, double* u1, double* u2,
                            size_t ld, double* zbig) {
    start_vec sv(j);
    double a = d[0] - lam;                 // current pivot-row diagonal
    double b = N > 1 ? e[0] : 0.0;         // current pivot-row superdiagonal
    double xk = gen ? sv.next() : s * x[0];
    for (int k0 = 0; k0 + 1 < N; k0 += CF) {
        double xb[CF];
        // branch-free loads (index clamped to the last row; the surplus values are not used),
        // so that all of them are in flight before the first is waited for
        if (gen) {
#pragma unroll
            for (int q = 0; q < CF; ++q) xb[q] = k0 + 1 + q < N ? sv.next() : 0.0;
        } else {
#pragma unroll
            for (int q = 0; q < CF; ++q) xb[q] = s * x[(size_t)min(k0 + 1 + q, N - 1) * ld];
        }
#pragma unroll
        for (int q = 0; q < CF; ++q) {
            const int k = k0 + q;
            if (k + 1 >= N) break;
            const double c = e[k];             // subdiagonal below the pivot
            const double dn = d[k + 1] - lam;  // next diagonal
            const double en = (k + 2 < N) ? e[k + 1] : 0.0;
            const double xn = xb[q];
            if (fabs(a) >= fabs(c)) {
                const double aa = a == 0.0 ? tiny : a;
                const double m = c / aa;
                u0[(size_t)k * ld] = aa; u1[(size_t)k * ld] = b; u2[(size_t)k * ld] = 0.0;
                x[(size_t)k * ld] = xk;
                a = dn - m * b;
                b = en;
                xk = xn - m * xk;
            } else {
                const double m = a / c;
                u0[(size_t)k * ld] = c; u1[(size_t)k * ld] = dn; u2[(size_t)k * ld] = en;
                x[(size_t)k * ld] = xn;
                a = b - m * dn;
                b = -m * en;
                xk = xk - m * xn;
            }
        }
    }
    if (a == 0.0) a = tiny;
    u0[(size_t)(N - 1) * ld] = a;
    x[(size_t)(N - 1) * ld] = xk;
    // back substitution with U (diagonals u0, u1, u2)
    double z1 = 0.0, z2 = 0.0;  // z[k+1], z[k+2]
    double nrm2 = 0.0, big = 0.0;
    for (int k0 = N - 1; k0 >= 0; k0 -= CB) {
        double xb[CB], p0[CB], p1[CB], p2[CB];
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            const size_t o = (size_t)max(k0 - q, 0) * ld;  // branch-free (row 0 repeated past the end)
            xb[q] = x[o];
            p0[q] = u0[o];
            p1[q] = u1[o];
            p2[q] = u2[o];
        }
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            const int k = k0 - q;
            if (k < 0) break;
            double t = xb[q];
            if (k + 1 < N) t -= p1[q] * z1;
            if (k + 2 < N) t -= p2[q] * z2;
            const double z = t / p0[q];
            x[(size_t)k * ld] = z;
            nrm2 = fma(z, z, nrm2);
            if (fabs(z) > fabs(big)) big = z;
            z2 = z1;
            z1 = z;
        }
    }
    *zbig = big;
    return nrm2;
}

// Orthogonalise x (column stride N) against the columns p in [p0, p1) of Z (unit vectors),
// modified Gram-Schmidt, loads in branch-free chunks.  Returns ||x||^2; *xbig gets the entry
// of largest magnitude.
template <int CV>
__device__ double orth_against(double* x, const double* Z, int p0, int p1, int N, double nrm2, double* xbig) {
    for (int p = p0; p < p1; ++p) {
        const double* y = Z + p;
        double dot = 0.0;
        for (int i0 = 0; i0 < N; i0 += CV) {
            double xb[CV], yb[CV];
#pragma unroll
            for (int q = 0; q < CV; ++q) {
                const size_t o = (size_t)min(i0 + q, N - 1) * N;
                xb[q] = x[o];
                yb[q] = y[o];
            }
#pragma unroll
            for (int q = 0; q < CV; ++q)
                if (i0 + q < N) dot = fma(xb[q], yb[q], dot);
        }
        nrm2 = 0.0;
        double big = 0.0;
        for (int i0 = 0; i0 < N; i0 += CV) {
            double xb[CV], yb[CV];
#pragma unroll
            for (int q = 0; q < CV; ++q) {
                const size_t o = (size_t)min(i0 + q, N - 1) * N;
                xb[q] = x[o];
                yb[q] = y[o];
            }
#pragma unroll
            for (int q = 0; q < CV; ++q) {
                if (i0 + q >= N) break;
                const double v = fma(-dot, yb[q], xb[q]);
                x[(size_t)(i0 + q) * N] = v;
                nrm2 = fma(v, v, nrm2);
                if (fabs(v) > fabs(big)) big = v;
            }
        }
        *xbig = big;
    }
    return nrm2;
}

// x <- x sign(big) / sqrt(nrm2): unit norm, largest-magnitude component positive.
template <int CV>
__device__ void finish_vec(double* x, int N, double nrm2, double big) {
    const double sc = nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 1.0;
    const double sg = big < 0.0 ? -sc : sc;
    for (int i0 = 0; i0 < N; i0 += CV) {
        double xb[CV];
#pragma unroll
        for (int q = 0; q < CV; ++q) xb[q] = x[(size_t)min(i0 + q, N - 1) * N];
#pragma unroll
        for (int q = 0; q < CV; ++q)
            if (i0 + q < N) x[(size_t)(i0 + q) * N] = sg * xb[q];
    }
}

// Gram-Schmidt step of a loose-cluster member, every member at once: x (column stride N, unit
// norm, largest-magnitude component positive) against the columns [p0, p1) of Z as inverse
// iteration left them (unit norm).  Inverse iteration has separated x from them already, up to
// delta = eps ||T|| / gap <= 2e-9 (gap > 1e-7 ||T|| outside the tight clusters; a tight-cluster
// member is orthogonal to its tight predecessors already), so projecting on these columns
// instead of their orthogonalised forms leaves overlaps of order r^2 delta^2 (r members before
// x: 2e-16 at r = 7), and no member has to wait for an earlier one.  Blocks of OL columns in
// turn: one pass forms a block's OL dot products, one subtracts them (classical Gram-Schmidt
// within a block, which at overlaps this small loses nothing against the modified form).  The
// result goes to w (stride N; the other members read Z meanwhile); returns its squared norm.
// The sign stays: a change of 2e-9 per component cannot turn the largest-magnitude component
// (>= N^-1/2) negative.
constexpr int OL = 4;
template <int CL>
__device__ double orth_loose(const double* x, const double* Z, int p0, int p1, int N, double* w) {
    double nrm2 = 0.0;
    for (int pb = p0; pb < p1; pb += OL) {
        const int nb = min(OL, p1 - pb);
        const double* src = pb == p0 ? x : w;
        double dot[OL];
#pragma unroll
        for (int m = 0; m < OL; ++m) dot[m] = 0.0;
        for (int i0 = 0; i0 < N; i0 += CL) {
            double xb[CL], yb[OL][CL];
            // branch-free loads: rows clamped to the last one, columns past the block's end
            // to its last column (their products are dropped below)
#pragma unroll
            for (int q = 0; q < CL; ++q) {
                const size_t o = (size_t)min(i0 + q, N - 1) * N;
                xb[q] = src[o];
#pragma unroll
                for (int m = 0; m < OL; ++m) yb[m][q] = Z[o + pb + min(m, nb - 1)];
            }
#pragma unroll
            for (int q = 0; q < CL; ++q) {
                if (i0 + q >= N) break;
#pragma unroll
                for (int m = 0; m < OL; ++m) dot[m] = fma(xb[q], yb[m][q], dot[m]);
            }
        }
#pragma unroll
        for (int m = 0; m < OL; ++m)
            if (m >= nb) dot[m] = 0.0;
        nrm2 = 0.0;
        for (int i0 = 0; i0 < N; i0 += CL) {
            double xb[CL], yb[OL][CL];
#pragma unroll
            for (int q = 0; q < CL; ++q) {
                const size_t o = (size_t)min(i0 + q, N - 1) * N;
                xb[q] = src[o];
#pragma unroll
                for (int m = 0; m < OL; ++m) yb[m][q] = Z[o + pb + min(m, nb - 1)];
            }
#pragma unroll
            for (int q = 0; q < CL; ++q) {
                if (i0 + q >= N) break;
                double v = xb[q];
#pragma unroll
                for (int m = 0; m < OL; ++m) v = fma(-dot[m], yb[m][q], v);
                w[(size_t)(i0 + q) * N] = v;
                nrm2 = fma(v, v, nrm2);
            }
        }
    }
    return nrm2;
}

// Grid (B, G): G workgroups per matrix, each taking the eigenvalues of whole clusters.  d, e,
// ev: B x N.  Z: B x N x N output.  scratch: B x 3 x N x N.
// Two cluster levels (by gaps between consecutive eigenvalues):
//   * tight (gap <= 1e-7 ||T||): inverse iteration cannot separate the members, so they are
//     computed in rounds (position in the cluster), three solves each, every iterate
//     orthogonalised against the earlier members (as LAPACK dstein);
//   * loose (gap <= 1e-3 ||T||, dstein's threshold): inverse iteration separates them by a
//     factor ~ eps ||T|| / gap <= 2e-9 per solve, so every eigenvalue outside a tight
//     cluster is computed at once (two solves), and the members of each loose cluster are
//     then projected off the earlier members, all at once (orth_loose).
template <int NTB, int NL, int CF, int CB, int CV, int CL>
__global__ void __launch_bounds__(NTB) k_tri_evecs(int B, int N, const double* __restrict__ dg,
                                                    const double* __restrict__ eg, const double* __restrict__ evg,
                                                    double* __restrict__ Zg, double* __restrict__ scr,
                                                    const int32_t* __restrict__ skip) {
    __shared__ double d[NL], e[NL], evs[NL];  // NL >= N
    __shared__ short cstart[NL], tstart[NL];
    __shared__ int maxpos_l, maxpos_t, jlo, jhi;
    const int b = blockIdx.x;
    const int tid = threadIdx.x, NT = blockDim.x;
    if (b >= B || (skip && skip[b])) return;
    // this workgroup's share of the eigenvalues: the loose clusters that start in
    // [c0, c1) (a tight cluster lies inside a loose one), a contiguous range [jlo, jhi)
    const int G = gridDim.y, chunk = (N + G - 1) / G;
    const int c0 = (int)blockIdx.y * chunk, c1 = c0 + chunk;
    const double* ev = evg + (size_t)b * N;
    double* Z = Zg + (size_t)b * N * N;
    const size_t NN = (size_t)N * N;
    double* u0 = scr + (size_t)b * 3 * NN;
    double* u1 = u0 + NN;
    double* u2 = u1 + NN;
    double tn = 0.0;
    for (int i = tid; i < N; i += NT) {
        d[i] = dg[(size_t)b * N + i];
        e[i] = eg[(size_t)b * N + i];
        evs[i] = ev[i];
    }
    __syncthreads();
    for (int i = 0; i < N; ++i) tn = fmax(tn, fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < N ? fabs(e[i]) : 0.0));
    const double eps = 2.220446049250313e-16;
    const double ortol = 1e-3 * tn;  // loose clusters (LAPACK dstein's threshold)
    const double ttol = 1e-7 * tn;   // tight clusters
    const double tiny = eps * fmax(tn, 1e-300);
    // cluster starts (sequential scan by one thread over eigenvalues staged in LDS)
    if (tid == 0) {
        int cs = 0, ts = 0, mpl = 0, mpt = 0, lo = 0, hi = 0;
        double prev = 0.0;
        for (int j = 0; j < N; ++j) {
            const double ej = evs[j];
            if (j > 0 && ej - prev > ortol) cs = j;
            if (j > 0 && ej - prev > ttol) ts = j;
            cstart[j] = (short)cs;
            tstart[j] = (short)ts;
            lo += cs < c0;
            hi += cs < c1;
            if (cs >= c0 && cs < c1) {
                mpl = max(mpl, j - cs);
                mpt = max(mpt, j - ts);
            }
            prev = ej;
        }
        maxpos_l = mpl;
        maxpos_t = mpt;
        jlo = lo;
        jhi = hi;
    }
    __syncthreads();
    const int j0 = jlo, j1 = jhi;
    // inverse iteration: tight clusters in rounds, everything else in round 0
    for (int round = 0; round <= maxpos_t; ++round) {
        for (int j = j0 + tid; j < j1; j += NT) {
            const int pos = j - tstart[j];
            if (pos != round) continue;
            const bool tight = pos > 0 || (j + 1 < N && tstart[j + 1] == j);
            double* x = Z + j;  // column j, stride N
            double nrm2 = 0.0, big = 0.0;
            const int its = tight ? 3 : 2;
            for (int it = 0; it < its; ++it) {
                const double sc = it == 0 ? 1.0 : (nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 1.0);
                nrm2 = tri_solve<CF, CB>(d, e, N, ev[j], tiny, x, sc, it == 0, j, u0 + j, u1 + j, u2 + j, N, &big);
                if (pos > 0) nrm2 = orth_against<CV>(x, Z, tstart[j], j, N, nrm2, &big);
            }
            finish_vec<CV>(x, N, nrm2, big);
        }
        __threadfence();
        __syncthreads();
    }
    // loose clusters: every member after the first projected off the earlier members at once
    // (orth_loose; members of a tight cluster are already orthogonal to their tight
    // predecessors, which the projection repeats harmlessly) into the free solve scratch u0,
    // its squared norm to u1[j]; copied back normalised once every member has read the columns
    // it needs
    if (maxpos_l > 0) {
        for (int j = j0 + tid; j < j1; j += NT) {
            if (j == cstart[j]) continue;
            u1[j] = orth_loose<CL>(Z + j, Z, cstart[j], j, N, u0 + j);
        }
        __syncthreads();
        for (int j = j0 + tid; j < j1; j += NT) {
            if (j == cstart[j]) continue;
            const double nrm2 = u1[j];
            const double sc = nrm2 > 0.0 ? 1.0 / sqrt(nrm2) : 1.0;
            for (int i0 = 0; i0 < N; i0 += CV) {
                double xb[CV];
#pragma unroll
                for (int q = 0; q < CV; ++q) xb[q] = u0[(size_t)min(i0 + q, N - 1) * N + j];
#pragma unroll
                for (int q = 0; q < CV; ++q)
                    if (i0 + q < N) Z[(size_t)(i0 + q) * N + j] = sc * xb[q];
            }
        }
    }
}

typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int BT_NB = 16;  // reflectors per block

__device__ __forceinline__ v4d mfma64(double a, double b, v4d c) {
    // 16x16x4: lane l supplies A[l%16][l/16] and B[l/16][l%16]; accumulator
    // register r of lane l is D[(l/16) + 4r][l%16]
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// T factor of each block: grid (B, nblk), 256 threads.  Reflector k (row k of vh) is
// zero for k >= N - 2.  T[a][c] (upper triangular) with H_k0 ... H_k0+15 = I - Y T Y^T:
// T[c][c] = tau_c, T[0:c, c] = -tau_c T[0:c, 0:c] (Y^T y_c)[0:c].
__global__ void __launch_bounds__(256) k_bt_tfac(int B, int N, int nblk, const double* __restrict__ vg,
                                                 const double* __restrict__ taug, double* __restrict__ Tg,
                                                 const int32_t* __restrict__ skip) {
    constexpr int CH = 256, LDY = CH + 1;
    __shared__ double ych[BT_NB * LDY];
    __shared__ double G[BT_NB][BT_NB + 1];
    __shared__ double T[BT_NB][BT_NB + 1];
    const int b = blockIdx.x, s = blockIdx.y, tid = threadIdx.x;
    if (b >= B || (skip && skip[b])) return;
    const int k0 = s * BT_NB;
    const double* V = vg + (size_t)b * N * N;
    const int a = tid / BT_NB, c = tid % BT_NB;
    double g = 0.0;
    for (int i0 = 0; i0 < N; i0 += CH) {
        __syncthreads();
        for (int idx = tid; idx < BT_NB * CH; idx += 256) {
            const int m = idx / CH, i = i0 + idx % CH, k = k0 + m;
            ych[m * LDY + idx % CH] = (k < N - 2 && i < N) ? V[(size_t)k * N + i] : 0.0;
        }
        __syncthreads();
        for (int i = 0; i < CH; ++i) g = fma(ych[a * LDY + i], ych[c * LDY + i], g);
    }
    G[a][c] = g;
    if (tid < BT_NB) {
        const int k = k0 + tid;
        for (int q = 0; q < BT_NB; ++q) T[q][tid] = 0.0;
        T[tid][tid] = k < N - 2 ? taug[(size_t)b * N + k] : 0.0;
    }
    __syncthreads();
    for (int cc = 1; cc < BT_NB; ++cc) {
        if (tid < cc) {
            double t = 0.0;
            for (int m = tid; m < cc; ++m) t = fma(T[tid][m], G[m][cc], t);
            T[tid][cc] = -T[cc][cc] * t;
        }
        __syncthreads();
    }
    Tg[((size_t)b * nblk + s) * BT_NB * BT_NB + tid] = T[tid / BT_NB][tid % BT_NB];
}

// Z <- Q Z for a panel of 16 CT columns: grid (B, ceil(N / (16 CT))), 1024 threads.
// Wave w holds the row tiles R = 16 w + 256 t (t < RT) of the panel as accumulators
// acc[t][c] (rows R + (l/16) + 4r, column 16 c + l%16).  Per block of reflectors
// (last first): P = Y^T Z over the wave's rows (the accumulator register r is the
// MFMA B operand of rows R + 4r .. R + 4r + 3), partials summed over the waves in LDS,
// W = T P, then Z -= Y W.  Tiles wholly above row k0 + 1 are skipped (Y is zero there).
template <int RT, int CT>
__global__ void __launch_bounds__(1024) k_bt_apply(int B, int N, int nblk, const double* __restrict__ vg,
                                                   const double* __restrict__ Tg, double* __restrict__ Zg,
                                                   const int32_t* __restrict__ skip) {
    constexpr int NW = 16;
    __shared__ double part[NW][CT][256];
    __shared__ double Ps[CT][BT_NB][BT_NB + 1];
    __shared__ double Ws[CT][BT_NB][BT_NB + 1];
    __shared__ double Ts[BT_NB][BT_NB + 1];
    const int b = blockIdx.x;
    if (b >= B || (skip && skip[b])) return;
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    const int c0 = blockIdx.y * 16 * CT;
    const double* V = vg + (size_t)b * N * N;
    double* Z = Zg + (size_t)b * N * N;
    const int lr = l >> 4, lc = l & 15;
    v4d acc[RT][CT];
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * w + 256 * t + lr + 4 * r, col = c0 + 16 * c + lc;
                acc[t][c][r] = (row < N && col < N) ? Z[(size_t)row * N + col] : 0.0;
            }
    for (int s = nblk - 1; s >= 0; --s) {
        const int k0 = s * BT_NB;
        // P = Y^T Z over this wave's rows
        v4d p[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) p[c] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int R = 16 * w + 256 * t;
            if (R + 15 <= k0) continue;  // uniform over the wave
            const int k = k0 + lc;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = R + 4 * r + lr;
                const double y = (k < N - 2 && row < N) ? V[(size_t)k * N + row] : 0.0;
#pragma unroll
                for (int c = 0; c < CT; ++c) p[c] = mfma64(y, acc[t][c][r], p[c]);
            }
            if constexpr (RT > 2) __builtin_amdgcn_sched_barrier(0);  // bound the loads in flight
        }
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) part[w][c][(lr + 4 * r) * BT_NB + lc] = p[c][r];
        if (tid < BT_NB * BT_NB) Ts[tid / BT_NB][tid % BT_NB] = Tg[((size_t)b * nblk + s) * BT_NB * BT_NB + tid];
        __syncthreads();
        // sum of the wave partials (fixed order)
        if (tid < CT * 256) {
            const int c = tid >> 8, e = tid & 255;
            double v = 0.0;
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) v += part[ww][c][e];
            Ps[c][e / BT_NB][e % BT_NB] = v;
        }
        __syncthreads();
        // W = T P (T upper triangular)
        if (tid < CT * 256) {
            const int c = tid >> 8, a = (tid >> 4) & 15, j = tid & 15;
            double v = 0.0;
            for (int m = a; m < BT_NB; ++m) v = fma(Ts[a][m], Ps[c][m][j], v);
            Ws[c][a][j] = v;
        }
        __syncthreads();
        // Z -= Y W: A = -Y[R + l%16][k0 + 4q + l/16], B = W[4q + l/16][l%16]
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int R = 16 * w + 256 * t;
            if (R + 15 <= k0) continue;
            const int row = R + lc;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = k0 + 4 * q + lr;
                const double y = (k < N - 2 && row < N) ? -V[(size_t)k * N + row] : 0.0;
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[t][c] = mfma64(y, Ws[c][4 * q + lr][lc], acc[t][c]);
            }
            if constexpr (RT > 2) __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * w + 256 * t + lr + 4 * r, col = c0 + 16 * c + lc;
                if (row < N && col < N) Z[(size_t)row * N + col] = acc[t][c][r];
            }
}

}  // namespace fkv

size_t fk_evecs_scratch_doubles(int B, int N) {
    // the inverse-iteration LU (3 N^2 per matrix), then the T factors (256 per block of 16 reflectors)
    const size_t nblk = N > 2 ? (size_t)(N - 2 + fkv::BT_NB - 1) / fkv::BT_NB : 0;
    return (size_t)B * std::max<size_t>(3 * (size_t)N * N, nblk * fkv::BT_NB * fkv::BT_NB);
}

hipError_t fk_launch_evecs(int B, int N, const double* d, const double* e, const double* ev, const double* vh,
                           const double* tau, double* Z, double* scratch, hipStream_t st, const int32_t* skip) {
    if (N > 2048) return hipErrorInvalidValue;
    // A batch of fewer matrices than CUs splits each matrix's eigenvalues over G workgroups
    // (whole clusters each, so the results do not depend on G): the solves of one matrix are
    // bound by its CU's memory throughput, not by their sequential length.
    // Such a workgroup holds few solves, each a sequence of dependent memory round trips, so it
    // runs with one wave a SIMD and stages longer chunks (fewer round trips) in the registers
    // that frees; the chunk sizes change how loads are grouped, not the arithmetic.
    // Up to N = 256 the LDS tables are sized for 256 eigenvalues (7 KB instead of 56), so that
    // LDS no longer caps a CU at two of these workgroups.
    const int G = B < 256 ? std::min((256 + B - 1) / B, std::max(1, N / 64)) : 1;
    if (G > 1 && N <= 256)
        hipLaunchKernelGGL((fkv::k_tri_evecs<256, 256, 16, 12, 16, 8>), dim3(B, G), dim3(256), 0, st, B, N, d, e, ev,
                           Z, scratch, skip);
    else if (G > 1)
        hipLaunchKernelGGL((fkv::k_tri_evecs<256, 2048, 16, 12, 16, 8>), dim3(B, G), dim3(256), 0, st, B, N, d, e, ev,
                           Z, scratch, skip);
    else if (N <= 256)
        hipLaunchKernelGGL((fkv::k_tri_evecs<256, 256, 8, 4, 8, 8>), dim3(B, 1), dim3(256), 0, st, B, N, d, e, ev, Z,
                           scratch, skip);
    else
        hipLaunchKernelGGL((fkv::k_tri_evecs<1024, 2048, 8, 4, 8, 8>), dim3(B, 1), dim3(1024), 0, st, B, N, d, e, ev,
                           Z, scratch, skip);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    if (N < 3) return hipSuccess;  // no reflectors
    // T factors go to the solve scratch, free once the vectors of T are formed
    const int nblk = (N - 2 + fkv::BT_NB - 1) / fkv::BT_NB;
    hipLaunchKernelGGL(fkv::k_bt_tfac, dim3(B, nblk), dim3(256), 0, st, B, N, nblk, vh, tau, scratch, skip);
    r = hipGetLastError();
    if (r != hipSuccess) return r;
    if (N <= 256)
        hipLaunchKernelGGL((fkv::k_bt_apply<1, 2>), dim3(B, (N + 31) / 32), dim3(1024), 0, st, B, N, nblk, vh, scratch,
                           Z, skip);
    else if (N <= 512)
        hipLaunchKernelGGL((fkv::k_bt_apply<2, 2>), dim3(B, (N + 31) / 32), dim3(1024), 0, st, B, N, nblk, vh, scratch,
                           Z, skip);
    else if (N <= 768)  // three row tiles: four would spill at this workgroup size
        hipLaunchKernelGGL((fkv::k_bt_apply<3, 2>), dim3(B, (N + 31) / 32), dim3(1024), 0, st, B, N, nblk, vh, scratch,
                           Z, skip);
    else if (N <= 1024)
        hipLaunchKernelGGL((fkv::k_bt_apply<4, 1>), dim3(B, (N + 15) / 16), dim3(1024), 0, st, B, N, nblk, vh, scratch,
                           Z, skip);
    else
        hipLaunchKernelGGL((fkv::k_bt_apply<8, 1>), dim3(B, (N + 15) / 16), dim3(1024), 0, st, B, N, nblk, vh, scratch,
                           Z, skip);
    return hipGetLastError();
}
