// fk_solve.h -- device building blocks of the per-move diagonalisation path.
//
// Replaces the body of configuration_t::calc_hamiltonian + calc_ed
// (src/configuration.cpp:79-91 and 208-246): H = hop + diag(-mu_c + U n_f),
// eigenvalues ascending, logZ = sum_k log(1 + e^{-beta eps_k}).
//
// Layout (one workgroup per matrix, P*P threads):
//   thread (r, c), r = row of the thread grid, c = column, owns the matrix
//   entries A[r + P*a][c + P*b] for the block pairs a >= b (block-cyclic lower
//   triangle).  For a == b it also owns the mirrored upper entry when r < c,
//   so every diagonal block is stored in full.  The whole lower triangle lives
//   in VGPRs: NB*(NB+1)/2 doubles per thread (36 for N <= 256, P = 32).
//   A Householder step (Golub & Van Loan 8.3.1) needs v and w broadcast to the
//   owners of each row/column and the symv partial sums reduced across the
//   grid; both go through LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fks {

constexpr int WAVE = 64;

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    return v;
}

// Knuth TwoSum based pair (s, c) addition: (s1,c1) + (s2,c2).
__device__ inline void two_sum_merge(double& s, double& c, double s2, double c2) {
    double t = s + s2;
    double bp = t - s;
    double err = (s - (t - bp)) + (s2 - bp);
    s = t;
    c = c + c2 + err;
}
__device__ inline void wave_sum_comp(double& s, double& c) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        double s2 = __shfl_xor(s, o, WAVE);
        double c2 = __shfl_xor(c, o, WAVE);
        // merge in a lane-order independent way so every lane ends identical
        double a = (threadIdx.x & o) ? s2 : s, ac = (threadIdx.x & o) ? c2 : c;
        double b = (threadIdx.x & o) ? s : s2, bc = (threadIdx.x & o) ? c : c2;
        two_sum_merge(a, ac, b, bc);
        s = a;
        c = ac;
    }
}

// log(1 + e^{-x}) without overflow.
__device__ inline double log1p_exp_neg(double x) {
    return x > 0.0 ? log1p(exp(-x)) : -x + log1p(exp(x));
}

// ---------------------------------------------------------------------------
// Register-resident Householder tridiagonalisation.
// ---------------------------------------------------------------------------
template <int P, int NB>
struct tri_shape {
    static constexpr int NT = P * P;             // threads per workgroup
    static constexpr int NP = P * NB;            // padded matrix size
    static constexpr int S = NB * (NB + 1) / 2;  // slots per thread
    static constexpr int TPR = NT / NP;          // threads per row in the reduction phase
    static constexpr int NW = NT / WAVE;
    static_assert(TPR >= 1 && (TPR & (TPR - 1)) == 0, "threads per row must be a power of two");
    static_assert(P <= WAVE && (WAVE % P) == 0, "a wave holds whole rows of the thread grid");
};

__host__ __device__ constexpr int slot(int a, int b) { return a * (a + 1) / 2 + b; }

template <int P, int NB>
struct tri_smem {
    static constexpr int NP = P * NB;
    double x[2][NP];          // column k (double-buffered by step parity)
    double p[NP];             // tau * A v
    double part[NP][2 * P + 1];  // symv partials: [i][c] row part, [i][P + r] column part
    double d[NP];             // tridiagonal diagonal
    double e[NP];             // tridiagonal off-diagonal
    double red[8 + 16];       // scalars; red[8 + w] = per-wave partial of p.v
};

// Assemble H(f) = hop + diag(-mu_c + U f) into the thread's slots.
template <int P, int NB>
__device__ inline void assemble(double (&A)[NB * (NB + 1) / 2], const double* __restrict__ hop,
                                const uint8_t* f_lds, double U, double mu_c, int N, int r, int c) {
#pragma unroll
    for (int a = 0; a < NB; ++a) {
#pragma unroll
        for (int b = 0; b <= a; ++b) {
            int i = r + P * a, j = c + P * b;
            double v = 0.0;
            if (i < N && j < N) {
                v = hop[(size_t)i * N + j];
                if (i == j) v += -mu_c + U * (double)f_lds[i];
            }
            A[slot(a, b)] = v;
        }
    }
}

// Write column kk (entries below the diagonal) to xbuf and the diagonal to d.
template <int P, int NB>
__device__ inline void emit_column(const double (&A)[NB * (NB + 1) / 2], tri_smem<P, NB>& sm, double* xbuf,
                                   int kk, int N, int r, int c) {
    if (c != kk % P) return;
    const int bb = kk / P;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (b != bb) continue;
#pragma unroll
        for (int a = b; a < NB; ++a) {
            int i = r + P * a;
            if (i > kk && i < N) xbuf[i] = A[slot(a, b)];
            if (i == kk) sm.d[kk] = A[slot(a, b)];
        }
    }
}

// Reduce the symmetric matrix held in A to tridiagonal form; sm.d[0..N),
// sm.e[0..N-1) receive T.  Must be called by all P*P threads.
// If vout != nullptr the Householder vectors are stored row-wise:
// vout[k*N + i] = v_k[i] (v_k[k+1] = 1, zero above), tauout[k] = tau_k.
template <int P, int NB>
__device__ void tridiagonalize(double (&A)[NB * (NB + 1) / 2], tri_smem<P, NB>& sm, int N,
                               double* __restrict__ vout, double* __restrict__ tauout) {
    using sh = tri_shape<P, NB>;
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int r = (tid / WAVE) * (WAVE / P) + lane / P;
    const int c = lane % P;

    // zero the column buffers (padding beyond N must read as 0)
    for (int i = tid; i < sh::NP; i += sh::NT) { sm.x[0][i] = 0.0; sm.x[1][i] = 0.0; sm.d[i] = 0.0; sm.e[i] = 0.0; }
    __syncthreads();
    emit_column<P, NB>(A, sm, sm.x[0], 0, N, r, c);
    __syncthreads();

    for (int k = 0; k + 2 < N; ++k) {
        const double* xk = sm.x[k & 1];
        double* xn = sm.x[(k + 1) & 1];
        // --- Householder reflector from x = A[k+1:N, k] (every wave redundantly)
        double alpha = xk[k + 1];
        double part = 0.0;
        for (int i = k + 2 + lane; i < N; i += WAVE) part += xk[i] * xk[i];
        double xn2 = wave_sum(part);
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (xn2 != 0.0) {
            beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        if (tid == 0) sm.e[k] = beta;
        if (vout) {
            for (int i = tid; i < N; i += sh::NT)
                vout[(size_t)k * N + i] = (i == k + 1) ? 1.0 : ((i > k + 1) ? xk[i] * scal : 0.0);
            if (tid == 0) tauout[k] = tau;
        }
        const int a0 = (k + 1) / P;
        auto vfun = [&](int i) -> double {
            return (i == k + 1) ? 1.0 : ((i > k + 1 && i < N) ? xk[i] * scal : 0.0);
        };
        double vc[NB];
        if (tau != 0.0) {
            // --- symv partials
            double pcol[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) { vc[b] = (b >= a0) ? vfun(c + P * b) : 0.0; pcol[b] = 0.0; }
#pragma unroll
            for (int a = 0; a < NB; ++a) {
                if (a < a0) continue;
                const double vra = vfun(r + P * a);
                double s = 0.0;
#pragma unroll
                for (int b = 0; b <= a; ++b) {
                    if (b < a0) continue;
                    const double av = A[slot(a, b)];
                    s = fma(av, vc[b], s);
                    if (b < a) pcol[b] = fma(av, vra, pcol[b]);
                }
                sm.part[r + P * a][c] = s;
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                if (b < a0) continue;
                sm.part[c + P * b][P + r] = pcol[b];
            }
        }
        __syncthreads();
        if (tau != 0.0) {
            // --- p = tau * A v: TPR threads per row, each summing 2P/TPR partials
            constexpr int CH = (2 * P) / sh::TPR;
            const int i = tid / sh::TPR, q = tid % sh::TPR;
            double s = 0.0;
            if (i < N && i > k) {
                const double* row = sm.part[i];
#pragma unroll
                for (int u = 0; u < CH; ++u) s += row[q * CH + u];
            }
#pragma unroll
            for (int o = sh::TPR / 2; o >= 1; o >>= 1) s += __shfl_xor(s, o, WAVE);
            double pi = (i < N && i > k) ? tau * s : 0.0;
            if (q == 0 && i < sh::NP) sm.p[i] = pi;
            double pv = (q == 0 && i < N) ? pi * vfun(i) : 0.0;
            pv = wave_sum(pv);
            if (lane == 0) sm.red[8 + tid / WAVE] = pv;
        }
        __syncthreads();
        if (tau != 0.0) {
            // --- rank-2 update A -= v w^T + w v^T, w = p - (tau/2)(p.v) v
            double pvt = 0.0;  // fixed summation order: bitwise reproducible
#pragma unroll
            for (int w = 0; w < sh::NW; ++w) pvt += sm.red[8 + w];
            const double coef = 0.5 * tau * pvt;
            double wc[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) wc[b] = (b >= a0) ? sm.p[c + P * b] - coef * vc[b] : 0.0;
#pragma unroll
            for (int a = 0; a < NB; ++a) {
                if (a < a0) continue;
                const double vra = vfun(r + P * a);
                const double wra = sm.p[r + P * a] - coef * vra;
#pragma unroll
                for (int b = 0; b <= a; ++b) {
                    if (b < a0) continue;
                    A[slot(a, b)] = fma(-vra, wc[b], fma(-wra, vc[b], A[slot(a, b)]));
                }
            }
        }
        emit_column<P, NB>(A, sm, xn, k + 1, N, r, c);
        __syncthreads();
    }
    // trailing 2x2 (or 1x1) block
    if (N >= 2) {
        if (tid == 0) sm.e[N - 2] = sm.x[(N - 2) & 1][N - 1];
        // owner of (N-1, N-1)
        const int kk = N - 1;
        if (r == kk % P && c == kk % P) {
            const int bb = kk / P;
#pragma unroll
            for (int b = 0; b < NB; ++b)
                if (b == bb) sm.d[kk] = A[slot(b, b)];
        }
        if (vout) {
            for (int i = tid; i < N; i += sh::NT) {
                vout[(size_t)(N - 2) * N + i] = 0.0;
                vout[(size_t)(N - 1) * N + i] = 0.0;
            }
            if (tid == 0) { tauout[N - 2] = 0.0; tauout[N - 1] = 0.0; }
        }
    } else if (vout) {
        if (tid == 0) { vout[0] = 0.0; tauout[0] = 0.0; }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Eigenvalues of the symmetric tridiagonal (d, e) by Sturm-count bisection,
// one eigenvalue per thread, ascending.  scratch: >= 2*NT/WAVE doubles.
// e2 (LDS, N entries) receives e^2.
// ---------------------------------------------------------------------------
__device__ inline int sturm_count(const double* __restrict__ d, const double* __restrict__ e2, int N, double sigma,
                                  double pivmin) {
    int cnt = 0;
    double q = d[0] - sigma;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += (q < 0.0);
    for (int i = 1; i < N; ++i) {
        q = (d[i] - sigma) - e2[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += (q < 0.0);
    }
    return cnt;
}

template <int NT>
__device__ void tridiag_eigvals(const double* d, const double* e, double* e2, int N, double* ev_out,
                                double* scratch) {
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wid = tid / WAVE;
    double lo = INFINITY, hi = -INFINITY, em = 0.0;
    for (int i = tid; i < N; i += NT) {
        double el = i > 0 ? fabs(e[i - 1]) : 0.0;
        double er = i + 1 < N ? fabs(e[i]) : 0.0;
        lo = fmin(lo, d[i] - el - er);
        hi = fmax(hi, d[i] + el + er);
        if (i + 1 < N) { e2[i] = e[i] * e[i]; em = fmax(em, e2[i]); }
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    em = wave_max(em);
    if (lane == 0) { scratch[wid] = lo; scratch[NT / WAVE + wid] = hi; scratch[2 * (NT / WAVE) + wid] = em; }
    __syncthreads();
    lo = scratch[0]; hi = scratch[NT / WAVE]; em = scratch[2 * (NT / WAVE)];
    for (int w = 1; w < NT / WAVE; ++w) {
        lo = fmin(lo, scratch[w]);
        hi = fmax(hi, scratch[NT / WAVE + w]);
        em = fmax(em, scratch[2 * (NT / WAVE) + w]);
    }
    const double eps = 2.220446049250313e-16;
    const double safemin = 2.2250738585072014e-308;
    const double tnorm = fmax(fabs(lo), fabs(hi));
    const double pivmin = safemin * fmax(1.0, em);
    lo -= 2.0 * eps * tnorm * N + 2.0 * pivmin;
    hi += 2.0 * eps * tnorm * N + 2.0 * pivmin;
    const double atol = 2.0 * eps * tnorm;  // absolute width at which bisection stops
    for (int j = tid; j < N; j += NT) {
        double l = lo, h = hi;
        for (int it = 0; it < 128; ++it) {
            double mid = 0.5 * (l + h);
            if (h - l <= atol + 2.0 * eps * fmax(fabs(l), fabs(h)) || mid == l || mid == h) break;
            int cnt = sturm_count(d, e2, N, mid, pivmin);
            if (cnt >= j + 1) h = mid; else l = mid;
        }
        ev_out[j] = 0.5 * (l + h);
    }
    __syncthreads();
}

// logZ = sum_k log(1 + e^{-beta eps_k}) with compensated (TwoSum) reduction;
// also cached_exp[k] = e^{beta eps_k}, cached_fermi[k] = 1/(1+e^{beta eps_k})
// (src/configuration.cpp:324-334).  Result returned to every thread.
template <int NT>
__device__ double logz_block(const double* ev, int N, double beta, double* cexp, double* cfermi, double* scratch) {
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wid = tid / WAVE;
    double s = 0.0, c = 0.0;
    for (int k = tid; k < N; k += NT) {
        double x = beta * ev[k];
        two_sum_merge(s, c, log1p_exp_neg(x), 0.0);
        if (cexp) {
            double ce = exp(x);
            cexp[k] = ce;
            cfermi[k] = 1.0 / (1.0 + ce);
        }
    }
    wave_sum_comp(s, c);
    __syncthreads();
    if (lane == 0) { scratch[2 * wid] = s; scratch[2 * wid + 1] = c; }
    __syncthreads();
    double S = scratch[0], Cc = scratch[1];
    for (int w = 1; w < NT / WAVE; ++w) two_sum_merge(S, Cc, scratch[2 * w], scratch[2 * w + 1]);
    __syncthreads();
    return S + Cc;
}

}  // namespace fks
