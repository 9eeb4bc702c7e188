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

// Cross-lane moves on the VALU (no LDS traffic): DPP within 16-lane rows,
// v_permlane16/32_swap across rows.  A double moves as two dwords.  bound_ctrl
// writes 0 into a lane whose source is invalid, so no "old" value has to be
// materialised first (two v_mov per move fewer than an update with old = 0, same
// result: every control used here reads a valid lane of the same row anyway).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    int2 w = *reinterpret_cast<int2*>(&v);
    w.x = __builtin_amdgcn_mov_dpp(w.x, CTRL, 0xF, 0xF, true);
    w.y = __builtin_amdgcn_mov_dpp(w.y, CTRL, 0xF, 0xF, true);
    return *reinterpret_cast<double*>(&w);
}
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <- 7-i within 8
constexpr int DPP_MIRROR = 0x140;      // lane i <- 15-i within 16
// (row r, row r^1) pair of values after v_permlane16_swap: .x holds the even
// row's value in both rows, .y the odd row's (same for the 32-lane halves).
__device__ __forceinline__ double2 swap_rows16(double v) {
    int2 w = *reinterpret_cast<int2*>(&v);
    auto a = __builtin_amdgcn_permlane16_swap(w.x, w.x, false, false);
    auto b = __builtin_amdgcn_permlane16_swap(w.y, w.y, false, false);
    int2 lo = make_int2(a[0], b[0]), hi = make_int2(a[1], b[1]);
    return make_double2(*reinterpret_cast<double*>(&lo), *reinterpret_cast<double*>(&hi));
}
__device__ __forceinline__ double2 swap_halves32(double v) {
    int2 w = *reinterpret_cast<int2*>(&v);
    auto a = __builtin_amdgcn_permlane32_swap(w.x, w.x, false, false);
    auto b = __builtin_amdgcn_permlane32_swap(w.y, w.y, false, false);
    int2 lo = make_int2(a[0], b[0]), hi = make_int2(a[1], b[1]);
    return make_double2(*reinterpret_cast<double*>(&lo), *reinterpret_cast<double*>(&hi));
}
// Two-register exchanges between lanes l and l^32 (l^16): v_permlane32_swap(x, y)
// leaves x = [x.lo, y.lo], y = [x.hi, y.hi] over the two halves (rows).
__device__ __forceinline__ double2 pl32_swap(double x, double y) {
    int2 a = *reinterpret_cast<int2*>(&x), b = *reinterpret_cast<int2*>(&y);
    auto lo = __builtin_amdgcn_permlane32_swap(a.x, b.x, false, false);
    auto hi = __builtin_amdgcn_permlane32_swap(a.y, b.y, false, false);
    int2 nx = make_int2(lo[0], hi[0]), ny = make_int2(lo[1], hi[1]);
    return make_double2(*reinterpret_cast<double*>(&nx), *reinterpret_cast<double*>(&ny));
}
__device__ __forceinline__ double2 pl16_swap(double x, double y) {
    int2 a = *reinterpret_cast<int2*>(&x), b = *reinterpret_cast<int2*>(&y);
    auto lo = __builtin_amdgcn_permlane16_swap(a.x, b.x, false, false);
    auto hi = __builtin_amdgcn_permlane16_swap(a.y, b.y, false, false);
    int2 nx = make_int2(lo[0], hi[0]), ny = make_int2(lo[1], hi[1]);
    return make_double2(*reinterpret_cast<double*>(&nx), *reinterpret_cast<double*>(&ny));
}
constexpr int DPP_ROR8 = 0x128;  // row_ror:8, lane i <- lane (i + 8) mod 16 within a 16-lane row

// Butterfly all-reduce over the 64 lanes; every lane ends with the same bits.
template <class Op>
__device__ __forceinline__ double wave_reduce(double v, Op op) {
    v = op(v, dpp_f64<DPP_XOR1>(v));
    v = op(v, dpp_f64<DPP_XOR2>(v));
    v = op(v, dpp_f64<DPP_HALF_MIRROR>(v));
    v = op(v, dpp_f64<DPP_MIRROR>(v));
    double2 t = swap_rows16(v);
    v = op(t.x, t.y);
    t = swap_halves32(v);
    return op(t.x, t.y);
}
__device__ __forceinline__ double wave_sum(double v) {
    return wave_reduce(v, [](double a, double b) { return a + b; });
}
__device__ __forceinline__ double wave_min(double v) {
    return wave_reduce(v, [](double a, double b) { return fmin(a, b); });
}
__device__ __forceinline__ double wave_max(double v) {
    return wave_reduce(v, [](double a, double b) { return fmax(a, b); });
}
// Sum over aligned groups of 4 lanes (all four lanes get the sum).
__device__ __forceinline__ double quad_sum(double v) {
    v += dpp_f64<DPP_XOR1>(v);
    return v + dpp_f64<DPP_XOR2>(v);
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

// Broadcast lane 0's value of a wave-uniform double into scalar registers.
__device__ inline double uniform(double v) {
    int2 w = *reinterpret_cast<int2*>(&v);
    w.x = __builtin_amdgcn_readfirstlane(w.x);
    w.y = __builtin_amdgcn_readfirstlane(w.y);
    return *reinterpret_cast<double*>(&w);
}

// 1/q from v_rcp_f64 refined by two Newton-Raphson steps (full double precision
// for normal q; callers keep |q| >= pivmin).
__device__ inline double rcp_refined(double q) {
    double r = __builtin_amdgcn_rcp(q);
    double e = fma(-q, r, 1.0);
    r = fma(r, e, r);
    e = fma(-q, r, 1.0);
    return fma(r, e, r);
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
__device__ __forceinline__ void assemble(double (&A)[NB * (NB + 1) / 2], const double* __restrict__ hop,
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
__device__ __forceinline__ void emit_column(const double (&A)[NB * (NB + 1) / 2], tri_smem<P, NB>& sm, double* xbuf,
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
__device__ __forceinline__ void tridiagonalize(double (&A)[NB * (NB + 1) / 2], tri_smem<P, NB>& sm, int N,
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
            tau = (beta - alpha) * rcp_refined(beta);
            scal = rcp_refined(alpha - beta);
        }
        if (tid == 0) sm.e[k] = beta;
        if (vout) {
            for (int i = tid; i < N; i += sh::NT)
                vout[(size_t)k * N + i] = (i == k + 1) ? 1.0 : ((i > k + 1) ? xk[i] * scal : 0.0);
            if (tid == 0) tauout[k] = tau;
        }
        // wave-uniform scalars live in SGPRs (scalar branches, fewer VGPRs)
        tau = uniform(tau);
        scal = uniform(scal);
        const int a0 = (k + 1) / P;
        auto vfun = [&](int i) -> double {
            return (i == k + 1) ? 1.0 : ((i > k + 1 && i < N) ? xk[i] * scal : 0.0);
        };
        if (tau != 0.0) {
            // --- symv partials, two passes so that one NB-vector is live at a time.
            // rows: part[i][c] = sum_b A[i][c+Pb] v[c+Pb] over the thread's slots
            {
                double vc[NB];
#pragma unroll
                for (int b = 0; b < NB; ++b) vc[b] = (b >= a0) ? vfun(c + P * b) : 0.0;
#pragma unroll
                for (int a = 0; a < NB; ++a) {
                    if (a < a0) continue;
                    double s = 0.0;
#pragma unroll
                    for (int b = 0; b <= a; ++b) {
                        if (b < a0) continue;
                        s = fma(A[slot(a, b)], vc[b], s);
                    }
                    sm.part[r + P * a][c] = s;
                }
            }
            // columns (strictly lower blocks only): part[j][P+r] = sum_a A[r+Pa][j] v[r+Pa]
            {
                double vr[NB];
#pragma unroll
                for (int a = 0; a < NB; ++a) vr[a] = (a > a0) ? vfun(r + P * a) : 0.0;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    if (b < a0) continue;
                    double s = 0.0;
#pragma unroll
                    for (int a = b + 1; a < NB; ++a) s = fma(A[slot(a, b)], vr[a], s);
                    sm.part[c + P * b][P + r] = s;
                }
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
                for (int u = 0; u < CH; ++u) s += row[q + sh::TPR * u];  // interleaved: conflict-free
            }
            if constexpr (sh::TPR == 4) {
                s = quad_sum(s);
            } else if constexpr (sh::TPR == 2) {
                s += dpp_f64<DPP_XOR1>(s);
            } else {
#pragma unroll
                for (int o = sh::TPR / 2; o >= 1; o >>= 1) s += __shfl_xor(s, o, WAVE);
            }
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
            const double coef = uniform(0.5 * tau * pvt);
            // two halves of the block columns keep 2*NB/2 vector values live
            constexpr int HB = (NB + 1) / 2;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int b_lo = h * HB, b_hi = (h == 0) ? HB : NB;
                double vc[HB], wc[HB];
#pragma unroll
                for (int bb = 0; bb < HB; ++bb) {
                    const int b = b_lo + bb;
                    const bool act = b < b_hi && b >= a0;
                    vc[bb] = act ? vfun(c + P * b) : 0.0;
                    wc[bb] = act ? sm.p[c + P * b] - coef * vc[bb] : 0.0;
                }
#pragma unroll
                for (int a = 0; a < NB; ++a) {
                    if (a < a0 || a < b_lo) continue;
                    const double vra = vfun(r + P * a);
                    const double wra = sm.p[r + P * a] - coef * vra;
#pragma unroll
                    for (int bb = 0; bb < HB; ++bb) {
                        const int b = b_lo + bb;
                        if (b >= b_hi || b > a || b < a0) continue;
                        A[slot(a, b)] = fma(-vra, wc[bb], fma(-wra, vc[bb], A[slot(a, b)]));
                    }
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
// Eigenvalues of the symmetric tridiagonal T = (d, e), ascending.
//
// Sturm recurrence q_0 = d_0 - s, q_i = (d_i - s) - e_{i-1}^2 / q_{i-1}: the
// number of negative q_i is the number of eigenvalues below s.  Carrying q'
// and q'' along gives G = sum_k 1/(s - l_k) = d/ds log|det(T - s)| and
// H = sum_k 1/(s - l_k)^2 = -dG/ds, the inputs of Laguerre's iteration for the
// real-rooted polynomial det(T - s) (cubic convergence, monotone towards the
// adjacent root in the chosen direction).
//   1. every thread evaluates the count at MS of MS*NT equally spaced shifts in
//      the Gershgorin interval (multisection; two shifts share one pass);
//   2. thread j takes the cell holding eigenvalue j, bisects until the cell
//      holds that eigenvalue alone, then takes Laguerre steps towards it
//      (the Sturm count says which side it is on); a step that leaves the
//      bracket is replaced by bisection;
//   3. once a step is tiny, one two-shift Sturm pass checks that l_j lies in
//      (s - t, s + t]; if not, the bracket is updated and the iteration goes on
//      from the window's edge on the eigenvalue's side.
// de: LDS, >= 2*N doubles; cnt: LDS, >= MS*NT+2 ints; scratch >= 3*NT/WAVE.
// ---------------------------------------------------------------------------
struct sturm_res {
    int count;
    double G, H;
};

__device__ __forceinline__ sturm_res sturm_laguerre_q(const double2* __restrict__ de, int N, double sigma,
                                                    double pivmin) {
    double q = de[0].x - sigma;
    if (fabs(q) < pivmin) q = -pivmin;
    double dq = -1.0, ddq = 0.0, G = 0.0, Hm = 0.0;
    int cnt = (q < 0.0);
    int i = 1;
    auto step = [&](double2 t) {
        const double r = rcp_refined(q);
        const double a = dq * r;
        G += a;
        Hm = fma(ddq, r, Hm) - a * a;
        const double u = t.y * r;
        const double ur = u * r;
        ddq = ur * fma(-2.0 * dq, a, ddq);
        dq = fma(u, a, -1.0);
        q = (t.x - sigma) - u;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += (q < 0.0);
    };
    // loads issued four at a time ahead of the dependent chain
    for (; i + 3 < N; i += 4) {
        const double2 t0 = de[i], t1 = de[i + 1], t2 = de[i + 2], t3 = de[i + 3];
        step(t0); step(t1); step(t2); step(t3);
    }
    for (; i < N; ++i) step(de[i]);
    const double r = rcp_refined(q);
    const double a = dq * r;
    G += a;
    Hm = fma(ddq, r, Hm) - a * a;
    return {cnt, G, -Hm};
}

// counts at two shifts in one pass (independent chains interleave)
__device__ __forceinline__ void sturm_count2_q(const double2* __restrict__ de, int N, double s1, double s2,
                                             double pivmin, int& c1, int& c2) {
    double q1 = de[0].x - s1, q2 = de[0].x - s2;
    if (fabs(q1) < pivmin) q1 = -pivmin;
    if (fabs(q2) < pivmin) q2 = -pivmin;
    c1 = (q1 < 0.0);
    c2 = (q2 < 0.0);
    for (int i = 1; i < N; ++i) {
        const double2 t = de[i];
        q1 = (t.x - s1) - t.y * rcp_refined(q1);
        q2 = (t.x - s2) - t.y * rcp_refined(q2);
        if (fabs(q1) < pivmin) q1 = -pivmin;
        if (fabs(q2) < pivmin) q2 = -pivmin;
        c1 += (q1 < 0.0);
        c2 += (q2 < 0.0);
    }
}

__device__ __forceinline__ int sturm_count_q(const double2* __restrict__ de, int N, double s, double pivmin) {
    double q = de[0].x - s;
    if (fabs(q) < pivmin) q = -pivmin;
    int c = (q < 0.0);
    int i = 1;
    for (; i + 3 < N; i += 4) {
        const double2 t0 = de[i], t1 = de[i + 1], t2 = de[i + 2], t3 = de[i + 3];
        q = (t0.x - s) - t0.y * rcp_refined(q); if (fabs(q) < pivmin) q = -pivmin; c += (q < 0.0);
        q = (t1.x - s) - t1.y * rcp_refined(q); if (fabs(q) < pivmin) q = -pivmin; c += (q < 0.0);
        q = (t2.x - s) - t2.y * rcp_refined(q); if (fabs(q) < pivmin) q = -pivmin; c += (q < 0.0);
        q = (t3.x - s) - t3.y * rcp_refined(q); if (fabs(q) < pivmin) q = -pivmin; c += (q < 0.0);
    }
    for (; i < N; ++i) {
        const double2 t = de[i];
        q = (t.x - s) - t.y * rcp_refined(q);
        if (fabs(q) < pivmin) q = -pivmin;
        c += (q < 0.0);
    }
    return c;
}


// The same passes in leading-minor form: P_i = det(T[0..i] - s), P_{-1} = 1,
//   P_i = (d_i - s) P_{i-1} - e_{i-1}^2 P_{i-2},
// so q_i = P_i / P_{i-1}, the count is the number of sign changes of P, and
// G = P'/P, H = G^2 - P''/P (derivatives in s, carried by the differentiated
// recurrence).  A step is one dependent FMA instead of a reciprocal, and its rounding
// is that of the ratio form (relative perturbations of d_i - s and of e^2, one each),
// so the counts are those of a nearby matrix as before.  The carried values are
// rescaled by a power of two every SPF steps (exact); a vanishing pivot (the test below)
// or a value out of range sends the whole pass to the ratio form above.
constexpr int SPF = 8;  // steps between rescalings

__device__ __forceinline__ void rescale2(double& a1, double& a2, int ex) {
    a1 = __builtin_amdgcn_ldexp(a1, -ex);
    a2 = __builtin_amdgcn_ldexp(a2, -ex);
}

// Per-step bookkeeping kept in vector registers: a comparison's lane mask would pass through
// scalar registers and stall the pass's in-order instruction stream once per step.
// Sign change of consecutive leading minors: their sign bits differ (no minor is zero while
// the pivot test below holds).
__device__ __forceinline__ int sign_change(double p, double p1) {
    return (int)(((unsigned)__double2hiint(p) ^ (unsigned)__double2hiint(p1)) >> 31);
}
// Pivot test of every step as the minimum of |P_i| over the pass, against a floor: every
// SPF steps the carried minors are rescaled to |P| in [1/2, 1), and in 8 steps a normal pass
// moves them by far less than 2^-900 (ratios |q| of 10^-3 to 10^2), so a minor below the floor
// is an exact or underflowing pivot (s on an eigenvalue of a leading block to the last bits),
// where the signs of the minors that follow stop counting (a split e_i = 0 then zeroes all of
// them); such a pass goes to the ratio form.  One min a step (the relative test
// |P_i| >= pivmin |P_{i-1}| took an FMA more).  A NaN leaves the minimum unchanged but reaches
// the final value, which the callers test for finiteness.
constexpr double PIVOT_FLOOR = 0x1p-900;
__device__ __forceinline__ double pivot_margin(double mn, double p, double p1, double pivmin) {
    (void)p1;
    (void)pivmin;
    return fmin(mn, fabs(p));
}

__device__ __forceinline__ sturm_res sturm_laguerre(const double2* __restrict__ de, int N, double sigma,
                                                    double pivmin) {
    double p1 = 1.0, p2 = 0.0, d1 = 0.0, d2 = 0.0, h1 = 0.0, h2 = 0.0;
    int cnt = 0;
    double mn = 1.0;
    auto step = [&](double2 t) {
        const double c = t.x - sigma;
        const double p = fma(c, p1, -t.y * p2);
        const double dp = fma(c, d1, -fma(t.y, d2, p1));
        const double hp = fma(c, h1, -fma(t.y, h2, 2.0 * d1));
        mn = pivot_margin(mn, p, p1, pivmin);
        cnt += sign_change(p, p1);
        p2 = p1; p1 = p;
        d2 = d1; d1 = dp;
        h2 = h1; h1 = hp;
    };
    int i = 0;
    for (; i + SPF <= N; i += SPF) {
        double2 t[SPF];
#pragma unroll
        for (int q = 0; q < SPF; ++q) t[q] = de[i + q];
#pragma unroll
        for (int q = 0; q < SPF; ++q) step(t[q]);
        const int ex = __builtin_amdgcn_frexp_exp(p1);
        rescale2(p1, p2, ex);
        rescale2(d1, d2, ex);
        rescale2(h1, h2, ex);
    }
    for (; i < N; ++i) step(de[i]);
    const double G = d1 / p1;
    const double H = fma(G, G, -h1 / p1);
    if (!(mn >= PIVOT_FLOOR) || !isfinite(G) || !isfinite(H)) return sturm_laguerre_q(de, N, sigma, pivmin);
    return {cnt, G, H};
}

__device__ __forceinline__ int sturm_count(const double2* __restrict__ de, int N, double s, double pivmin) {
    double p1 = 1.0, p2 = 0.0;
    int c = 0;
    double mn = 1.0;
    auto step = [&](double2 t) {
        const double p = fma(t.x - s, p1, -t.y * p2);
        mn = pivot_margin(mn, p, p1, pivmin);
        c += sign_change(p, p1);
        p2 = p1; p1 = p;
    };
    int i = 0;
    for (; i + SPF <= N; i += SPF) {
        double2 t[SPF];
#pragma unroll
        for (int q = 0; q < SPF; ++q) t[q] = de[i + q];
#pragma unroll
        for (int q = 0; q < SPF; ++q) step(t[q]);
        rescale2(p1, p2, __builtin_amdgcn_frexp_exp(p1));
    }
    for (; i < N; ++i) step(de[i]);
    if (!(mn >= PIVOT_FLOOR) || !isfinite(p1)) return sturm_count_q(de, N, s, pivmin);
    return c;
}

// counts at two shifts in one pass (independent chains interleave)
__device__ __forceinline__ void sturm_count2(const double2* __restrict__ de, int N, double s1, double s2,
                                             double pivmin, int& c1, int& c2) {
    double a1 = 1.0, a2 = 0.0, b1 = 1.0, b2 = 0.0;
    int ca = 0, cb = 0;
    double mna = 1.0, mnb = 1.0;
    auto step = [&](double2 t) {
        const double pa = fma(t.x - s1, a1, -t.y * a2);
        const double pb = fma(t.x - s2, b1, -t.y * b2);
        mna = pivot_margin(mna, pa, a1, pivmin);
        mnb = pivot_margin(mnb, pb, b1, pivmin);
        ca += sign_change(pa, a1);
        cb += sign_change(pb, b1);
        a2 = a1; a1 = pa;
        b2 = b1; b1 = pb;
    };
    int i = 0;
    for (; i + SPF <= N; i += SPF) {
        double2 t[SPF];
#pragma unroll
        for (int q = 0; q < SPF; ++q) t[q] = de[i + q];
#pragma unroll
        for (int q = 0; q < SPF; ++q) step(t[q]);
        rescale2(a1, a2, __builtin_amdgcn_frexp_exp(a1));
        rescale2(b1, b2, __builtin_amdgcn_frexp_exp(b1));
    }
    for (; i < N; ++i) step(de[i]);
    if (!(mna >= PIVOT_FLOOR) || !(mnb >= PIVOT_FLOOR) || !isfinite(a1) || !isfinite(b1)) {
        sturm_count2_q(de, N, s1, s2, pivmin, c1, c2);
        return;
    }
    c1 = ca;
    c2 = cb;
}

// counts at four shifts in one pass (four independent chains)
__device__ __forceinline__ void sturm_count4(const double2* __restrict__ de, int N, const double (&s)[4], double pivmin,
                                             int (&c)[4]) {
    double a1[4], a2[4], mn[4];
    int cc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { a1[u] = 1.0; a2[u] = 0.0; mn[u] = 1.0; cc[u] = 0; }
    auto step = [&](double2 t) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double p = fma(t.x - s[u], a1[u], -t.y * a2[u]);
            mn[u] = pivot_margin(mn[u], p, a1[u], pivmin);
            cc[u] += sign_change(p, a1[u]);
            a2[u] = a1[u];
            a1[u] = p;
        }
    };
    int i = 0;
    for (; i + SPF <= N; i += SPF) {
        double2 t[SPF];
#pragma unroll
        for (int q = 0; q < SPF; ++q) t[q] = de[i + q];
#pragma unroll
        for (int q = 0; q < SPF; ++q) step(t[q]);
#pragma unroll
        for (int u = 0; u < 4; ++u) rescale2(a1[u], a2[u], __builtin_amdgcn_frexp_exp(a1[u]));
    }
    for (; i < N; ++i) step(de[i]);
    bool ok = true;
#pragma unroll
    for (int u = 0; u < 4; ++u) ok &= (mn[u] >= PIVOT_FLOOR) && isfinite(a1[u]);
    if (!ok) {
        sturm_count2_q(de, N, s[0], s[1], pivmin, c[0], c[1]);
        sturm_count2_q(de, N, s[2], s[3], pivmin, c[2], c[3]);
        return;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = cc[u];
}

template <int NT, int MS = 1>
__device__ __forceinline__ void tridiag_eigvals(const double* d, const double* e, double* de_buf, int* cnt, int N,
                                                double* ev_out, double* scratch) {
    static_assert(MS == 1 || MS == 2 || MS == 4, "one, two or four multisection shifts per thread");
    constexpr int NC = MS * NT;  // interior multisection points
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1), wid = tid / WAVE;
    double2* de = reinterpret_cast<double2*>(de_buf);
    double lo = INFINITY, hi = -INFINITY, em = 0.0;
    for (int i = tid; i < N; i += NT) {
        double el = i > 0 ? fabs(e[i - 1]) : 0.0;
        double er = i + 1 < N ? fabs(e[i]) : 0.0;
        lo = fmin(lo, d[i] - el - er);
        hi = fmax(hi, d[i] + el + er);
        double e2 = i > 0 ? e[i - 1] * e[i - 1] : 0.0;
        de[i] = make_double2(d[i], e2);
        em = fmax(em, e2);
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
    const double atol = 2.0 * eps * tnorm;
    const double vtol = 8.0 * eps * fmax(tnorm, 1.0);   // verification half-window
    // Laguerre step below which to verify: from a step of 1e-6 |T| the cubic convergence still
    // leaves errors of a window width (8 eps |T|) near a neighbour, and a failed test costs two passes
    const double small = 1e-7 * fmax(tnorm, 1.0);
    const double h = (hi - lo) / (double)(NC + 1);
    // 1. multisection: cnt[m] = count at lo + m*h, m = 0..NC+1
    if (MS == 1) {
        cnt[tid + 1] = sturm_count(de, N, lo + (double)(tid + 1) * h, pivmin);
    } else if (MS == 2) {
        int c1, c2;
        sturm_count2(de, N, lo + (double)(2 * tid + 1) * h, lo + (double)(2 * tid + 2) * h, pivmin, c1, c2);
        cnt[2 * tid + 1] = c1;
        cnt[2 * tid + 2] = c2;
    } else {
        const double sh[4] = {lo + (double)(4 * tid + 1) * h, lo + (double)(4 * tid + 2) * h,
                              lo + (double)(4 * tid + 3) * h, lo + (double)(4 * tid + 4) * h};
        int c4[4];
        sturm_count4(de, N, sh, pivmin, c4);
#pragma unroll
        for (int u = 0; u < 4; ++u) cnt[4 * tid + u + 1] = c4[u];
    }
    if (tid == 0) { cnt[0] = 0; cnt[NC + 1] = N; }
    __syncthreads();
    const double nd = (double)N;
    for (int j = tid; j < N; j += NT) {
        int ml = 0, mh = NC + 1;  // invariant: cnt[ml] <= j < cnt[mh]
        while (mh - ml > 1) {
            int mid = (ml + mh) >> 1;
            if (cnt[mid] > j) mh = mid; else ml = mid;
        }
        double l = lo + (double)ml * h, u = (mh == NC + 1) ? hi : lo + (double)mh * h;
        int cl = cnt[ml], cu = cnt[mh];
        double s = 0.5 * (l + u);
        double res = s;
        double prev = INFINITY;  // the last Laguerre step taken
        for (int it = 0; it < 200; ++it) {
            if (u - l <= atol + 2.0 * eps * fmax(fabs(l), fabs(u))) { res = 0.5 * (l + u); break; }
            const sturm_res sr = sturm_laguerre(de, N, s, pivmin);
            const bool left = sr.count > j;  // eigenvalue j lies below s
            if (left) { u = s; cu = sr.count; } else { l = s; cl = sr.count; }
            double next = 0.5 * (l + u);
            bool took = false;  // a Laguerre step
            if (cu - cl == 1) {
                const double disc = fmax((nd - 1.0) * (nd * sr.H - sr.G * sr.G), 0.0);
                const double sq = sqrt(disc);
                // the two Laguerre candidates lie on either side of s; take the one towards eigenvalue j
                const double den = left ? (sr.G + sq) : (sr.G - sq);
                if (den != 0.0 && isfinite(den)) {
                    const double sn = s - nd / den;
                    if (sn > l && sn < u) {
                        if (fabs(sn - s) <= small) {
                            int ca, cb;
                            sturm_count2(de, N, sn - vtol, sn + vtol, pivmin, ca, cb);
                            if (ca <= j && cb > j) { res = sn; break; }
                            // the eigenvalue lies just outside the window: go on from the window's
                            // edge on its side (a restart from the bracket midpoint costs 2-3 passes)
                            if (ca > j) { u = sn - vtol; cu = ca; next = u; }
                            else if (cb <= j) { l = sn + vtol; cl = cb; next = l; }
                            else next = 0.5 * (l + u);
                        } else if (!(fabs(sn - s) > 0.5 * prev)) {
                            // (steps that do not shrink: s sits next to a neighbouring eigenvalue,
                            // which holds Laguerre's steps towards eigenvalue j to a doubling
                            // sequence; those bisect instead)
                            next = sn;
                            took = true;
                        }
                    } else if (fabs(sn - s) <= small && (sn >= u || sn <= l)) {
                        // a short step onto or past a bracket edge: the eigenvalue lies within
                        // rounding of that edge (typically the edge a failed window test left
                        // just beside it), where Laguerre's step no longer resolves it; one count
                        // tests the window of width 2 vtol inside the edge (bisection from here
                        // takes some 30 passes)
                        if (sn >= u) {
                            const int c = sturm_count(de, N, u - 2.0 * vtol, pivmin);
                            if (c <= j) { res = u - vtol; break; }
                            u = u - 2.0 * vtol; cu = c;
                        } else {
                            const int c = sturm_count(de, N, l + 2.0 * vtol, pivmin);
                            if (c > j) { res = l + vtol; break; }
                            l = l + 2.0 * vtol; cl = c;
                        }
                        next = 0.5 * (l + u);
                    }
                }
            }
            if (took) prev = fabs(next - s);
            else if (next == 0.5 * (l + u)) prev = INFINITY;
            if (next == s) { res = s; break; }
            s = next;
            res = s;
        }
        ev_out[j] = res;
    }
    __syncthreads();
}

// logZ = sum_k log(1 + e^{-beta eps_k}) with compensated (TwoSum) reduction;
// also cached_exp[k] = e^{beta eps_k}, cached_fermi[k] = 1/(1+e^{beta eps_k})
// (src/configuration.cpp:324-334).  Result returned to every thread.
template <int NT>
__device__ __noinline__ double logz_block(const double* ev, int N, double beta, double* cexp, double* cfermi, double* scratch) {
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
