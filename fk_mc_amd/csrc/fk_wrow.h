// fk_wrow.h -- register-resident Householder tridiagonalisation, "wave-row" layout,
// for N <= 256: one workgroup of NWV waves (16 or 8) per matrix.
//
// Wave w holds the rows i = w + NWV a (a = 0 .. 256/NWV - 1), lane l the columns j = l + 64b
// (b = 0..3).  Thread (w, l) keeps A[w + NWV a][l + 64b] for b <= (row block of a) in VGPRs
// (40 doubles with 16 waves, 80 with 8): every 64 x 64 diagonal block is stored in full, the
// blocks below it once.  For a Householder step (Golub & Van Loan 8.3.1):
//   * a lane's column values v_j, w_j (j = l + 64b) are its own registers; the row values
//     v_i, p_i (i = w + NWV a, uniform over the wave) come from LDS by broadcast reads;
//   * the row sums of p = A v are reduced across the 64 lanes by a reduce-scatter on
//     v_permlane32_swap / v_permlane16_swap / DPP (no LDS); the column sums of the blocks
//     below the diagonal go through LDS (NWV partials per column);
//   * each step has two workgroup barriers (see step()).
// With 16 waves four share a SIMD and the matrix takes 40 of a lane's 128 VGPRs; with 8 two
// share a SIMD, 80 of 256: half the waves repeat the per-wave work (reflector, reductions)
// and each has twice the rows to interleave.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fk_solve.h"

namespace fkw {

using namespace fks;

__device__ __forceinline__ double readlane_d(double v, int lane) {
    int2 w = *reinterpret_cast<int2*>(&v);
    w.x = __builtin_amdgcn_readlane(w.x, lane);
    w.y = __builtin_amdgcn_readlane(w.y, lane);
    return *reinterpret_cast<double*>(&w);
}
// Reduce-scatter of 4 values per lane over the 64 lanes: afterwards every lane holds
// the full sum of value number 2*bit5 + bit4 of its lane id.
__device__ __forceinline__ double reduce4(double (&r)[4]) {
    // lanes l and l^32: v_permlane32_swap(x, y) leaves x = [x.lo, y.lo], y = [x.hi, y.hi]
    double s[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const double2 t = pl32_swap(r[a], r[a + 2]);
        s[a] = t.x + t.y;
    }
    // lanes l and l^16: v_permlane16_swap exchanges the odd rows of x with the even rows of y
    const double2 t = pl16_swap(s[0], s[1]);
    double u = t.x + t.y;
    // the 16 lanes sharing bits 5, 4 hold partial sums of the same value: all-reduce them
    // (row_ror:8 pairs lanes l, l^8; quad_perm l^1, l^2; row_half_mirror l^7)
    u += dpp_f64<DPP_ROR8>(u);
    u += dpp_f64<DPP_XOR1>(u);
    u += dpp_f64<DPP_XOR2>(u);
    u += dpp_f64<DPP_HALF_MIRROR>(u);
    return u;
}

// Reduce-scatter of 8 values per lane over the 64 lanes: afterwards every lane holds the
// full sum of value number 4*bit5 + 2*bit4 + bit3 of its lane id (34 VALU instructions
// against 42 for two reduce4).
__device__ __forceinline__ double reduce8(double (&r)[8]) {
    // lanes l and l^32: the low half sums value a, the high half value a + 4
    double s[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double2 t = pl32_swap(r[a], r[a + 4]);
        s[a] = t.x + t.y;
    }
    // lanes l and l^16: rows with bit 4 clear sum value a, the others a + 2
    double u[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const double2 t = pl16_swap(s[a], s[a + 2]);
        u[a] = t.x + t.y;
    }
    // lanes l and l^8 (row_ror:8): a lane keeps u[bit3] and passes its partner the other value
    const bool hi = (threadIdx.x & 8) != 0;
    const double keep = hi ? u[1] : u[0], send = hi ? u[0] : u[1];
    double v = keep + dpp_f64<DPP_ROR8>(send);
    // the 8 lanes sharing bits 5..3 hold partials of the same value: all-reduce them
    v += dpp_f64<DPP_XOR1>(v);
    v += dpp_f64<DPP_XOR2>(v);
    v += dpp_f64<DPP_HALF_MIRROR>(v);
    return v;
}

template <int NWV_>
struct wrow_solver_t {
    static constexpr int NWV = NWV_;       // waves = row residues
    static constexpr int NT = NWV * WAVE;
    static constexpr int NP = 256;
    static constexpr int NMAX = NP;
    static constexpr int NA = NP / NWV;    // rows per thread
    static constexpr int RB = WAVE / NWV;  // rows per thread in each 64-row block
    static constexpr int NBC = 4;          // column blocks per thread
    static constexpr int NS = 10 * RB;     // slots per thread: RB (1 + 2 + 3 + 4)
    static_assert(NWV == 16 || NWV == 8, "16 or 8 waves per matrix");
    static constexpr int EMS = NT >= 1024 ? 2 : 4;  // multisection shifts per thread: 2048 cells
    struct eig_area {
        double de[2 * NP];
        double ev[NP];
        int cnt[EMS * NT + 2];
        double scratch[3 * NWV + 8];
    };
    struct smem {
        double x[2][NP];       // x[(k + 1) & 1]: column k + 1 of A before step k's update
        double y[NP];          // step k's column below the diagonal (y_j, j > k), unscaled
        double rs[NP];
        double p[2][NP];       // p and v of step k in [k & 1] (read again by step k + 1)
        double v[2][NP];
        double d[NP], e[NP];
        double red[NWV + 8];
        union {
            double cpart[NWV][NP];
            eig_area ea;
        } u;
        double beta_s[2];      // beta across the tridiagonalisation (kept out of registers)
    };
    // 16-byte aligned buffers keep the 128-bit LDS accesses whole (one 8-byte shift of the
    // partials made the solve 11 % slower)
    static_assert(offsetof(smem, u) % 16 == 0 && offsetof(smem, rs) % 16 == 0 && offsetof(smem, v) % 16 == 0 &&
                      offsetof(smem, y) % 16 == 0,
                  "LDS buffers 16-byte aligned");
    // slot of row a (block g = a / RB, g + 1 column blocks) and column block b <= g
    __host__ __device__ static constexpr int sl(int a, int b) {
        return RB * ((a / RB) * (a / RB + 1) / 2) + (a % RB) * (a / RB + 1) + b;
    }
    __device__ static const double* dptr(const smem& sm) { return sm.d; }
    __device__ static const double* eptr(const smem& sm) { return sm.e; }
    __device__ static double* evptr(smem& sm) { return sm.u.ea.ev; }

    // A[i][j] (a stored slot: i / 64 >= j / 64) from the thread that holds it into *dst
    __device__ __forceinline__ static void put_elem(const double (&A)[NS], int i, int j, int w, int l, double* dst) {
        if (w != (i % NWV) || l != (j & 63)) return;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b <= a / RB; ++b)
                if (a == i / NWV && b == (j >> 6)) *dst = A[sl(a, b)];
    }

    // The step as 16 waves take it (128 registers a lane leave no room to overlap the sums with
    // the reflector's chain): the reflector first, then A v on v.  Same phases, barriers and
    // buffers as step(), which describes them; the column handed to the next step is stored
    // after the reflector and d, e right after it.
    template <int BM, bool CROSS>
    __device__ __forceinline__ static void step_v(double (&A)[NS], smem& sm, int N, int k, int w, int l, int tid,
                                                double* __restrict__ vout, double* __restrict__ tauout,
                                                double& tprev, double& cprev) {
        static_assert(!CROSS || BM < 3, "column k + 1 lies in block BM or BM + 1 <= 3");
        constexpr int CB = CROSS ? BM + 1 : BM;
        const int pk = k & 1, pq = pk ^ 1;
        const double* xe = sm.x[pk];
        const double* vq = sm.v[pq];
        const double* prq = sm.p[pq];
        // A[i][k] after step k - 1: without an update there (tau = 0) v_k = w_k = 0 leave the
        // stored column as it is (v, p hold finite older values), so no branch keeps the loads
        // of the lane's four entries from going out together
        const bool upd = tprev != 0.0;  // wave-uniform
        const double cq = cprev;
        const double vk = upd ? vq[k] : 0.0;
        const double wk = upd ? fma(-cq, vk, prq[k]) : 0.0;
        auto col = [&](int i) {
            const double vi = vq[i];
            const double wi = fma(-cq, vi, prq[i]);
            return fma(-vi, wk, fma(-wi, vk, xe[i]));
        };
        // reflector from x = A[k+1:N, k] (every wave redundantly), the lane's entries l + 64b
        const double alpha = col(k + 1);
        double xb[NBC], part = 0.0;
#pragma unroll
        for (int b = BM; b < NBC; ++b) {
            const int j = l + 64 * b;
            xb[b] = col(j);
            part += (j >= k + 2 && j < N) ? xb[b] * xb[b] : 0.0;
        }
        const double xn2 = wave_sum(part);
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (xn2 != 0.0) {
            beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
            tau = (beta - alpha) * rcp_refined(beta);
            scal = rcp_refined(alpha - beta);
        }
        if (tid == 0) {
            sm.e[k] = beta;
            sm.d[k] = col(k);  // A[k][k] is final after step k - 1
        }
        tau = uniform(tau);
        scal = uniform(scal);
        // v at the lane's columns; every wave stores the same values to sm.v, from where the
        // wave reads v at its rows (its own stores are ordered before its loads)
        double* vs = sm.v[pk];
        double vc[NBC];
#pragma unroll
        for (int b = 0; b < NBC; ++b) {
            const int j = l + 64 * b;
            vc[b] = b < BM ? 0.0 : ((j == k + 1) ? 1.0 : ((j > k + 1 && j < N) ? xb[b] * scal : 0.0));
            if (b >= BM) vs[j] = vc[b];
        }
        if (vout) {
            if (w == 0) {
#pragma unroll
                for (int b = 0; b < NBC; ++b)
                    if (l + 64 * b < N) vout[(size_t)k * N + l + 64 * b] = vc[b];
            }
            if (tid == 0) tauout[k] = tau;
        }
        // column k + 1 before this step's update, for the next step
        if (l == ((k + 1) & 63)) {
            double* xn = sm.x[pq];
#pragma unroll
            for (int a = RB * CB; a < NA; ++a) xn[w + NWV * a] = A[sl(a, CB)];
        }
        if (tau != 0.0) {
            // p = A v as column sums over the diagonal blocks and the blocks below them (the
            // diagonal blocks are stored in full, so their column sums are complete; partials
            // per wave through LDS) plus row sums over the blocks left of the diagonal blocks
            // (reduced over the lanes).  Rows of a finished block are not needed (p is zero
            // there), and the rows of block BM have no blocks left of theirs, so only rows
            // in blocks BM + 1 .. 3 are reduced across lanes.
            // The groups' row sums are stored together after all reductions: a store per group
            // under a lane condition would split the code into blocks and keep the groups'
            // reductions from interleaving.
            constexpr int H0 = RB * (BM + 1) / 4, NH = NA / 4 - H0;
            double rsum[NH > 0 ? NH : 1];
#pragma unroll
            for (int h = H0; h < NA / 4; ++h) {
                const int g = 4 * h / RB;
                double rp[4];
#pragma unroll
                for (int aa = 0; aa < 4; ++aa) {
                    const int a = 4 * h + aa;
                    double s = 0.0;
#pragma unroll
                    for (int b = BM; b < g; ++b) s = fma(A[sl(a, b)], vc[b], s);
                    rp[aa] = s;
                }
                rsum[h - H0] = reduce4(rp);
            }
            if (NH > 0 && (l & 15) == 0) {
                const int ar = 2 * ((l >> 5) & 1) + ((l >> 4) & 1);
#pragma unroll
                for (int h = H0; h < NA / 4; ++h) sm.rs[w + NWV * (4 * h + ar)] = rsum[h - H0];
            }
            double cs[NBC];
#pragma unroll
            for (int b = 0; b < NBC; ++b) cs[b] = 0.0;
#pragma unroll
            for (int g = BM; g < NBC; ++g) {
                double vr[RB];  // a 64-row block's rows of the wave loaded together
#pragma unroll
                for (int aa = 0; aa < RB; ++aa) vr[aa] = vs[w + NWV * (RB * g + aa)];
#pragma unroll
                for (int aa = 0; aa < RB; ++aa)
#pragma unroll
                    for (int b = BM; b <= g; ++b) cs[b] = fma(A[sl(RB * g + aa, b)], vr[aa], cs[b]);
            }
#pragma unroll
            for (int b = BM; b < NBC; ++b) sm.u.cpart[w][l + 64 * b] = cs[b];
        }
        __syncthreads();
        double* ps = sm.p[pk];
        constexpr int PW = NP / WAVE;  // waves forming p: one row a lane, one wave a SIMD
        if (tau != 0.0 && w < PW) {
            // p_i = tau (row sum + NWV column partials, pairwise in a fixed order); entries of
            // finished rows and columns hold older partials and are masked.  The other waves go
            // on to the barrier: this phase is a latency chain, which fewer waves do not lengthen.
            const int i = tid;
            double t[NWV / 4];
#pragma unroll
            for (int g = 0; g < NWV / 4; ++g)
                t[g] = (sm.u.cpart[4 * g][i] + sm.u.cpart[4 * g + 1][i]) +
                       (sm.u.cpart[4 * g + 2][i] + sm.u.cpart[4 * g + 3][i]);
            const double tt = NWV == 16 ? (t[0] + t[1]) + (t[NWV / 4 - 2] + t[NWV / 4 - 1]) : t[0] + t[1];
            const double s = ((i >> 6) > BM ? sm.rs[i] : 0.0) + tt;
            const double pi = (i > k && i < N) ? tau * s : 0.0;
            ps[i] = pi;
            const double vi = (i > k) ? vs[i] : 0.0;
            const double pv = wave_sum(pi * vi);
            if (l == 0) sm.red[w] = pv;
        }
        __syncthreads();
        double coef = 0.0;
        if (tau != 0.0) {
            double pvt = 0.0;  // fixed order: bitwise reproducible
#pragma unroll
            for (int q = 0; q < PW; ++q) pvt += sm.red[q];
            coef = uniform(0.5 * tau * pvt);
            double wc[NBC];
#pragma unroll
            for (int b = BM; b < NBC; ++b) wc[b] = fma(-coef, vc[b], ps[l + 64 * b]);
            // row values by broadcast LDS reads, a 64-row block's rows of the wave loaded together
            // (one LDS round trip a block), w_i formed as the lanes form w_j; a finished row
            // (v_i = p_i = 0) is left as it is
#pragma unroll
            for (int g = BM; g < NBC; ++g) {
                double vra[RB], pra[RB];
#pragma unroll
                for (int aa = 0; aa < RB; ++aa) {
                    const int i = w + NWV * (RB * g + aa);
                    vra[aa] = vs[i];
                    pra[aa] = ps[i];
                }
#pragma unroll
                for (int aa = 0; aa < RB; ++aa) {
                    const int a = RB * g + aa;
                    const double wra = fma(-coef, vra[aa], pra[aa]);
#pragma unroll
                    for (int b = BM; b <= g; ++b)
                        A[sl(a, b)] = fma(-vra[aa], wc[b], fma(-wra, vc[b], A[sl(a, b)]));
                }
            }
        }
        tprev = tau;
        cprev = coef;
        // no barrier here: this step's buffers that the next step writes before its first
        // barrier (x[k & 1], v[(k + 1) & 1], rs, cpart) were last read before this step's
        // second barrier, and the next step writes p, red only after its first barrier
    }

    // Householder step k with BM = k / 64 finished column blocks known at compile time
    // (slots of those blocks take no part) and CROSS when column k + 1 lies in block BM + 1.
    //
    // Look-ahead: column k of A after step k - 1 is not handed over after that step's update
    // (which would need a barrier between the update and this step's reflector); every wave
    // forms it itself from column k before step k - 1's update (stored during step k - 1,
    // before its updates) and step k - 1's v, p and coef, as the update forms the registers.
    // So a wave that has finished step k - 1's update goes on into step k while others still
    // update, and a step has two barriers.  tprev / cprev carry step k - 1's tau and coef.
    //
    // The product A v is formed from the unscaled column y (y_j = A[j][k], j > k), known before
    // the reflector's norm, square root and reciprocals: with v = scal y + (1 - scal alpha)
    // e_{k+1}, A v = scal (A y) + (1 - scal alpha) A[:, k+1], and A[:, k+1] is the column this
    // step stores for the next one.  The reflector's latency chain and the product's FMAs are
    // one branch-free stretch of code, which the compiler interleaves.
    template <int BM, bool CROSS>
    __device__ __forceinline__ static void step(double (&A)[NS], smem& sm, int N, int k, int w, int l, int tid,
                                                double* __restrict__ vout, double* __restrict__ tauout,
                                                double& tprev, double& cprev) {
        if constexpr (NWV == 16) {
            step_v<BM, CROSS>(A, sm, N, k, w, l, tid, vout, tauout, tprev, cprev);
        } else {
            static_assert(!CROSS || BM < 3, "column k + 1 lies in block BM or BM + 1 <= 3");
            constexpr int CB = CROSS ? BM + 1 : BM;
            const int pk = k & 1, pq = pk ^ 1;
            // column k + 1 before this step's update, for the next step and for A v
            if (l == ((k + 1) & 63)) {
                double* xn = sm.x[pq];
                // rows <= k of block CB are not read (a wave-uniform test per row)
#pragma unroll
                for (int a = RB * CB; a < RB * CB + RB; ++a)
                    if (w + NWV * a > k) xn[w + NWV * a] = A[sl(a, CB)];
#pragma unroll
                for (int a = RB * CB + RB; a < NA; ++a) xn[w + NWV * a] = A[sl(a, CB)];
            }
            const double* xe = sm.x[pk];
            const double* vq = sm.v[pq];
            const double* prq = sm.p[pq];
            // A[i][k] after step k - 1: without an update there (tau = 0) v_k = w_k = 0 leave the
            // stored column as it is (v, p hold finite older values), so no branch keeps the loads
            // of the lane's four entries from going out together
            const bool upd = tprev != 0.0;  // wave-uniform
            const double cq = cprev;
            const double vk = upd ? vq[k] : 0.0;
            const double wk = upd ? fma(-cq, vk, prq[k]) : 0.0;
            auto col = [&](int i) {
                const double vi = vq[i];
                const double wi = fma(-cq, vi, prq[i]);
                return fma(-vi, wk, fma(-wi, vk, xe[i]));
            };
            // y at the lane's entries l + 64b (every wave redundantly; every wave stores the same
            // values to sm.y, from where the wave reads y at its rows: its own stores are ordered
            // before its loads)
            const double alpha = col(k + 1);
            const double dk = col(k);  // A[k][k] is final after step k - 1
            double* ys = sm.y;
            double yc[NBC], part = 0.0;
#pragma unroll
            for (int b = 0; b < NBC; ++b) {
                const int j = l + 64 * b;
                if (b < BM) {
                    yc[b] = 0.0;
                } else {
                    const double x = col(j);
                    yc[b] = (j > k && j < N) ? x : 0.0;
                    part += (j >= k + 2 && j < N) ? x * x : 0.0;
                    ys[j] = yc[b];
                }
            }
            // reflector (selects instead of branches)
            const double xn2 = wave_sum(part);
            const bool refl = xn2 != 0.0;
            const double bt = -copysign(sqrt(alpha * alpha + xn2), alpha);
            const double tau = uniform(refl ? (bt - alpha) * rcp_refined(bt) : 0.0);
            const double scal = uniform(refl ? rcp_refined(alpha - bt) : 0.0);
            const double beta = refl ? bt : alpha;
            // A y as column sums over the diagonal blocks and the blocks below them (the diagonal
            // blocks are stored in full, so their column sums are complete; partials per wave
            // through LDS) plus row sums over the blocks left of the diagonal blocks (reduced over
            // the lanes).  Rows of a finished block are not needed (p is zero there), and the rows
            // of block BM have no blocks left of theirs, so only rows in blocks BM + 1 .. 3 are
            // reduced across lanes.  The groups' row sums are stored together after all
            // reductions: a store per group under a lane condition would split the code into
            // blocks and keep the groups' reductions from interleaving.
            // (with 8 waves a wave has 8 rows in a block: one 8-row reduce-scatter a block)
            static_assert(RB == 8, "8 rows of a wave in a block");
            constexpr int G0 = BM + 1, NG = NBC - G0;
            double rsum[NG > 0 ? NG : 1];
#pragma unroll
            for (int g = G0; g < NBC; ++g) {
                double rp[8];
#pragma unroll
                for (int aa = 0; aa < 8; ++aa) {
                    const int a = RB * g + aa;
                    double t = 0.0;
#pragma unroll
                    for (int b = BM; b < g; ++b) t = fma(A[sl(a, b)], yc[b], t);
                    rp[aa] = t;
                }
                rsum[g - G0] = reduce8(rp);
            }
            double cs[NBC];
#pragma unroll
            for (int b = 0; b < NBC; ++b) cs[b] = 0.0;
#pragma unroll
            for (int g = BM; g < NBC; ++g) {
                double vr[RB];  // a 64-row block's rows of the wave loaded together
#pragma unroll
                for (int aa = 0; aa < RB; ++aa) vr[aa] = ys[w + NWV * (RB * g + aa)];
#pragma unroll
                for (int aa = 0; aa < RB; ++aa)
#pragma unroll
                    for (int b = BM; b <= g; ++b) cs[b] = fma(A[sl(RB * g + aa, b)], vr[aa], cs[b]);
            }
#pragma unroll
            for (int b = BM; b < NBC; ++b) sm.u.cpart[w][l + 64 * b] = cs[b];
            // v at the lane's columns, to sm.v (read at the wave's rows after the first barrier)
            double* vs = sm.v[pk];
            double vc[NBC];
#pragma unroll
            for (int b = 0; b < NBC; ++b) {
                const int j = l + 64 * b;
                vc[b] = b < BM ? 0.0 : ((j == k + 1) ? 1.0 : ((j > k + 1 && j < N) ? yc[b] * scal : 0.0));
                if (b >= BM) vs[j] = vc[b];
            }
            if (NG > 0 && (l & 7) == 0) {
                const int ar = (l >> 3) & 7;
#pragma unroll
                for (int g = G0; g < NBC; ++g) sm.rs[w + NWV * (RB * g + ar)] = rsum[g - G0];
            }
            if (vout) {
                if (w == 0) {
#pragma unroll
                    for (int b = 0; b < NBC; ++b)
                        if (l + 64 * b < N) vout[(size_t)k * N + l + 64 * b] = vc[b];
                }
                if (tid == 0) tauout[k] = tau;
            }
            __syncthreads();
            double* ps = sm.p[pk];
            const double c1 = fma(-scal, alpha, 1.0);  // v_{k+1} - scal y_{k+1}
            constexpr int PW = NP / WAVE;  // waves forming p: one row a lane, one wave a SIMD
            if (w < PW) {
                // p_i = tau (scal (A y)_i + (1 - scal alpha) A[i][k+1]), (A y)_i = row sum + NWV
                // column partials, pairwise in a fixed order; entries of finished rows and columns
                // hold older partials and are masked, and without a reflector p = 0.  The other
                // waves go on to the barrier: this phase is a latency chain, which fewer waves do
                // not lengthen.
                const int i = tid;
                double t[NWV / 4];
#pragma unroll
                for (int g = 0; g < NWV / 4; ++g)
                    t[g] = (sm.u.cpart[4 * g][i] + sm.u.cpart[4 * g + 1][i]) +
                           (sm.u.cpart[4 * g + 2][i] + sm.u.cpart[4 * g + 3][i]);
                const double tt = NWV == 16 ? (t[0] + t[1]) + (t[NWV / 4 - 2] + t[NWV / 4 - 1]) : t[0] + t[1];
                const double s = ((i >> 6) > BM ? sm.rs[i] : 0.0) + tt;
                const double pi = (tau != 0.0 && i > k && i < N) ? tau * fma(scal, s, c1 * sm.x[pq][i]) : 0.0;
                ps[i] = pi;
                const double vi = (i > k) ? vs[i] : 0.0;
                const double pv = wave_sum(pi * vi);
                if (l == 0) sm.red[w] = pv;
                if (tid == 0) {  // read after the tridiagonalisation only
                    sm.e[k] = beta;
                    sm.d[k] = dk;
                }
            }
            __syncthreads();
            double coef = 0.0;
            if (tau != 0.0) {
                double pvt = 0.0;  // fixed order: bitwise reproducible
#pragma unroll
                for (int q = 0; q < PW; ++q) pvt += sm.red[q];
                coef = uniform(0.5 * tau * pvt);
                double wc[NBC];
#pragma unroll
                for (int b = BM; b < NBC; ++b) wc[b] = fma(-coef, vc[b], ps[l + 64 * b]);
                // row values by broadcast LDS reads, a 64-row block's rows of the wave loaded together
                // (one LDS round trip a block), w_i formed as the lanes form w_j; a finished row
                // (v_i = p_i = 0) is left as it is
#pragma unroll
                for (int g = BM; g < NBC; ++g) {
                    double vra[RB], pra[RB];
#pragma unroll
                    for (int aa = 0; aa < RB; ++aa) {
                        const int i = w + NWV * (RB * g + aa);
                        vra[aa] = vs[i];
                        pra[aa] = ps[i];
                    }
#pragma unroll
                    for (int aa = 0; aa < RB; ++aa) {
                        const int a = RB * g + aa;
                        const double wra = fma(-coef, vra[aa], pra[aa]);
#pragma unroll
                        for (int b = BM; b <= g; ++b)
                            A[sl(a, b)] = fma(-vra[aa], wc[b], fma(-wra, vc[b], A[sl(a, b)]));
                    }
                }
            }
            tprev = tau;
            cprev = coef;
            // no barrier here: this step's buffers that the next step writes before its first
            // barrier (x[k & 1], y, v[(k + 1) & 1], rs, cpart) were last read before this step's
            // second barrier, and the next step writes p, red only after its first barrier
        }
    }

    __device__ __forceinline__ static void tridiagonalize(double (&A)[NS], smem& sm, int N, double* __restrict__ vout,
                                                          double* __restrict__ tauout) {
        const int tid = threadIdx.x;
        const int l = tid & (WAVE - 1);
        const int w = __builtin_amdgcn_readfirstlane(tid / WAVE);
        for (int i = tid; i < NP; i += NT) {
            sm.x[0][i] = 0.0; sm.x[1][i] = 0.0; sm.d[i] = 0.0; sm.e[i] = 0.0; sm.rs[i] = 0.0;
            sm.v[0][i] = 0.0; sm.v[1][i] = 0.0; sm.p[0][i] = 0.0; sm.p[1][i] = 0.0;
        }
        for (int i = tid; i < NWV * NP; i += NT) (&sm.u.cpart[0][0])[i] = 0.0;
        __syncthreads();
        // column 0 for step 0
        if (l == 0) {
#pragma unroll
            for (int a = 0; a < NA; ++a) sm.x[0][w + NWV * a] = A[sl(a, 0)];
        }
        __syncthreads();
        double tprev = 0.0, cprev = 0.0;
        int k = 0;
        for (; k + 2 < N && k < 63; ++k) step<0, false>(A, sm, N, k, w, l, tid, vout, tauout, tprev, cprev);
        if (k + 2 < N && k == 63) step<0, true>(A, sm, N, k++, w, l, tid, vout, tauout, tprev, cprev);
        for (; k + 2 < N && k < 127; ++k) step<1, false>(A, sm, N, k, w, l, tid, vout, tauout, tprev, cprev);
        if (k + 2 < N && k == 127) step<1, true>(A, sm, N, k++, w, l, tid, vout, tauout, tprev, cprev);
        for (; k + 2 < N && k < 191; ++k) step<2, false>(A, sm, N, k, w, l, tid, vout, tauout, tprev, cprev);
        if (k + 2 < N && k == 191) step<2, true>(A, sm, N, k++, w, l, tid, vout, tauout, tprev, cprev);
        for (; k + 2 < N; ++k) step<3, false>(A, sm, N, k, w, l, tid, vout, tauout, tprev, cprev);
        __syncthreads();
        // the last two diagonal entries and the last off-diagonal one, from their owners
        if (N >= 2) {
            put_elem(A, N - 2, N - 2, w, l, &sm.d[N - 2]);
            put_elem(A, N - 1, N - 2, w, l, &sm.e[N - 2]);
            put_elem(A, N - 1, N - 1, w, l, &sm.d[N - 1]);
            if (vout) {
                for (int i = tid; i < N; i += NT) {
                    vout[(size_t)(N - 2) * N + i] = 0.0;
                    vout[(size_t)(N - 1) * N + i] = 0.0;
                }
                if (tid == 0) { tauout[N - 2] = 0.0; tauout[N - 1] = 0.0; }
            }
        } else if (N == 1) {
            put_elem(A, 0, 0, w, l, &sm.d[0]);
        }
        __syncthreads();
    }

    __device__ __forceinline__ static double solve(smem& sm, const double* __restrict__ hop, const uint8_t* f_lds,
                                                   int N, double beta, double U, double mu_c, double* ws,
                                                   double* vout, double* tauout, double* cexp, double* cfermi) {
        const int tid = threadIdx.x;
        const int l = tid & (WAVE - 1);
        const int w = __builtin_amdgcn_readfirstlane(tid / WAVE);
        double A[NS];
        // all template loads in flight at once (indices clamped, out-of-range entries zeroed
        // afterwards), then the diagonal
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b <= a / RB; ++b) {
                const int i = min(w + NWV * a, N - 1), j = min(l + 64 * b, N - 1);
                A[sl(a, b)] = hop[(size_t)i * N + j];
            }
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b <= a / RB; ++b) {
                const int i = w + NWV * a, j = l + 64 * b;
                double v = (i < N && j < N) ? A[sl(a, b)] : 0.0;
                if (i == j && i < N) v += -mu_c + U * (double)f_lds[i];
                A[sl(a, b)] = v;
            }
        if (tid == 0) sm.beta_s[0] = beta;
        tridiagonalize(A, sm, N, vout, tauout);
        tridiag_eigvals<NT, EMS>(sm.d, sm.e, sm.u.ea.de, sm.u.ea.cnt, N, sm.u.ea.ev, sm.u.ea.scratch);
        return logz_block<NT>(sm.u.ea.ev, N, sm.beta_s[0], cexp, cfermi, sm.u.ea.scratch);
    }
};

typedef wrow_solver_t<16> wrow_solver;
typedef wrow_solver_t<8> wrow8_solver;

}  // namespace fkw
