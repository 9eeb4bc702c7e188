// fk_kpm_local.hip -- dlogZ of a proposal that changes one or two sites, for the
// chains on the Chebyshev path (the move weights of src/moves_chebyshev.cpp:6-72),
// from exact local Chebyshev moments and a Matsubara sum.  No stochastic probes.
//
// A move changes the diagonal of H at sites p (and q): H' = H + Delta,
// Delta = rho_p e_p e_p^T + rho_q e_q e_q^T.  From
//   log(1 + e^{-beta x}) = -beta x / 2 + log 2 + sum_{n>=0} log(1 + x^2 / w_n^2),
//   w_n = (2n + 1) pi / beta,
// and det(1 + H^2/w^2) = |det(i w - H)|^2 / w^{2N}:
//   dlogZ = -beta Tr(Delta) / 2 + sum_{n>=0} log |D(i w_n)|^2,
//   D(z) = det(z - H') / det(z - H) = det(1 - G_P(z) Delta_P),
// with G_P the 1x1 or 2x2 block of G(z) = (z - H)^{-1} on the changed sites:
//   D = 1 - rho_p G_pp                                        (one site)
//   D = (1 - rho_p G_pp)(1 - rho_q G_qq) - rho_p rho_q G_pq^2 (two sites).
// With X = (H - b)/a (spectrum inside [-1, 1]) and z' = (z - b)/a,
//   G_pq(z) = (1/a) sum_m c_m(z') mu^{pq}_m,   mu^{pq}_m = <e_p| T_m(X) |e_q>,
//   c_m(z') = (2 - delta_m0) r^m / s,  s = sqrt(z'^2 - 1),  r = z' - s  (|r| < 1),
// the Chebyshev series of 1/(z' - x), exact for every z' off [-1, 1] (no kernel
// damping is involved; the truncation error after M terms is |r|^M, largest at the
// lowest frequency, where |r| ~ 1 - pi / (beta a)).  The moments are entries of the
// Chebyshev vectors t_m = T_m(X) e_p (t_{m+1} = 2 X t_m - t_{m-1}): mu^{pp}_m = t_m[p],
// mu^{pq}_m = t_m[q], and mu^{qq}_m from the vector started at e_q.  The frequencies
// after the first n0 follow log|D(iw)|^2 = 2 c2 / w^2 + d4 / w^4 + O(w^-6), where
// c2 = Tr(H Delta) + Tr(Delta^2) / 2 exactly (the 1/z^2 term of Tr log(1 - G Delta)) and
// d4 is fitted at the last computed frequency; the sums of w_n^-2 and w_n^-4 over
// n >= n0 are constants of n0 (computed on the host).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_internal.h"

namespace fkl {

constexpr int KL_NT = 1024;                 // threads of the moment kernel
constexpr int KL_NMAX = FK_KPM_LOCAL_NMAX;  // sites: two FP64 vectors in LDS
constexpr int KL_SPT = KL_NMAX / KL_NT;     // sites per thread
constexpr int KF_NT = 1024;                 // threads of the frequency kernel

struct hop_slots {
    double v[8];
};

// Changed sites of a proposal and the changes of their diagonal entries.
struct delta_sites {
    int p, q;
    double rp, rq;
};

__device__ inline delta_sites changed_sites(const move_prop& pr, const uint8_t* f, const int32_t* fprop, double U) {
    delta_sites d{-1, -1, 0.0, 0.0};
    if (pr.sa >= 0) {
        d.p = pr.sa;
        d.rp = U * (double)(fprop[pr.sa] - (int)f[pr.sa]);
    }
    if (pr.sb >= 0) {
        d.q = pr.sb;
        d.rq = U * (double)(fprop[pr.sb] - (int)f[pr.sb]);
    }
    return d;
}

// Chebyshev vectors of one start vector e_s of the current configuration (grid: chains x 2;
// y = 0 starts at the first changed site and records mu^pp and mu^pq, y = 1 at the second
// and records mu^qq).  One barrier per step: step m reads buffer m & 1 and writes the other.
// UH: the hopping value of a neighbour slot is the same at every site (hq), else per site.
template <int ZM, bool UH>
__global__ void __launch_bounds__(KL_NT)
    k_local_moments(int N, int M, const uint8_t* __restrict__ f, const int32_t* __restrict__ fprop,
                    const move_prop* __restrict__ props, const int32_t* __restrict__ skip,
                    const double* __restrict__ U, const double* __restrict__ mu_c, const double* __restrict__ ab,
                    const int32_t* __restrict__ nbr, const int32_t* __restrict__ z, int zmax,
                    const double* __restrict__ hopv, hop_slots hq, double* __restrict__ mom) {
    __shared__ double buf[2][KL_NMAX];
    const int c = blockIdx.x, y = blockIdx.y, tid = threadIdx.x;
    if (skip[c]) return;  // uniform over the workgroup
    const uint8_t* fc = f + (size_t)c * N;
    const delta_sites d = changed_sites(props[c], fc, fprop + (size_t)c * N, U[c]);
    const int s = y == 0 ? d.p : d.q;
    if (s < 0) return;
    const int other = y == 0 ? d.q : -1;
    const double ainv = ab[2 * c], b = ab[2 * c + 1], Uc = U[c], mu = mu_c[c];
    double* mo = mom + (size_t)c * 3 * M;
    double* rec_s = mo + (size_t)(y == 0 ? 0 : 1) * M;  // mu^pp or mu^qq
    double* rec_o = mo + 2 * (size_t)M;                 // mu^pq
    double dg[KL_SPT], tp[KL_SPT], tc[KL_SPT];
    int nb[KL_SPT][ZM];
#pragma unroll
    for (int u = 0; u < KL_SPT; ++u) {
        const int i = tid + KL_NT * u;
        const bool ok = i < N;
        dg[u] = ok ? ((-mu + Uc * (double)fc[i]) - b) * ainv : 0.0;
#pragma unroll
        for (int q = 0; q < ZM; ++q) {
            // unused slots point at the site itself with a zero hopping value
            const bool has = ok && q < z[i];
            nb[u][q] = has ? nbr[(size_t)i * zmax + q] : (ok ? i : 0);
        }
        tc[u] = (i == s) ? 1.0 : 0.0;
        tp[u] = 0.0;
        if (ok) buf[0][i] = tc[u];
    }
    double hs[ZM];
#pragma unroll
    for (int q = 0; q < ZM; ++q) hs[q] = UH ? hq.v[q] * ainv : 0.0;
    __syncthreads();
    for (int m = 0; m < M; ++m) {
#pragma unroll
        for (int u = 0; u < KL_SPT; ++u) {
            const int i = tid + KL_NT * u;
            if (i == s) rec_s[m] = tc[u];
            if (i == other) rec_o[m] = tc[u];
        }
        if (m + 1 == M) break;
        const double* cur = buf[m & 1];
        double* nxt = buf[(m + 1) & 1];
        double nx[KL_SPT];
#pragma unroll
        for (int u = 0; u < KL_SPT; ++u) {
            double acc;
            if (UH) {
                acc = dg[u] * tc[u];
#pragma unroll
                for (int q = 0; q < ZM; ++q) acc = fma(hs[q], cur[nb[u][q]], acc);
            } else {
                // per-site hopping values from the (cached) neighbour table, zero past z[i]
                const int i = min(tid + KL_NT * u, N - 1);
                const double* hrow = hopv + (size_t)i * zmax;
                double h = 0.0;
#pragma unroll
                for (int q = 0; q < ZM; ++q) h = fma(q < zmax ? hrow[q] : 0.0, cur[nb[u][q]], h);
                acc = fma(ainv, h, dg[u] * tc[u]);
            }
            nx[u] = m == 0 ? acc : fma(2.0, acc, -tp[u]);
        }
#pragma unroll
        for (int u = 0; u < KL_SPT; ++u) {
            const int i = tid + KL_NT * u;
            tp[u] = tc[u];
            tc[u] = nx[u];
            if (i < N) nxt[i] = nx[u];
        }
        __syncthreads();
    }
}

// sqrt of x + iy (y != 0 or x >= 0), principal branch
__device__ inline double2 csqrt_d(double x, double y) {
    const double r = hypot(x, y);
    if (x >= 0.0) {
        const double t = sqrt(0.5 * (r + x));
        return make_double2(t, y / (2.0 * t));
    }
    const double t = sqrt(0.5 * (r - x));
    return make_double2(fabs(y) / (2.0 * t), copysign(t, y));
}

// Matsubara sum of one chain (one workgroup): thread t takes the frequencies n = t (mod KF_NT)
// below n0; the lowest frequencies need the most Chebyshev terms, so the sum stops where
// |r|^m < 1e-17 (the remaining terms are below 1e-17 / (1 - |r|) of the leading one).
__global__ void __launch_bounds__(KF_NT)
    k_local_finish(int N, int M, int n0, const uint8_t* __restrict__ f, const int32_t* __restrict__ fprop,
                   const move_prop* __restrict__ props, const int32_t* __restrict__ skip,
                   const double* __restrict__ U, const double* __restrict__ mu_c, const double* __restrict__ beta,
                   const double* __restrict__ ab, const double* __restrict__ mom, const double* __restrict__ logz,
                   double s2c, double s4c, double* __restrict__ lz_out) {
    __shared__ double mu_s[3][FK_KPM_LOCAL_MMAX];
    __shared__ double red[KF_NT / 64];
    __shared__ double flast;
    const int c = blockIdx.x, tid = threadIdx.x;
    if (skip[c]) return;
    const uint8_t* fc = f + (size_t)c * N;
    const delta_sites d = changed_sites(props[c], fc, fprop + (size_t)c * N, U[c]);
    if (d.p < 0) return;  // not a one- or two-site move (refused on the host)
    const bool two = d.q >= 0;
    const double* mo = mom + (size_t)c * 3 * M;
    for (int m = tid; m < M; m += KF_NT) {
        mu_s[0][m] = mo[m];
        mu_s[1][m] = two ? mo[M + m] : 0.0;
        mu_s[2][m] = two ? mo[2 * M + m] : 0.0;
    }
    __syncthreads();
    const double bt = beta[c], ainv = ab[2 * c], b = ab[2 * c + 1];
    const double zr = -b * ainv;
    double fsum = 0.0;
    for (int n = tid; n < n0; n += KF_NT) {
        const double w = (2.0 * n + 1.0) * M_PI / bt;
        const double zi = w * ainv;
        // s = sqrt(z'^2 - 1), r = z' - s with |r| < 1
        double2 sq = csqrt_d(zr * zr - zi * zi - 1.0, 2.0 * zr * zi);
        double rr = zr - sq.x, ri = zi - sq.y;
        if (rr * rr + ri * ri > 1.0) {
            sq.x = -sq.x;
            sq.y = -sq.y;
            rr = zr - sq.x;
            ri = zi - sq.y;
        }
        // c0 = 2 / s
        const double sden = sq.x * sq.x + sq.y * sq.y;
        double tr = 2.0 * sq.x / sden, ti = -2.0 * sq.y / sden;
        const double c0a = tr * tr + ti * ti;
        double gpr = 0.5 * tr * mu_s[0][0], gpi = 0.5 * ti * mu_s[0][0];
        double gqr = 0.5 * tr * mu_s[1][0], gqi = 0.5 * ti * mu_s[1][0];
        double gxr = 0.5 * tr * mu_s[2][0], gxi = 0.5 * ti * mu_s[2][0];
        for (int m = 1; m < M; ++m) {
            const double nr = tr * rr - ti * ri, ni = tr * ri + ti * rr;
            tr = nr;
            ti = ni;
            gpr = fma(tr, mu_s[0][m], gpr);
            gpi = fma(ti, mu_s[0][m], gpi);
            gqr = fma(tr, mu_s[1][m], gqr);
            gqi = fma(ti, mu_s[1][m], gqi);
            gxr = fma(tr, mu_s[2][m], gxr);
            gxi = fma(ti, mu_s[2][m], gxi);
            if (tr * tr + ti * ti < 1e-34 * c0a) break;
        }
        gpr *= ainv; gpi *= ainv; gqr *= ainv; gqi *= ainv; gxr *= ainv; gxi *= ainv;
        // D = 1 - rp Gpp (one site), (1 - rp Gpp)(1 - rq Gqq) - rp rq Gpq^2 (two sites)
        const double ar = 1.0 - d.rp * gpr, ai = -d.rp * gpi;
        double Dr = ar, Di = ai;
        if (two) {
            const double br = 1.0 - d.rq * gqr, bi = -d.rq * gqi;
            const double x2r = gxr * gxr - gxi * gxi, x2i = 2.0 * gxr * gxi;
            Dr = ar * br - ai * bi - d.rp * d.rq * x2r;
            Di = ar * bi + ai * br - d.rp * d.rq * x2i;
        }
        const double F = log(Dr * Dr + Di * Di);
        fsum += F;
        if (n == n0 - 1) flast = F;
    }
    // fixed-order reduction: bitwise reproducible
    for (int o = 32; o > 0; o >>= 1) fsum += __shfl_down(fsum, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = fsum;
    __syncthreads();
    if (tid == 0) {
        double sum = 0.0;
        for (int wv = 0; wv < KF_NT / 64; ++wv) sum += red[wv];
        const double hpp = -mu_c[c] + U[c] * (double)fc[d.p];
        const double hqq = two ? -mu_c[c] + U[c] * (double)fc[d.q] : 0.0;
        const double c2 = d.rp * hpp + d.rq * hqq + 0.5 * (d.rp * d.rp + d.rq * d.rq);
        const double wl = (2.0 * (n0 - 1) + 1.0) * M_PI / bt;
        const double d4 = (flast - 2.0 * c2 / (wl * wl)) * (wl * wl) * (wl * wl);
        const double bp = bt / M_PI;
        const double dl = -0.5 * bt * (d.rp + d.rq) + sum + 2.0 * c2 * bp * bp * s2c + d4 * (bp * bp) * (bp * bp) * s4c;
        lz_out[c] = logz[c] + dl;
    }
}

}  // namespace fkl

using namespace fkl;

hipError_t fk_launch_kpm_local(int B, int N, int M, int n0, const uint8_t* f, const int32_t* fprop,
                               const move_prop* props, const int32_t* skip, const double* U, const double* mu_c,
                               const double* beta, const double* ab, const int32_t* nbr, const int32_t* z, int zmax,
                               const double* hopv, const double* hopq, double s2c, double s4c, double* mom,
                               const double* logz, double* lz_out, hipStream_t st) {
    if (N > KL_NMAX || M < 2 || M > FK_KPM_LOCAL_MMAX || zmax > 8 || n0 < 2) return hipErrorInvalidValue;
    hop_slots hq{};
    if (hopq)
        for (int q = 0; q < zmax; ++q) hq.v[q] = hopq[q];
    const dim3 grid(B, 2);
#define FK_LOCAL_LAUNCH(ZM, UH)                                                                                 \
    hipLaunchKernelGGL((k_local_moments<ZM, UH>), grid, dim3(KL_NT), 0, st, N, M, f, fprop, props, skip, U, mu_c, \
                       ab, nbr, z, zmax, hopv, hq, mom)
    if (zmax <= 4) {
        if (hopq) FK_LOCAL_LAUNCH(4, true); else FK_LOCAL_LAUNCH(4, false);
    } else if (zmax <= 6) {
        if (hopq) FK_LOCAL_LAUNCH(6, true); else FK_LOCAL_LAUNCH(6, false);
    } else {
        if (hopq) FK_LOCAL_LAUNCH(8, true); else FK_LOCAL_LAUNCH(8, false);
    }
#undef FK_LOCAL_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_local_finish, dim3(B), dim3(KF_NT), 0, st, N, M, n0, f, fprop, props, skip, U, mu_c, beta,
                       ab, mom, logz, s2c, s4c, lz_out);
    return hipGetLastError();
}
