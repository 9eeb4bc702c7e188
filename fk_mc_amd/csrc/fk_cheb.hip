// fk_cheb.hip -- Chebyshev paths for logZ.
//
// 1. Exact trace, the reference's configuration_t::calc_chebyshev
//    (src/configuration.cpp:94-206): X = (H - b)/a with a, b from the extreme
//    eigenvalues; the matrices T_m(X) are built by T_{m+1} = 2 X T_m - T_{m-1}
//    (X is the lattice stencil plus the diagonal, so one step is a stencil
//    product over N x N dense matrices, HBM-streaming) and traced:
//    mu_m = Tr T_m / N, with the doubling identities for the upper half
//    (mu_{2m-1} = (2 Tr T_{m-1} T_m - Tr X)/N, mu_{2m} = 2 Tr T_m^2 / N - 1).
// 2. Stochastic trace (kernel polynomial method): R Rademacher vectors r,
//    v_{m+1} = 2 X v_m - v_{m-1} in FP32 held in LDS (one workgroup per
//    configuration x vector, N <= 4096), FP64 accumulation of <v_m|v_m> and
//    <v_{m+1}|v_m>; mu_{2m} = 2<v_m|v_m> - mu_0, mu_{2m+1} = 2<v_{m+1}|v_m> - mu_1.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_internal.h"
#include "fk_solve.h"

namespace fkc {

using namespace fks;

// ---- exact trace --------------------------------------------------------------
// X_ij = (hop_ij + delta_ij (diag_i - b)) / a, diag_i = -mu_c + U f_i.
// out = 2 X * cur - prev (prev overwritten), row-major N x N.
__global__ void k_cheb_step(int N, const int32_t* __restrict__ nbr, const int32_t* __restrict__ z, int zmax,
                            const double* __restrict__ hopv, const double* __restrict__ diag, double inv_a, double b,
                            const double* __restrict__ cur, double* __restrict__ prev_out, int first) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t NN = (size_t)N * N;
    if (idx >= NN) return;
    const int i = (int)(idx / N), j = (int)(idx % N);
    double s = (diag[i] - b) * cur[idx];
    for (int q = 0; q < z[i]; ++q) {
        const int k = nbr[(size_t)i * zmax + q];
        s = fma(hopv[(size_t)i * zmax + q], cur[(size_t)k * N + j], s);
    }
    s *= inv_a;
    // first step: T_1 = X T_0 (no factor 2, nothing subtracted)
    prev_out[idx] = first ? s : 2.0 * s - prev_out[idx];
}

// per-block partial sums of the trace (diag) or of sum_ij A_ij B_ij
__global__ void k_trace_parts(int N, const double* __restrict__ A, const double* __restrict__ Bm, int mode,
                              double* __restrict__ parts) {
    __shared__ double red[256];
    const size_t NN = (size_t)N * N;
    double s = 0.0;
    if (mode == 0) {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)N; i += (size_t)gridDim.x * blockDim.x)
            s += A[i * N + i];
    } else {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < NN; i += (size_t)gridDim.x * blockDim.x)
            s = fma(A[i], Bm[i], s);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}

__global__ void k_sum_parts(int n, const double* __restrict__ parts, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += parts[i];
    *out = s;
}

__global__ void k_identity(int N, double* __restrict__ A) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t NN = (size_t)N * N;
    if (idx >= NN) return;
    A[idx] = (idx / N == idx % N) ? 1.0 : 0.0;
}

// ---- stochastic trace (KPM) ------------------------------------------------------
__device__ __forceinline__ float rademacher(uint64_t seed, int r, int i) {
    // splitmix64 of (seed, vector, site): deterministic and identical for every configuration
    uint64_t x = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(r + 1)) ^ (0xBF58476D1CE4E5B9ull * (uint64_t)(i + 1));
    x ^= x >> 30; x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27; x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (x & 1ull) ? 1.0f : -1.0f;
}

constexpr int KPM_NT = 1024;
constexpr int KPM_NMAX = 4096;
constexpr int KPM_MMAX = 512;

// grid (R, B).  f: B x N occupations, scale[b] = (inv_a, b_shift), diag from f.
// mom_out[(b*R + r)*M + m] = <r|T_m(X)|r> (before the 1/N), FP64.
__global__ void __launch_bounds__(KPM_NT) k_kpm(int N, int R, int M, const int32_t* __restrict__ f,
                                                const double* __restrict__ U, const double* __restrict__ mu_c,
                                                const double* __restrict__ ab, const int32_t* __restrict__ nbr,
                                                const int32_t* __restrict__ z, int zmax,
                                                const double* __restrict__ hopv, uint64_t seed0,
                                                int per_item_seed, const int32_t* __restrict__ skip,
                                                double* __restrict__ mom_out) {
    __shared__ float vbuf[3][KPM_NMAX];
    __shared__ double dots[2][KPM_MMAX / 2 + 2][KPM_NT / WAVE];
    const int r = blockIdx.x, b = blockIdx.y;
    if (skip && skip[b]) return;  // uniform over the workgroup
    const uint64_t seed = per_item_seed ? seed0 + 0x9E3779B97F4A7C15ull * (uint64_t)(b + 1) : seed0;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const double inv_a = ab[2 * b], shift = ab[2 * b + 1];
    constexpr int PER = KPM_NMAX / KPM_NT;
    // per-thread sites and stencil (kept in registers)
    float dg[PER];
    int nb[PER][8];
    float hv[PER][8];
    int zz[PER];
    for (int u = 0; u < PER; ++u) {
        const int i = tid + u * KPM_NT;
        zz[u] = 0;
        dg[u] = 0.0f;
        if (i < N) {
            dg[u] = (float)((-mu_c[b] + U[b] * (double)f[(size_t)b * N + i] - shift) * inv_a);
            zz[u] = z[i] < 8 ? z[i] : 8;
            for (int q = 0; q < 8; ++q) {
                nb[u][q] = q < zz[u] ? nbr[(size_t)i * zmax + q] : 0;
                hv[u][q] = q < zz[u] ? (float)(hopv[(size_t)i * zmax + q] * inv_a) : 0.0f;
            }
            const float x0 = rademacher(seed, r, i);
            vbuf[0][i] = x0;
        }
    }
    __syncthreads();
    // v1 = X v0
    for (int u = 0; u < PER; ++u) {
        const int i = tid + u * KPM_NT;
        if (i >= N) continue;
        float s = dg[u] * vbuf[0][i];
        for (int q = 0; q < zz[u]; ++q) s = fmaf(hv[u][q], vbuf[0][nb[u][q]], s);
        vbuf[1][i] = s;
    }
    __syncthreads();
    int p0 = 0, p1 = 1, p2 = 2;  // v_{m-1}, v_m, scratch
    const int half = M / 2;
    for (int m = 1; m <= half; ++m) {
        // dots of step m: <v_m|v_m>, <v_m|v_{m-1}>  (m = 1 also gives mu_1 = <v_1|v_0>)
        double d_mm = 0.0, d_mp = 0.0;
        for (int u = 0; u < PER; ++u) {
            const int i = tid + u * KPM_NT;
            if (i >= N) continue;
            const double a1 = vbuf[p1][i], a0 = vbuf[p0][i];
            d_mm = fma(a1, a1, d_mm);
            d_mp = fma(a1, a0, d_mp);
        }
        d_mm = wave_sum(d_mm);
        d_mp = wave_sum(d_mp);
        if (lane == 0) { dots[0][m][wid] = d_mm; dots[1][m][wid] = d_mp; }
        if (m == half) break;
        // v_{m+1} = 2 X v_m - v_{m-1}
        for (int u = 0; u < PER; ++u) {
            const int i = tid + u * KPM_NT;
            if (i >= N) continue;
            float s = dg[u] * vbuf[p1][i];
            for (int q = 0; q < zz[u]; ++q) s = fmaf(hv[u][q], vbuf[p1][nb[u][q]], s);
            vbuf[p2][i] = 2.0f * s - vbuf[p0][i];
        }
        __syncthreads();
        const int t = p0; p0 = p1; p1 = p2; p2 = t;
    }
    __syncthreads();
    if (tid < M) {
        // mu_0 = <v0|v0> = N exactly; mu_1 = <v1|v0>
        const int m = tid;
        double val;
        auto sumw = [&](int k, int h) { double s = 0.0; for (int w = 0; w < KPM_NT / WAVE; ++w) s += dots[k][h][w]; return s; };
        const double mu1 = sumw(1, 1);
        if (m == 0) val = (double)N;
        else if (m == 1) val = mu1;
        else if ((m & 1) == 0) val = 2.0 * sumw(0, m / 2) - (double)N;      // 2<v_k|v_k> - mu_0, k = m/2
        else val = 2.0 * sumw(1, (m + 1) / 2) - mu1;                         // 2<v_{k}|v_{k-1}> - mu_1, k=(m+1)/2
        mom_out[((size_t)b * R + r) * M + m] = val;
    }
}

// Average over the R vectors (fixed order), Jackson damping, logZ = sum_m g_m c_m mu_m.
// coef[b*M + m] = c_m (Chebyshev coefficient of log(1+e^{-beta(a x + b)})) computed on the host.
__global__ void k_kpm_finish(int B, int N, int R, int M, const double* __restrict__ mom, const double* __restrict__ coef,
                             int jackson, const int32_t* __restrict__ skip, double* __restrict__ mu_out,
                             double* __restrict__ logz) {
    const int b = blockIdx.x;
    if (b >= B || threadIdx.x != 0) return;
    if (skip && skip[b]) return;
    const double PI = 3.14159265358979323846;
    double s = 0.0;
    for (int m = 0; m < M; ++m) {
        double mu = 0.0;
        for (int r = 0; r < R; ++r) mu += mom[((size_t)b * R + r) * M + m];
        mu /= (double)R * (double)N;
        double g = 1.0;
        if (jackson) {
            const double q = PI / (double)(M + 1);
            g = ((double)(M - m + 1) * cos(q * m) + sin(q * m) / tan(q)) / (double)(M + 1);
        }
        if (mu_out) mu_out[(size_t)b * M + m] = mu;
        s += g * coef[(size_t)b * M + m] * mu;
    }
    logz[b] = s * (double)N;
}

}  // namespace fkc

using namespace fkc;

hipError_t fk_cheb_exact_moments(int N, const int32_t* nbr, const int32_t* z, int zmax, const double* hopv,
                                 const double* diag, double a, double b, int M, double* T0, double* T1,
                                 double* parts, double* scal, double* traces, hipStream_t st) {
    // traces (host-visible device buffer, 3*M doubles): [0..M) Tr T_m, [M..2M) Tr T_{m-1}T_m, [2M..3M) Tr T_m^2
    const size_t NN = (size_t)N * N;
    const int blocks = (int)((NN + 255) / 256);
    const int pb = 512;
    hipError_t r;
    auto trace = [&](const double* A, const double* Bm, int mode, double* out) -> hipError_t {
        hipLaunchKernelGGL(k_trace_parts, dim3(pb), dim3(256), 0, st, N, A, Bm, mode, parts);
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1), 0, st, pb, parts, out);
        return hipGetLastError();
    };
    hipLaunchKernelGGL(k_identity, dim3(blocks), dim3(256), 0, st, N, T0);
    // T1 = X T0
    hipLaunchKernelGGL(k_cheb_step, dim3(blocks), dim3(256), 0, st, N, nbr, z, zmax, hopv, diag, 1.0 / a, b, T0, T1, 1);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    if ((r = trace(T1, nullptr, 0, traces + 1)) != hipSuccess) return r;
    double* prev = T0;  // T_{m-1}
    double* cur = T1;   // T_m
    for (int m = 2; m <= M / 2; ++m) {
        // prev <- 2 X cur - prev = T_m; then swap so cur = T_m, prev = T_{m-1}
        hipLaunchKernelGGL(k_cheb_step, dim3(blocks), dim3(256), 0, st, N, nbr, z, zmax, hopv, diag, 1.0 / a, b, cur,
                           prev, 0);
        if ((r = hipGetLastError()) != hipSuccess) return r;
        double* t = prev; prev = cur; cur = t;
        if ((r = trace(cur, nullptr, 0, traces + m)) != hipSuccess) return r;
        if ((r = trace(prev, cur, 1, traces + M + m)) != hipSuccess) return r;
        if ((r = trace(cur, cur, 1, traces + 2 * M + m)) != hipSuccess) return r;
    }
    (void)scal;
    return hipSuccess;
}

hipError_t fk_launch_kpm(int B, int N, int R, int M, const int32_t* f, const double* U, const double* mu_c,
                         const double* ab, const int32_t* nbr, const int32_t* z, int zmax, const double* hopv,
                         uint64_t seed, int per_item_seed, const int32_t* skip, const double* coef, int jackson,
                         double* mom, double* mu_out, double* logz, hipStream_t st) {
    if (N > KPM_NMAX || M > KPM_MMAX || M < 2 || (M & 1) || zmax > 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_kpm, dim3(R, B), dim3(KPM_NT), 0, st, N, R, M, f, U, mu_c, ab, nbr, z, zmax, hopv, seed,
                       per_item_seed, skip, mom);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_kpm_finish, dim3(B), dim3(64), 0, st, B, N, R, M, mom, coef, jackson, skip, mu_out, logz);
    return hipGetLastError();
}
