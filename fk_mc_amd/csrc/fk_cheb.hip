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
constexpr int KPM_P = 4;  // probe vectors per workgroup, packed as float4 per site

// Reduce-scatter of 8 values per lane over the 64 lanes: afterwards lane l holds the
// full sum of value number l >> 3 (pairs l^32, l^16 by lane permutes, l^8 by DPP, then
// an all-reduce over the 8 lanes of a group).
__device__ __forceinline__ double reduce8(const double (&r)[8]) {
    const int lane = threadIdx.x & (WAVE - 1);
    double s1[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const double2 t = pl32_swap(r[a], r[a + 4]);
        s1[a] = t.x + t.y;  // value a + 4*bit5
    }
    double s2[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const double2 t = pl16_swap(s1[a], s1[a + 2]);
        s2[a] = t.x + t.y;  // value a + 2*bit4 + 4*bit5
    }
    const bool hi = (lane >> 3) & 1;
    const double keep = hi ? s2[1] : s2[0], send = hi ? s2[0] : s2[1];
    double u = keep + dpp_f64<DPP_ROR8>(send);  // value bit3 + 2*bit4 + 4*bit5
    u += dpp_f64<DPP_XOR1>(u);
    u += dpp_f64<DPP_XOR2>(u);
    u += dpp_f64<DPP_HALF_MIRROR>(u);
    return u;
}

__device__ __forceinline__ float4 f4_fma(float h, float4 x, float4 s) {
    return make_float4(fmaf(h, x.x, s.x), fmaf(h, x.y, s.y), fmaf(h, x.z, s.z), fmaf(h, x.w, s.w));
}
__device__ __forceinline__ float f4_get(const float4& v, int p) {
    return p == 0 ? v.x : p == 1 ? v.y : p == 2 ? v.z : v.w;
}

// grid (ceil(R / KPM_P), B): workgroup (g, b) runs the probes r = KPM_P g + p of item b.
// Each thread owns the sites i = tid + u KPM_NT (u < PER) for the whole recursion: their
// stencil rows (ZM neighbours, zero-padded), diagonal, v_{m-1} and v_m live in VGPRs; v_m
// of all sites is shared through a double-buffered float4 array in LDS (one barrier per
// step).  Per step a wave reduces its 2 KPM_P dot products by a reduce-scatter, and wave 0
// sums the per-wave partials of the previous step in a fixed order.
// mom_out[(b*R + r)*M + m] = <r|T_m(X)|r> (before the 1/N), FP64.
struct kpm_hopq {
    double v[8];
};

template <int ZM, bool UH>
__global__ void __launch_bounds__(KPM_NT) k_kpm(int N, int R, int M, const int32_t* __restrict__ f,
                                                const double* __restrict__ U, const double* __restrict__ mu_c,
                                                const double* __restrict__ ab, const int32_t* __restrict__ nbr,
                                                const int32_t* __restrict__ z, int zmax,
                                                const double* __restrict__ hopv, kpm_hopq hq, uint64_t seed0,
                                                int per_item_seed, const int32_t* __restrict__ skip,
                                                double* __restrict__ mom_out) {
    constexpr int NW = KPM_NT / WAVE;
    constexpr int PER = KPM_NMAX / KPM_NT;
    constexpr int NV = 2 * KPM_P;  // dot products per step: <v_m|v_m>, <v_m|v_{m-1}> per probe
    constexpr int HV = UH ? 1 : PER;  // per-site hopping values unless they are the same at every site
    // The two buffers of v_m interleave in blocks of 64 sites: site i of buffer t is element
    // vb[lidx(i) + 64 t], so that the buffer is a constant offset (1 KB) in the LDS
    // instruction and one address register per neighbour serves both.
    __shared__ float4 vb[2 * KPM_NMAX];
    __shared__ double dots[8][NV][NW];
    __shared__ double sums[NV][KPM_MMAX / 2 + 2];
    auto lidx = [](int i) { return ((i >> 6) << 7) + (i & 63); };
    const int g = blockIdx.x, b = blockIdx.y;
    if (skip && skip[b]) return;  // uniform over the workgroup
    const uint64_t seed = per_item_seed ? seed0 + 0x9E3779B97F4A7C15ull * (uint64_t)(b + 1) : seed0;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const double inv_a = ab[2 * b], shift = ab[2 * b + 1];
    float dg[PER], hv[HV][ZM];
    int nb[PER][ZM];
    float4 prev[PER], cur[PER];
    // The recursion runs on 2X (dg, hv hold twice the diagonal and hopping values, exact in
    // binary); v1 = X v0 is formed as 0.5 (2X v0), which rounds the same as X v0.
    if constexpr (UH) {
#pragma unroll
        for (int q = 0; q < ZM; ++q) {
            // wave-uniform: kept in SGPRs (scalar operand of the FMAs)
            const float h = q < zmax ? 2.0f * (float)(hq.v[q] * inv_a) : 0.0f;
            hv[0][q] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(h)));
        }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int i = tid + u * KPM_NT;
        dg[u] = 0.0f;
        prev[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        // unused slots point at the site itself: a site i >= N holds 0 in both buffers for the
        // whole recursion (its hopping values may be nonzero when they are shared by all sites)
#pragma unroll
        for (int q = 0; q < ZM; ++q) {
            nb[u][q] = lidx(i);
            if constexpr (!UH) hv[u][q] = 0.0f;
        }
        if (i < N) {
            dg[u] = 2.0f * (float)((-mu_c[b] + U[b] * (double)f[(size_t)b * N + i] - shift) * inv_a);
            const int zi = z[i];
#pragma unroll
            for (int q = 0; q < ZM; ++q)
                if (q < zi) {
                    nb[u][q] = lidx(nbr[(size_t)i * zmax + q]);
                    if constexpr (!UH) hv[u][q] = 2.0f * (float)(hopv[(size_t)i * zmax + q] * inv_a);
                }
            float x[KPM_P];
#pragma unroll
            for (int p = 0; p < KPM_P; ++p) {
                const int r = g * KPM_P + p;
                x[p] = r < R ? rademacher(seed, r, i) : 0.0f;
            }
            prev[u] = make_float4(x[0], x[1], x[2], x[3]);
        }
        vb[lidx(i)] = prev[u];
    }
    __syncthreads();
    // v1 = X v0
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        float4 s = make_float4(dg[u] * prev[u].x, dg[u] * prev[u].y, dg[u] * prev[u].z, dg[u] * prev[u].w);
#pragma unroll
        for (int q = 0; q < ZM; ++q) s = f4_fma(hv[UH ? 0 : u][q], vb[nb[u][q]], s);
        s = make_float4(0.5f * s.x, 0.5f * s.y, 0.5f * s.z, 0.5f * s.w);
        cur[u] = s;
        vb[lidx(tid + u * KPM_NT) + 64] = s;
    }
    __syncthreads();
    const int half = M / 2;
    // One step m: dots of (x = v_m, y = v_{m-1}), then y <- v_{m+1} from buffer offset src
    // (holding v_m) into buffer offset dst.  Per-wave dot sums go to ring slot m & 7; after
    // the barrier that closes a block of 4 steps (or the last step) one wave folds the
    // block's 16 partials per value in a fixed order.  A slot is rewritten 8 steps later,
    // after 4 more barriers.
    auto step = [&](float4 (&x)[PER], float4 (&y)[PER], const int src, const int dst, int m) {
        double d[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) d[k] = 0.0;
        // per site slot: neighbour loads first, the FP64 dot products while they are in
        // flight, then the stencil, so that LDS and VALU work overlap within the step (the
        // last step also forms v_{M/2+1}, which is not used: one extra step of 128)
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            float4 nv[ZM];
#pragma unroll
            for (int q = 0; q < ZM; ++q) nv[q] = vb[nb[u][q] + src];
            // keep the loads ahead of the FP64 work (where the registers allow it)
            if constexpr (UH && ZM <= 4) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int p = 0; p < KPM_P; ++p) {
                const double a1 = f4_get(x[u], p), a0 = f4_get(y[u], p);
                d[p] = fma(a1, a1, d[p]);
                d[KPM_P + p] = fma(a1, a0, d[KPM_P + p]);
            }
            if constexpr (UH && ZM <= 4) __builtin_amdgcn_sched_barrier(0);
            const float dgu = dg[u];
            float4 s = make_float4(fmaf(dgu, x[u].x, -y[u].x), fmaf(dgu, x[u].y, -y[u].y),
                                   fmaf(dgu, x[u].z, -y[u].z), fmaf(dgu, x[u].w, -y[u].w));
#pragma unroll
            for (int q = 0; q < ZM; ++q) s = f4_fma(hv[UH ? 0 : u][q], nv[q], s);
            y[u] = s;
            vb[lidx(tid + u * KPM_NT) + dst] = s;
        }
        const double red = reduce8(d);
        if ((lane & 7) == 0) dots[m & 7][lane >> 3][wid] = red;
        __syncthreads();
        if ((m & 3) == 0 || m == half) {
            const int m0 = ((m - 1) & ~3) + 1;
            const int ms = m0 + (lane >> 3), v = lane & 7;
            if (wid == ((m >> 2) & (NW - 1)) && lane < 32 && ms <= m) {
                double t = 0.0;
#pragma unroll
                for (int w = 0; w < NW; ++w) t += dots[ms & 7][v][w];
                sums[v][ms] = t;
            }
        }
    };
    // steps alternate the roles of the register sets and of the two buffers
    for (int m = 1; m <= half; m += 2) {
        step(cur, prev, 64, 0, m);          // prev <- v_{m+1}
        if (m + 1 > half) break;
        step(prev, cur, 0, 64, m + 1);      // cur <- v_{m+2}
    }
    __syncthreads();
    for (int idx = tid; idx < KPM_P * M; idx += KPM_NT) {
        const int p = idx / M, m = idx % M, r = g * KPM_P + p;
        if (r >= R) continue;
        // mu_0 = <v0|v0> = N exactly; mu_1 = <v1|v0>
        const double mu1 = sums[KPM_P + p][1];
        double val;
        if (m == 0) val = (double)N;
        else if (m == 1) val = mu1;
        else if ((m & 1) == 0) val = 2.0 * sums[p][m / 2] - (double)N;           // 2<v_k|v_k> - mu_0, k = m/2
        else val = 2.0 * sums[KPM_P + p][(m + 1) / 2] - mu1;                     // 2<v_k|v_{k-1}> - mu_1
        mom_out[((size_t)b * R + r) * M + m] = val;
    }
}

// Average over the R vectors (fixed order), Jackson damping, logZ = sum_m g_m c_m mu_m.
// coef[b*M + m] = c_m (Chebyshev coefficient of log(1+e^{-beta(a x + b)})) computed on the host.
// One workgroup per item; thread t takes the moments m = t, t + 256, ...; the partial sums are
// combined by a fixed tree, so the result does not depend on scheduling.
constexpr int KF_NT = 256;
__global__ void __launch_bounds__(KF_NT) k_kpm_finish(int B, int N, int R, int M, const double* __restrict__ mom,
                                                      const double* __restrict__ coef, int jackson,
                                                      const int32_t* __restrict__ skip, double* __restrict__ mu_out,
                                                      double* __restrict__ logz) {
    __shared__ double part[KF_NT];
    const int b = blockIdx.x, tid = threadIdx.x;
    if (b >= B) return;
    if (skip && skip[b]) return;  // uniform over the workgroup
    const double PI = 3.14159265358979323846;
    double s = 0.0;
    for (int m = tid; m < M; m += KF_NT) {
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
    part[tid] = s;
    __syncthreads();
    for (int h = KF_NT / 2; h > 0; h >>= 1) {
        if (tid < h) part[tid] += part[tid + h];
        __syncthreads();
    }
    if (tid == 0) logz[b] = part[0] * (double)N;
}

// ---- Lanczos for the extreme eigenvalues (the ARPACK calls of src/configuration.cpp:99-100) ----
// One workgroup runs `iters` Lanczos steps on H = hop + diag(d) from its saved state
// (vectors in global memory, L2-resident; a workgroup's barrier orders them):
//   w = H v_j - beta_{j-1} v_{j-1};  alpha_j = <w, v_j>;  w -= alpha_j v_j;
//   beta_j = |w|;  v_{j+1} = w / beta_j
// state: v_prev [N], v_cur [N], w [N], and st[0] = beta_{j-1}, st[1] = j (as double).
// alpha[j], beta[j] are appended for j = st[1] ... st[1] + iters - 1.  Without
// reorthogonalisation the extreme Ritz values still converge to the extreme eigenvalues
// (lost orthogonality only repeats converged values).  Fixed-order reductions.
constexpr int LZ_NT = 1024;
__device__ __forceinline__ double block_sum_lz(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x / WAVE;
    __syncthreads();
    if ((threadIdx.x & (WAVE - 1)) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < LZ_NT / WAVE; ++i) s += red[i];
    return s;
}
__global__ void __launch_bounds__(LZ_NT)
    k_lanczos(int N, const int32_t* __restrict__ nbr, const int32_t* __restrict__ z, int zmax,
              const double* __restrict__ hopv, const double* __restrict__ diag, int iters, double* vp, double* vc,
              double* w, double* st, double* __restrict__ alpha, double* __restrict__ beta) {
    __shared__ double red[LZ_NT / WAVE];
    const int tid = threadIdx.x;
    double bprev = st[0];
    int j = (int)st[1];
    if (j == 0) {  // normalise the start vector
        double pn = 0.0;
        for (int i = tid; i < N; i += LZ_NT) pn = fma(vc[i], vc[i], pn);
        const double inv = 1.0 / sqrt(block_sum_lz(pn, red));
        for (int i = tid; i < N; i += LZ_NT) vc[i] *= inv;
        __syncthreads();
    }
    for (int it = 0; it < iters; ++it, ++j) {
        double pa = 0.0;
        for (int i = tid; i < N; i += LZ_NT) {
            double s = diag[i] * vc[i];
            for (int q = 0; q < z[i]; ++q) s = fma(hopv[(size_t)i * zmax + q], vc[nbr[(size_t)i * zmax + q]], s);
            s = fma(-bprev, vp[i], s);
            w[i] = s;
            pa = fma(s, vc[i], pa);
        }
        const double a = block_sum_lz(pa, red);
        double pb = 0.0;
        for (int i = tid; i < N; i += LZ_NT) {
            const double s = fma(-a, vc[i], w[i]);
            w[i] = s;
            pb = fma(s, s, pb);
        }
        const double bn = sqrt(block_sum_lz(pb, red));
        const double inv = bn > 0.0 ? 1.0 / bn : 0.0;
        for (int i = tid; i < N; i += LZ_NT) {
            vp[i] = vc[i];
            vc[i] = w[i] * inv;
        }
        if (tid == 0) {
            alpha[j] = a;
            beta[j] = bn;
        }
        bprev = bn;
        __syncthreads();
    }
    if (tid == 0) {
        st[0] = bprev;
        st[1] = (double)j;
    }
}

__global__ void k_lanczos_start(int N, double* vp, double* vc, double* st) {
    // a fixed pseudo-random start vector (normalised on the host side of the first step)
    const double s = rsqrt((double)N);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)i * 2654435761u;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        vc[i] = s * (0.5 + (double)(h & 0xFFFF) / 65536.0);
        vp[i] = 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[0] = 0.0;
        st[1] = 0.0;
    }
}

}  // namespace fkc

using namespace fkc;

hipError_t fk_launch_lanczos(int N, const int32_t* nbr, const int32_t* z, int zmax, const double* hopv,
                             const double* diag, int first, int iters, double* work, double* alpha, double* beta,
                             hipStream_t st) {
    double* vp = work;
    double* vc = work + N;
    double* w = work + 2 * (size_t)N;
    double* s = work + 3 * (size_t)N;
    if (first) {
        hipLaunchKernelGGL(k_lanczos_start, dim3(64), dim3(256), 0, st, N, vp, vc, s);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_lanczos, dim3(1), dim3(LZ_NT), 0, st, N, nbr, z, zmax, hopv, diag, iters, vp, vc, w, s, alpha,
                       beta);
    return hipGetLastError();
}

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
                         const double* hopq, uint64_t seed, int per_item_seed, const int32_t* skip,
                         const double* coef, int jackson, double* mom, double* mu_out, double* logz, hipStream_t st) {
    if (N > KPM_NMAX || M > KPM_MMAX || M < 2 || (M & 1) || zmax > 8 || R < 1) return hipErrorInvalidValue;
    const dim3 grid((R + KPM_P - 1) / KPM_P, B);
    kpm_hopq hq{};
    if (hopq)
        for (int q = 0; q < zmax; ++q) hq.v[q] = hopq[q];
#define FK_KPM_LAUNCH(ZM, UH)                                                                                  \
    hipLaunchKernelGGL((k_kpm<ZM, UH>), grid, dim3(KPM_NT), 0, st, N, R, M, f, U, mu_c, ab, nbr, z, zmax, hopv, hq, \
                       seed, per_item_seed, skip, mom)
    if (zmax <= 4) {
        if (hopq) FK_KPM_LAUNCH(4, true); else FK_KPM_LAUNCH(4, false);
    } else if (zmax <= 6) {
        if (hopq) FK_KPM_LAUNCH(6, true); else FK_KPM_LAUNCH(6, false);
    } else {
        if (hopq) FK_KPM_LAUNCH(8, true); else FK_KPM_LAUNCH(8, false);
    }
#undef FK_KPM_LAUNCH
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_kpm_finish, dim3(B), dim3(KF_NT), 0, st, B, N, R, M, mom, coef, jackson, skip, mu_out, logz);
    return hipGetLastError();
}
