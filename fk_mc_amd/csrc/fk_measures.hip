// fk_measures.hip -- eigenvector measures and the adjoint of Omega.
//
//   k_ipr            measure_ipr::accumulate (include/fk_mc/measures/ipr.hpp:39-56):
//                    ipr_k = ||v_k||_4 / ||v_k||_2^2 (lpNorm<4>, not its 4th power).
//   k_stiff_prep     measure_stiffness (include/fk_mc/measures/stiffness.hpp:69-153):
//                    (J V) with J_{i,left} = -1, J_{i,right} = +1 along direction 0,
//                    and the kinetic diagonal (V^T T_x V)_kk.
//   k_gemm_tn        C = A^T B (N x N, FP64, LDS-tiled), used for V^T (J V).
//   k_stiff_pairs    the pair sums of stiffness.hpp:156-185: V = sum 2 sigma/(e_k - e_l)
//                    and the Lorentzian conductivity on a frequency grid.
//   k_dos            DOS rho(w) = (1/N) sum_k (eta/pi)/((w - e_k)^2 + eta^2)
//                    (prog/data_save.hxx:264-345, dos0_f).
//   k_adjoint        dOmega/dH = rho = V f(Lambda) V^T restricted to the diagonal and
//                    the hopping bonds: local densities, dOmega/dU, dOmega/dt,
//                    dOmega/dtp, dOmega/dmu_c, and dN_c/dmu_c = beta sum f(1-f).
//   k_chain_measure  one measurement of every chain of a chain set, recorded on the
//                    device: measure_energy (src/measures/energy.cpp:6-26),
//                    measure_spectrum (src/measures/spectrum.cpp:13-21),
//                    measure_spectrum_history (src/measures/spectrum_history.cpp:13-19),
//                    measure_focc (src/measures/focc_history.cpp:8-13), the IPR
//                    history of measure_ipr, the stiffness / conductivity history of
//                    measure_stiffness and the eigenvectors of measure_eigenfunctions.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_internal.h"
#include "fk_moves.h"
#include "fk_solve.h"

namespace fkm {

using namespace fks;

__global__ void k_ipr(int B, int N, const double* __restrict__ V, double* __restrict__ out) {
    const int b = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || k >= N) return;
    const double* v = V + (size_t)b * N * N;
    double s2 = 0.0, s4 = 0.0;
    for (int i = 0; i < N; ++i) {
        const double x = v[(size_t)i * N + k];
        const double x2 = x * x;
        s2 += x2;
        s4 = fma(x2, x2, s4);
    }
    out[(size_t)b * N + k] = sqrt(sqrt(s4)) / s2;
}

// JV[i][k] = V[right(i)][k] - V[left(i)][k];  tdiag[k] = sum_i V[i][k] * (-t)(V[left][k] + V[right][k])
__global__ void k_stiff_prep(int B, int N, int stride0, int L0, double t, const double* __restrict__ V,
                             double* __restrict__ JV, double* __restrict__ tdiag) {
    const int b = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || k >= N) return;
    const double* v = V + (size_t)b * N * N;
    double* jv = JV + (size_t)b * N * N;
    double td = 0.0;
    for (int i = 0; i < N; ++i) {
        const int x = i / stride0, rest = i % stride0;
        const int il = ((x - 1 + L0) % L0) * stride0 + rest;
        const int ir = ((x + 1) % L0) * stride0 + rest;
        const double vl = v[(size_t)il * N + k], vr = v[(size_t)ir * N + k];
        jv[(size_t)i * N + k] = vr - vl;
        td = fma(v[(size_t)i * N + k], -t * vl - t * vr, td);
    }
    tdiag[(size_t)b * N + k] = td;
}

// C[b] = A[b]^T * B[b], all N x N row-major.  16x16 output tiles, k-tiles in LDS.
constexpr int GT = 16;
__global__ void __launch_bounds__(GT * GT) k_gemm_tn(int B, int N, const double* __restrict__ A,
                                                     const double* __restrict__ Bm, double* __restrict__ C) {
    __shared__ double As[GT][GT + 1], Bs[GT][GT + 1];
    const int b = blockIdx.z;
    const int r0 = blockIdx.y * GT, c0 = blockIdx.x * GT;
    const int tx = threadIdx.x % GT, ty = threadIdx.x / GT;
    const double* a = A + (size_t)b * N * N;
    const double* bb = Bm + (size_t)b * N * N;
    double acc = 0.0;
    for (int k0 = 0; k0 < N; k0 += GT) {
        // As[kk][rr] = A[k0+kk][r0+rr]; Bs[kk][cc] = B[k0+kk][c0+cc]
        const int kk = ty, cc = tx;
        As[kk][cc] = (k0 + kk < N && r0 + cc < N) ? a[(size_t)(k0 + kk) * N + r0 + cc] : 0.0;
        Bs[kk][cc] = (k0 + kk < N && c0 + cc < N) ? bb[(size_t)(k0 + kk) * N + c0 + cc] : 0.0;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < GT; ++u) acc = fma(As[u][ty], Bs[u][tx], acc);
        __syncthreads();
    }
    if (r0 + ty < N && c0 + tx < N) C[(size_t)b * N * N + (size_t)(r0 + ty) * N + c0 + tx] = acc;
}

// Pair sums, one workgroup per matrix.  Threads stride over rows i; each row's
// partial V is summed over j < i sequentially; partials reduced in fixed order.
// cond[w] is accumulated by thread w (loop over all pairs, fixed order).
__global__ void __launch_bounds__(256) k_stiff_pairs(int B, int N, const double* __restrict__ mJg,
                                                     const double* __restrict__ evg, const double* __restrict__ fg,
                                                     const double* __restrict__ tdg, int nw,
                                                     const double* __restrict__ wgrid, double offset,
                                                     double* __restrict__ stiff, double* __restrict__ cond) {
    __shared__ double red[256];
    const int b = blockIdx.x;
    if (b >= B) return;
    const int tid = threadIdx.x;
    const double* mJ = mJg + (size_t)b * N * N;
    const double* ev = evg + (size_t)b * N;
    const double* fe = fg + (size_t)b * N;
    const double PI = 3.14159265358979323846;
    double vsum = 0.0, tsum = 0.0;
    for (int i = tid; i < N; i += blockDim.x) {
        tsum = fma(tdg[(size_t)b * N + i], fe[i], tsum);
        for (int j = 0; j < i; ++j) {
            if (!(fabs(ev[i] - ev[j]) > 1e-12)) continue;
            const double sigma = PI * (fe[j] - fe[i]) * mJ[(size_t)j * N + i] * mJ[(size_t)i * N + j];
            if (fabs(sigma) > 1e-13) vsum += 2.0 * sigma / (ev[i] - ev[j]);
        }
    }
    red[tid] = vsum;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int u = 0; u < (int)blockDim.x; ++u) s += red[u];
        red[0] = s;
    }
    __syncthreads();
    const double V = red[0];
    __syncthreads();
    red[tid] = tsum;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int u = 0; u < (int)blockDim.x; ++u) s += red[u];
        stiff[b] = (V - PI * s) / (double)N;
    }
    for (int w = tid; w < nw; w += blockDim.x) {
        const double om = wgrid[w];
        double acc = 0.0;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < i; ++j) {
                if (!(fabs(ev[i] - ev[j]) > 1e-12)) continue;
                const double sigma = PI * (fe[j] - fe[i]) * mJ[(size_t)j * N + i] * mJ[(size_t)i * N + j];
                if (!(fabs(sigma) > 1e-13)) continue;
                const double x1 = om + (ev[j] - ev[i]), x2 = om + (ev[i] - ev[j]);
                acc += offset / PI / (x1 * x1 + offset * offset) * sigma;
                acc += offset / PI / (x2 * x2 + offset * offset) * (-sigma);
            }
        cond[(size_t)b * nw + w] = acc;
    }
}

__global__ void k_dos(int S, int N, const double* __restrict__ spec, int nw, const double* __restrict__ wgrid,
                      double eta, double* __restrict__ out) {
    const int s = blockIdx.y;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || w >= nw) return;
    const double PI = 3.14159265358979323846;
    const double om = wgrid[w];
    const double* e = spec + (size_t)s * N;
    double acc = 0.0;
    for (int k = 0; k < N; ++k) {
        const double x = om - e[k];
        acc += (eta / PI) / (x * x + eta * eta);
    }
    out[(size_t)s * nw + w] = acc / (double)N;
}

// ipr_f of save_ipr (prog/data_save.hxx:490-501) for spectrum s at frequency w: the
// Lorentzian-weighted mean sum_k L_k ipr_k^4 / sum_k L_k, L_k = eta / ((w - e_k)^2 + eta^2);
// ipr_k^4 undoes the 4th root of the lpNorm<4> the records hold.
__global__ void k_ipr_w(int S, int N, const double* __restrict__ spec, const double* __restrict__ ipr, int nw,
                        const double* __restrict__ wgrid, double eta, double* __restrict__ out) {
    const int s = blockIdx.y;
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || w >= nw) return;
    const double om = wgrid[w];
    const double* e = spec + (size_t)s * N;
    const double* r = ipr + (size_t)s * N;
    double nom = 0.0, den = 0.0;
    for (int k = 0; k < N; ++k) {
        const double x = om - e[k];
        const double l = eta / (x * x + eta * eta);
        const double r2 = r[k] * r[k];
        den += l;
        nom += l * (r2 * r2);
    }
    out[(size_t)s * nw + w] = nom / den;
}

// One wave per site i: rho_ii and the bond sums of row i, rho_ij = sum_k V_ik f_k V_jk.
// The lanes stride over k, so row i of V and its neighbours' rows are read coalesced
// (row-major V: row i holds component i of every vector); the lane partial sums are
// reduced across the wave.  Per-matrix reductions of the per-site values are done by
// k_adjoint_reduce.
constexpr int ADJ_SITES = 4;  // sites (waves) per workgroup
__global__ void __launch_bounds__(WAVE * ADJ_SITES) k_adjoint_sites(int B, int N, const double* __restrict__ V,
                                                                    const double* __restrict__ fermi,
                                                                    const int32_t* __restrict__ nbr,
                                                                    const int32_t* __restrict__ z, int zmax,
                                                                    int nt_bonds, double* __restrict__ dens,
                                                                    double* __restrict__ bond_t,
                                                                    double* __restrict__ bond_tp) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * ADJ_SITES + (int)(threadIdx.x / WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    if (b >= B || i >= N) return;  // uniform over the wave: the reductions below need all 64 lanes
    const double* v = V + (size_t)b * N * N;
    const double* f = fermi + (size_t)b * N;
    const double* vi = v + (size_t)i * N;
    double rii = 0.0;
    for (int k = lane; k < N; k += WAVE) rii = fma(vi[k] * vi[k], f[k], rii);
    double bt = 0.0, btp = 0.0;
    const int zi = z[i];
    for (int s = 0; s < zi; ++s) {
        const double* vj = v + (size_t)nbr[(size_t)i * zmax + s] * N;
        double rij = 0.0;
        for (int k = lane; k < N; k += WAVE) rij = fma(vi[k] * vj[k], f[k], rij);
        if (s < nt_bonds) bt += rij; else btp += rij;
    }
    rii = wave_sum(rii);
    bt = wave_sum(bt);
    btp = wave_sum(btp);
    if (lane == 0) {
        dens[(size_t)b * N + i] = rii;
        bond_t[(size_t)b * N + i] = bt;
        bond_tp[(size_t)b * N + i] = btp;
    }
}

// out[b*8 + ...] = {dOmega/dU, dOmega/dt, dOmega/dtp, dOmega/dmu_c, N_c, dN_c/dmu_c}
__global__ void __launch_bounds__(256) k_adjoint_reduce(int B, int N, const int32_t* __restrict__ f,
                                                        const double* __restrict__ fermi,
                                                        const double* __restrict__ beta,
                                                        const double* __restrict__ dens,
                                                        const double* __restrict__ bond_t,
                                                        const double* __restrict__ bond_tp,
                                                        double* __restrict__ out) {
    __shared__ double red[5][256];
    const int b = blockIdx.x;
    if (b >= B) return;
    const int tid = threadIdx.x;
    double su = 0, st = 0, stp = 0, snc = 0, sdn = 0;
    for (int i = tid; i < N; i += blockDim.x) {
        su += f[(size_t)b * N + i] ? dens[(size_t)b * N + i] : 0.0;
        st += bond_t[(size_t)b * N + i];
        stp += bond_tp[(size_t)b * N + i];
        const double fk = fermi[(size_t)b * N + i];
        snc += fk;
        sdn += fk * (1.0 - fk);
    }
    red[0][tid] = su; red[1][tid] = st; red[2][tid] = stp; red[3][tid] = snc; red[4][tid] = sdn;
    __syncthreads();
    if (tid < 5) {
        double s = 0.0;
        for (int u = 0; u < (int)blockDim.x; ++u) s += red[tid][u];
        red[tid][0] = s;
    }
    __syncthreads();
    if (tid == 0) {
        double* o = out + (size_t)b * 8;
        o[0] = red[0][0];                // dOmega/dU = sum_i n_f,i rho_ii
        o[1] = -red[1][0];               // dOmega/dt: H_ij = -t on bonds
        o[2] = -red[2][0];               // dOmega/dtp
        o[3] = -red[3][0];               // dOmega/dmu_c = -sum_k f_k
        o[4] = red[3][0];                // N_c
        o[5] = beta[b] * red[4][0];      // dN_c/dmu_c
        o[6] = 0.0;
        o[7] = 0.0;
    }
}

// rho = V diag(f) V^T (full matrix), 16x16 tiles.
__global__ void __launch_bounds__(GT * GT) k_rho_full(int B, int N, const double* __restrict__ V,
                                                      const double* __restrict__ fermi, double* __restrict__ R) {
    __shared__ double As[GT][GT + 1], Bs[GT][GT + 1];
    const int b = blockIdx.z;
    const int r0 = blockIdx.y * GT, c0 = blockIdx.x * GT;
    const int tx = threadIdx.x % GT, ty = threadIdx.x / GT;
    const double* v = V + (size_t)b * N * N;
    const double* f = fermi + (size_t)b * N;
    double acc = 0.0;
    for (int k0 = 0; k0 < N; k0 += GT) {
        // As[rr][kk] = V[r0+rr][k0+kk] f[k0+kk], Bs[cc][kk] = V[c0+cc][k0+kk]
        As[ty][tx] = (r0 + ty < N && k0 + tx < N) ? v[(size_t)(r0 + ty) * N + k0 + tx] * f[k0 + tx] : 0.0;
        Bs[ty][tx] = (c0 + ty < N && k0 + tx < N) ? v[(size_t)(c0 + ty) * N + k0 + tx] : 0.0;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < GT; ++u) acc = fma(As[ty][u], Bs[tx][u], acc);
        __syncthreads();
    }
    if (r0 + ty < N && c0 + tx < N) R[(size_t)b * N * N + (size_t)(r0 + ty) * N + c0 + tx] = acc;
}

// One workgroup per chain.  Record m = count[0]: E_c = sum_k e_k f(e_k),
// d2E = sum_k e_k^2 f(1-f) (written, as the reference does, as
// e^2 / (1 + (x + 1/x)/2) / 2 with x = exp(beta e)), E = E_c - mu_f N_f + E_ff; the
// spectrum's running mean (avg Z + e) / (Z + 1) in the reference's order of operations
// (no fused multiply-add, so the bits equal a host loop's); occupation counts; and the
// histories when their buffers are given.
constexpr int CM_T = 256;
__device__ inline double running_mean(double avg, double z, double e) {
#pragma clang fp contract(off)
    return (avg * z + e) / (z + 1.0);
}

__global__ void __launch_bounds__(CM_T) k_chain_measure(chain_meas_args a) {
    __shared__ double red[2][CM_T / WAVE];
    const int b = blockIdx.x, N = a.N, tid = threadIdx.x;
    const long long m = a.count[0];
    const bool rec = m < (long long)a.cap;
    const double beta = a.beta[b];
    const double z = (double)m;
    const size_t row = (size_t)b * N;
    const size_t hrow = ((size_t)b * a.cap + (size_t)(rec ? m : 0)) * N;
    double ec = 0.0, d2 = 0.0;
    const bool spec = a.spectrum != nullptr;  // uniform: null when only occupations are measured
    for (int k = tid; k < N; k += CM_T) {
        if (spec) {
            const double e = a.spectrum[row + k];
            const double x = exp(beta * e);
            ec += e / (1.0 + x);
            d2 += e * e / (1.0 + 0.5 * (x + 1.0 / x));
            a.spec_mean[row + k] = running_mean(a.spec_mean[row + k], z, e);
            if (rec && a.spec_hist) a.spec_hist[hrow + k] = e;
        }
        if (a.ipr && rec && a.ipr_hist) a.ipr_hist[hrow + k] = a.ipr[row + k];
        const uint8_t fi = a.f[row + k];
        a.focc_sum[row + k] += fi;
        if (rec && a.focc_hist) a.focc_hist[hrow + k] = fi;
    }
    if (a.V && rec) {
        const size_t NN = (size_t)N * N;
        const double* src = a.V + (size_t)b * NN;
        double* dst = a.evec_hist + ((size_t)b * a.cap + (size_t)m) * NN;
        for (size_t q = tid; q < NN; q += CM_T) dst[q] = src[q];
    }
    if (a.stiff) {
        const size_t crow = (size_t)b * a.nw;
        const size_t chrow = ((size_t)b * a.cap + (size_t)(rec ? m : 0)) * a.nw;
        for (int q = tid; q < a.nw; q += CM_T) {
            const double c = a.cond[crow + q];
            a.cond_sum[crow + q] += c;
            if (rec) a.cond_hist[chrow + q] = c;
        }
        if (tid == 0) {
            a.stiff_sum[b] += a.stiff[b];
            if (rec) a.stiff_hist[(size_t)b * a.cap + m] = a.stiff[b];
        }
    }
    ec = wave_sum(ec);
    d2 = wave_sum(d2);
    const int w = tid / WAVE;
    if ((tid & (WAVE - 1)) == 0) {
        red[0][w] = ec;
        red[1][w] = d2;
    }
    __syncthreads();
    if (tid == 0 && !spec && rec) {  // occupations only: the energy records are zero
        a.e[(size_t)b * a.cap + m] = 0.0;
        a.d2e[(size_t)b * a.cap + m] = 0.0;
        a.ce[(size_t)b * a.cap + m] = 0.0;
    }
    if (tid == 0 && spec) {
        ec = 0.0;
        d2 = 0.0;
        for (int i = 0; i < CM_T / WAVE; ++i) {
            ec += red[0][i];
            d2 += red[1][i];
        }
        const double ff = a.nW > 0 ? fkk::ff_energy_dev(a.f + row, N, a.nW, a.W) : 0.0;
        const double e = ec - a.mu_f[b] * (double)a.nf[b] + ff;
        d2 *= 0.5;
        if (rec) {
            a.e[(size_t)b * a.cap + m] = e;
            a.d2e[(size_t)b * a.cap + m] = d2;
            a.ce[(size_t)b * a.cap + m] = ec;
        }
        a.esum[b] += e;
        a.d2sum[b] += d2;
    }
}

__global__ void k_fermi(int B, int N, const double* __restrict__ ev, const double* __restrict__ beta,
                        double* __restrict__ fermi) {
    const int b = blockIdx.y, k = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || k >= N) return;
    fermi[(size_t)b * N + k] = 1.0 / (1.0 + exp(beta[b] * ev[(size_t)b * N + k]));
}

// count += 1, after every chain has recorded the measurement.  One thread.
__global__ void k_chain_measure_tick(long long* count) {
    if (threadIdx.x == 0) count[0] += 1;
}

}  // namespace fkm

using namespace fkm;

hipError_t fk_launch_ipr(int B, int N, const double* V, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_ipr, dim3((N + 255) / 256, B), dim3(256), 0, st, B, N, V, out);
    return hipGetLastError();
}

hipError_t fk_launch_stiffness(int B, int N, int stride0, int L0, double t, const double* V, const double* ev,
                               const double* fermi, int nw, const double* wgrid, double offset, double* JV,
                               double* mJ, double* tdiag, double* stiff, double* cond, hipStream_t st) {
    hipLaunchKernelGGL(k_stiff_prep, dim3((N + 255) / 256, B), dim3(256), 0, st, B, N, stride0, L0, t, V, JV, tdiag);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_gemm_tn, dim3((N + GT - 1) / GT, (N + GT - 1) / GT, B), dim3(GT * GT), 0, st, B, N, V, JV, mJ);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    hipLaunchKernelGGL(k_stiff_pairs, dim3(B), dim3(256), 0, st, B, N, mJ, ev, fermi, tdiag, nw, wgrid, offset, stiff,
                       cond);
    return hipGetLastError();
}

// spectra in chunks of at most 65535 (the grid's y extent)
hipError_t fk_launch_dos(int S, int N, const double* spec, int nw, const double* wgrid, double eta, double* out,
                         hipStream_t st) {
    for (int s0 = 0; s0 < S; s0 += 65535) {
        const int n = S - s0 < 65535 ? S - s0 : 65535;
        hipLaunchKernelGGL(k_dos, dim3((nw + 255) / 256, n), dim3(256), 0, st, n, N, spec + (size_t)s0 * N, nw, wgrid,
                           eta, out + (size_t)s0 * nw);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t fk_launch_ipr_w(int S, int N, const double* spec, const double* ipr, int nw, const double* wgrid,
                           double eta, double* out, hipStream_t st) {
    for (int s0 = 0; s0 < S; s0 += 65535) {
        const int n = S - s0 < 65535 ? S - s0 : 65535;
        hipLaunchKernelGGL(k_ipr_w, dim3((nw + 255) / 256, n), dim3(256), 0, st, n, N, spec + (size_t)s0 * N,
                           ipr + (size_t)s0 * N, nw, wgrid, eta, out + (size_t)s0 * nw);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t fk_launch_adjoint(int B, int N, const double* V, const double* fermi, const int32_t* f,
                             const double* beta, const int32_t* nbr, const int32_t* z, int zmax, int nt_bonds,
                             double* dens, double* bond_t, double* bond_tp, double* out, double* rho_full,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_adjoint_sites, dim3((N + ADJ_SITES - 1) / ADJ_SITES, B), dim3(WAVE * ADJ_SITES), 0, st, B,
                       N, V, fermi, nbr, z, zmax, nt_bonds, dens, bond_t, bond_tp);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_adjoint_reduce, dim3(B), dim3(256), 0, st, B, N, f, fermi, beta, dens, bond_t, bond_tp, out);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    if (rho_full) {
        hipLaunchKernelGGL(k_rho_full, dim3((N + GT - 1) / GT, (N + GT - 1) / GT, B), dim3(GT * GT), 0, st, B, N, V,
                           fermi, rho_full);
        r = hipGetLastError();
    }
    return r;
}

hipError_t fk_launch_chain_measure(const chain_meas_args& a, long long* count, hipStream_t st) {
    if (a.B <= 0 || a.N <= 0 || a.cap < 0 || a.nW < 0 || a.nW > 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_chain_measure, dim3(a.B), dim3(CM_T), 0, st, a);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_chain_measure_tick, dim3(1), dim3(64), 0, st, count);
    return hipGetLastError();
}

hipError_t fk_launch_fermi(int B, int N, const double* ev, const double* beta, double* fermi, hipStream_t st) {
    hipLaunchKernelGGL(k_fermi, dim3((N + 255) / 256, B), dim3(256), 0, st, B, N, ev, beta, fermi);
    return hipGetLastError();
}
