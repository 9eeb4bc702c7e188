// fk_green.hip -- lattice Green's function and local densities of states averaged over the
// eigen-decompositions a chain set recorded (measure_eigenfunctions with the spectrum
// history), the quantities of data_saver::save_gwr (prog/data_save.hxx:535-708):
//
//   G(z)_{ij} = (1/n) sum_r sum_k V_r[i][k] V_r[j][k] / (z - e_rk),   z = w + i eta
//             Re: (w - e) / ((w - e)^2 + eta^2),  Im: -eta / ((w - e)^2 + eta^2)
//             (data_save.hxx:620-621, gwr_re and gwr_im)
//   rho_i(w)  = -Im G(z)_{ii} / pi
//             = (1/n) sum_r sum_k V_r[i][k]^2 (eta / pi) / ((w - e_rk)^2 + eta^2)
//             (data_save.hxx:625, gw_diag)
//
// Both are products over K = (record, eigenvalue) pairs with a weight per K: G is
// V diag(g) V^T summed over records, rho is L(w, K) (V o V)^T.  A 64 x 64 output tile per
// workgroup of 256 threads, 4 x 4 outputs per thread, K staged through LDS 16 at a time;
// the records are split over grid.z so that small lattices still fill the chip, with a
// deterministic reduction of the partial tiles afterwards.  FP64 vector FMA: on gfx950
// the FP64 matrix-core rate equals the vector rate, so FMA tiles lose nothing.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_internal.h"

namespace fkgr {

constexpr int GT = 64;   // output tile
constexpr int GK = 16;   // K chunk
constexpr int NTH = 256;

// record r of the chain set: chain r / n, measurement r % n
__device__ __forceinline__ size_t rec_index(int r, int n, int cap) { return (size_t)(r / n) * cap + (r % n); }

__global__ void __launch_bounds__(NTH) k_green(int n, int cap, int N, const double* __restrict__ V,
                                               const double* __restrict__ ev, double w, double eta, int rps,
                                               int nrec, double* __restrict__ part_re, double* __restrict__ part_im) {
    __shared__ double Ai[GK][GT + 1], Aj[GK][GT + 1];
    __shared__ double gre[GK], gim[GK];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
    const int r0 = blockIdx.z * rps, r1 = min(nrec, r0 + rps);
    double are[4][4], aim[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) are[a][b] = aim[a][b] = 0.0;
    for (int r = r0; r < r1; ++r) {
        const size_t ri = rec_index(r, n, cap);
        const double* Vr = V + ri * N * N;
        const double* er = ev + ri * N;
        for (int k0 = 0; k0 < N; k0 += GK) {
#pragma unroll
            for (int q = 0; q < (GT * GK) / NTH; ++q) {
                const int idx = tid + NTH * q, row = idx / GK, kk = idx % GK, gk = k0 + kk;
                const int gi = i0 + row, gj = j0 + row;
                Ai[kk][row] = (gi < N && gk < N) ? Vr[(size_t)gi * N + gk] : 0.0;
                Aj[kk][row] = (gj < N && gk < N) ? Vr[(size_t)gj * N + gk] : 0.0;
            }
            if (tid < GK) {
                double re = 0.0, im = 0.0;
                if (k0 + tid < N) {
                    const double d = w - er[k0 + tid];
                    const double den = d * d + eta * eta;
                    re = d / den;
                    im = -eta / den;
                }
                gre[tid] = re;
                gim[tid] = im;
            }
            __syncthreads();
#pragma unroll 4
            for (int kk = 0; kk < GK; ++kk) {
                double bj[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) bj[b] = Aj[kk][tx + 16 * b];
                const double g_re = gre[kk], g_im = gim[kk];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const double ai = Ai[kk][ty + 16 * a];
                    const double xr = ai * g_re, xi = ai * g_im;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        are[a][b] = fma(xr, bj[b], are[a][b]);
                        aim[a][b] = fma(xi, bj[b], aim[a][b]);
                    }
                }
            }
            __syncthreads();
        }
    }
    const size_t base = (size_t)blockIdx.z * N * N;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
            if (i < N && j < N) {
                part_re[base + (size_t)i * N + j] = are[a][b];
                part_im[base + (size_t)i * N + j] = aim[a][b];
            }
        }
}

// rows of the output are frequencies, columns sites
__global__ void __launch_bounds__(NTH) k_ldos(int n, int cap, int N, const double* __restrict__ V,
                                              const double* __restrict__ ev, int nw, const double* __restrict__ wg,
                                              double eta, int rps, int nrec, double* __restrict__ part) {
    __shared__ double Lw[GK][GT + 1], V2[GK][GT + 1];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int w0 = blockIdx.y * GT, i0 = blockIdx.x * GT;
    const int r0 = blockIdx.z * rps, r1 = min(nrec, r0 + rps);
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    const double eta2 = eta * eta;
    for (int r = r0; r < r1; ++r) {
        const size_t ri = rec_index(r, n, cap);
        const double* Vr = V + ri * N * N;
        const double* er = ev + ri * N;
        for (int k0 = 0; k0 < N; k0 += GK) {
#pragma unroll
            for (int q = 0; q < (GT * GK) / NTH; ++q) {
                const int idx = tid + NTH * q, row = idx / GK, kk = idx % GK, gk = k0 + kk;
                const int gi = i0 + row, gw = w0 + row;
                double v = (gi < N && gk < N) ? Vr[(size_t)gi * N + gk] : 0.0;
                V2[kk][row] = v * v;
                double l = 0.0;
                if (gw < nw && gk < N) {
                    const double d = wg[gw] - er[gk];
                    l = 1.0 / (d * d + eta2);
                }
                Lw[kk][row] = l;
            }
            __syncthreads();
#pragma unroll 4
            for (int kk = 0; kk < GK; ++kk) {
                double vb[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) vb[b] = V2[kk][tx + 16 * b];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const double la = Lw[kk][ty + 16 * a];
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = fma(la, vb[b], acc[a][b]);
                }
            }
            __syncthreads();
        }
    }
    const size_t base = (size_t)blockIdx.z * nw * N;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int wi = w0 + ty + 16 * a, i = i0 + tx + 16 * b;
            if (wi < nw && i < N) part[base + (size_t)wi * N + i] = acc[a][b];
        }
}

// out[x] = scale * sum_s part[s][x], in split order
__global__ void k_reduce_parts(size_t len, int ns, const double* __restrict__ part, double scale,
                               double* __restrict__ out) {
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= len) return;
    double s = 0.0;
    for (int q = 0; q < ns; ++q) s += part[(size_t)q * len + x];
    out[x] = s * scale;
}

}  // namespace fkgr

using namespace fkgr;

int fk_green_splits(int nrec, int tiles) {
    // about 1024 workgroups in all, at least one record per split
    int s = (1024 + tiles - 1) / tiles;
    if (s > nrec) s = nrec;
    return s < 1 ? 1 : s;
}

hipError_t fk_launch_green(int n, int cap, int nchains, int N, const double* V, const double* ev, double w,
                           double eta, int splits, double* part_re, double* part_im, double* g_re, double* g_im,
                           hipStream_t st) {
    const int nrec = n * nchains;
    if (nrec < 1 || N < 1 || splits < 1) return hipErrorInvalidValue;
    const int rps = (nrec + splits - 1) / splits;
    const int ns = (nrec + rps - 1) / rps;
    const int tb = (N + GT - 1) / GT;
    hipLaunchKernelGGL(k_green, dim3(tb, tb, ns), dim3(NTH), 0, st, n, cap, N, V, ev, w, eta, rps, nrec, part_re,
                       part_im);
    const size_t len = (size_t)N * N;
    const unsigned nb = (unsigned)((len + 255) / 256);
    hipLaunchKernelGGL(k_reduce_parts, dim3(nb), dim3(256), 0, st, len, ns, part_re, 1.0 / nrec, g_re);
    hipLaunchKernelGGL(k_reduce_parts, dim3(nb), dim3(256), 0, st, len, ns, part_im, 1.0 / nrec, g_im);
    return hipGetLastError();
}

hipError_t fk_launch_ldos(int n, int cap, int nchains, int N, const double* V, const double* ev, int nw,
                          const double* wgrid, double eta, int splits, double* part, double* ldos, hipStream_t st) {
    const int nrec = n * nchains;
    if (nrec < 1 || N < 1 || nw < 1 || splits < 1) return hipErrorInvalidValue;
    const int rps = (nrec + splits - 1) / splits;
    const int ns = (nrec + rps - 1) / rps;
    hipLaunchKernelGGL(k_ldos, dim3((N + GT - 1) / GT, (nw + GT - 1) / GT, ns), dim3(NTH), 0, st, n, cap, N, V, ev,
                       nw, wgrid, eta, rps, nrec, part);
    const size_t len = (size_t)nw * N;
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len, ns, part,
                       eta / (M_PI * nrec), ldos);
    return hipGetLastError();
}
