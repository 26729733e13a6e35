// FP64 complex Cholesky factorisation and solve on gfx950 for the shifted
// S-matrix (S + eps I) x = g of a stochastic-reconfiguration step.  The
// reference has no counterpart (its get_S_matrix, S_matrix.cu:10-58, only
// builds S).  Matrices are row-major, n x n complex<double>; the factor
// overwrites the lower triangle (A = L L^H).
//
// Blocked right-looking factorisation with NB = 32:
//   k_potrf_diag   factor the diagonal block in LDS (one 1024-thread workgroup)
//   k_trsm_panel   rows below: X L_kk^H = A_ik (one thread per row)
//   k_herk_update  trailing lower triangle: A_il -= X_i X_l^H (32 x 32 tiles)
// then forward and backward substitution in one workgroup.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "linalg_kernels.hpp"

namespace rbmhip {

constexpr int NB = 32;

RBM_DI cx<double> cmul_conj(cx<double> a, cx<double> b) {   // a * conj(b)
    return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}

// One workgroup of NB x NB threads factors A[k0:k0+nb, k0:k0+nb] in LDS.
// A non-positive pivot sets *info = global pivot index + 1.
__global__ __launch_bounds__(1024) void k_potrf_diag(int n, int k0, cx<double>* __restrict__ A, int* __restrict__ info) {
    __shared__ cx<double> L[NB][NB + 1];
    __shared__ int bad;
    const int i = threadIdx.x / NB, l = threadIdx.x % NB;
    const int nb = n - k0 < NB ? n - k0 : NB;
    if (threadIdx.x == 0) bad = 0;
    L[i][l] = (i < nb && l < nb && l <= i) ? A[(int64_t)(k0 + i) * n + k0 + l] : mk(0.0, 0.0);
    __syncthreads();
    for (int j = 0; j < nb; j++) {
        if (i == j && l == j) {
            const double d = L[j][j].re;
            if (!(d > 0.0)) {
                bad = k0 + j + 1;
                L[j][j] = mk(1.0, 0.0);
            } else {
                L[j][j] = mk(sqrt(d), 0.0);
            }
        }
        __syncthreads();
        if (l == j && i > j && i < nb) {
            const double r = 1.0 / L[j][j].re;
            L[i][j] = mk(L[i][j].re * r, L[i][j].im * r);
        }
        __syncthreads();
        if (i > j && i < nb && l > j && l <= i) {
            const cx<double> u = cmul_conj(L[i][j], L[l][j]);
            L[i][l] = mk(L[i][l].re - u.re, L[i][l].im - u.im);
        }
        __syncthreads();
    }
    if (i < nb && l < nb && l <= i) A[(int64_t)(k0 + i) * n + k0 + l] = L[i][l];
    if (threadIdx.x == 0 && bad && *info == 0) *info = bad;
}

// Rows r >= k0 + nb: solve X L_kk^H = A[r, k0:k0+nb] in place.
__global__ __launch_bounds__(256) void k_trsm_panel(int n, int k0, cx<double>* __restrict__ A) {
    __shared__ cx<double> L[NB][NB + 1];
    const int nb = n - k0 < NB ? n - k0 : NB;
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
        const int i = t / NB, l = t % NB;
        L[i][l] = (i < nb && l <= i) ? A[(int64_t)(k0 + i) * n + k0 + l] : mk(0.0, 0.0);
    }
    __syncthreads();
    const int64_t r = (int64_t)k0 + nb + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    cx<double> x[NB];
    cx<double>* row = A + r * n + k0;
#pragma unroll
    for (int j = 0; j < NB; j++) x[j] = j < nb ? row[j] : mk(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j < nb) {
            cx<double> a = x[j];
#pragma unroll
            for (int m = 0; m < NB; m++) {
                if (m < j) {
                    const cx<double> u = cmul_conj(x[m], L[j][m]);
                    a = mk(a.re - u.re, a.im - u.im);
                }
            }
            const double r_ = 1.0 / L[j][j].re;
            x[j] = mk(a.re * r_, a.im * r_);
        }
    }
#pragma unroll
    for (int j = 0; j < NB; j++)
        if (j < nb) row[j] = x[j];
}

// Trailing update of the lower triangle below the panel: tile (bi, bj) with
// bi >= bj of 32 x 32 entries, rows/cols offset by k0 + nb.
__global__ __launch_bounds__(256) void k_herk_update(int n, int k0, cx<double>* __restrict__ A) {
    __shared__ cx<double> Pi[NB][NB + 1];
    __shared__ cx<double> Pl[NB][NB + 1];
    const int nb = n - k0 < NB ? n - k0 : NB;
    const int base = k0 + nb;
    // blockIdx.x enumerates lower-triangle tiles: t = bi (bi + 1) / 2 + bj
    const int t = blockIdx.x;
    int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= t) bi++;
    while (bi * (bi + 1) / 2 > t) bi--;
    const int bj = t - bi * (bi + 1) / 2;
    const int r0 = base + bi * NB, c0 = base + bj * NB;
    for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) {
        const int a = q / NB, m = q % NB;
        const int ri = r0 + a, rl = c0 + a;
        Pi[a][m] = (ri < n && m < nb) ? A[(int64_t)ri * n + k0 + m] : mk(0.0, 0.0);
        Pl[a][m] = (rl < n && m < nb) ? A[(int64_t)rl * n + k0 + m] : mk(0.0, 0.0);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) {
        const int a = q / NB, b = q % NB;
        const int ri = r0 + a, cl = c0 + b;
        if (ri >= n || cl >= n || cl > ri) continue;
        double sr = 0.0, si = 0.0;
#pragma unroll 8
        for (int m = 0; m < NB; m++) {
            const cx<double> u = cmul_conj(Pi[a][m], Pl[b][m]);
            sr += u.re;
            si += u.im;
        }
        cx<double>& e = A[(int64_t)ri * n + cl];
        e = mk(e.re - sr, e.im - si);
    }
}

// Solve L y = b then L^H x = y in one workgroup (x may alias b).
__global__ __launch_bounds__(1024) void k_chol_solve(int n, const cx<double>* __restrict__ A, cx<double>* __restrict__ x) {
    __shared__ cx<double> blk[NB];
    // forward: y_i = (b_i - sum_{m<i} L_im y_m) / L_ii
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = n - k0 < NB ? n - k0 : NB;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int j = 0; j < nb; j++) {
                cx<double> a = x[k0 + j];
                for (int m = 0; m < j; m++) {
                    const cx<double> L = A[(int64_t)(k0 + j) * n + k0 + m];
                    const cx<double> y = blk[m];
                    a = mk(a.re - (L.re * y.re - L.im * y.im), a.im - (L.re * y.im + L.im * y.re));
                }
                const double r = 1.0 / A[(int64_t)(k0 + j) * n + k0 + j].re;
                blk[j] = mk(a.re * r, a.im * r);
                x[k0 + j] = blk[j];
            }
        }
        __syncthreads();
        for (int i = k0 + nb + threadIdx.x; i < n; i += blockDim.x) {
            cx<double> a = x[i];
            for (int m = 0; m < nb; m++) {
                const cx<double> L = A[(int64_t)i * n + k0 + m];
                const cx<double> y = blk[m];
                a = mk(a.re - (L.re * y.re - L.im * y.im), a.im - (L.re * y.im + L.im * y.re));
            }
            x[i] = a;
        }
    }
    // backward: x_i = (y_i - sum_{m>i} conj(L_mi) x_m) / L_ii
    const int nblk = (n + NB - 1) / NB;
    for (int kb = nblk - 1; kb >= 0; kb--) {
        const int k0 = kb * NB;
        const int nb = n - k0 < NB ? n - k0 : NB;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int j = nb - 1; j >= 0; j--) {
                cx<double> a = x[k0 + j];
                for (int m = j + 1; m < nb; m++) {
                    const cx<double> u = cmul_conj(blk[m], A[(int64_t)(k0 + m) * n + k0 + j]);
                    a = mk(a.re - u.re, a.im - u.im);
                }
                const double r = 1.0 / A[(int64_t)(k0 + j) * n + k0 + j].re;
                blk[j] = mk(a.re * r, a.im * r);
                x[k0 + j] = blk[j];
            }
        }
        __syncthreads();
        // rows above the block: x_i -= sum_{m in block} conj(L_mi) x_m
        for (int i = threadIdx.x; i < k0; i += blockDim.x) {
            cx<double> a = x[i];
            for (int m = 0; m < nb; m++) {
                const cx<double> u = cmul_conj(blk[m], A[(int64_t)(k0 + m) * n + i]);
                a = mk(a.re - u.re, a.im - u.im);
            }
            x[i] = a;
        }
    }
}

hipError_t launch_cholesky(int n, cx<double>* A, int* info, hipStream_t stream) {
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = n - k0 < NB ? n - k0 : NB;
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(NB * NB), 0, stream, n, k0, A, info);
        const int rest = n - k0 - nb;
        if (rest <= 0) break;
        hipLaunchKernelGGL(k_trsm_panel, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, stream, n, k0, A);
        const int tiles = (rest + NB - 1) / NB;
        hipLaunchKernelGGL(k_herk_update, dim3((unsigned)(tiles * (tiles + 1) / 2)), dim3(256), 0, stream, n, k0, A);
    }
    return hipGetLastError();
}

hipError_t launch_cholesky_solve(int n, const cx<double>* L, cx<double>* x, hipStream_t stream) {
    hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(1024), 0, stream, n, L, x);
    return hipGetLastError();
}

}  // namespace rbmhip
