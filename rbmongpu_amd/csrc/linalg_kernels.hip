// FP64 complex Cholesky factorisation and solve on gfx950 for the shifted
// S-matrix (S + eps I) x = g of a stochastic-reconfiguration step.  The
// reference has no counterpart (its get_S_matrix, S_matrix.cu:10-58, only
// builds S).  Matrices are row-major, n x n complex<double>; the factor
// overwrites the lower triangle (A = L L^H).
//
// Blocked right-looking factorisation, panels of NB = 32 in block columns of
// NB2 = 128:
//   k_potrf_diag   factor the diagonal block in LDS (one 1024-thread workgroup)
//   k_trsm_panel   rows below: X L_kk^H = A_ik (32 lanes per row)
//   k_herk_update  lower-triangle update A_il -= X_i X_l^H (64 x 64 tiles):
//                  per panel within its block column, per block column
//                  (rank 128) on the trailing matrix
// then blocked forward and backward substitution (k_trsv_block, k_trsv_update).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "linalg_kernels.hpp"

namespace rbmhip {

constexpr int NB = 32;

RBM_DI cx<double> cmul_conj(cx<double> a, cx<double> b) {   // a * conj(b)
    return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}

// One workgroup of NB x NB threads factors A[k0:k0+nb, k0:k0+nb] in LDS, one
// barrier per column: at step j every thread reads the pivot d = A_jj,
// thread (i, j) writes the finished L_ij (A_ij / sqrt(d), or sqrt(d) on the
// diagonal) straight to global memory, and threads (i, l) with j < l <= i
// subtract (A_ij / sqrt d) conj(A_lj / sqrt d) in LDS.  Column j is only read
// during step j, so the scaled values never need to go back to LDS.
// A non-positive pivot sets *info = global pivot index + 1 (and is replaced
// by 1 so that the factorisation runs to the end).
__global__ __launch_bounds__(1024) void k_potrf_diag(int n, int k0, cx<double>* __restrict__ A, int* __restrict__ info) {
    __shared__ cx<double> L[NB][NB + 1];
    const int i = threadIdx.x / NB, l = threadIdx.x % NB;
    const int nb = n - k0 < NB ? n - k0 : NB;
    L[i][l] = (i < nb && l < nb && l <= i) ? A[(int64_t)(k0 + i) * n + k0 + l] : mk(0.0, 0.0);
    __syncthreads();
    int bad = 0;
    for (int j = 0; j < nb; j++) {
        double d = L[j][j].re;
        if (!(d > 0.0)) {
            if (!bad) bad = k0 + j + 1;
            d = 1.0;
        }
        const double sd = sqrt(d), rs = 1.0 / sd;
        if (i < nb && i >= j) {
            if (l == j) {
                A[(int64_t)(k0 + i) * n + k0 + j] = i == j ? mk(sd, 0.0) : mk(L[i][j].re * rs, L[i][j].im * rs);
            } else if (l > j && l <= i) {
                const cx<double> a = L[i][j], c = L[l][j];
                const cx<double> u = cmul_conj(mk(a.re * rs, a.im * rs), mk(c.re * rs, c.im * rs));
                L[i][l] = mk(L[i][l].re - u.re, L[i][l].im - u.im);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && bad && *info == 0) *info = bad;
}

// Lane j's value of a complex double, broadcast to the whole wavefront (j is
// wave-uniform).
RBM_DI cx<double> lane_bcast(cx<double> v, int j) {
    const int2 r = *reinterpret_cast<int2*>(&v.re), i = *reinterpret_cast<int2*>(&v.im);
    int2 rr, ii;
    rr.x = __builtin_amdgcn_readlane(r.x, j); rr.y = __builtin_amdgcn_readlane(r.y, j);
    ii.x = __builtin_amdgcn_readlane(i.x, j); ii.y = __builtin_amdgcn_readlane(i.y, j);
    return mk(*reinterpret_cast<double*>(&rr), *reinterpret_cast<double*>(&ii));
}

// Rows r >= k0 + nb: solve X L_kk^H = A[r, k0:k0+nb] in place, i.e. for each
// row x_j = (a_j - sum_{m<j} x_m conj(L_jm)) / L_jj.  32 lanes per row (two
// rows per wavefront), lane j holding x_j: at step m the row's x_m is final and
// lanes j > m subtract x_m conj(L_jm), so each x_j receives its terms in the
// order m = 0, 1, ... as in a sequential solve.
__global__ __launch_bounds__(256) void k_trsm_panel(int n, int k0, cx<double>* __restrict__ A) {
    __shared__ cx<double> L[NB][NB + 1];
    const int nb = n - k0 < NB ? n - k0 : NB;
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
        const int i = t / NB, l = t % NB;
        L[i][l] = (i < nb && l <= i) ? A[(int64_t)(k0 + i) * n + k0 + l] : mk(0.0, 0.0);
    }
    __syncthreads();
    const int j = threadIdx.x & 31, half = (threadIdx.x >> 5) & 1;
    const int64_t r = (int64_t)k0 + nb + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 32;
    const bool live = r < n && j < nb;
    cx<double>* row = A + (r < n ? r : n - 1) * n + k0;
    cx<double> x = live ? row[j] : mk(0.0, 0.0);
    for (int m = 0; m < nb; m++) {
        if (j == m) {
            const double rd = 1.0 / L[m][m].re;
            x = mk(x.re * rd, x.im * rd);
        }
        const cx<double> x0 = lane_bcast(x, m), x1 = lane_bcast(x, 32 + m);
        const cx<double> xm = half ? x1 : x0;
        if (j > m) {
            const cx<double> u = cmul_conj(xm, L[j][m]);
            x = mk(x.re - u.re, x.im - u.im);
        }
    }
    if (live) row[j] = x;
}

// Trailing update A_il -= sum_m A_im conj(A_lm) over the panel columns
// m = kc .. kc + kw - 1, for entries with row i >= column l >= base and
// l < col_end.  One 256-thread workgroup per 64 x 64 tile: thread
// (tx, ty) = (t % 16, t / 16) owns rows ty + 16 a and columns tx + 16 b
// (a, b < 4), so each panel value read from LDS serves four products.  The
// panel is staged 32 columns at a time.  ctiles == 0: the lower-triangle tiles
// of [base, n)^2, t = bi (bi + 1) / 2 + bj; ctiles > 0: a 2-D grid over a band
// of ctiles column tiles (bj = blockIdx.x, bi = blockIdx.y), tiles above the
// diagonal exit.
constexpr int HT = 64;
__global__ __launch_bounds__(256) void k_herk_update(int n, int kc, int kw, int base, int col_end, int ctiles,
                                                     cx<double>* __restrict__ A) {
    __shared__ cx<double> Pi[HT][NB + 1];
    __shared__ cx<double> Pl[HT][NB + 1];
    int bi, bj;
    if (ctiles == 0) {
        const int t = blockIdx.x;
        bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((bi + 1) * (bi + 2) / 2 <= t) bi++;
        while (bi * (bi + 1) / 2 > t) bi--;
        bj = t - bi * (bi + 1) / 2;
    } else {
        bj = blockIdx.x;
        bi = blockIdx.y;
        if (bj > bi) return;
    }
    const int r0 = base + bi * HT, c0 = base + bj * HT;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double accr[4][4], acci[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) accr[a][b] = acci[a][b] = 0.0;
    for (int k0 = 0; k0 < kw; k0 += NB) {
        for (int q = threadIdx.x; q < HT * NB; q += 256) {
            const int a = q / NB, m = q % NB;
            const int ri = r0 + a, rl = c0 + a;
            const bool live = k0 + m < kw;
            Pi[a][m] = (ri < n && live) ? A[(int64_t)ri * n + kc + k0 + m] : mk(0.0, 0.0);
            Pl[a][m] = (rl < n && live) ? A[(int64_t)rl * n + kc + k0 + m] : mk(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll 2
        for (int m = 0; m < NB; m++) {
            cx<double> pi[4], pl[4];
#pragma unroll
            for (int a = 0; a < 4; a++) pi[a] = Pi[ty + 16 * a][m];
#pragma unroll
            for (int b = 0; b < 4; b++) pl[b] = Pl[tx + 16 * b][m];
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    // pi conj(pl)
                    accr[a][b] = __builtin_fma(pi[a].re, pl[b].re, __builtin_fma(pi[a].im, pl[b].im, accr[a][b]));
                    acci[a][b] = __builtin_fma(pi[a].im, pl[b].re, __builtin_fma(-pi[a].re, pl[b].im, acci[a][b]));
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int ri = r0 + ty + 16 * a;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int cl = c0 + tx + 16 * b;
            if (ri < n && cl <= ri && cl < col_end) {
                cx<double>& e = A[(int64_t)ri * n + cl];
                e = mk(e.re - accr[a][b], e.im - acci[a][b]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Products C = X^T Y of row-major K x ld operands on FP32 MFMA: the per-layer
// gradient products of PsiDeep (all tiles, split over K).
// A 256-thread workgroup computes a 128 x 128 tile as 2 x 2 waves of 64 x 64,
// each 2 x 2 v_mfma_f32_32x32x2f32 tiles (lane l supplies A[l % 32][l / 32]
// and B[l / 32][l % 32]; accumulator v of lane l is row
// 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32).  K advances 16 rows per LDS
// stage; the next stage's rows are loaded into registers while the current
// one is multiplied.
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int GR_T = 128, GR_K = 16, GR_LD = GR_T + 4;

// C[m0.., n0..] = X[k_begin..k_end)^T Y[k_begin..k_end) for one 128 x 128 tile
RBM_DI void tile_tn(const float* __restrict__ X, int ldx, int m0, const float* __restrict__ Y, int ldy, int n0,
                    int k_begin, int k_end, float* __restrict__ C, int ldc) {
    __shared__ float As[GR_K][GR_LD];
    __shared__ float Bs[GR_K][GR_LD];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    // staging: 16 rows x 128 columns per operand = 512 float4; thread tid loads
    // float4 q = tid and tid + 256 of each (row q / 32, column 4 (q % 32))
    const int q0 = tid, q1 = tid + 256;
    const int r0 = q0 >> 5, c0 = (q0 & 31) * 4, r1 = q1 >> 5, c1 = (q1 & 31) * 4;
    float4 ra0, ra1, rb0, rb1;
#define GR_FETCH(k0)                                                                        \
    ra0 = *reinterpret_cast<const float4*>(X + (int64_t)((k0) + r0) * ldx + m0 + c0);        \
    ra1 = *reinterpret_cast<const float4*>(X + (int64_t)((k0) + r1) * ldx + m0 + c1);        \
    rb0 = *reinterpret_cast<const float4*>(Y + (int64_t)((k0) + r0) * ldy + n0 + c0);        \
    rb1 = *reinterpret_cast<const float4*>(Y + (int64_t)((k0) + r1) * ldy + n0 + c1);
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x16{};
    GR_FETCH(k_begin)
    for (int k0 = k_begin; k0 < k_end; k0 += GR_K) {
        __syncthreads();
        *reinterpret_cast<float4*>(&As[r0][c0]) = ra0;
        *reinterpret_cast<float4*>(&As[r1][c1]) = ra1;
        *reinterpret_cast<float4*>(&Bs[r0][c0]) = rb0;
        *reinterpret_cast<float4*>(&Bs[r1][c1]) = rb1;
        __syncthreads();
        if (k0 + GR_K < k_end) { GR_FETCH(k0 + GR_K) }
#pragma unroll
        for (int kk = 0; kk < GR_K; kk += 2) {
            float a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) a[i] = As[kk + (lane >> 5)][wm + 32 * i + (lane & 31)];
#pragma unroll
            for (int j = 0; j < 2; j++) b[j] = Bs[kk + (lane >> 5)][wn + 32 * j + (lane & 31)];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#undef GR_FETCH
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int row = m0 + wm + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int col = n0 + wn + 32 * j + (lane & 31);
                C[(int64_t)row * ldc + col] = acc[i][j][v];
            }
}

// Split-K product: workgroup (x, z) covers tile x of the (ldx / 128) x (ldy / 128)
// grid over rows [z Kper, (z + 1) Kper); out[z] is ldx x ldy.
__global__ __launch_bounds__(256) void k_gemm_tn(int ldx, int ldy, int Kper, const float* __restrict__ X,
                                                 const float* __restrict__ Y, float* __restrict__ out) {
    const int tn = ldy / GR_T;
    const int bm = blockIdx.x / tn, bn = blockIdx.x % tn;
    const int k0 = (int)blockIdx.z * Kper;
    tile_tn(X, ldx, bm * GR_T, Y, ldy, bn * GR_T, k0, k0 + Kper, out + (int64_t)blockIdx.z * ldx * ldy, ldy);
}

// acc[i] += sum_z part[z n + i], the split partials added in FP64 in a fixed order
__global__ __launch_bounds__(256) void k_accum_splits_f64(int64_t n, int splits, const float* __restrict__ part,
                                                          double* __restrict__ acc) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int z = 0; z < splits; z++) s += (double)part[(int64_t)z * n + i];
        acc[i] += s;
    }
}

hipError_t launch_accum_splits_f64(int64_t n, int splits, const float* part, double* acc, hipStream_t stream) {
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_accum_splits_f64, dim3((unsigned)g), dim3(256), 0, stream, n, splits, part, acc);
    return hipGetLastError();
}

hipError_t launch_gemm_tn_splitk(int ldx, int ldy, int64_t K, int splits, const float* X, const float* Y, float* out,
                                 hipStream_t stream) {
    if (ldx <= 0 || ldy <= 0 || ldx % GR_T || ldy % GR_T || splits < 1 || K <= 0 || K % ((int64_t)GR_K * splits) ||
        K > ((int64_t)1 << 30))
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)((ldx / GR_T) * (ldy / GR_T)), 1, (unsigned)splits);
    hipLaunchKernelGGL(k_gemm_tn, grid, dim3(256), 0, stream, ldx, ldy, (int)(K / splits), X, Y, out);
    return hipGetLastError();
}

// Two-level blocking: block columns of NB2 = 128 are factored panel by panel
// (NB = 32), each panel's update confined to the rest of its block column;
// then one rank-128 update covers the trailing matrix, which is read and
// written once per 128 columns instead of once per 32.
constexpr int NB2 = 128;
hipError_t launch_cholesky(int n, cx<double>* A, int* info, hipStream_t stream) {
    for (int K0 = 0; K0 < n; K0 += NB2) {
        const int bend = n - K0 < NB2 ? n : K0 + NB2;
        for (int c = K0; c < bend; c += NB) {
            const int nb = n - c < NB ? n - c : NB;
            hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(NB * NB), 0, stream, n, c, A, info);
            const int rest = n - c - nb;
            if (rest <= 0) break;
            hipLaunchKernelGGL(k_trsm_panel, dim3((unsigned)((rest + 7) / 8)), dim3(256), 0, stream, n, c, A);
            const int base = c + nb;
            if (base < bend) {
                const int ct = (bend - base + HT - 1) / HT, rt = (n - base + HT - 1) / HT;
                hipLaunchKernelGGL(k_herk_update, dim3((unsigned)ct, (unsigned)rt), dim3(256), 0, stream, n, c, nb,
                                   base, bend, ct, A);
            }
        }
        const int rest = n - bend;
        if (rest <= 0) break;
        const int tiles = (rest + HT - 1) / HT;
        hipLaunchKernelGGL(k_herk_update, dim3((unsigned)(tiles * (tiles + 1) / 2)), dim3(256), 0, stream, n, K0,
                           bend - K0, bend, n, 0, A);
    }
    return hipGetLastError();
}

// Blocked substitution over all CUs.  Per diagonal block k0 (NB rows):
//   k_trsv_block   one wavefront solves the block with its 32 x 32 factor,
//                  one lane per row
//                  (forward: y_b = L_bb^{-1} x_b; backward: x_b = L_bb^{-H} y_b)
//   k_trsv_update  the rows outside the block subtract the block's products,
//                  one wavefront per row pair with coalesced row reads
//                  (forward: rows below, x_i -= L_ib y_b; backward: rows above,
//                  x_i -= sum_m conj(L_mi) x_m over the block's rows m).
__global__ __launch_bounds__(64) void k_trsv_block(int n, int k0, int backward, const cx<double>* __restrict__ A,
                                                   cx<double>* __restrict__ x) {
    // lane j < nb holds x_j; at step m the finished x_m is broadcast and the
    // lanes still to finish subtract its term (forward: L_jm x_m for j > m;
    // backward: conj(L_mj) x_m for j < m)
    __shared__ cx<double> Lb[NB][NB + 1];
    const int nb = n - k0 < NB ? n - k0 : NB;
    for (int t = threadIdx.x; t < NB * NB; t += 64) {
        const int i = t / NB, l = t % NB;
        Lb[i][l] = (i < nb && l < nb && l <= i) ? A[(int64_t)(k0 + i) * n + k0 + l] : mk(0.0, 0.0);
    }
    __syncthreads();
    const int j = threadIdx.x;
    cx<double> v = j < nb ? x[k0 + j] : mk(0.0, 0.0);
    if (!backward) {
        for (int m = 0; m < nb; m++) {
            if (j == m) {
                const double r = 1.0 / Lb[m][m].re;
                v = mk(v.re * r, v.im * r);
            }
            const cx<double> y = lane_bcast(v, m);
            if (j > m && j < nb) {
                const cx<double> L = Lb[j][m];
                v = mk(v.re - (L.re * y.re - L.im * y.im), v.im - (L.re * y.im + L.im * y.re));
            }
        }
    } else {
        for (int m = nb - 1; m >= 0; m--) {
            if (j == m) {
                const double r = 1.0 / Lb[m][m].re;
                v = mk(v.re * r, v.im * r);
            }
            const cx<double> y = lane_bcast(v, m);
            if (j < m) {
                const cx<double> u = cmul_conj(y, Lb[m][j]);
                v = mk(v.re - u.re, v.im - u.im);
            }
        }
    }
    if (j < nb) x[k0 + j] = v;
}

__global__ __launch_bounds__(256) void k_trsv_update(int n, int k0, int backward, const cx<double>* __restrict__ A,
                                                     cx<double>* __restrict__ x) {
    __shared__ cx<double> xb[NB];
    const int nb = n - k0 < NB ? n - k0 : NB;
    if (threadIdx.x < NB) xb[threadIdx.x] = threadIdx.x < nb ? x[k0 + threadIdx.x] : mk(0.0, 0.0);
    __syncthreads();
    const int lane = threadIdx.x & 31, half = threadIdx.x >> 5;   // 8 rows per workgroup, 32 lanes per row
    const int64_t row = (int64_t)blockIdx.x * 8 + half;
    // forward: rows k0 + nb .. n-1; backward: rows 0 .. k0-1
    const int64_t i = backward ? row : k0 + nb + row;
    const bool live = backward ? row < k0 : i < n;
    cx<double> u = mk(0.0, 0.0);
    if (live && lane < nb) {
        if (!backward) {
            const cx<double> L = A[i * n + k0 + lane], y = xb[lane];
            u = mk(L.re * y.re - L.im * y.im, L.re * y.im + L.im * y.re);
        } else {
            u = cmul_conj(xb[lane], A[(int64_t)(k0 + lane) * n + i]);
        }
    }
    // sum over the 32 lanes of the row (lanes 0-31 and 32-63 of a wavefront hold two rows)
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
        u.re += __shfl_xor(u.re, off, 32);
        u.im += __shfl_xor(u.im, off, 32);
    }
    if (live && lane == 0) x[i] = mk(x[i].re - u.re, x[i].im - u.im);
}

hipError_t launch_cholesky_solve(int n, const cx<double>* L, cx<double>* x, hipStream_t stream) {
    const int nblk = (n + NB - 1) / NB;
    for (int kb = 0; kb < nblk; kb++) {
        const int k0 = kb * NB, nb = n - k0 < NB ? n - k0 : NB;
        hipLaunchKernelGGL(k_trsv_block, dim3(1), dim3(64), 0, stream, n, k0, 0, L, x);
        const int rest = n - k0 - nb;
        if (rest > 0) hipLaunchKernelGGL(k_trsv_update, dim3((unsigned)((rest + 7) / 8)), dim3(256), 0, stream, n, k0, 0, L, x);
    }
    for (int kb = nblk - 1; kb >= 0; kb--) {
        const int k0 = kb * NB;
        hipLaunchKernelGGL(k_trsv_block, dim3(1), dim3(64), 0, stream, n, k0, 1, L, x);
        if (k0 > 0) hipLaunchKernelGGL(k_trsv_update, dim3((unsigned)((k0 + 7) / 8)), dim3(256), 0, stream, n, k0, 1, L, x);
    }
    return hipGetLastError();
}

}  // namespace rbmhip
