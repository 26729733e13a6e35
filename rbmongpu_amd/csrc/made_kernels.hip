// gfx950 kernels of the masked autoregressive proposal (MADE):
//   k_gemm<EPI>           FP32 MFMA GEMM (v_mfma_f32_16x16x4f32) with fused epilogues
//                         for the training forward and backward passes
//   k_made_gather         spin words -> 0/1 input rows
//   k_made_batch_index    Philox-drawn minibatch indices
//   k_colsum_part/_finish, k_adam   bias gradients and the Adam update
//   k_made_inf_weights    bfloat16 inference copies of W1 (transposed) and W2
//   k_made_sample<HPL,S,W1F>  autoregressive sampling / scoring, weights in LDS
//   k_imh                 independent Metropolis-Hastings over precomputed proposals
#include "made_kernels.hpp"

namespace rbmhip {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
// GEMM.  A 256-thread workgroup computes a 64 x 64 tile of C as 2 x 2 waves of
// 32 x 32, each 2 x 2 MFMA 16x16x4 tiles; K advances 16 at a time through LDS.
// MFMA operand layout (16x16x4 f32): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator element v of lane l is D[4 (l / 16) + v][l % 16].
// ---------------------------------------------------------------------------
constexpr int GT = 64, GK = 16;

RBM_DI float sigmoidf_acc(float l) { return 1.0f / (1.0f + expf(-l)); }
RBM_DI float softplusf_acc(float l) { return fmaxf(l, 0.0f) + log1pf(expf(-fabsf(l))); }

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* __restrict__ A, long sam, long sak,
                                              const float* __restrict__ B, long sbk, long sbn, float* __restrict__ C,
                                              long ldc, GemmEpiArgs e) {
    __shared__ float As[GK][GT + 4];
    __shared__ float Bs[GK][GT + 4];
    __shared__ double red[256];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    if (e.lower_only && n0 > m0) return;   // uniform per workgroup, before any barrier
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    // split K: workgroup z covers [z K / gridDim.z, (z + 1) K / gridDim.z) and
    // writes its partial product at C + z M ldc
    const int kspan = K / gridDim.z;
    const int kbeg = blockIdx.z * kspan, kend = kbeg + kspan;
    C += (long)blockIdx.z * M * ldc;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int idx = tid + 256 * r;
            int mm, kk;
            if (sak == 1) { kk = idx & 15; mm = idx >> 4; } else { mm = idx & 63; kk = idx >> 6; }
            As[kk][mm] = A[(long)(m0 + mm) * sam + (long)(k0 + kk) * sak];
            int nn, kb;
            if (sbn == 1) { nn = idx & 63; kb = idx >> 6; } else { kb = idx & 15; nn = idx >> 4; }
            Bs[kb][nn] = B[(long)(k0 + kb) * sbk + (long)(n0 + nn) * sbn];
        }
        __syncthreads();
#pragma unroll
        for (int kq = 0; kq < GK; kq += 4) {
            float af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; i++) af[i] = As[kq + (lane >> 4)][wm + 16 * i + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; j++) bf[j] = Bs[kq + (lane >> 4)][wn + 16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    double part = 0.0;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                const int n = n0 + wn + 16 * j + (lane & 15);
                const float x = acc[i][j][v];
                float* out = C + (long)m * ldc + n;
                if (EPI == EPI_STORE) {
                    *out = x;
                } else if (EPI == EPI_BIAS_RELU) {
                    const float a = x + e.bias[n];
                    *out = a;
                    e.C2[(long)m * ldc + n] = fmaxf(a, 0.0f);
                } else if (EPI == EPI_LOGITS) {
                    const float l = x + e.bias[n];
                    const float xv = e.aux[(long)m * e.ld_aux + n];
                    const bool live = m < e.m_live && n < e.n_live;
                    *out = live ? (sigmoidf_acc(l) - xv) * e.inv_batch : 0.0f;
                    if (live) part += (double)(xv * l) - (double)softplusf_acc(l);
                } else if (EPI == EPI_MASK_W2) {
                    const bool ok = m < e.m_live && n < e.n_live && made_degree(n, e.m_live) <= m;
                    *out = ok ? x : 0.0f;
                } else if (EPI == EPI_MASK_W1) {
                    const bool ok = m < e.m_live && n < e.n_live && n < made_degree(m, e.n_live);
                    *out = ok ? x : 0.0f;
                } else if (EPI == EPI_RELU_GRAD) {
                    *out = e.aux[(long)m * e.ld_aux + n] > 0.0f ? x : 0.0f;
                }
            }
    if (EPI == EPI_LOGITS) {
        red[tid] = part;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int k = 0; k < 256; k++) s += red[k];
            e.partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
        }
    }
}

// Split-K reduction: C = epilogue(sum over the splits of the partial products).
template <int EPI>
__global__ __launch_bounds__(256) void k_splitk_reduce(int M, int N, int splits, const float* __restrict__ part,
                                                       float* __restrict__ C, long ldc, GemmEpiArgs e) {
    const int64_t total = (int64_t)M * N;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(t / N), n = (int)(t % N);
        float x = 0.0f;
        for (int z = 0; z < splits; z++) x += part[(int64_t)z * total + t];
        bool ok = true;
        if (EPI == EPI_MASK_W2) ok = m < e.m_live && n < e.n_live && made_degree(n, e.m_live) <= m;
        if (EPI == EPI_MASK_W1) ok = m < e.m_live && n < e.n_live && n < made_degree(m, e.n_live);
        C[(long)m * ldc + n] = ok ? x : 0.0f;
    }
}

hipError_t launch_gemm_splitk(int epi, int M, int N, int K, const float* A, long sam, long sak, const float* B,
                              long sbk, long sbn, float* C, long ldc, const GemmEpiArgs& e, int splits, float* scratch,
                              hipStream_t stream) {
    if (M % GT || N % GT || M <= 0 || N <= 0 || K <= 0 || splits < 1 || K % (GK * splits)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_gemm<EPI_STORE>), dim3(N / GT, M / GT, splits), dim3(256), 0, stream, M, N, K, A, sam, sak, B,
                       sbk, sbn, scratch, (long)N, e);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    const int64_t total = (int64_t)M * N;
    const unsigned blocks = (unsigned)((total + 255) / 256 < 2048 ? (total + 255) / 256 : 2048);
    switch (epi) {
        case EPI_STORE: hipLaunchKernelGGL((k_splitk_reduce<EPI_STORE>), dim3(blocks), dim3(256), 0, stream, M, N, splits, scratch, C, ldc, e); break;
        case EPI_MASK_W1: hipLaunchKernelGGL((k_splitk_reduce<EPI_MASK_W1>), dim3(blocks), dim3(256), 0, stream, M, N, splits, scratch, C, ldc, e); break;
        case EPI_MASK_W2: hipLaunchKernelGGL((k_splitk_reduce<EPI_MASK_W2>), dim3(blocks), dim3(256), 0, stream, M, N, splits, scratch, C, ldc, e); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gemm(int epi, int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk,
                       long sbn, float* C, long ldc, const GemmEpiArgs& e, hipStream_t stream) {
    if (M % GT || N % GT || K % GK || M <= 0 || N <= 0 || K <= 0) return hipErrorInvalidValue;
    const dim3 grid(N / GT, M / GT), block(256);
#define RBM_GEMM_CASE(E) \
    case E: hipLaunchKernelGGL((k_gemm<E>), grid, block, 0, stream, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, e); break;
    switch (epi) {
        RBM_GEMM_CASE(EPI_STORE)
        RBM_GEMM_CASE(EPI_BIAS_RELU)
        RBM_GEMM_CASE(EPI_LOGITS)
        RBM_GEMM_CASE(EPI_MASK_W2)
        RBM_GEMM_CASE(EPI_MASK_W1)
        RBM_GEMM_CASE(EPI_RELU_GRAD)
        default: return hipErrorInvalidValue;
    }
#undef RBM_GEMM_CASE
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Training helpers
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_made_gather(int64_t rows, int n_live, int N, int Np,
                                                     const uint64_t* __restrict__ spins,
                                                     const int64_t* __restrict__ idx, float* __restrict__ X) {
    const int64_t total = rows * Np;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = t / Np;
        const int i = (int)(t - s * Np);
        float v = 0.0f;
        if (s < n_live && i < N) {
            const int64_t src = idx ? idx[s] : s;
            v = (float)bit_of(spins[2 * src], spins[2 * src + 1], i);
        }
        X[t] = v;
    }
}

hipError_t launch_made_gather(int64_t rows, int n_live, int N, int Np, const uint64_t* spins, const int64_t* idx,
                              float* X, hipStream_t stream) {
    const int64_t total = rows * Np;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_made_gather, dim3((unsigned)blocks), dim3(256), 0, stream, rows, n_live, N, Np, spins, idx, X);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_made_batch_index(int B, int64_t n, uint32_t step, uint32_t seed_lo,
                                                          uint32_t seed_hi, int64_t* __restrict__ idx) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    const u32x4 r = philox(step, (uint32_t)s, 0u, 4u, seed_lo, seed_hi);
    const double u = ((double)r.x + 0.5) * (1.0 / 4294967296.0);
    int64_t k = (int64_t)(u * (double)n);
    idx[s] = k < n ? k : n - 1;
}

hipError_t launch_made_batch_index(int B, int64_t n, uint32_t step, uint32_t seed_lo, uint32_t seed_hi, int64_t* idx,
                                   hipStream_t stream) {
    hipLaunchKernelGGL(k_made_batch_index, dim3((B + 255) / 256), dim3(256), 0, stream, B, n, step, seed_lo, seed_hi,
                       idx);
    return hipGetLastError();
}

// Column sums of the rows x cols matrix G (the bias gradients: sums over the
// batch), columns >= n_live zero.  Two deterministic stages: the 256-thread
// workgroup (x, y) owns columns [64 x, 64 x + 64) of row slice y (rows
// [y R, (y + 1) R), R = ceil(rows / COLSUM_SLICES)); its 4 row groups of 64
// threads stride over the slice with coalesced 64-wide reads and are added in
// a fixed order into part[y][c]; k_colsum_finish then adds the slices in
// order.  part holds COLSUM_SLICES x cols.
__global__ __launch_bounds__(256) void k_colsum_part(int rows, int cols, int n_live, const float* __restrict__ G,
                                                     float* __restrict__ part) {
    __shared__ float red[4][64];
    const int cl = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + cl;
    const int per = (rows + (int)gridDim.y - 1) / (int)gridDim.y;
    const int r0 = (int)blockIdx.y * per, r1 = r0 + per < rows ? r0 + per : rows;
    float s = 0.0f;
    if (c < cols && c < n_live)
        for (int r = r0 + grp; r < r1; r += 4) s += G[(long)r * cols + c];
    red[grp][cl] = s;
    __syncthreads();
    if (grp == 0 && c < cols) part[(long)blockIdx.y * cols + c] = ((red[0][cl] + red[1][cl]) + red[2][cl]) + red[3][cl];
}

__global__ __launch_bounds__(256) void k_colsum_finish(int cols, int slices, const float* __restrict__ part,
                                                       float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float t = 0.0f;
    for (int y = 0; y < slices; y++) t += part[(long)y * cols + c];
    out[c] = t;
}

hipError_t launch_colsum(int rows, int cols, int n_live, const float* G, float* out, float* part, hipStream_t stream) {
    if (rows <= 0 || cols <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_colsum_part, dim3((cols + 63) / 64, COLSUM_SLICES), dim3(256), 0, stream, rows, cols, n_live, G,
                       part);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_colsum_finish, dim3((cols + 255) / 256), dim3(256), 0, stream, cols, COLSUM_SLICES, part, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_zero(int64_t n, float* __restrict__ p) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) p[k] = 0.0f;
}

hipError_t launch_zero(int64_t n, float* p, hipStream_t stream) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, stream, n, p);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float c1, float c2,
                                              float lr, float b1, float b2, float eps) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const float gk = g[k];
        const float mk_ = b1 * m[k] + (1.0f - b1) * gk;
        const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
        m[k] = mk_;
        v[k] = vk;
        p[k] -= lr * (mk_ * c1) / (sqrtf(vk * c2) + eps);
    }
}

hipError_t launch_adam(int64_t n, float* p, const float* g, float* m, float* v, int t, float lr, float b1, float b2,
                       float eps, hipStream_t stream) {
    const float c1 = (float)(1.0 / (1.0 - pow((double)b1, t)));
    const float c2 = (float)(1.0 / (1.0 - pow((double)b2, t)));
    int64_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)blocks), dim3(256), 0, stream, n, p, g, m, v, c1, c2, lr, b1, b2, eps);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_made_inf_weights(MadeShape s, const float* __restrict__ W1,
                                                          const float* __restrict__ b1, const float* __restrict__ W2,
                                                          uint16_t* __restrict__ W1t, uint16_t* __restrict__ W2b,
                                                          float* __restrict__ b1s) {
    const int64_t total = (int64_t)s.N * s.Hp;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / s.Hp), e = (int)(t % s.Hp);
        const int w = e >> 1, half = e & 1;                        // word w = 32 t + 2 l + b
        const int v = 2 * (w >> 5) + (w & 1), l = (w >> 1) & 15;   // pair v = 2 t + b of lane l
        const int p = 32 * v + 16 * half + l;                      // sorted position
        const int k = made_sorted_unit(p, s.N, s.H);
        const __bf16 a = (__bf16)W1[(long)k * s.Np + i];
        const __bf16 b = (__bf16)W2[(long)i * s.Hp + k];
        W1t[t] = __builtin_bit_cast(uint16_t, a);
        W2b[t] = __builtin_bit_cast(uint16_t, b);
    }
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < s.Hp; p += gridDim.x * blockDim.x)
        b1s[p] = b1[made_sorted_unit(p, s.N, s.H)];
}

hipError_t launch_made_inference_weights(const MadeShape& s, const float* W1, const float* b1, const float* W2,
                                         uint16_t* W1t_bf, uint16_t* W2_bf, float* b1_sorted, hipStream_t stream) {
    hipLaunchKernelGGL(k_made_inf_weights, dim3(256), dim3(256), 0, stream, s, W1, b1, W2, W1t_bf, W2_bf, b1_sorted);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Autoregressive sampling (SCORE = false) and scoring (SCORE = true).
// A 16-lane row owns two configurations (one above 256 hidden units),
// interleaved so that their serial per-position chains (dot product, row sum,
// sigmoid, update) overlap and each weight word read from LDS serves both.
// Hidden units are in degree-sorted order: lane l holds sorted units 16 u + l
// (u < HPL) as FP32 pre-activations a[u] and bfloat16 pairs hb[v] of
// relu(a[2v]), relu(a[2v + 1]).  With D_j units of degree <= j, the masks make
// output j read units [0, D_j) only and input j feed units [D_j, Hp) only, so
// per position j (D_j is the same for every lane, and whole 64-unit groups of
// zero weights are skipped):
//   l_j = b2_j + sum_{k < D_j} W2[j][k] relu(a_k)   (v_dot2c_f32_bf16, DPP row sum)
//   x_j drawn (or read), log q += log sigmoid(+-l_j)
//   a_k += x_j W1[k][j] for k >= D_j (an FMA with x_j in {0, 1}: a + W exactly
//   when x_j = 1, a unchanged when x_j = 0), relu(a) repacked in the groups
//   that output j + 1 reads
// Both weight matrices sit in LDS as bfloat16 (N x Hp each, the layout of
// launch_made_inference_weights), W1 expanded to FP32 when it fits (W1F).
// Uniforms: lane l of the row computes Philox blocks l and 16 + l of each of
// its configurations; position j reads word j % 4 of block j / 4 from its lane.
// ---------------------------------------------------------------------------
RBM_DI float dot2_bf16(uint32_t w, uint32_t h, float c) {
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w), __builtin_bit_cast(bf16x2, h), c, false);
}
// relu as one v_max_f32.  fmaxf(x, 0) compiles to two (a v_max_f32 x, x that
// quiets a possible signalling NaN, then the max with 0); the pre-activations
// here are FMA results, never NaN, and both forms agree on every other value.
RBM_DI float relu(float x) {
    float r;
    asm("v_max_f32 %0, 0, %1" : "=v"(r) : "v"(x));
    return r;
}
// log(1 + e) for e = exp(-|l|) in (0, 1]: the argument lies in (1, 2], so
// v_log_f32 (log2, about 1 ulp there) times ln 2 needs none of the
// denormal-range handling of logf.
RBM_DI float log_1p_e(float e) { return __builtin_amdgcn_logf(1.0f + e) * 0.6931471805599453f; }
RBM_DI uint32_t pack_relu_bf16(float lo, float hi) {
    bf16x2 v;
    v.x = (__bf16)relu(lo);
    v.y = (__bf16)relu(hi);
    return __builtin_bit_cast(uint32_t, v);
}

int made_hidden_per_lane(int Hp) { return Hp / 16; }

// configurations per 16-lane row: two while both fit in the 128 registers a
// 1024-thread workgroup allows (up to 256 hidden units), else one
constexpr int made_cpr(int hpl) { return hpl <= 16 ? 2 : 1; }

RBM_DI float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
RBM_DI float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }

// W1F: W1 sits in LDS as FP32, expanded from its bfloat16 copy at the start
// (the same values), so the hidden-unit updates need no unpacking; used when
// N x Hp x 6 bytes fit in LDS.  Otherwise W1 stays bfloat16 like W2.
template <int HPL, bool SCORE, bool W1F>
__global__ __launch_bounds__(1024) void k_made_sample(MadeInf m, int64_t n, uint32_t seed_lo, uint32_t seed_hi,
                                                      uint32_t call, const uint64_t* __restrict__ in_spins,
                                                      uint64_t* __restrict__ out_spins, double* __restrict__ log_q) {
    static_assert(HPL % 4 == 0, "hidden units per lane must be a multiple of 4");
    constexpr int C = made_cpr(HPL);
    constexpr int V = HPL / 2;   // bfloat16 pairs per lane
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N = m.N, Hp = m.Hp;
    uint16_t* sW2 = reinterpret_cast<uint16_t*>(smem_raw + (size_t)N * Hp * (W1F ? 4 : 2));
    {
        const uint4* g1 = reinterpret_cast<const uint4*>(m.W1t);
        const uint4* g2 = reinterpret_cast<const uint4*>(m.W2);
        uint4* l2 = reinterpret_cast<uint4*>(sW2);
        const int nv = N * Hp / 8;
        for (int k = threadIdx.x; k < nv; k += blockDim.x) {
            const uint4 w = g1[k];
            if (W1F) {
                float4* l1 = reinterpret_cast<float4*>(smem_raw) + 2 * k;
                l1[0] = make_float4(bf16_lo(w.x), bf16_hi(w.x), bf16_lo(w.y), bf16_hi(w.y));
                l1[1] = make_float4(bf16_lo(w.z), bf16_hi(w.z), bf16_lo(w.w), bf16_hi(w.w));
            } else {
                reinterpret_cast<uint4*>(smem_raw)[k] = w;
            }
            l2[k] = g2[k];
        }
    }
    __syncthreads();
    const uint32_t* sW1w = reinterpret_cast<const uint32_t*>(smem_raw);
    const float* sW1f = reinterpret_cast<const float*>(smem_raw);
    const uint32_t* sW2w = reinterpret_cast<const uint32_t*>(sW2);
    const int lane = threadIdx.x & 15, rows = blockDim.x >> 4;
    const int row_base = (threadIdx.x & 63) & ~15;

    for (int64_t g0 = ((int64_t)blockIdx.x * rows + (threadIdx.x >> 4)) * C; g0 < n; g0 += (int64_t)gridDim.x * rows * C) {
        int64_t gc[C];
        uint32_t uw[C][8];
        uint64_t x0[C], x1[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            gc[c] = g0 + c < n ? g0 + c : g0;      // a missing partner repeats the first (not written)
            x0[c] = 0; x1[c] = 0;
            if (SCORE) {
                x0[c] = in_spins[2 * gc[c]];
                x1[c] = in_spins[2 * gc[c] + 1];
            } else {
                const int64_t g = gc[c];
                const u32x4 r0 = philox((uint32_t)g, (uint32_t)(g >> 32), call, 0x100u + lane, seed_lo, seed_hi);
                uw[c][0] = r0.x; uw[c][1] = r0.y; uw[c][2] = r0.z; uw[c][3] = r0.w;
                if (N > 64) {
                    const u32x4 r1 = philox((uint32_t)g, (uint32_t)(g >> 32), call, 0x110u + lane, seed_lo, seed_hi);
                    uw[c][4] = r1.x; uw[c][5] = r1.y; uw[c][6] = r1.z; uw[c][7] = r1.w;
                } else {
                    uw[c][4] = uw[c][5] = uw[c][6] = uw[c][7] = 0u;
                }
                // each word becomes its uniform u = (w + 1/2) 2^-32 once, here,
                // as FP32 bits, instead of at every position that reads it
#pragma unroll
                for (int k = 0; k < 8; k++)
                    uw[c][k] = __float_as_uint(__builtin_fmaf((float)uw[c][k], 2.3283064365386963e-10f,
                                                              1.1641532182693481e-10f));
            }
        }
        float a[C][HPL];
        uint32_t hb[C][V];
        double lq[C];
        // b1 is re-read (cache hits) rather than held in registers for the whole loop
#pragma unroll
        for (int c = 0; c < C; c++) {
            lq[c] = 0.0;
#pragma unroll
            for (int u = 0; u < HPL; u++) a[c][u] = m.b1[16 * u + lane];
#pragma unroll
            for (int v = 0; v < V; v++) hb[c][v] = pack_relu_bf16(a[c][2 * v], a[c][2 * v + 1]);
        }
        // C == 2 sampling: each lane records its own half's configuration
        // (lanes 0-7 configuration 0, lanes 8-15 configuration 1), bits of the
        // current 32-position word in wacc, whole words in own0 / own1
        uint32_t wacc = 0;
        uint64_t own0 = 0, own1 = 0;
        for (int jb = 0; jb < N; jb += 4) {
            uint32_t cur[C][4];
            if (!SCORE) {
                const int b = jb >> 2;
                const int src = row_base | (b & 15);
#pragma unroll
                for (int c = 0; c < C; c++)
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        cur[c][q] = (uint32_t)__shfl((int)((b >> 4) ? uw[c][4 + q] : uw[c][q]), src, 64);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int j = jb + q;
                if (j >= N) break;
                const int D = made_prefix(j, N, m.H);
                // pairs 2t, 2t + 1 (64 sorted units) per branch and LDS read
                const uint2* w2 = reinterpret_cast<const uint2*>(sW2w + (size_t)j * (Hp / 2)) + lane;
                const uint2* w1 = reinterpret_cast<const uint2*>(sW1w + (size_t)j * (Hp / 2)) + lane;
                const float4* w1f = reinterpret_cast<const float4*>(sW1f + (size_t)j * Hp) + lane;
                float s0[C], s1[C];
#pragma unroll
                for (int c = 0; c < C; c++) { s0[c] = 0.0f; s1[c] = 0.0f; }
#pragma unroll
                for (int t = 0; t < V / 2; t++) {
                    if (64 * t < D) {
                        const uint2 w = w2[16 * t];
#pragma unroll
                        for (int c = 0; c < C; c++) {
                            s0[c] = dot2_bf16(w.x, hb[c][2 * t], s0[c]);
                            s1[c] = dot2_bf16(w.y, hb[c][2 * t + 1], s1[c]);
                        }
                    }
                }
                const float b2j = m.b2[j];
                float xf[C];
                if constexpr (C == 2) {
                    // The two configurations share one logit / sampling pass: lanes 0-7
                    // of the row carry configuration 0 and lanes 8-15 configuration 1.
                    // Half sums: a = t0[l] + t0[l + 8] (l < 8) or t1[l] + t1[l - 8], then
                    // half-row mirror and two quad permutations add each half's 8 lanes.
                    const bool lo = lane < 8;
                    const float t0 = s0[0] + s1[0], t1 = s0[1] + s1[1];
                    float a = (lo ? t0 : t1) + dpp<0x128>(lo ? t1 : t0);   // row_ror:8
                    a += dpp<0x141>(a);                                     // row_half_mirror
                    a += dpp<0xB1>(a);                                      // quad_perm 1,0,3,2
                    a += dpp<0x4E>(a);                                      // quad_perm 2,3,0,1
                    const float l = a + b2j;
                    const float e = __expf(-fabsf(l));
                    const float L1 = log_1p_e(e);
                    bool x;
                    if (SCORE) {
                        x = lo ? bit_of(x0[0], x1[0], j) != 0 : bit_of(x0[1], x1[1], j) != 0;
                    } else {
                        const float r = __builtin_amdgcn_rcpf(1.0f + e);
                        const float p = l >= 0.0f ? r : e * r;
                        const float u = __uint_as_float(lo ? cur[0][q] : cur[1][q]);
                        x = u < p;
                    }
                    // log sigmoid(l) = -(max(-l, 0) + log(1 + e^-|l|)), log(1 - sigmoid(l)) = -(max(l, 0) + ...)
                    lq[0] -= (double)((x ? fmaxf(-l, 0.0f) : fmaxf(l, 0.0f)) + L1);
                    // each half holds one decision: swap halves to give every lane both
                    const float xm = x ? 1.0f : 0.0f, xo = dpp<0x128>(xm);
                    xf[0] = lo ? xm : xo;
                    xf[1] = lo ? xo : xm;
                    if (!SCORE) wacc |= x ? 1u << (j & 31) : 0u;
                } else {
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        const float l = row_sum16(s0[c] + s1[c]) + b2j;
                        const float e = __expf(-fabsf(l));
                        const float L1 = log_1p_e(e);
                        bool x;
                        if (SCORE) {
                            x = bit_of(x0[c], x1[c], j) != 0;
                        } else {
                            const float r = __builtin_amdgcn_rcpf(1.0f + e);
                            const float p = l >= 0.0f ? r : e * r;
                            const float u = __uint_as_float(cur[c][q]);
                            x = u < p;
                            const uint64_t bit = x ? (uint64_t)1 << (j & 63) : 0u;
                            if (j < 64) x0[c] |= bit; else x1[c] |= bit;
                        }
                        lq[c] -= (double)((x ? fmaxf(-l, 0.0f) : fmaxf(l, 0.0f)) + L1);
                        xf[c] = x ? 1.0f : 0.0f;
                    }
                }
                // Only the groups that output j + 1 reads ([0, D_{j+1})) are
                // repacked: a group above that bound is packed by a later
                // update before any output reads it, and a group below D_j
                // receives no more updates.
                const int Dn = made_prefix(j + 1, N, m.H);
#pragma unroll
                for (int t = 0; t < V / 2; t++) {
                    if (64 * t + 64 > D) {
                        // the 4 weights of pairs 2t, 2t + 1 of this lane
                        float wv[4];
                        if (W1F) {
                            const float4 w = w1f[16 * t];
                            wv[0] = w.x; wv[1] = w.y; wv[2] = w.z; wv[3] = w.w;
                        } else {
                            const uint2 w = w1[16 * t];
                            wv[0] = bf16_lo(w.x); wv[1] = bf16_hi(w.x); wv[2] = bf16_lo(w.y); wv[3] = bf16_hi(w.y);
                        }
                        const bool repack = 64 * t < Dn;
#pragma unroll
                        for (int b = 0; b < 2; b++) {
                            const int v = 2 * t + b;
                            const float wlo = wv[2 * b], whi = wv[2 * b + 1];
#pragma unroll
                            for (int c = 0; c < C; c++) {
                                a[c][2 * v] = __builtin_fmaf(xf[c], wlo, a[c][2 * v]);
                                a[c][2 * v + 1] = __builtin_fmaf(xf[c], whi, a[c][2 * v + 1]);
                            }
                        }
                        if (repack) {
#pragma unroll
                            for (int b = 0; b < 2; b++)
#pragma unroll
                                for (int c = 0; c < C; c++)
                                    hb[c][2 * t + b] = pack_relu_bf16(a[c][4 * t + 2 * b], a[c][4 * t + 2 * b + 1]);
                        }
                    }
                }
            }
            // positions jb .. jb + 3 share one 32-bit word (jb is a multiple of 4)
            if (C == 2 && !SCORE && (((jb + 4) & 31) == 0 || jb + 4 >= N)) {
                const int wi = jb >> 5;
                if (wi == 0) own0 |= wacc;
                else if (wi == 1) own0 |= (uint64_t)wacc << 32;
                else if (wi == 2) own1 |= wacc;
                else own1 |= (uint64_t)wacc << 32;
                wacc = 0;
            }
        }
        // configuration c is written by lane 8 c (with two configurations per row,
        // lanes 0-7 accumulated configuration 0's log q in lq[0] and its spins in
        // own0 / own1, lanes 8-15 configuration 1's)
#pragma unroll
        for (int c = 0; c < C; c++) {
            if (lane == 8 * c && g0 + c < n) {
                if (!SCORE) {
                    out_spins[2 * (g0 + c)] = C == 2 ? own0 : x0[c];
                    out_spins[2 * (g0 + c) + 1] = C == 2 ? own1 : x1[c];
                }
                log_q[g0 + c] = C == 2 ? lq[0] : lq[c];
            }
        }
    }
}

#define RBM_HPL_LIST(X) X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32)

size_t made_sampler_lds(const MadeInf& m) { return (size_t)2 * m.N * m.Hp * sizeof(uint16_t); }

static hipError_t launch_made_common(const MadeInf& m, int64_t n, uint32_t seed_lo, uint32_t seed_hi, uint32_t call,
                                     const uint64_t* in_spins, uint64_t* out_spins, double* log_q, bool score,
                                     hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (made_sampler_lds(m) > 160 * 1024) return hipErrorInvalidValue;
    // sampling holds W1 as FP32 when it fits (scoring uses the bfloat16 layout;
    // both read the same weight values)
    const size_t lds_f = (size_t)m.N * m.Hp * 6;
    const bool w1f = !score && lds_f <= 160 * 1024;
    const size_t lds = w1f ? lds_f : made_sampler_lds(m);
    const int hpl = made_hidden_per_lane(m.Hp);
    const int rows = 64;
    int64_t blocks = (n + rows * made_cpr(hpl) - 1) / (rows * made_cpr(hpl));
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (blocks > cus) blocks = cus;
    const dim3 grid((unsigned)blocks), block(rows * 16);
#define RBM_MADE_CASE(H)                                                                                           \
    case H:                                                                                                        \
        if (score) {                                                                                               \
            (void)hipFuncSetAttribute((const void*)&k_made_sample<H, true, false>,                                 \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                        \
            hipLaunchKernelGGL((k_made_sample<H, true, false>), grid, block, lds, stream, m, n, seed_lo, seed_hi,    \
                               call, in_spins, out_spins, log_q);                                                   \
        } else if (w1f) {                                                                                          \
            (void)hipFuncSetAttribute((const void*)&k_made_sample<H, false, true>,                                 \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                        \
            hipLaunchKernelGGL((k_made_sample<H, false, true>), grid, block, lds, stream, m, n, seed_lo, seed_hi,   \
                               call, in_spins, out_spins, log_q);                                                   \
        } else {                                                                                                   \
            (void)hipFuncSetAttribute((const void*)&k_made_sample<H, false, false>,                                \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                        \
            hipLaunchKernelGGL((k_made_sample<H, false, false>), grid, block, lds, stream, m, n, seed_lo, seed_hi,  \
                               call, in_spins, out_spins, log_q);                                                   \
        }                                                                                                          \
        break;
    switch (hpl) {
        RBM_HPL_LIST(RBM_MADE_CASE)
        default: return hipErrorInvalidValue;
    }
#undef RBM_MADE_CASE
    return hipGetLastError();
}

hipError_t launch_made_sample(const MadeInf& m, int64_t n, uint32_t seed_lo, uint32_t seed_hi, uint32_t call,
                              uint64_t* spins, double* log_q, hipStream_t stream) {
    return launch_made_common(m, n, seed_lo, seed_hi, call, nullptr, spins, log_q, false, stream);
}

hipError_t launch_made_log_q(const MadeInf& m, int64_t n, const uint64_t* spins, double* log_q, hipStream_t stream) {
    return launch_made_common(m, n, 0, 0, 0, spins, nullptr, log_q, true, stream);
}

// ---------------------------------------------------------------------------
// Independent Metropolis-Hastings, one thread per chain.  A rejected proposal
// repeats the held configuration, so the measured samples of a chain form runs
// of equal configurations.  Pass 0 counts the runs of each chain; pass 1 (with
// run offsets) writes each run's configuration once with its length as the
// multiplicity, and the run of every sample.  Both passes replay the same
// decisions.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_imh(int chains, int therm, int steps, const uint64_t* __restrict__ prop,
                                             const cx<double>* __restrict__ lp, const double* __restrict__ lq,
                                             uint32_t seed_lo, uint32_t seed_hi, uint32_t call, int pass,
                                             uint32_t* __restrict__ runs, const int64_t* __restrict__ run_offset,
                                             uint64_t* __restrict__ run_spins, double* __restrict__ run_mult,
                                             int64_t* __restrict__ sample_run, uint32_t* __restrict__ counts) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= chains) return;
    int64_t cur = c;
    double lpc = lp[cur].re, lqc = lq[cur];
    uint32_t acc = 0, nrun = 0, len = 0;
    const int total = therm + steps;
    const int64_t off = pass ? run_offset[c] : 0;
    for (int t = 1; t <= total; t++) {
        const int64_t g = (int64_t)t * chains + c;
        const double lpg = lp[g].re, lqg = lq[g];
        const double r = 2.0 * (lpg - lpc) + lqc - lqg;
        const u32x4 R = philox((uint32_t)c, (uint32_t)t, call, 2u, seed_lo, seed_hi);
        const double u = ((double)R.x + 0.5) * (1.0 / 4294967296.0);
        const bool accept = u <= exp(r);
        if (accept) {
            cur = g;
            lpc = lpg;
            lqc = lqg;
            acc++;
        }
        if (t > therm) {
            const int k = t - therm - 1;
            if (k == 0 || accept) {
                // a new run starts: close the previous one
                if (pass && nrun > 0) run_mult[off + nrun - 1] = (double)len;
                if (pass) {
                    run_spins[2 * (off + nrun)] = prop[2 * cur];
                    run_spins[2 * (off + nrun) + 1] = prop[2 * cur + 1];
                }
                nrun++;
                len = 0;
            }
            len++;
            if (pass) sample_run[(int64_t)k * chains + c] = off + nrun - 1;
        }
    }
    if (pass) {
        if (nrun > 0) run_mult[off + nrun - 1] = (double)len;
    } else {
        runs[c] = nrun;
        if (counts) {
            counts[2 * c] = acc;
            counts[2 * c + 1] = (uint32_t)total - acc;
        }
    }
}

hipError_t launch_imh(int chains, int therm, int steps, const uint64_t* prop_spins, const cx<double>* log_psi,
                      const double* log_q, uint32_t seed_lo, uint32_t seed_hi, uint32_t call, int pass, uint32_t* runs,
                      const int64_t* run_offset, uint64_t* run_spins, double* run_mult, int64_t* sample_run,
                      uint32_t* counts, hipStream_t stream) {
    hipLaunchKernelGGL(k_imh, dim3((chains + 255) / 256), dim3(256), 0, stream, chains, therm, steps, prop_spins,
                       log_psi, log_q, seed_lo, seed_hi, call, pass, runs, run_offset, run_spins, run_mult, sample_run,
                       counts);
    return hipGetLastError();
}

}  // namespace rbmhip
