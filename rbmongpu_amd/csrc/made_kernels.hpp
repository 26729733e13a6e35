// Masked autoregressive proposal (MADE, Germain et al. 2015) on gfx950:
// FP32 MFMA training kernels and a bfloat16-weight sampler.  The reference
// has no counterpart; the model is defined in the C ABI header.
//
// Device layout (padded to multiples of 64 so every GEMM tile is full):
//   Np = round_up(N, 64), Hp = round_up(H, 64)
//   W1  Hp x Np  FP32 (zero outside the mask and the padding)
//   b1  Hp,  W2  Np x Hp,  b2  Np
// Inference copies: W1t_bf  N x Hp bfloat16 (row i = column i of W1),
//                   W2_bf   N x Hp bfloat16 (row j of W2).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbm_device.hpp"

namespace rbmhip {

struct MadeShape {
    int N, H, Np, Hp;
};

inline int made_pad(int x) { return (x + 63) / 64 * 64; }

// degree of hidden unit k: 1 + k mod (N - 1) (0 for N = 1)
__host__ __device__ inline int made_degree(int k, int N) { return N > 1 ? 1 + k % (N - 1) : 0; }

// Number of hidden units of degree <= j: the units output j reads.  Degree d
// (1 <= d <= N - 1) holds q + [d <= r] units, q = H / (N - 1), r = H mod (N - 1).
__host__ __device__ inline int made_prefix(int j, int N, int H) {
    if (N <= 1 || j >= N - 1) return H;
    if (j <= 0) return 0;
    const int q = H / (N - 1), r = H % (N - 1);
    return j * q + (j < r ? j : r);
}

// Hidden unit at position p of the degree-sorted order (ties by unit index);
// positions >= H are the zero padding units and map to themselves.
__host__ __device__ inline int made_sorted_unit(int p, int N, int H) {
    if (N <= 1 || p >= H) return p;
    const int q = H / (N - 1), r = H % (N - 1);
    int d1, t;   // degree - 1 and rank among the units of that degree
    if (p < r * (q + 1)) { d1 = p / (q + 1); t = p % (q + 1); }
    else { const int p2 = p - r * (q + 1); d1 = r + p2 / q; t = p2 % q; }
    return d1 + t * (N - 1);
}

enum GemmEpi {
    EPI_STORE = 0,       // C = acc
    EPI_BIAS_RELU = 1,   // C = acc + bias[n] (pre-activation), C2 = relu(C)
    EPI_LOGITS = 2,      // l = acc + bias[n]; C = (sigmoid(l) - x) / B on live entries, 0 elsewhere; log-q partials
    EPI_MASK_W2 = 3,     // C (Np x Hp) = acc where deg(n) <= m < N and n < H, else 0
    EPI_MASK_W1 = 4,     // C (Hp x Np) = acc where n < deg(m), m < H and n < N, else 0
    EPI_RELU_GRAD = 5,   // C = acc where aux[m][n] > 0, else 0
};

struct GemmEpiArgs {
    const float* bias;   // EPI_BIAS_RELU, EPI_LOGITS
    float* C2;           // EPI_BIAS_RELU
    const float* aux;    // EPI_LOGITS: X (B x Np); EPI_RELU_GRAD: pre-activation
    long ld_aux;
    int n_live, m_live;  // live columns / rows (N or H, real batch)
    float inv_batch;
    double* partials;    // EPI_LOGITS: one log-q partial per workgroup
    int lower_only;      // compute only the 64 x 64 tiles with n0 <= m0 (symmetric products)
};

// C[m][n] = sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn];
// M, N multiples of 64, K a multiple of 16.
hipError_t launch_gemm(int epi, int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk,
                       long sbn, float* C, long ldc, const GemmEpiArgs& e, hipStream_t stream);

// The same product with K split over `splits` workgroups (K a multiple of
// 16 * splits): partial products go to scratch (splits x M x N floats) and a
// second kernel adds them in order and applies EPI_STORE, EPI_MASK_W1 or
// EPI_MASK_W2.  For the weight gradients, whose M x N is small and K (the batch)
// long.
hipError_t launch_gemm_splitk(int epi, int M, int N, int K, const float* A, long sam, long sak, const float* B,
                              long sbk, long sbn, float* C, long ldc, const GemmEpiArgs& e, int splits, float* scratch,
                              hipStream_t stream);

// X[s][i] = bit i of spins[idx[s]] (or of spins[s] without idx) for i < N, 0 for the padding
// and for rows s >= n_live.
hipError_t launch_made_gather(int64_t rows, int n_live, int N, int Np, const uint64_t* spins, const int64_t* idx,
                              float* X, hipStream_t stream);
// batch indices: idx[s] = floor(u * n) with u from Philox(seed; ctr = (step, s, 0, 4))
hipError_t launch_made_batch_index(int B, int64_t n, uint32_t step, uint32_t seed_lo, uint32_t seed_hi, int64_t* idx,
                                   hipStream_t stream);
// out[n] = sum_m G[m][n] over rows m < rows (ld = cols), fixed order
constexpr int COLSUM_SLICES = 32;   // row slices of launch_colsum; part holds COLSUM_SLICES x cols floats
hipError_t launch_colsum(int rows, int cols, int n_live, const float* G, float* out, float* part, hipStream_t stream);
// p[0 .. n) = 0 (a fill kernel, so the optimiser state can be cleared inside
// captured work)
hipError_t launch_zero(int64_t n, float* p, hipStream_t stream);
// Adam (Kingma & Ba 2015) over n parameters; t is the 1-based step
hipError_t launch_adam(int64_t n, float* p, const float* g, float* m, float* v, int t, float lr, float b1, float b2,
                       float eps, hipStream_t stream);
// bfloat16 inference copies of W1 (transposed) and W2 in the sampler's layout:
// hidden units in degree-sorted order (made_sorted_unit), row j of either
// matrix as Hp / 2 words, word 32 t + 2 l + b holding pair v = 2 t + b of
// lane l: sorted units 32 v + l (low half) and 32 v + 16 + l (high half);
// b1 permuted to the sorted order.
hipError_t launch_made_inference_weights(const MadeShape& s, const float* W1, const float* b1, const float* W2,
                                         uint16_t* W1t_bf, uint16_t* W2_bf, float* b1_sorted, hipStream_t stream);

struct MadeInf {
    int N, H, Hp;
    const uint16_t* W1t;   // N x Hp, sampler layout
    const uint16_t* W2;    // N x Hp, sampler layout
    const float* b1;       // Hp, degree-sorted (zero padded)
    const float* b2;       // >= N
};
size_t made_sampler_lds(const MadeInf& m);
int made_hidden_per_lane(int Hp);
// Autoregressive sampling of n configurations: proposal g draws its uniforms
// from Philox(seed; ctr = (g_lo, g_hi, call, 0x100 + j / 4)), word j % 4,
// u = (w + 0.5) 2^-32 in FP32, x_j = 1 iff u < sigmoid(l_j).
hipError_t launch_made_sample(const MadeInf& m, int64_t n, uint32_t seed_lo, uint32_t seed_hi, uint32_t call,
                              uint64_t* spins, double* log_q, hipStream_t stream);
// log q of given configurations with the sampler's arithmetic
hipError_t launch_made_log_q(const MadeInf& m, int64_t n, const uint64_t* spins, double* log_q, hipStream_t stream);

// Independent Metropolis-Hastings over precomputed proposals: chain c uses
// proposals t * chains + c for t = 0 .. therm + steps; proposal 0 starts the
// chain and t >= 1 is accepted iff u <= exp(2 (Re lp' - Re lp) + lq - lq'),
// u from Philox(seed; ctr = (c, t, call, 2)) in FP64.  Measured sample k of
// chain c is the configuration held after proposal therm + 1 + k.
// pass 0: runs[c] = number of runs of equal consecutive samples, counts[2c],
//         counts[2c + 1] = acceptances, rejections;
// pass 1: with run_offset[c] (exclusive prefix of runs), run r of chain c is
//         written at run_offset[c] + r: its configuration and its length
//         (run_mult); sample_run[k * chains + c] = the run holding sample k.
hipError_t launch_imh(int chains, int therm, int steps, const uint64_t* prop_spins, const cx<double>* log_psi,
                      const double* log_q, uint32_t seed_lo, uint32_t seed_hi, uint32_t call, int pass, uint32_t* runs,
                      const int64_t* run_offset, uint64_t* run_spins, double* run_mult, int64_t* sample_run,
                      uint32_t* counts, hipStream_t stream);

}  // namespace rbmhip
