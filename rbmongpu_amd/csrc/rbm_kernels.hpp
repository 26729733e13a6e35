// Kernel-side data views and host launchers shared by rbm_kernels.hip and
// the C-ABI host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbm_device.hpp"

namespace rbmhip {

// Device view of an RBM: W row-major N x ld with ld = rbm_padded_width(M)
// (columns M..ld-1 are zero), b (ld, zero padded), optional visible bias a (N).
template <typename T>
struct RbmDev {
    int N, M, ld;
    const cx<T>* W;
    const cx<T>* b;
    const cx<T>* a;
};

// Per-sample outputs; any pointer may be null.
struct Records {
    uint64_t* spins;      // 2 words per sample
    cx<double>* log_psi;  // per sample
    cx<double>* e_loc;    // per sample
    void* tanh;           // cx<T> per sample x M
};

// move: MC_SINGLE_FLIP, MC_EXCHANGE (total S_z conserved; k_up spins up) or
// MC_FAST_SWEEP (tries random masks ANDed; one proposal per sweep)
enum { MC_SINGLE_FLIP = 0, MC_EXCHANGE = 1, MC_FAST_SWEEP = 2 };
struct McParams {
    int chains;
    int steps_per_chain;
    int sweeps;
    int therm;
    uint32_t seed_lo, seed_hi, call;
    int move = MC_SINGLE_FLIP;
    int tries = 1;
    int k_up = 0;
    // RBM samplers only:
    //   start      chain start configurations (2 words per chain) instead of
    //              the Philox random start (persistent chains), or null;
    //   final_out  where each chain leaves its last configuration, or null
    //              (may alias start: a chain reads its slot before writing it);
    //   angles_out each chain's final angles (chains x M complex of the
    //              kernel's type), or null;
    //   resync     FP32 chains recompute theta = b + W^T s from the
    //              configuration at the first draw-block end after every
    //              `resync` proposals (0: never), which bounds the drift of
    //              the incremental updates.
    const uint64_t* start = nullptr;
    uint64_t* final_out = nullptr;
    void* angles_out = nullptr;
    int resync = 0;
    //   k0         index of the chain's first proposal in this launch (single-flip
    //              RBM chains): a chain continued from `start` at k0 draws what
    //              one launch would draw from proposal k0 on.  A multiple of
    //              MC_DRAW_BLOCK and, with resync, of resync.
    uint64_t k0 = 0;
};

// proposals per Philox draw block of the single-flip RBM sampler
constexpr int MC_DRAW_BLOCK = 32;

// FP32 Metropolis chains recompute their angles from the configuration every
// MC_RESYNC_SWEEPS sweeps (McParams::resync = MC_RESYNC_SWEEPS * N proposals)
constexpr int MC_RESYNC_SWEEPS = 16;

constexpr int SC_STRIDE = 4;
constexpr int OK_ROWS_PER_THREAD = 8;

struct CoefArgs {
    int64_t n;
    int mode;            // 0 energy, 1 Hilbert-space distance, 2 centred energy, 3 centred distance, 4 weight,
                         // 5 affine w (alpha E + e_mean), 6 corrected, 7 difference, 8 probability ratio
    int exact;           // ExactSummation weights
    const double* mult;  // per-configuration multiplicity weights (runs of equal samples), may be null
    int is_unitary;
    double num_configs;
    double log_prefactor;
    const cx<double>* lp;
    const cx<double>* el;
    const cx<double>* lp2;
    const cx<double>* el2;  // modes 7: local energy of the second wavefunction
    cx<double> alpha;       // mode 5
    cx<double>* coef;    // n (modes 2, 3), may be null
    double* sc;          // SC_STRIDE x n (modes 0, 1)
    cx<double> e_mean;   // mode 2: sum w E / n
    cx<double> omega_mean;  // mode 3
    double ratio_mean;      // mode 3
};

int max_hidden_units();
// hidden units rounded up to 16 lanes x the compiled units-per-lane count
int rbm_padded_width(int M);

// Metropolis chains with the exchange or fast-sweep move (rbm_kernels.hip)
template <typename T>
hipError_t launch_mc_moves(const RbmDev<T>& md, const McParams& mp, uint64_t* spins, uint32_t* counts, hipStream_t stream);
template <typename T>
hipError_t launch_mc(const RbmDev<T>& md, const McParams& mp, uint64_t* spins, uint32_t* counts, bool wlds,
                     size_t lds_bytes, hipStream_t stream);
template <typename T>
hipError_t launch_eval(const RbmDev<T>& md, const OpDev& op, int has_op, int64_t n, const uint64_t* configs,
                       uint64_t index_base, const Records& rec, bool wlds, size_t lds_bytes, int max_blocks,
                       hipStream_t stream);
// lp2 = lp + log psi'(s) - log psi(s) for two RBMs of one shape, from the
// angle difference (dW = W' - W as N x ld, db = b' - b as ld, both cx<T>;
// da = a' - a as N complex128 or null)
template <typename T>
hipError_t launch_rbm_log_ratio(const RbmDev<T>& md, const void* dW, const void* db, const cx<double>* da, int64_t n,
                                const uint64_t* configs, uint64_t index_base, const cx<double>* lp, cx<double>* lp2,
                                int max_blocks, hipStream_t stream);
template <typename T>
hipError_t launch_ok_reduce(int64_t n, int N, int M, const void* Trec, const uint64_t* spins, const cx<double>* coef,
                            int nc, int64_t chunk, int nblocks, cx<double>* partials, hipStream_t stream);
// FP32 MFMA contraction (complex64 records): P[m][n] = sum_s A_m(s) B_n(s)
// with m = 2r + part (rows {1, s_i}, Re/Im of the coefficient) and
// n = 2j + part (columns {T_j, 1}, Re/Im); per-chunk FP32 partials.
// padded row / column counts of the MFMA contraction for an N x M RBM
void ok_mfma_shape(int N, int M, int* rows_p, int* cols_p);
constexpr int MFMA_SUM_SPLIT = 16;   // FP64 partial-sum groups (Psum holds MFMA_SUM_SPLIT x rows_p x cols_p)
// N >= 32 uses the bfloat16 kernel, which reads the spins as 32-sample
// bit-planes: planes must hold ceil(n / 32) x N uint32 and chunk be a multiple
// of 32; smaller N ignore planes (may be null).
bool ok_mfma_needs_planes(int N);
hipError_t launch_ok_mfma(int64_t n, int N, int M, const void* Trec, const uint64_t* spins, const cx<double>* coef,
                          int64_t chunk, int nchunks, int rows_p, int cols_p, float* partials, uint32_t* planes,
                          hipStream_t stream);
// FP64 sum of FP32 partials over chunks, then R[r][j] (complex) from P.
hipError_t launch_mfma_finish(int nchunks, int N, int M, int rows_p, int cols_p, const float* partials, double* Psum,
                              cx<double>* R, hipStream_t stream);
hipError_t launch_partials_sum(int nblocks, int64_t len, const cx<double>* partials, cx<double>* out, hipStream_t stream);
hipError_t launch_coef(const CoefArgs& a, hipStream_t stream);
// E_k estimator: virtual configurations of samples [i0, i0 + cnt) and their
// contraction coefficients (rbm_kernels.hip)
hipError_t launch_ek_expand(const OpDev& op, int64_t i0, int64_t cnt, const uint64_t* spins, uint64_t* vspins,
                            cx<double>* vfac, hipStream_t stream);
hipError_t launch_ek_coef(int64_t i0, int64_t cnt, int T1, const cx<double>* w, const cx<double>* lp, const cx<double>* el,
                          cx<double> gamma, cx<double> delta, const cx<double>* vlp, const cx<double>* vfac,
                          cx<double>* vcoef, hipStream_t stream);
// Free quantum axis of HilbertSpaceDistance: configurations with one spin
// flipped, the first-order axis term of log psi' (flip ratios kept in r, flip
// i at r + i n), and the alpha / beta rows of sum_s coef conj(O_k) as
// slices x 2N partials (rbm_kernels.hip).
hipError_t launch_flip_configs(int64_t n, const uint64_t* configs, uint64_t index_base, int i0, int cnt, uint64_t* out,
                               hipStream_t stream);
hipError_t launch_axis_log_psi(int64_t n, int N, const uint64_t* configs, uint64_t index_base, const double* axis,
                               cx<double>* lp, cx<double>* r, hipStream_t stream);
hipError_t launch_axis_contract(int64_t n, int N, const uint64_t* configs, uint64_t index_base, const double* axis,
                                const cx<double>* r, const cx<double>* coef, int slices, cx<double>* partials,
                                hipStream_t stream);
hipError_t launch_scalar_partials(int64_t n, const double* sc, int nblocks, double* partials, hipStream_t stream);
hipError_t launch_lse_partials(int64_t n, const cx<double>* lp, double log_prefactor, int nblocks, double* partials,
                               hipStream_t stream);
template <typename T>
hipError_t launch_s_matrix_ref(int64_t n, int N, int M, int has_a, const void* Trec, const uint64_t* spins, double inv,
                               cx<double>* S, hipStream_t stream);
hipError_t launch_activation(int64_t n, const cx<double>* z, cx<double>* out, int deriv, hipStream_t stream);
// Centred S-matrix pieces: realified, weighted, centred O rows of samples
// [s0, s0 + nb) (rows nb .. nbp zero; columns [Re (P) | Im (P) | 0 .. ld2)),
// split into three bfloat16 parts and stored per 32-sample block as
// [block][part][column][32 samples] (nbp a multiple of 32; out holds
// nbp * ld2 * 3 bfloat16); assembly of S (P x P complex, scaled by inv, eps
// added on the diagonal) from the lower triangle of the FP64 Gram matrix.
template <typename T>
hipError_t launch_o_split3(int64_t s0, int64_t nb, int64_t nbp, int N, int M, int P, int ld2, const void* Trec,
                           const uint64_t* spins, const cx<double>* w, const cx<float>* obar, uint32_t* out,
                           hipStream_t stream);
hipError_t launch_s_assemble(int P, int ld2, double inv, double eps, const double* G, cx<double>* S, hipStream_t stream);
// log psi of n configurations of an FP32 RBM without visible bias (N <= 128):
// theta = b + W^T s on bfloat16 MFMA with W split exactly into three bfloat16
// parts (exact products, FP32 sums), my_logcosh summed in the epilogue.
hipError_t launch_theta_logpsi(const RbmDev<float>& md, int64_t n, const uint64_t* spins, cx<double>* lp,
                               hipStream_t stream);

}  // namespace rbmhip
