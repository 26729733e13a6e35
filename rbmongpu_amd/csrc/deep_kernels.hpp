// Device view and launchers of the PsiDeep kernels (deep_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbm_device.hpp"
#include "rbm_kernels.hpp"

namespace rbmhip {

constexpr int DEEP_MAX_LAYERS = 4;
constexpr int DEEP_MAX_WIDTH = 256;
constexpr int DEEP_MAX_UNITS = 512;

// Layer l: size[l] units, conn[l] lhs connections per unit; its C x size
// connection and weight tables start at w_off[l] of conns / weights, its biases
// at b_off[l].  The transpose (rhs) of layer l+1's connections, used by the
// backward pass, lists for unit i of layer l the weight slots of layer l+1
// reading it: rhs_entries[rhs_begin[rb_off[l] + i] .. rhs_begin[rb_off[l] + i + 1])
// with entry = c * size[l+1] + j.
template <typename T>
struct DeepDev {
    int N;
    int num_layers;
    int width;   // max(N, sizes)
    int units;   // sum of sizes
    int size[DEEP_MAX_LAYERS];
    int conn[DEEP_MAX_LAYERS];
    int64_t w_off[DEEP_MAX_LAYERS];
    int64_t b_off[DEEP_MAX_LAYERS];
    int rb_off[DEEP_MAX_LAYERS];
    const uint32_t* conns;
    const cx<T>* weights;
    const cx<T>* biases;
    const cx<T>* final_w;
    const int* rhs_begin;
    const int* rhs_entries;
    int translational_invariance;
    int ref_compat_final;
};

template <typename T>
hipError_t launch_deep_eval(const DeepDev<T>& md, const OpDev& op, int has_op, int64_t n, const uint64_t* configs,
                            uint64_t index_base, const Records& rec, int max_blocks, hipStream_t stream);
template <typename T>
hipError_t launch_deep_mc(const DeepDev<T>& md, const OpDev& op, int has_op, const McParams& mp, const Records& rec,
                          uint32_t* counts, hipStream_t stream);
// R[q][k] = sum_s coef_q[s] conj(O[s][k]) over materialised O rows (K per sample)
template <typename T>
hipError_t launch_rows_reduce(int64_t n, int64_t K, const void* O, const cx<double>* coef, int nc, int64_t chunk,
                              int nblocks, cx<double>* partials, hipStream_t stream);
// PsiDeep as dense BF16 MFMA layers (deep_mfma.hip).  Layer l has in_p[l]
// inputs (padded to 32; layer 0's inputs are the spins) and out_p[l] units
// (padded to 32, the last layer to 16).  Its weights are bfloat16 fragments
// at frag[frag_off[l] ...]: block (mt, kb) of 16 units x 32 inputs, part 0
// (real) and 1 (imaginary), is 64 lanes x 8 values at
// frag_off[l] + ((mt * KB + kb) * 2 + part) * 64 + lane, where lane holds unit
// mt*16 + lane%16 and the inputs of K block kb in the order
//   layer 0:  kb*32 + 8*(lane/16) + v
//   layer >0: kb*32 + 4*(lane/16) + v (v < 4),  kb*32 + 16 + 4*(lane/16) + v-4 (v >= 4).
// vec holds per layer the FP32 biases (real parts at vec_off[l], imaginary
// parts out_p[l] later) and the final weights (final_off, same split).
//
// The backward pass (O_k by back-propagation, PsiDeep.hpp:214-293) computes
// delta_l = (W_{l+1} delta_{l+1}) * my_tanh(a_l) with fragments of W_{l+1}
// itself at bfrag_off[l] (l < L-1): block (mt, kb) of 16 units of layer l x 32
// units of layer l+1, in the "layer > 0" K order and KB = ceil(out_p[l+1]/32).
// Gradient records (one row of rec_stride complex per configuration): x_l at
// rec_x[l], delta_l at rec_d[l] (out_p[l] entries each) and log psi at rec_lp.
constexpr int DM_MAX_KB = 8;   // at most 256 inputs or units per layer
struct DeepMfmaDev {
    int N, L;
    int in_p[DEEP_MAX_LAYERS];
    int out_p[DEEP_MAX_LAYERS];
    int frag_off[DEEP_MAX_LAYERS];   // in 16-byte units
    int bfrag_off[DEEP_MAX_LAYERS];  // in 16-byte units
    int vec_off[DEEP_MAX_LAYERS];    // in floats
    int final_off;
    int fwd_total;                   // 16-byte units of the forward fragments
    int frag_total;                  // forward + backward fragments
    int vec_total;                   // floats (a multiple of 4)
    int rec_x[DEEP_MAX_LAYERS], rec_d[DEEP_MAX_LAYERS], rec_lp, rec_stride;
    const uint4* frag;
    const float* vec;
};

// LDS bytes of the forward kernels (grad = false) or the gradient kernel
size_t deep_mfma_lds_bytes(const DeepMfmaDev& md, bool grad = false);
// Gradient records of n configurations (rec_stride complex each)
hipError_t launch_deep_mfma_grad(const DeepMfmaDev& md, int64_t n, const uint64_t* configs, cx<float>* records,
                                 int num_cus, hipStream_t stream);
// Real rows of one contraction GEMM for samples [s0, s0 + cnt) of the records:
// X (rows x ldx) and Y (rows x ldy), rows = a multiple of 64 (zero padded).
// layer < L: X = [Re x_{l-1}, Im x_{l-1}, 1] (spins for layer 0),
//            Y = [Re, Im] of conj(coef) delta_l;
// layer = L: X = [Re x_{L-1}, Im x_{L-1}], Y = [Re, Im] of conj(coef) (log psi if lp_factor).
hipError_t launch_deep_grad_rows(const DeepMfmaDev& md, int layer, int in, int out, int64_t s0, int64_t cnt, int64_t rows,
                                 const uint64_t* configs, const cx<float>* records, const cx<double>* coef,
                                 int lp_factor, float* X, int ldx, float* Y, int ldy, hipStream_t stream);
hipError_t launch_deep_mfma_eval(const DeepMfmaDev& md, const OpDev& op, int has_op, int64_t n, const uint64_t* configs,
                                 uint64_t index_base, const Records& rec, int num_cus, hipStream_t stream);
hipError_t launch_deep_mfma_mc(const DeepMfmaDev& md, const McParams& mp, uint64_t* spins, uint32_t* counts, int num_cus,
                               hipStream_t stream);
hipError_t launch_exp_vector(int64_t n, const cx<double>* lp, double log_prefactor, cx<double>* out, hipStream_t stream);

}  // namespace rbmhip
