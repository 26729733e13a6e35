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
constexpr int DM_MAX_KB = 8;   // at most 256 inputs or units per layer
struct DeepMfmaDev {
    int N, L;
    int in_p[DEEP_MAX_LAYERS];
    int out_p[DEEP_MAX_LAYERS];
    int frag_off[DEEP_MAX_LAYERS];   // in 16-byte units
    int vec_off[DEEP_MAX_LAYERS];    // in floats
    int final_off;
    int frag_total;                  // 16-byte units
    int vec_total;                   // floats
    const uint4* frag;
    const float* vec;
};

size_t deep_mfma_lds_bytes(const DeepMfmaDev& md);
hipError_t launch_deep_mfma_eval(const DeepMfmaDev& md, const OpDev& op, int has_op, int64_t n, const uint64_t* configs,
                                 uint64_t index_base, const Records& rec, int num_cus, hipStream_t stream);
hipError_t launch_deep_mfma_mc(const DeepMfmaDev& md, const McParams& mp, uint64_t* spins, uint32_t* counts, int num_cus,
                               hipStream_t stream);
hipError_t launch_exp_vector(int64_t n, const cx<double>* lp, double log_prefactor, cx<double>* out, hipStream_t stream);

}  // namespace rbmhip
