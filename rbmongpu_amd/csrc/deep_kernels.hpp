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
hipError_t launch_exp_vector(int64_t n, const cx<double>* lp, double log_prefactor, cx<double>* out, hipStream_t stream);

}  // namespace rbmhip
