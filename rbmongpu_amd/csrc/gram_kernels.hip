// Gram product of the S-matrix rows on the bfloat16 matrix cores (the
// centred S of a stochastic-reconfiguration step; the reference only builds
// S, S_matrix.cu:10-58).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "linalg_kernels.hpp"

namespace rbmhip {

// ---------------------------------------------------------------------------
// Gram product G = A^T A of the S-matrix rows (lower 128 x 128 tiles) on the
// bfloat16 matrix cores with FP32 accuracy.  Every FP32 value of A arrives
// split exactly into three bfloat16 parts, x = hi + mid + lo (hi = bf16(x),
// mid = bf16(x - hi), lo = x - hi - mid), stored per block of 32 samples as
// [block][part][column][32 samples].  A product x y is the sum of the six part
// products hi hi, hi mid, mid hi, hi lo, lo hi, mid mid, which leaves out
// terms below 2^-24 |x y|, and v_mfma_f32_16x16x32_bf16 sums them in FP32
// like an FP32 GEMM; six bfloat16 products cost 6 / 16 of one FP32 product on
// the matrix cores.  The batch's FP32 tile is added into the FP64 accumulator
// C, so the FP32 error does not grow with the number of batches.
//
// A 128-thread workgroup owns the lower tile (bm, bn) that order[blockIdx.x]
// names (gram_tile_order below), as two waves of 128 x 64 (8 x 4 MFMA tiles; lane l supplies A[l % 16][8 (l /
// 16) + q] and B[8 (l / 16) + q][l % 16], accumulator v of lane l is row 4 (l /
// 16) + v, column l % 16).  The stored layout is the MFMA operand layout: the
// fragment of lane l for a column tile is the 16 bytes of column c0 + l % 16 at
// samples 8 (l / 16) .. + 7, so every fragment is one coalesced 16-byte load
// and no LDS or barrier is needed.  The B fragments of the next block load
// into a second register set at the start of a block and are moved into place
// at its end; the A fragments of row tile i are reloaded for the next block as
// soon as row tile i's 24 MFMAs of this block are issued.  With 128 FP32
// accumulators the wave runs alone on its SIMD.
// Measured against other forms on the 6x6 Heisenberg SR step (21 batches of
// 12 480 samples and one of 64, ld 10 752), in the row-major tile order:
// 150 ms of Gram product per step; staging both column stripes through LDS
// for 2 x 2 waves of 64 x 64, 165-171 ms; 64 x 64 per wave loaded straight
// from memory (two waves per SIMD), 244 ms.  The grouped order below: 147 ms.
// ---------------------------------------------------------------------------
constexpr int GM_T = 128, GB_K = 32;
typedef float gbf32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 gbbf16x8 __attribute__((ext_vector_type(8)));

constexpr int RT = 8;   // row tiles per wave

__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_gram_lower_bf3(int ld, int K, const uint4* __restrict__ A, double* __restrict__ C,
                      const int2* __restrict__ order) {
    const int2 tile = order[blockIdx.x];
    if (tile.x < 0) return;
    const int bm = tile.x, bn = tile.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = bm * GM_T, n0 = bn * GM_T + 64 * wave;
    const int nblk = K / GB_K;
    // uint4 index of (block, part, column, chunk): ((blk * 3 + q) * ld + col) * 4 + chunk
    const uint4* pa = A + (int64_t)(m0 + (lane & 15)) * 4 + (lane >> 4);
    const uint4* pb = A + (int64_t)(n0 + (lane & 15)) * 4 + (lane >> 4);
    const int64_t part_stride = (int64_t)ld * 4, blk_stride = 3 * part_stride;
    uint4 fa[RT][3], fb[4][3], fn[4][3];
    gbf32x4 acc[RT][4];
#pragma unroll
    for (int i = 0; i < RT; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = gbf32x4{0.f, 0.f, 0.f, 0.f};
    // first loads in the order the loop issues them (B, then A by row tile),
    // so that the waits at the loop head hold for the first block too
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int j = 0; j < 4; j++) fb[j][q] = pb[q * part_stride + 64 * j];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < RT; i++) {
#pragma unroll
        for (int q = 0; q < 3; q++) fa[i][q] = pa[q * part_stride + 64 * i];
        __builtin_amdgcn_sched_barrier(0);
    }
    for (int blk = 0; blk < nblk; blk++) {
        // the last block reloads itself, so that no load is conditional
        const int64_t o = (int64_t)(blk + 1 < nblk ? blk + 1 : blk) * blk_stride;
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int j = 0; j < 4; j++) fn[j][q] = pb[o + q * part_stride + 64 * j];
        // the loads stay where they are issued (the scheduler would move them
        // next to their first use)
        __builtin_amdgcn_sched_barrier(0);
        constexpr int PA[6] = {1, 0, 2, 0, 1, 0}, PB[6] = {1, 2, 0, 1, 0, 0};
#pragma unroll
        for (int i = 0; i < RT; i++) {
#pragma unroll
            for (int p = 0; p < 6; p++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(gbbf16x8, fa[i][PA[p]]),
                        __builtin_bit_cast(gbbf16x8, fb[j][PB[p]]), acc[i][j], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 3; q++) fa[i][q] = pa[o + q * part_stride + 64 * i];
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 3; q++) fb[j][q] = fn[j][q];
    }
#pragma unroll
    for (int i = 0; i < RT; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = m0 + 16 * i + 4 * (lane >> 4) + v;
                const int col = n0 + 16 * j + (lane & 15);
                C[(int64_t)row * ld + col] += (double)acc[i][j][v];
            }
}

// Workgroup order of the lower 128 x 128 tiles.  Workgroup b runs on XCD b mod
// 8, after the earlier workgroups of that XCD, and an XCD holds 64 of these
// workgroups at once (two per CU).  The tiles are listed by 8 x 8 groups (the
// lower triangle of groups row by row, the tiles of a group row by row) and
// XCD x takes the x-th eighth of the list in order, so the 64 workgroups it
// runs at once mostly form one group and read 16 column stripes between them,
// which its L2 holds, instead of up to 128.  Entries past the list are (-1, -1).
int gram_tile_order(int ld, std::vector<int2>& order) {
    const int T = ld / GM_T, G = 8;
    std::vector<int2> list;
    list.reserve((size_t)T * (T + 1) / 2);
    for (int gm = 0; gm * G < T; gm++)
        for (int gn = 0; gn <= gm; gn++)
            for (int bm = gm * G; bm < std::min(T, gm * G + G); bm++)
                for (int bn = gn * G; bn < std::min(bm + 1, gn * G + G); bn++) list.push_back(make_int2(bm, bn));
    const int n = (int)list.size(), per = (n + 7) / 8;
    order.assign((size_t)8 * per, make_int2(-1, -1));
    for (int b = 0; b < 8 * per; b++) {
        const int k = (b % 8) * per + b / 8;
        if (k < n) order[(size_t)b] = list[(size_t)k];
    }
    return n;
}

hipError_t launch_gram_lower(int ld, int K, const uint32_t* A, double* C, const int2* order, int n_order,
                             hipStream_t stream) {
    if (ld <= 0 || ld % GM_T || K <= 0 || K % (2 * GB_K) || n_order <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gram_lower_bf3, dim3((unsigned)n_order), dim3(128), 0, stream, ld, K,
                       reinterpret_cast<const uint4*>(A), C, order);
    return hipGetLastError();
}

}  // namespace rbmhip
