// FP64 complex Cholesky factorisation and solve (row-major n x n; the factor
// overwrites the lower triangle).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "rbm_device.hpp"

namespace rbmhip {

// A = L L^H in place (lower triangle); *info (zeroed by the caller) receives
// the 1-based index of the first non-positive pivot, 0 if none.
hipError_t launch_cholesky(int n, cx<double>* A, int* info, hipStream_t stream);
// x <- (L L^H)^{-1} x with the factor from launch_cholesky
hipError_t launch_cholesky_solve(int n, const cx<double>* L, cx<double>* x, hipStream_t stream);
// C += A^T A on the lower 128 x 128 tiles of the FP64 matrix C (ld x ld) for
// a K x ld FP32 matrix A given split into three bfloat16 parts per 32-row
// block ([block][part][column][32 rows], as launch_o_split3 writes it), the
// product formed on the bfloat16 matrix cores with FP32 accuracy; ld a
// multiple of 128, K of 64.  Tiles above the diagonal are not touched.
// order (n_order entries, from gram_tile_order) assigns the tiles to
// workgroups.
hipError_t launch_gram_lower(int ld, int K, const uint32_t* A, double* C, const int2* order, int n_order,
                             hipStream_t stream);
// Fills order with the workgroup-to-tile map of launch_gram_lower for ld
// (entries (bm, bn) with bn <= bm, or (-1, -1)); returns the number of tiles.
int gram_tile_order(int ld, std::vector<int2>& order);
// out[z] = X_z^T Y_z (ldx x ldy) for the row ranges z = 0 .. splits - 1 of
// length K / splits of row-major K x ldx and K x ldy operands (same MFMA
// tiles); ldx and ldy multiples of 128, K a multiple of 16 splits.
hipError_t launch_gemm_tn_splitk(int ldx, int ldy, int64_t K, int splits, const float* X, const float* Y, float* out,
                                 hipStream_t stream);
// acc[i] += sum over z of part[z n + i] (FP64, fixed order)
hipError_t launch_accum_splits_f64(int64_t n, int splits, const float* part, double* acc, hipStream_t stream);

}  // namespace rbmhip
