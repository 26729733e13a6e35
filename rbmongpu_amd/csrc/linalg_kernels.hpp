// FP64 complex Cholesky factorisation and solve (row-major n x n; the factor
// overwrites the lower triangle).
#pragma once

#include <hip/hip_runtime.h>

#include "rbm_device.hpp"

namespace rbmhip {

// A = L L^H in place (lower triangle); *info (zeroed by the caller) receives
// the 1-based index of the first non-positive pivot, 0 if none.
hipError_t launch_cholesky(int n, cx<double>* A, int* info, hipStream_t stream);
// x <- (L L^H)^{-1} x with the factor from launch_cholesky
hipError_t launch_cholesky_solve(int n, const cx<double>* L, cx<double>* x, hipStream_t stream);

}  // namespace rbmhip
