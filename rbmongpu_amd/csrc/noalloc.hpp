// Scratch of the allocation-free entry point: a fixed device array per GPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace rbmhip {

// 16 MiB per device: the model and operator tables (at most 4 MiB) and three
// records of 16 bytes per configuration (spins, log psi, E_loc) for at most
// NOALLOC_MAX_CONFIGS configurations.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
constexpr size_t NOALLOC_TABLE_BYTES = (size_t)4 << 20;
constexpr int64_t NOALLOC_MAX_CONFIGS = (int64_t)1 << 18;

// Device address of the scratch array on the current device (once per device).
hipError_t noalloc_scratch_address(void** ptr);

}  // namespace rbmhip
