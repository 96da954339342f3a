// Launchers of the batched persistent kernel (k_batched.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sep_kernels.hpp"

namespace dipamd {

hipError_t launch_batched(InType ti, bool is_double, int cfg, bool fast, const void *args, int grid, size_t lds,
                          hipStream_t stream);
int batched_occupancy(InType ti, bool is_double, int cfg, bool fast, size_t lds);
// The radius-class batched kernel (fused_rc.hpp) for x radius class rxc (4..24); with
// occ non-null it only reports the workgroups per CU for `lds` bytes of dynamic LDS.
hipError_t launch_batched_rc(InType ti, bool is_double, int cfg, bool fast, int rxc, const void *args, int grid,
                             size_t lds, hipStream_t stream, int *occ);
// Clears `n` consecutive work-queue counters with a kernel (no memset node under capture).
hipError_t launch_reset_counter(int32_t *counter, int n, hipStream_t stream);

}  // namespace dipamd
