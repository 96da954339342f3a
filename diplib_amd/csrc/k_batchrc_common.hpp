// Launch helpers of the radius-class batched kernel (fused_rc.hpp), included by the
// per-type translation units k_batchrc_*.hip.
#pragma once
#include "fused_rc.hpp"

namespace dipamd {
namespace {

template <typename TI, typename TB, int TXC, int P, bool FAST, int RXC>
hipError_t batched_rc_t(const void *args, int grid, size_t lds, hipStream_t st, int *occ) {
    if (occ) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batched_rc<TI, TB, TXC, P, FAST, RXC>, kBlock, lds) !=
            hipSuccess)
            n = 1;
        *occ = n > 0 ? n : 1;
        return hipSuccess;
    }
    const auto &a = *static_cast<const BatchArgs<TB> *>(args);
    hipLaunchKernelGGL((k_batched_rc<TI, TB, TXC, P, FAST, RXC>), dim3(static_cast<unsigned>(grid)), dim3(kBlock), lds,
                       st, a);
    return hipGetLastError();
}

template <typename TI, typename TB, int TXC, int P, bool FAST>
hipError_t batched_rc_class(int rxc, const void *args, int grid, size_t lds, hipStream_t st, int *occ) {
    switch (rxc) {
        case 4: return batched_rc_t<TI, TB, TXC, P, FAST, 4>(args, grid, lds, st, occ);
        case 8: return batched_rc_t<TI, TB, TXC, P, FAST, 8>(args, grid, lds, st, occ);
        case 12: return batched_rc_t<TI, TB, TXC, P, FAST, 12>(args, grid, lds, st, occ);
        case 16: return batched_rc_t<TI, TB, TXC, P, FAST, 16>(args, grid, lds, st, occ);
        case 20: return batched_rc_t<TI, TB, TXC, P, FAST, 20>(args, grid, lds, st, occ);
        case 24: return batched_rc_t<TI, TB, TXC, P, FAST, 24>(args, grid, lds, st, occ);
    }
    return hipErrorNotSupported;
}

template <typename TI, typename TB, bool FAST>
hipError_t batched_rc_cfg(int cfg, int rxc, const void *args, int grid, size_t lds, hipStream_t st, int *occ) {
    switch (cfg) {
        case 0: return batched_rc_class<TI, TB, 64, 4, FAST>(rxc, args, grid, lds, st, occ);
        case 1: return batched_rc_class<TI, TB, 128, 8, FAST>(rxc, args, grid, lds, st, occ);
        case 2:
            if constexpr (sizeof(TB) == 4) return batched_rc_class<TI, TB, 256, 16, FAST>(rxc, args, grid, lds, st, occ);
            return hipErrorNotSupported;
    }
    return hipErrorNotSupported;
}

}  // namespace
}  // namespace dipamd
