// Batched persistent kernel instantiations and launchers.
#include "batch_kernels.hpp"
#include "sep_kernels_impl.hpp"

namespace dipamd {

// Clears the work-queue counters with a kernel (graph-safe: no memset node).
__global__ void k_reset_counter(int32_t *counter, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) counter[i] = 0;
}

namespace {

template <typename TI, typename TB, int TXC, int P, bool FAST>
hipError_t batched_t(const void *args, int grid, size_t lds, hipStream_t st) {
    const auto &a = *static_cast<const BatchArgs<TB> *>(args);
    hipLaunchKernelGGL((k_batched<TI, TB, TXC, P, FAST>), dim3(static_cast<unsigned>(grid)), dim3(kBlock), lds, st, a);
    return hipGetLastError();
}

template <typename TI, typename TB, int TXC, int P, bool FAST>
int occupancy_t(size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batched<TI, TB, TXC, P, FAST>, kBlock, lds) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

template <typename TI, typename TB, bool FAST>
hipError_t batched_cfg(int cfg, const void *args, int grid, size_t lds, hipStream_t st) {
    switch (cfg) {
        case 0: return batched_t<TI, TB, 64, 4, FAST>(args, grid, lds, st);
        case 1: return batched_t<TI, TB, 128, 8, FAST>(args, grid, lds, st);
        case 2:
            if constexpr (sizeof(TB) == 4) return batched_t<TI, TB, 256, 16, FAST>(args, grid, lds, st);
            return hipErrorInvalidValue;
    }
    return hipErrorInvalidValue;
}

template <typename TI, typename TB, bool FAST>
int occupancy_cfg(int cfg, size_t lds) {
    switch (cfg) {
        case 0: return occupancy_t<TI, TB, 64, 4, FAST>(lds);
        case 1: return occupancy_t<TI, TB, 128, 8, FAST>(lds);
        case 2:
            if constexpr (sizeof(TB) == 4) return occupancy_t<TI, TB, 256, 16, FAST>(lds);
            return 1;
    }
    return 1;
}

}  // namespace

hipError_t launch_batched(InType ti, bool is_double, int cfg, bool fast, const void *args, int grid, size_t lds,
                          hipStream_t stream) {
    if (is_double) return fast ? batched_cfg<double, double, true>(cfg, args, grid, lds, stream)
                               : batched_cfg<double, double, false>(cfg, args, grid, lds, stream);
    if (ti == IN_U8) return fast ? batched_cfg<uint8_t, float, true>(cfg, args, grid, lds, stream)
                                 : batched_cfg<uint8_t, float, false>(cfg, args, grid, lds, stream);
    return fast ? batched_cfg<float, float, true>(cfg, args, grid, lds, stream)
                : batched_cfg<float, float, false>(cfg, args, grid, lds, stream);
}

int batched_occupancy(InType ti, bool is_double, int cfg, bool fast, size_t lds) {
    if (is_double) return fast ? occupancy_cfg<double, double, true>(cfg, lds) : occupancy_cfg<double, double, false>(cfg, lds);
    if (ti == IN_U8) return fast ? occupancy_cfg<uint8_t, float, true>(cfg, lds) : occupancy_cfg<uint8_t, float, false>(cfg, lds);
    return fast ? occupancy_cfg<float, float, true>(cfg, lds) : occupancy_cfg<float, float, false>(cfg, lds);
}

hipError_t batched_rc_entry_u8(int cfg, bool fast, int rxc, const void *args, int grid, size_t lds, hipStream_t st,
                               int *occ);
hipError_t batched_rc_entry_f32(int cfg, bool fast, int rxc, const void *args, int grid, size_t lds, hipStream_t st,
                                int *occ);
hipError_t batched_rc_entry_f64(int cfg, bool fast, int rxc, const void *args, int grid, size_t lds, hipStream_t st,
                                int *occ);

hipError_t launch_batched_rc(InType ti, bool is_double, int cfg, bool fast, int rxc, const void *args, int grid,
                             size_t lds, hipStream_t stream, int *occ) {
    if (is_double) return batched_rc_entry_f64(cfg, fast, rxc, args, grid, lds, stream, occ);
    if (ti == IN_U8) return batched_rc_entry_u8(cfg, fast, rxc, args, grid, lds, stream, occ);
    return batched_rc_entry_f32(cfg, fast, rxc, args, grid, lds, stream, occ);
}

hipError_t launch_reset_counter(int32_t *counter, int n, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(256), 0, stream, counter, n);
    return hipGetLastError();
}

}  // namespace dipamd
