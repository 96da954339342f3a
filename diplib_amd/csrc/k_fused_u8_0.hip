// Fused 2-D kernel instantiations: input uint8_t, buffer float, fast=false.
#include "k_fused_common.hpp"

namespace dipamd {

hipError_t fused_entry_u8_0(int cfg, int sx, int sy, const void *args, int64_t nx, int64_t nc, int64_t no,
                                   int rx, int ry, hipStream_t st) {
    const auto &a = *static_cast<const FusedArgs<float> *>(args);
    switch (cfg) {
        case 0: return fused_sym<uint8_t, float, 64, 4, false>(sx, sy, a, nx, nc, no, rx, ry, st);
        case 1: return fused_sym<uint8_t, float, 128, 8, false>(sx, sy, a, nx, nc, no, rx, ry, st);
        case 2: return fused_sym<uint8_t, float, 256, 16, false>(sx, sy, a, nx, nc, no, rx, ry, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace dipamd
