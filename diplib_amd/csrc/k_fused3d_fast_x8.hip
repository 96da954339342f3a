// Fused 3-D kernel instances: fast arithmetic, x radius class 8.
#include "k_fused3d_inst.hpp"

namespace dipamd {

bool fused3d_launch_fast_x8(int sx, int sy, int rz, const Fused3Args &a, int64_t nblocks, size_t lds,
                          hipStream_t st, int *occ) {
    return fused3d_launch_x<8, true>(sx, sy, rz, a, nblocks, lds, st, occ);
}

}  // namespace dipamd
