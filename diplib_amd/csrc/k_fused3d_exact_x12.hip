// Fused 3-D kernel instances: exact arithmetic, x radius class 12.
#include "k_fused3d_inst.hpp"

namespace dipamd {

bool fused3d_launch_exact_x12(int sx, int sy, int rz, const Fused3Args &a, int64_t nblocks, size_t lds,
                          hipStream_t st, int *occ) {
    return fused3d_launch_x<12, false>(sx, sy, rz, a, nblocks, lds, st, occ);
}

}  // namespace dipamd
