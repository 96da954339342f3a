// Instances of the fused 2-D backward kernel (k_bwd2d.hpp) whose x and y radius
// classes are adjacent and unequal -- a first derivative along x with a Gaussian along
// y at the same sigma (radius ceil(3.5 sigma) against ceil(3 sigma)), and the reverse:
// each stage then runs its own window length instead of the larger of the two.
#include "k_bwd2d.hpp"

namespace dipamd {

#define DIPAMD_BWD_PAIRS(X) X(5, 3) X(3, 5) X(7, 5) X(5, 7) X(9, 7) X(7, 9) X(11, 9) X(9, 11) \
    X(13, 11) X(11, 13) X(16, 13) X(13, 16)

bool bwd2d_launch_xy(int cx, int cy, int symx, int symy, const BwdArgs &a, int64_t nblocks, hipStream_t stream) {
    switch (32 * cx + cy) {
#define DIPAMD_CASE(RX, RY) \
    case 32 * RX + RY: return bwd2d_launch_pair<RX, RY>(symx, symy, a, nblocks, stream);
        DIPAMD_BWD_PAIRS(DIPAMD_CASE)
#undef DIPAMD_CASE
        default: return false;
    }
}

size_t bwd2d_lds_xy(int cx, int cy) {
    switch (32 * cx + cy) {
#define DIPAMD_CASE(RX, RY) \
    case 32 * RX + RY: return bwd2d_lds_of<RX, RY>();
        DIPAMD_BWD_PAIRS(DIPAMD_CASE)
#undef DIPAMD_CASE
        default: return 0;
    }
}

#undef DIPAMD_BWD_PAIRS

}  // namespace dipamd
