// Streaming backward instances with y radius class 9.
#include "k_bwd2s_inst.hpp"

namespace dipamd {

bool bwd2s_launch_y9(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream) {
    return bwd2s_launch_y<9>(cx, symx, symy, a, nblocks, stream);
}

}  // namespace dipamd
