// Instances of the streaming backward kernel (k_bwd2s.hpp): one column per lane, equal
// odd radius classes 3..13 along x and y, one translation unit per class
// (k_bwd2s_y<class>.hip).
#pragma once

#include "k_bwd2s.hpp"

namespace dipamd {

// Launches k_bwd2s for equal x and y classes CY; false for other pairs (those take the
// tiled kernel, which measured faster for them) or symmetries without an instance.
template <int CY>
bool bwd2s_launch_y(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream) {
    if (cx == CY) return bwd2s_launch_sym<CY, CY, 1>(symx, symy, a, nblocks, stream);
    return false;
}

bool bwd2s_launch_y3(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y5(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y7(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y9(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y11(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y13(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);

}  // namespace dipamd
