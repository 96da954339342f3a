// Instances of the streaming backward kernel (k_bwd2s.hpp): one column per lane, odd
// radius classes 3..13 along x and y, equal or adjacent (x class = y class +- 2, a first
// derivative along one axis), one translation unit per y class (k_bwd2s_y<class>.hip).
#pragma once

#include "k_bwd2s.hpp"

namespace dipamd {

// Launches k_bwd2s for y class CY and an equal or adjacent x class; false for other
// pairs (those take the tiled kernel) or symmetries without an instance.
template <int CY>
bool bwd2s_launch_y(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream) {
    if (cx == CY) return bwd2s_launch_sym<CY, CY, 1>(symx, symy, a, nblocks, stream);
    if constexpr (CY + 2 <= 13) {
        if (cx == CY + 2) return bwd2s_launch_sym<CY + 2, CY, 1>(symx, symy, a, nblocks, stream);
    }
    if constexpr (CY - 2 >= 3) {
        if (cx == CY - 2) return bwd2s_launch_sym<CY - 2, CY, 1>(symx, symy, a, nblocks, stream);
    }
    return false;
}

bool bwd2s_launch_y3(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y5(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y7(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y9(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y11(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y13(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);

}  // namespace dipamd
