// Launch helpers of the fused 3-D kernel (fused3d.hpp), included by the translation
// units k_fused3d_<arithmetic>_x<class>.hip (one per arithmetic and x radius class).
#pragma once

#include "fused3d.hpp"

namespace dipamd {

template <int SX, int SY, int RXC, int RZ, bool FAST>
bool fused3d_launch_t(const Fused3Args &a, int64_t nblocks, size_t lds, hipStream_t st, int *occ) {
    const auto k = k_fused3d<float, SX, SY, SYM_EVEN, RXC, RZ, FAST>;
    if (occ) {  // resident workgroups per CU at `lds` bytes, no launch
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kF3Threads, lds) != hipSuccess) n = 1;
        *occ = n > 0 ? n : 1;
        return true;
    }
    hipLaunchKernelGGL(k, dim3(static_cast<unsigned>(nblocks)), dim3(kF3Threads), lds, st, a);
    return true;
}

template <int SX, int SY, int RXC, bool FAST>
bool fused3d_launch_rz(int rz, const Fused3Args &a, int64_t nblocks, size_t lds, hipStream_t st, int *occ) {
    switch (rz) {
        case 1: return fused3d_launch_t<SX, SY, RXC, 1, FAST>(a, nblocks, lds, st, occ);
        case 2: return fused3d_launch_t<SX, SY, RXC, 2, FAST>(a, nblocks, lds, st, occ);
        case 3: return fused3d_launch_t<SX, SY, RXC, 3, FAST>(a, nblocks, lds, st, occ);
        case 4: return fused3d_launch_t<SX, SY, RXC, 4, FAST>(a, nblocks, lds, st, occ);
    }
    return false;
}

template <int RXC, bool FAST>
bool fused3d_launch_x(int sx, int sy, int rz, const Fused3Args &a, int64_t nblocks, size_t lds, hipStream_t st,
                      int *occ) {
    if (sx == SYM_EVEN && sy == SYM_EVEN) return fused3d_launch_rz<SYM_EVEN, SYM_EVEN, RXC, FAST>(rz, a, nblocks, lds, st, occ);
    if (sx == SYM_EVEN && sy == SYM_ODD) return fused3d_launch_rz<SYM_EVEN, SYM_ODD, RXC, FAST>(rz, a, nblocks, lds, st, occ);
    if (sx == SYM_ODD && sy == SYM_EVEN) return fused3d_launch_rz<SYM_ODD, SYM_EVEN, RXC, FAST>(rz, a, nblocks, lds, st, occ);
    if (sx == SYM_ODD && sy == SYM_ODD) return fused3d_launch_rz<SYM_ODD, SYM_ODD, RXC, FAST>(rz, a, nblocks, lds, st, occ);
    return false;
}

}  // namespace dipamd
