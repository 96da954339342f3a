// Fused 2-D launch helpers (included by k_fused_*.hip).
#pragma once
#include "sep_kernels_impl.hpp"
#include "sep_config.hpp"

namespace dipamd {
namespace {

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST>
hipError_t fused_t(const FusedArgs<TB> &a, int64_t nx, int64_t nchunks, int64_t nouter, int rx, int ry,
                   hipStream_t st) {
    const int64_t ntiles = (nx + TXC - 1) / TXC;
    const int64_t nblk = ntiles * nchunks * nouter;
    const int S = kFusedStep;
    const int A = (2 * ry + S - 1) / S;
    const size_t lds =
        (static_cast<size_t>(S) * fused_pitch(TXC + 2 * rx, P) + static_cast<size_t>((A + 1) * S) * fused_pitch(TXC, P)) *
        sizeof(TB);
    hipLaunchKernelGGL((k_fused2d<TI, TB, TXC, P, SX, SY, FAST>), dim3(static_cast<unsigned>(nblk)), dim3(kBlock),
                       lds, st, a, ntiles, nchunks);
    return hipGetLastError();
}

template <typename TI, typename TB, int TXC, int P, bool FAST>
hipError_t fused_sym(int sx, int sy, const FusedArgs<TB> &a, int64_t nx, int64_t nc, int64_t no, int rx, int ry,
                     hipStream_t st) {
    if (sx == SYM_EVEN && sy == SYM_EVEN) return fused_t<TI, TB, TXC, P, SYM_EVEN, SYM_EVEN, FAST>(a, nx, nc, no, rx, ry, st);
    if (sx == SYM_EVEN && sy == SYM_ODD) return fused_t<TI, TB, TXC, P, SYM_EVEN, SYM_ODD, FAST>(a, nx, nc, no, rx, ry, st);
    if (sx == SYM_ODD && sy == SYM_EVEN) return fused_t<TI, TB, TXC, P, SYM_ODD, SYM_EVEN, FAST>(a, nx, nc, no, rx, ry, st);
    if (sx == SYM_ODD && sy == SYM_ODD) return fused_t<TI, TB, TXC, P, SYM_ODD, SYM_ODD, FAST>(a, nx, nc, no, rx, ry, st);
    if (sx == SYM_NONE && sy == SYM_EVEN) return fused_t<TI, TB, TXC, P, SYM_NONE, SYM_EVEN, FAST>(a, nx, nc, no, rx, ry, st);
    if (sx == SYM_NONE && sy == SYM_ODD) return fused_t<TI, TB, TXC, P, SYM_NONE, SYM_ODD, FAST>(a, nx, nc, no, rx, ry, st);
    return hipErrorInvalidValue;
}

template <typename TI, typename TB>
hipError_t fused_cfg(int cfg, int sx, int sy, bool fast, const FusedArgs<TB> &a, int64_t nx, int64_t nc, int64_t no,
                     int rx, int ry, hipStream_t st) {
    switch (cfg) {
        case 0:
            return fast ? fused_sym<TI, TB, 64, 4, true>(sx, sy, a, nx, nc, no, rx, ry, st)
                        : fused_sym<TI, TB, 64, 4, false>(sx, sy, a, nx, nc, no, rx, ry, st);
        case 1:
            return fast ? fused_sym<TI, TB, 128, 8, true>(sx, sy, a, nx, nc, no, rx, ry, st)
                        : fused_sym<TI, TB, 128, 8, false>(sx, sy, a, nx, nc, no, rx, ry, st);
        case 2:
            if constexpr (sizeof(TB) == 4) {
                return fast ? fused_sym<TI, TB, 256, 16, true>(sx, sy, a, nx, nc, no, rx, ry, st)
                            : fused_sym<TI, TB, 256, 16, false>(sx, sy, a, nx, nc, no, rx, ry, st);
            }
            return hipErrorInvalidValue;
    }
    return hipErrorInvalidValue;
}

}  // namespace
}  // namespace dipamd
