// Launch helpers of the radius-class fused 2-D kernel (fused_rc.hpp), included by the
// per-type translation units k_fusedrc_*.hip.
#pragma once
#include "fused_rc.hpp"

namespace dipamd {
namespace {

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST, int RXC>
hipError_t fused_rc_t(const FusedArgs<TB> &a, int64_t nx, int64_t nchunks, int64_t nouter, int ry, hipStream_t st) {
    const int64_t ntiles = (nx + TXC - 1) / TXC;
    const int64_t nblk = ntiles * nchunks * nouter;
    const size_t lds = rc_lds_bytes<TB>(TXC, RXC, ry);
    hipLaunchKernelGGL((k_fused2d_rc<TI, TB, TXC, P, SX, SY, FAST, RXC>), dim3(static_cast<unsigned>(nblk)),
                       dim3(kBlock), lds, st, a, ntiles, nchunks);
    return hipGetLastError();
}

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST>
hipError_t fused_rc_class(int rxc, const FusedArgs<TB> &a, int64_t nx, int64_t nc, int64_t no, int ry,
                          hipStream_t st) {
    if constexpr (SX == SYM_NONE) {
        return rxc == 0 ? fused_rc_t<TI, TB, TXC, P, SX, SY, FAST, 0>(a, nx, nc, no, ry, st) : hipErrorNotSupported;
    } else {
        switch (rxc) {
            case 4: return fused_rc_t<TI, TB, TXC, P, SX, SY, FAST, 4>(a, nx, nc, no, ry, st);
            case 8: return fused_rc_t<TI, TB, TXC, P, SX, SY, FAST, 8>(a, nx, nc, no, ry, st);
            case 12: return fused_rc_t<TI, TB, TXC, P, SX, SY, FAST, 12>(a, nx, nc, no, ry, st);
            case 16: return fused_rc_t<TI, TB, TXC, P, SX, SY, FAST, 16>(a, nx, nc, no, ry, st);
            case 20: return fused_rc_t<TI, TB, TXC, P, SX, SY, FAST, 20>(a, nx, nc, no, ry, st);
            case 24: return fused_rc_t<TI, TB, TXC, P, SX, SY, FAST, 24>(a, nx, nc, no, ry, st);
        }
        return hipErrorNotSupported;
    }
}

template <typename TI, typename TB, int TXC, int P, bool FAST>
hipError_t fused_rc_sym(int sx, int sy, int rxc, const FusedArgs<TB> &a, int64_t nx, int64_t nc, int64_t no, int ry,
                        hipStream_t st) {
#define DIPAMD_RC_SYM(X, Y) \
    if (sx == X && sy == Y) return fused_rc_class<TI, TB, TXC, P, X, Y, FAST>(rxc, a, nx, nc, no, ry, st);
    DIPAMD_RC_SYM(SYM_EVEN, SYM_EVEN)
    DIPAMD_RC_SYM(SYM_EVEN, SYM_ODD)
    DIPAMD_RC_SYM(SYM_ODD, SYM_EVEN)
    DIPAMD_RC_SYM(SYM_ODD, SYM_ODD)
    DIPAMD_RC_SYM(SYM_NONE, SYM_EVEN)
    DIPAMD_RC_SYM(SYM_NONE, SYM_ODD)
#undef DIPAMD_RC_SYM
    return hipErrorNotSupported;
}

template <typename TI, typename TB, bool FAST>
hipError_t fused_rc_cfg(int cfg, int sx, int sy, int rxc, const void *args, int64_t nx, int64_t nc, int64_t no,
                        int ry, hipStream_t st) {
    const auto &a = *static_cast<const FusedArgs<TB> *>(args);
    switch (cfg) {
        case 0: return fused_rc_sym<TI, TB, 64, 4, FAST>(sx, sy, rxc, a, nx, nc, no, ry, st);
        case 1: return fused_rc_sym<TI, TB, 128, 8, FAST>(sx, sy, rxc, a, nx, nc, no, ry, st);
        case 2:
            if constexpr (sizeof(TB) == 4) return fused_rc_sym<TI, TB, 256, 16, FAST>(sx, sy, rxc, a, nx, nc, no, ry, st);
            return hipErrorNotSupported;
    }
    return hipErrorNotSupported;
}

}  // namespace
}  // namespace dipamd
