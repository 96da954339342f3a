// Launch tables (see sep_kernels_impl.hpp).
#include "sep_kernels_impl.hpp"

#include "fused_rc.hpp"
#include "sep_config.hpp"

namespace dipamd {
size_t line_lds_bytes(int border, bool is_double) {
    const int P = is_double ? kLinePD : kLinePF;
    return static_cast<size_t>(4) * lpad_size(64 * P + 2 * border) * (is_double ? 8 : 4);
}

size_t fused_lds_bytes(int cfg, int rx, int ry, bool is_double) {
    const FusedConfig c = kFusedConfigs[cfg];
    const int S = kFusedStep;
    const int A = (2 * ry + S - 1) / S;
    const int RC = (A + 1) * S;
    const size_t words = static_cast<size_t>(S) * fused_pitch(c.txc + 2 * rx, c.p) +
                         static_cast<size_t>(RC) * fused_pitch(c.txc, c.p);
    return words * (is_double ? 8 : 4);
}

size_t fused_rc_lds_bytes(int cfg, int rxc, int ry, bool is_double) {
    const int txc = kFusedConfigs[cfg].txc;
    return is_double ? rc_lds_bytes<double>(txc, rxc, ry) : rc_lds_bytes<float>(txc, rxc, ry);
}

namespace {

template <typename TI, typename TB, int P, int SYM, bool FAST>
hipError_t line_t(const PassArgs<TB> &a, int64_t n, int64_t nlines, int border, hipStream_t st) {
    const int64_t nseg = (n + 64 * P - 1) / (64 * P);
    const int64_t nblk = nseg * ((nlines + 3) / 4);
    const size_t lds = static_cast<size_t>(4) * lpad_size(64 * P + 2 * border) * sizeof(TB);
    hipLaunchKernelGGL((k_line<TI, TB, P, SYM, FAST>), dim3(static_cast<unsigned>(nblk)), dim3(kBlock), lds, st, a,
                       nseg, static_cast<const TI *>(a.in), static_cast<TB *>(a.out));
    return hipGetLastError();
}

template <typename TI, typename TB, int P>
hipError_t line_sym(int sym, bool fast, const PassArgs<TB> &a, int64_t n, int64_t nl, int b, hipStream_t st) {
    switch (sym) {
        case SYM_EVEN:
            return fast ? line_t<TI, TB, P, SYM_EVEN, true>(a, n, nl, b, st)
                        : line_t<TI, TB, P, SYM_EVEN, false>(a, n, nl, b, st);
        case SYM_ODD:
            return fast ? line_t<TI, TB, P, SYM_ODD, true>(a, n, nl, b, st)
                        : line_t<TI, TB, P, SYM_ODD, false>(a, n, nl, b, st);
        case SYM_GENERAL: return line_t<TI, TB, P, SYM_GENERAL, false>(a, n, nl, b, st);
        case SYM_D_EVEN: return line_t<TI, TB, P, SYM_D_EVEN, false>(a, n, nl, b, st);
        case SYM_D_ODD: return line_t<TI, TB, P, SYM_D_ODD, false>(a, n, nl, b, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_line(InType ti, bool is_double, int sym, bool fast, const void *args, int64_t n, int64_t nlines,
                       int border, hipStream_t stream) {
    if (is_double) {
        return line_sym<double, double, kLinePD>(sym, fast, *static_cast<const PassArgs<double> *>(args), n, nlines,
                                                 border, stream);
    }
    const auto &a = *static_cast<const PassArgs<float> *>(args);
    return ti == IN_U8 ? line_sym<uint8_t, float, kLinePF>(sym, fast, a, n, nlines, border, stream)
                       : line_sym<float, float, kLinePF>(sym, fast, a, n, nlines, border, stream);
}

}  // namespace dipamd
