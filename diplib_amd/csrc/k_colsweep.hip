// Launch table of the column-tile sweep (k_colsweep, see sep_kernels_impl.hpp).
#include "sep_kernels_impl.hpp"

#include "sep_config.hpp"

#include <cstdlib>

namespace dipamd {

namespace {
// waves per workgroup: eight, so that a workgroup computing beside one that stages
// its tile still has two waves per SIMD to hide the LDS latency
int colsweep_waves() {
    static int nw = [] {
        const char *v = std::getenv("DIPAMD_COLSWEEP_WAVES");
        return (v && std::atoi(v) == 4) ? 4 : 8;
    }();
    return nw;
}
}  // namespace

int colsweep_rows(int border, bool is_double) {
    // staged samples per column: R + 2*border, 64 columns, within the LDS limit
    const int P = is_double ? kColPD : kColPF;
    const size_t es = is_double ? 8 : 4;
    const int hmax = static_cast<int>(65536 / (64 * es)) - 1;  // odd pitch <= hmax+1
    int R = hmax - 2 * border;
    R = (R / P) * P;
    return R >= P ? R : 0;
}

namespace {

template <typename TI, typename TB, int P, int SYM, bool FAST>
hipError_t col_t(const PassArgs<TB> &a, int64_t e, int64_t nchunks, int64_t nouter, hipStream_t st) {
    const int64_t ntiles = (e + 63) / 64;
    const int64_t nblk = ntiles * nchunks * nouter;
    const int H = static_cast<int>(a.chunk) + 2 * a.border;
    const size_t lds = static_cast<size_t>(64) * (H | 1) * sizeof(TB);
    if (colsweep_waves() == 4) {
        hipLaunchKernelGGL((k_colsweep<TI, TB, P, SYM, FAST, 4>), dim3(static_cast<unsigned>(nblk)), dim3(256), lds,
                           st, a, ntiles, nchunks, static_cast<const TI *>(a.in), static_cast<TB *>(a.out));
    } else {
        hipLaunchKernelGGL((k_colsweep<TI, TB, P, SYM, FAST, 8>), dim3(static_cast<unsigned>(nblk)), dim3(512), lds,
                           st, a, ntiles, nchunks, static_cast<const TI *>(a.in), static_cast<TB *>(a.out));
    }
    return hipGetLastError();
}

template <typename TI, typename TB, int P>
hipError_t col_sym(int sym, bool fast, const PassArgs<TB> &a, int64_t e, int64_t nc, int64_t no, hipStream_t st) {
    switch (sym) {
        case SYM_EVEN:
            return fast ? col_t<TI, TB, P, SYM_EVEN, true>(a, e, nc, no, st)
                        : col_t<TI, TB, P, SYM_EVEN, false>(a, e, nc, no, st);
        case SYM_ODD:
            return fast ? col_t<TI, TB, P, SYM_ODD, true>(a, e, nc, no, st)
                        : col_t<TI, TB, P, SYM_ODD, false>(a, e, nc, no, st);
        case SYM_GENERAL: return col_t<TI, TB, P, SYM_GENERAL, false>(a, e, nc, no, st);
        case SYM_D_EVEN: return col_t<TI, TB, P, SYM_D_EVEN, false>(a, e, nc, no, st);
        case SYM_D_ODD: return col_t<TI, TB, P, SYM_D_ODD, false>(a, e, nc, no, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_colsweep(InType ti, bool is_double, int sym, bool fast, const void *args, int64_t e,
                           int64_t nchunks, int64_t nouter, hipStream_t stream) {
    if (is_double) {
        return col_sym<double, double, kColPD>(sym, fast, *static_cast<const PassArgs<double> *>(args), e, nchunks,
                                               nouter, stream);
    }
    const auto &a = *static_cast<const PassArgs<float> *>(args);
    return ti == IN_U8 ? col_sym<uint8_t, float, kColPF>(sym, fast, a, e, nchunks, nouter, stream)
                       : col_sym<float, float, kColPF>(sym, fast, a, e, nchunks, nouter, stream);
}

}  // namespace dipamd
