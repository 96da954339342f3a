// Launch tables (see sep_kernels_impl.hpp).
#include "sep_kernels_impl.hpp"

#include "sep_config.hpp"

namespace dipamd {
namespace {

template <typename TI, typename TB, int P, int SYM, bool FAST>
hipError_t sweep_t(const PassArgs<TB> &a, int64_t e, int64_t nchunks, int64_t nouter, hipStream_t st) {
    const int bs = e >= 256 ? 256 : 64;
    const int64_t ntiles = (e + bs - 1) / bs;
    const int64_t nblk = ntiles * nchunks * nouter;
    hipLaunchKernelGGL((k_sweep<TI, TB, P, SYM, FAST>), dim3(static_cast<unsigned>(nblk)), dim3(bs), 0, st, a, ntiles,
                       nchunks, static_cast<const TI *>(a.in), static_cast<TB *>(a.out));
    return hipGetLastError();
}

template <typename TI, typename TB>
hipError_t sweep_sym(int sym, bool fast, const PassArgs<TB> &a, int64_t e, int64_t nc, int64_t no, hipStream_t st) {
    constexpr int P = kSweepP;
    switch (sym) {
        case SYM_EVEN:
            return fast ? sweep_t<TI, TB, P, SYM_EVEN, true>(a, e, nc, no, st)
                        : sweep_t<TI, TB, P, SYM_EVEN, false>(a, e, nc, no, st);
        case SYM_ODD:
            return fast ? sweep_t<TI, TB, P, SYM_ODD, true>(a, e, nc, no, st)
                        : sweep_t<TI, TB, P, SYM_ODD, false>(a, e, nc, no, st);
        case SYM_GENERAL: return sweep_t<TI, TB, P, SYM_GENERAL, false>(a, e, nc, no, st);
        case SYM_D_EVEN: return sweep_t<TI, TB, P, SYM_D_EVEN, false>(a, e, nc, no, st);
        case SYM_D_ODD: return sweep_t<TI, TB, P, SYM_D_ODD, false>(a, e, nc, no, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_sweep(InType ti, bool is_double, int sym, bool fast, const void *args, int64_t e, int64_t nchunks,
                        int64_t nouter, hipStream_t stream) {
    if (is_double) {
        return sweep_sym<double, double>(sym, fast, *static_cast<const PassArgs<double> *>(args), e, nchunks, nouter,
                                         stream);
    }
    const auto &a = *static_cast<const PassArgs<float> *>(args);
    return ti == IN_U8 ? sweep_sym<uint8_t, float>(sym, fast, a, e, nchunks, nouter, stream)
                       : sweep_sym<float, float>(sym, fast, a, e, nchunks, nouter, stream);
}

}  // namespace dipamd
