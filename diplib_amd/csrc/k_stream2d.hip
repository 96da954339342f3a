// Launch table of the register-streaming fused kernel (stream2d.hpp): one launch
// for the interior strips, one for the (at most two) edge strips.
#include "stream2d.hpp"

namespace dipamd {

namespace {
template <int SX, int SY, bool FAST>
hipError_t stream_t(const StreamArgs &a, hipStream_t st) {
    const int64_t per = a.nwaves / a.nstrips;  // waves per strip (chunks x planes)
    const int64_t ninterior = a.nstrips > 2 ? (a.nstrips - 2) * per : 0;
    const int64_t nedge = (a.nstrips > 1 ? 2 : 1) * per;
    if (ninterior > 0)
        hipLaunchKernelGGL((k_stream2d<SX, SY, FAST, false>), dim3(static_cast<unsigned>((ninterior + 3) / 4)),
                           dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL((k_stream2d<SX, SY, FAST, true>), dim3(static_cast<unsigned>((nedge + 3) / 4)), dim3(kBlock),
                       0, st, a);
    return hipGetLastError();
}
template <int SX, int SY>
hipError_t stream_f(bool fast, const StreamArgs &a, hipStream_t st) {
    return fast ? stream_t<SX, SY, true>(a, st) : stream_t<SX, SY, false>(a, st);
}
}  // namespace

hipError_t launch_stream2d(int symx, int symy, bool fast, const void *args, hipStream_t stream) {
    const StreamArgs &a = *static_cast<const StreamArgs *>(args);
    if (symx == SYM_EVEN && symy == SYM_EVEN) return stream_f<SYM_EVEN, SYM_EVEN>(fast, a, stream);
    if (symx == SYM_EVEN && symy == SYM_ODD) return stream_f<SYM_EVEN, SYM_ODD>(fast, a, stream);
    if (symx == SYM_ODD && symy == SYM_EVEN) return stream_f<SYM_ODD, SYM_EVEN>(fast, a, stream);
    if (symx == SYM_ODD && symy == SYM_ODD) return stream_f<SYM_ODD, SYM_ODD>(fast, a, stream);
    return hipErrorInvalidValue;
}

}  // namespace dipamd
