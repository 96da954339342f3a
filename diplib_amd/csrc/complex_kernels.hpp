// The complex separable pass (k_complex.hip): arguments and launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "sep_common.hpp"

namespace dipamd {

// The conjugate-symmetric forms of a complex filter (DIPAMD_SYM_CONJ / D_CONJ).
constexpr int kSymConj = 3, kSymDConj = 6;

// One pass along one axis over interleaved (re, im) samples of TS; strides count
// complex samples.  `taps` (device) holds the reversed (correlation-order) taps: ntaps
// values, or ntaps (re, im) pairs for a complex filter.
template <typename TS>
struct CxPassArgs {
    const TS *in;
    TS *out;
    int64_t n, sin, sout;  // along the axis
    int64_t nlines;
    Outer outer;           // the other dims (input / output strides)
    int32_t ntaps, origin, bc, zlo;
    TS cval;
    const TS *taps;
};

hipError_t launch_cpass(bool is_double, bool complex_filter, int sym, const void *args, hipStream_t stream);
// Any sample type into the complex buffer type (geo: sizes, input and output strides in
// samples of each side's type).
hipError_t launch_to_complex(int dt_in, bool to_double, const void *in, void *out, const void *geo,
                             hipStream_t stream);
// Complex to complex, changing the precision if asked.
hipError_t launch_complex_copy(bool from_double, bool to_double, const void *in, void *out, const void *geo,
                               hipStream_t stream);

}  // namespace dipamd
