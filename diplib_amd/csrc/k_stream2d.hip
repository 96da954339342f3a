// Dispatch of the register-streaming fused kernel (stream2d.hpp) over the symmetry
// pairs; each pair's kernels are compiled in their own translation unit.
#include "stream2d.hpp"

namespace dipamd {

extern template hipError_t stream2d_launch<SYM_EVEN, SYM_EVEN>(const StreamArgs &, bool, hipStream_t);
extern template hipError_t stream2d_launch<SYM_EVEN, SYM_ODD>(const StreamArgs &, bool, hipStream_t);
extern template hipError_t stream2d_launch<SYM_ODD, SYM_EVEN>(const StreamArgs &, bool, hipStream_t);
extern template hipError_t stream2d_launch<SYM_ODD, SYM_ODD>(const StreamArgs &, bool, hipStream_t);

hipError_t launch_stream2d(int symx, int symy, bool fast, const void *args, hipStream_t stream) {
    const StreamArgs &a = *static_cast<const StreamArgs *>(args);
    if (symx == SYM_EVEN && symy == SYM_EVEN) return stream2d_launch<SYM_EVEN, SYM_EVEN>(a, fast, stream);
    if (symx == SYM_EVEN && symy == SYM_ODD) return stream2d_launch<SYM_EVEN, SYM_ODD>(a, fast, stream);
    if (symx == SYM_ODD && symy == SYM_EVEN) return stream2d_launch<SYM_ODD, SYM_EVEN>(a, fast, stream);
    if (symx == SYM_ODD && symy == SYM_ODD) return stream2d_launch<SYM_ODD, SYM_ODD>(a, fast, stream);
    return hipErrorInvalidValue;
}

}  // namespace dipamd
