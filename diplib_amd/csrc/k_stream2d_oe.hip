// Register-streaming fused kernels (stream2d.hpp) for one symmetry pair: x odd, y even.
#include "stream2d.hpp"

namespace dipamd {
template hipError_t stream2d_launch<SYM_ODD, SYM_EVEN>(const StreamArgs &a, bool fast, hipStream_t st);
}  // namespace dipamd
