// Register-streaming fused kernels (stream2d.hpp) for one symmetry pair: x even, y even.
#include "stream2d.hpp"

namespace dipamd {
template hipError_t stream2d_launch<SYM_EVEN, SYM_EVEN>(const StreamArgs &a, bool fast, hipStream_t st);
}  // namespace dipamd
