// The derivative family's one-launch form (stream2d_deriv.hpp) for the radius triple
// (3, 4, 4) of the derivative orders 0, 1, 2, as straight-line code.
#include "stream2d_deriv.hpp"

namespace dipamd {
namespace deriv {
template hipError_t deriv_dispatch<3, 4, 4>(int, const DerivStreamArgs &, bool, hipStream_t);
}  // namespace deriv
}  // namespace dipamd
