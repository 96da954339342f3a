// The derivative family's one-launch form (stream2d_deriv.hpp): dispatch over the
// radius triples with straight-line kernels (k_stream2d_deriv_<r0><r1><r2>.hip) and the
// kernels taking the radii from their arguments (instantiated here).
#include "stream2d_deriv.hpp"

namespace dipamd {

namespace deriv {
extern template hipError_t deriv_dispatch<3, 4, 4>(int, const DerivStreamArgs &, bool, hipStream_t);
extern template hipError_t deriv_dispatch<3, 3, 4>(int, const DerivStreamArgs &, bool, hipStream_t);
template hipError_t deriv_dispatch<0, 0, 0>(int, const DerivStreamArgs &, bool, hipStream_t);
}  // namespace deriv

hipError_t launch_stream2d_deriv(int mode, bool fast, const void *args, hipStream_t stream) {
    const DerivStreamArgs &a = *static_cast<const DerivStreamArgs *>(args);
    for (int o = 0; o < 3; ++o)
        if (a.rx[o] < 1 || a.rx[o] > kStreamRMax || a.ry[o] < 1 || a.ry[o] > kStreamRMax) return hipErrorInvalidValue;
    // the orders the function uses: Gradient and GradientMagnitude 0 and 1, Laplace 0
    // and 2, the others all three
    const bool u1 = mode != kDerivLaplace;
    const bool u2 = mode != kDerivGradient && mode != kDerivGradientMagnitude;
    auto is = [&](int r0, int r1, int r2) {
        return a.rx[0] == r0 && a.ry[0] == r0 && (!u1 || (a.rx[1] == r1 && a.ry[1] == r1)) &&
               (!u2 || (a.rx[2] == r2 && a.ry[2] == r2));
    };
    if (is(3, 4, 4)) return deriv::deriv_dispatch<3, 4, 4>(mode, a, fast, stream);  // sigma 0.86 .. 1
    if (is(3, 3, 4)) return deriv::deriv_dispatch<3, 3, 4>(mode, a, fast, stream);  // sigma 0.75 .. 0.85
    return deriv::deriv_dispatch<0, 0, 0>(mode, a, fast, stream);
}

}  // namespace dipamd
