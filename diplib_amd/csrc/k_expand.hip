// Extended lines: every line of an image along one dim written with its boundary
// extension into a buffer 2*border longer along that dim.  The filter kernels read
// index-map conditions through the map, but the conditions computed from the edge
// values (sep_extend.hpp) have none, so a pass with one of them filters this buffer;
// the line buffers of the plug-in framework (lines.cpp) use it for every condition
// (index-map ones fold the sample, negate it or give the constant).
#include "sep_extend.hpp"
#include "sep_kernels_impl.hpp"

namespace dipamd {

namespace {

template <typename TB>
__device__ __forceinline__ TB type_max();
template <>
__device__ __forceinline__ float type_max<float>() {
    return __FLT_MAX__;
}
template <>
__device__ __forceinline__ double type_max<double>() {
    return __DBL_MAX__;
}

// One thread per sample of the extended buffer (contiguous, dim `d` of size
// n + 2 (border + zpad)); the zpad outermost samples of each side are 0.
template <typename TI, typename TB>
__global__ __launch_bounds__(kBlock) void k_expand(const TI *in, TB *out, const CastGeo g, int d, int n,
                                                   int border, int zpad, int bc) {
    for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < g.total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t r = t, oline = 0, oo = 0;
        int p = 0;
        for (int e = 0; e < g.ndim; ++e) {
            const int64_t c = r % g.size[e];
            r /= g.size[e];
            oo += c * g.sout[e];
            if (e == d) {
                p = static_cast<int>(c);
            } else {
                oline += c * g.sin[e];
            }
        }
        const TI *line = in + oline;
        const int64_t st = g.sin[d];
        const int i = p - border - zpad;
        TB v;
        if (i >= 0 && i < n) {
            v = to_buf<TB>(line[static_cast<int64_t>(i) * st]);
        } else if (i < -border || i > n - 1 + border) {
            v = TB(0);
        } else if (bc != DIPAMD_BC_ANTISYMMETRIC_REFLECT && bc != DIPAMD_BC_FIRST_ORDER_EXTRAPOLATE &&
                   bc != DIPAMD_BC_SECOND_ORDER_EXTRAPOLATE && bc != DIPAMD_BC_THIRD_ORDER_EXTRAPOLATE) {
            // an index-map condition: the folded sample, its negation, or the constant
            const int f = bc_fold_packed(i, n, bc);
            if (f < 0) {
                v = bc == DIPAMD_BC_ADD_MAX_VALUE ? type_max<TB>() : (bc == DIPAMD_BC_ADD_MIN_VALUE ? -type_max<TB>() : TB(0));
            } else {
                const TB s = to_buf<TB>(line[static_cast<int64_t>(f >> 1) * st]);
                v = (f & 1) ? -s : s;
            }
        } else if (i < 0) {
            v = extension_value<TI, TB>(line, st, n, 0, -i, border, bc);
        } else {
            v = extension_value<TI, TB>(line, st, n, 1, i - (n - 1), border, bc);
        }
        out[oo] = v;
    }
}

template <typename TI, typename TB>
hipError_t expand_t(const void *in, void *out, const CastGeo &g, int d, int n, int border, int zpad, int bc,
                    hipStream_t st) {
    int64_t nblk = (g.total + kBlock - 1) / kBlock;
    if (nblk > 65536) nblk = 65536;
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL((k_expand<TI, TB>), dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, st,
                       static_cast<const TI *>(in), static_cast<TB *>(out), g, d, n, border, zpad, bc);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_expand(InType ti, bool is_double, const void *in, void *out, const void *geo, int d, int n,
                         int border, int zpad, int bc, hipStream_t stream) {
    const CastGeo &g = *static_cast<const CastGeo *>(geo);
    if (is_double) return expand_t<double, double>(in, out, g, d, n, border, zpad, bc, stream);
    return ti == IN_U8 ? expand_t<uint8_t, float>(in, out, g, d, n, border, zpad, bc, stream)
                       : expand_t<float, float>(in, out, g, d, n, border, zpad, bc, stream);
}

}  // namespace dipamd
