// Boundary conditions whose extension is computed from the edge values instead of
// copied: antisymmetric reflection (copy_buffer.cpp:551-589) and first to third
// order extrapolation (copy_buffer.cpp:380-513, with the fall-backs of :791-808 for
// lines of one sample).  The filter kernels read boundary extensions through an
// index map, which these conditions do not have, so a pass with one of them first
// writes the extended lines into a buffer that is 2*border longer along the pass
// axis (this kernel), filters that buffer, and keeps the interior.
//
// Values follow the reference's arithmetic: the line is in the buffer type, the
// extension is computed in double (no contraction) and stored with a plain cast.
#include "sep_kernels_impl.hpp"

namespace dipamd {

namespace {

// Sample at distance i from the edge (side 0: from the first sample, 1: from the last).
template <typename TI, typename TB>
__device__ __forceinline__ double edge_sample(const TI *line, int64_t st, int n, int side, int i) {
    const int j = side == 0 ? i : n - 1 - i;
    return static_cast<double>(to_buf<TB>(line[static_cast<int64_t>(j) * st]));
}

// Extension sample at distance k >= 1 outside the edge `side` of a line of n samples
// for a border of `border` samples.
template <typename TI, typename TB>
__device__ TB extension_value(const TI *line, int64_t st, int n, int side, int k, int border, int bc) {
    if (bc == DIPAMD_BC_ANTISYMMETRIC_REFLECT) {
        const double x0 = edge_sample<TI, TB>(line, st, n, side, 0);
        if (n == 1) return static_cast<TB>(x0);
        // the walk of the reference: the source moves inward and turns at either end,
        // and each turn re-centres the reflection on the sample just written
        int i = 0, sign = -1;
        double offset = x0 - sign * x0;
        TB v = TB(0);
        for (int it = 1; it <= k; ++it) {
            i -= sign;
            const double xi = edge_sample<TI, TB>(line, st, n, side, i);
            v = static_cast<TB>(offset + sign * xi);
            if (i == n - 1 || i == 0) {
                sign = -sign;
                offset = static_cast<double>(v) - sign * xi;
            }
        }
        return v;
    }
    if (bc == DIPAMD_BC_FIRST_ORDER_EXTRAPOLATE || n == 1) {
        double d0 = edge_sample<TI, TB>(line, st, n, side, 0);
        const double d1 = d0 / static_cast<double>(border + 1);
        for (int it = 1; it <= k; ++it) d0 -= d1;
        return static_cast<TB>(d0);
    }
    const double b = static_cast<double>(border) + 1.0;
    const double d0 = edge_sample<TI, TB>(line, st, n, side, 0);
    const double f1 = edge_sample<TI, TB>(line, st, n, side, 1);
    const double x = static_cast<double>(k);
    if (bc == DIPAMD_BC_SECOND_ORDER_EXTRAPOLATE) {
        const double d1 = (b - 1.0) / b * d0 - b / (b + 1.0) * f1;
        const double d2 = -1.0 / b * d0 + 1.0 / (b + 1.0) * f1;
        return static_cast<TB>(d0 + x * d1 + x * x * d2);
    }
    const double b12 = (b + 1) * (b + 1);
    const double d1 = -(2.0 * d0) / b + d0 - (b * b * f1) / b12;
    const double d2 = (2.0 * b * f1) / b12 - (d0 * (2.0 * b - 1.0)) / (b * b);
    const double d3 = d0 / (b * b) - f1 / b12;
    return static_cast<TB>(d0 + x * d1 + x * x * d2 + x * x * x * d3);
}

// One thread per sample of the extended buffer (contiguous, dim `d` of size n+2B).
template <typename TI, typename TB>
__global__ __launch_bounds__(kBlock) void k_expand(const TI *in, TB *out, const CastGeo g, int d, int n,
                                                   int border, int bc) {
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
        const int i = p - border;
        TB v;
        if (i >= 0 && i < n) {
            v = to_buf<TB>(line[static_cast<int64_t>(i) * st]);
        } else if (i < 0) {
            v = extension_value<TI, TB>(line, st, n, 0, -i, border, bc);
        } else {
            v = extension_value<TI, TB>(line, st, n, 1, i - (n - 1), border, bc);
        }
        out[oo] = v;
    }
}

template <typename TI, typename TB>
hipError_t expand_t(const void *in, void *out, const CastGeo &g, int d, int n, int border, int bc, hipStream_t st) {
    int64_t nblk = (g.total + kBlock - 1) / kBlock;
    if (nblk > 65536) nblk = 65536;
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL((k_expand<TI, TB>), dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, st,
                       static_cast<const TI *>(in), static_cast<TB *>(out), g, d, n, border, bc);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_expand(InType ti, bool is_double, const void *in, void *out, const void *geo, int d, int n,
                         int border, int bc, hipStream_t stream) {
    const CastGeo &g = *static_cast<const CastGeo *>(geo);
    if (is_double) return expand_t<double, double>(in, out, g, d, n, border, bc, stream);
    return ti == IN_U8 ? expand_t<uint8_t, float>(in, out, g, d, n, border, bc, stream)
                       : expand_t<float, float>(in, out, g, d, n, border, bc, stream);
}

}  // namespace dipamd
