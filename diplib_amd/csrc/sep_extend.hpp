// Boundary conditions whose extension is computed from the edge values instead of
// copied: antisymmetric reflection (copy_buffer.cpp:551-589) and first to third
// order extrapolation (copy_buffer.cpp:380-513, with the fall-backs of :791-808 for
// lines of one sample).  Values follow the reference's arithmetic: the line is in the
// buffer type, the extension is computed in double (no contraction) and stored with
// a plain cast.
#pragma once

#include "sep_common.hpp"

namespace dipamd {

// Sample at distance i from the edge (side 0: from the first sample, 1: from the last).
template <typename TI, typename TB>
__device__ __forceinline__ double edge_sample(const TI *line, int64_t st, int n, int side, int i) {
    const int j = side == 0 ? i : n - 1 - i;
    return static_cast<double>(to_buf<TB>(line[static_cast<int64_t>(j) * st]));
}

// Extension sample at distance k >= 1 outside the edge `side` of a line of n samples
// for a border of `border` samples.
template <typename TI, typename TB>
__device__ inline TB extension_value(const TI *line, int64_t st, int n, int side, int k, int border, int bc) {
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

}  // namespace dipamd
