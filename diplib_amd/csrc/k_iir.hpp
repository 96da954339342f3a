// K5 k_iir: the recursive Gaussian of one axis (GaussIIRLineFilter::Filter,
// gaussiir.cpp:468-738), one thread per image line, in double.  Both recursions
// run exactly as the reference runs them: the same initial values, the same
// positions set from them, the same operation order in every step (moving average
// summed from 0.0 in index order, then the feedback terms subtracted one by one,
// nearest first), no contraction.  The causal result goes to a [position][line]
// scratch, the anti-causal one keeps its last nn values in registers and writes
// the interior positions to the destination.
//
// Throughput: each step is a chain of dependent double operations, so a pass costs
// (n + 2 border) steps of that chain per line; lines run in parallel.  Loads of the
// next P samples are issued before the current P steps, so they are in flight
// while the chain runs.
#pragma once

#include <type_traits>

#include "iir_common.hpp"
#include "sep_extend.hpp"

namespace dipamd {

constexpr int kIirBlock = 64;  // one wave per workgroup: few lines spread over many CUs
constexpr int kIirP = 32;      // samples per prefetch block

__device__ __forceinline__ bool iir_value_bc(int bc) {
    return bc == DIPAMD_BC_ANTISYMMETRIC_REFLECT || bc == DIPAMD_BC_FIRST_ORDER_EXTRAPOLATE ||
           bc == DIPAMD_BC_SECOND_ORDER_EXTRAPOLATE || bc == DIPAMD_BC_THIRD_ORDER_EXTRAPOLATE;
}

// Sample at position ii of the extended line (ii - border is the image index), in
// double, with the boundary condition outside the image.
template <typename TI>
__device__ __noinline__ double iir_ext(const TI *src, int64_t sin, int n, int border, int bc, double cval, int ii) {
    const int j = ii - border;
    if (j >= 0 && j < n) return to_buf<double>(src[static_cast<int64_t>(j) * sin]);
    if (iir_value_bc(bc))
        return extension_value<TI, double>(src, sin, n, j < 0 ? 0 : 1, j < 0 ? -j : j - (n - 1), border, bc);
    int mul;
    bool is_const;
    const int64_t idx = bc_map(j, n, bc, &mul, &is_const);
    if (is_const) return cval;
    const double v = to_buf<double>(src[idx * sin]);
    return mul < 0 ? -v : v;
}

// Moving-average ranges per derivative order (gaussiir.cpp:351-387): the forward
// part covers x[i - j] for j in [0, FM2], the backward part y[i + j] for j in
// [BM4, BM5]; order 0 copies (no sum).
template <int ORDER>
struct IirMa {
    static constexpr int FM2 = ORDER == 0 ? 0 : (ORDER == 2 ? 1 : 2);
    static constexpr int BM4 = ORDER == 1 ? 1 : 0;
    static constexpr int BM5 = ORDER == 0 ? 0 : (ORDER <= 2 ? 1 : 2);
};

template <typename TI, typename TO, int NN, int ORDER>
__global__ __launch_bounds__(kIirBlock) void k_iir(const IirArgs a) {
    using MA = IirMa<ORDER>;
    // loaded samples stay one per 32-bit register (a byte array would be packed, which
    // forces the loads to complete where they are issued)
    using TL = typename std::conditional<(sizeof(TI) < 4), uint32_t, TI>::type;
    constexpr int P = kIirP;
    const int64_t line = static_cast<int64_t>(blockIdx.x) * kIirBlock + threadIdx.x;
    if (line >= a.nlines) return;
    const int64_t lane = line % a.e;
    int64_t oin, oout;
    outer_offsets(a.outer, line / a.e, &oin, &oout);
    const TI *src = static_cast<const TI *>(a.in) + oin + lane * a.ein;
    TO *dst = static_cast<TO *>(a.out) + oout + lane * a.eout;
    const int64_t ps = a.nlines;
    const int n = static_cast<int>(a.n), B = a.border, L = static_cast<int>(a.L);
    // the causal result is kept for positions B..L-1 only: the anti-causal recursion
    // stops at the first interior sample, the positions before it feed nothing
    double *p1 = a.p1 + line - static_cast<int64_t>(B) * ps;
    const int64_t sin = a.sin, sout = a.sout;
    const double S = a.S;
    double b[NN + 1];
#pragma unroll
    for (int k = 1; k <= NN; ++k) b[k] = a.b[k];
    double fa[3], ba[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        fa[k] = a.a1[k];
        ba[k] = a.a2[k];
    }
    auto ext = [&](int ii) { return iir_ext<TI>(src, sin, n, B, a.bc, a.cval, ii); };

    // ---- causal recursion: y[i] = MA(x)[i] - b1 y[i-1] - ... - bnn y[i-nn] ----
    double h[NN];  // h[k] = y[i - 1 - k]
    double x1 = 0.0, x2 = 0.0;  // x[i - 1], x[i - 2]
    {
        double init;
        if (a.finit_diff == 2) {
            init = ext(0) / S;
        } else if (a.finit_diff == 1) {
            init = (ext(1) - ext(0)) / S;
        } else {
            double v = 0.0;
#pragma unroll
            for (int j = 0; j <= MA::FM2; ++j) v = v + fa[j] * ext(MA::FM2 - j);
            init = v / S;
        }
#pragma unroll
        for (int k = 0; k < NN; ++k) h[k] = init;
        if (MA::FM2 >= 1 && a.fnset >= 1) x1 = ext(a.fnset - 1);
        if (MA::FM2 >= 2 && a.fnset >= 2) x2 = ext(a.fnset - 2);
    }
    auto fstep = [&](double x) {
        double v;
        if (ORDER == 0) {
            v = x;
        } else {
            v = 0.0;
            v = v + fa[0] * x;
            v = v + fa[1] * x1;
            if (MA::FM2 >= 2) v = v + fa[2] * x2;
        }
#pragma unroll
        for (int k = 1; k <= NN; ++k) v = v - b[k] * h[k - 1];
#pragma unroll
        for (int k = NN - 1; k > 0; --k) h[k] = h[k - 1];
        h[0] = v;
        x2 = x1;
        x1 = x;
        return v;
    };
    for (int ii = a.fnset; ii < B; ++ii) fstep(ext(ii));
    // interior: image samples 0..n-1 in blocks of P; the next block's loads are issued
    // (unconditionally, clamped to the line) before this block's steps
    {
        const int nfull = n / P;
        TL cur[P], nxt[P];
        if (nfull > 0) {
#pragma unroll
            for (int p = 0; p < P; ++p) cur[p] = src[static_cast<int64_t>(p) * sin];
        }
        for (int blk = 0; blk < nfull; ++blk) {
            const int j0 = blk * P;
#pragma unroll
            for (int p = 0; p < P; ++p) nxt[p] = src[static_cast<int64_t>(min(j0 + P + p, n - 1)) * sin];
            __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the steps
#pragma unroll
            for (int p = 0; p < P; ++p)
                p1[static_cast<int64_t>(B + j0 + p) * ps] = fstep(to_buf<double>(static_cast<TI>(cur[p])));
#pragma unroll
            for (int p = 0; p < P; ++p) cur[p] = nxt[p];
        }
        for (int j = nfull * P; j < n; ++j)
            p1[static_cast<int64_t>(B + j) * ps] = fstep(to_buf<double>(src[static_cast<int64_t>(j) * sin]));
    }
    for (int ii = B + n; ii < L; ++ii) p1[static_cast<int64_t>(ii) * ps] = fstep(ext(ii));

    // ---- anti-causal recursion: z[i] = c MA(y)[i] - b1 z[i+1] - ... - bnn z[i+nn] ----
    const double c = a.c;
    double g[NN];  // g[k] = z[i + 1 + k]
    double y1 = 0.0, y2 = 0.0;  // y[i + 1], y[i + 2]
    const int top = L - 1 - a.bnset;
    {
        const double yl = p1[static_cast<int64_t>(L - 1) * ps];
        const double yl2 = p1[static_cast<int64_t>(L - 2) * ps];
        double init;
        if (a.binit == 2) {
            init = (c * (-yl2 + yl)) / S;
        } else if (a.binit == 1) {
            init = (c * yl) / S;
        } else {
            double v = 0.0;
#pragma unroll
            for (int j = MA::BM4; j <= MA::BM5; ++j)
                v = v + ba[j] * p1[static_cast<int64_t>(L - 1 - MA::BM5 + j) * ps];
            init = c * (v / S);
        }
#pragma unroll
        for (int k = 0; k < NN; ++k) g[k] = init;
        if (MA::BM5 >= 1) y1 = p1[static_cast<int64_t>(top + 1) * ps];
        if (MA::BM5 >= 2) y2 = p1[static_cast<int64_t>(top + 2) * ps];
    }
    auto bstep = [&](double y) {
        double v;
        if (ORDER == 0) {
            v = c * y;
        } else {
            double m = 0.0;
            if (MA::BM4 == 0) m = m + ba[0] * y;
            m = m + ba[1] * y1;
            if (MA::BM5 >= 2) m = m + ba[2] * y2;
            v = c * m;
        }
#pragma unroll
        for (int k = 1; k <= NN; ++k) v = v - b[k] * g[k - 1];
#pragma unroll
        for (int k = NN - 1; k > 0; --k) g[k] = g[k - 1];
        g[0] = v;
        y2 = y1;
        y1 = y;
        return v;
    };
    // trailing border: state only
    for (int i = top; i >= B + n; --i) bstep(p1[static_cast<int64_t>(i) * ps]);
    // interior, from the last sample down, in blocks of P with the next block in flight
    {
        const int nfull = n / P;
        double cur[P], nxt[P];
        if (nfull > 0) {
#pragma unroll
            for (int p = 0; p < P; ++p) cur[p] = p1[static_cast<int64_t>(B + n - 1 - p) * ps];
        }
        for (int blk = 0; blk < nfull; ++blk) {
            const int j0 = n - 1 - blk * P;  // image index of the block's first (highest) sample
#pragma unroll
            for (int p = 0; p < P; ++p) nxt[p] = p1[static_cast<int64_t>(B + max(j0 - P - p, 0)) * ps];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int p = 0; p < P; ++p)
                dst[static_cast<int64_t>(j0 - p) * sout] = static_cast<TO>(bstep(cur[p]));
#pragma unroll
            for (int p = 0; p < P; ++p) cur[p] = nxt[p];
        }
        for (int j = n - 1 - nfull * P; j >= 0; --j)
            dst[static_cast<int64_t>(j) * sout] = static_cast<TO>(bstep(p1[static_cast<int64_t>(B + j) * ps]));
    }
}

}  // namespace dipamd
