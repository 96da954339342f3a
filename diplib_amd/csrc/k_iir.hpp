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
// about 2 (n + border) steps of that chain per line; lines run in parallel.  With
// few lines (one wave per SIMD or less) the chain latency and the memory latency
// bound a pass, so memory operations per step are kept low: the next P samples are
// loaded before the current P steps run; the causal scratch is written and read in
// pairs (16-byte accesses); on a contiguous filter axis with aligned lines, samples
// are loaded and stored V at a time.
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

// Vector access to V consecutive samples of a line (V = 1: one sample).  Loaded words
// stay one per 32-bit register: a byte array would be packed, which forces the loads
// to complete where they are issued.
template <typename T, int V>
struct IirVec;
template <>
struct IirVec<uint8_t, 1> {
    using W = uint32_t;
    static __device__ __forceinline__ W load(const uint8_t *p) { return *p; }
    template <int K>
    static __device__ __forceinline__ double get(W w) { return static_cast<double>(w); }
};
template <>
struct IirVec<uint8_t, 4> {
    using W = uint32_t;
    static __device__ __forceinline__ W load(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }
    template <int K>
    static __device__ __forceinline__ double get(W w) { return static_cast<double>((w >> (8 * K)) & 0xffu); }
};
template <>
struct IirVec<float, 1> {
    using W = float;
    static __device__ __forceinline__ W load(const float *p) { return *p; }
    template <int K>
    static __device__ __forceinline__ double get(W w) { return static_cast<double>(w); }
};
template <>
struct IirVec<float, 4> {
    using W = float4;
    static __device__ __forceinline__ W load(const float *p) { return *reinterpret_cast<const float4 *>(p); }
    template <int K>
    static __device__ __forceinline__ double get(W w) {
        return static_cast<double>(K == 0 ? w.x : (K == 1 ? w.y : (K == 2 ? w.z : w.w)));
    }
};
template <>
struct IirVec<double, 1> {
    using W = double;
    static __device__ __forceinline__ W load(const double *p) { return *p; }
    template <int K>
    static __device__ __forceinline__ double get(W w) { return w; }
};
template <>
struct IirVec<double, 2> {
    using W = double2;
    static __device__ __forceinline__ W load(const double *p) { return *reinterpret_cast<const double2 *>(p); }
    template <int K>
    static __device__ __forceinline__ double get(W w) { return K == 0 ? w.x : w.y; }
};

// Input vector width on a contiguous axis, and the output store width.
template <typename TI>
struct IirVin {
    static constexpr int V = sizeof(TI) == 8 ? 2 : 4;
};
template <typename TO, int V>
struct IirStore {
    static constexpr int W = V == 1 ? 1 : (sizeof(TO) == 8 ? 2 : 4);
    static __device__ __forceinline__ void store(TO *p, const double *v) {
        if (W == 1) {
            *p = static_cast<TO>(v[0]);
        } else if (sizeof(TO) == 8) {
            *reinterpret_cast<double2 *>(p) = make_double2(v[0], v[1]);
        } else {
            *reinterpret_cast<float4 *>(p) = make_float4(static_cast<float>(v[0]), static_cast<float>(v[1]),
                                                         static_cast<float>(v[2]), static_cast<float>(v[3]));
        }
    }
};

template <int K, int P, typename F>
__device__ __forceinline__ void iir_unroll(F &&f) {
    if constexpr (K < P) {
        f(std::integral_constant<int, K>());
        iir_unroll<K + 1, P>(f);
    }
}

// VEC: the filter axis is contiguous (stride 1) in input and output and every line
// starts aligned to the vector widths.
template <typename TI, typename TO, int NN, int ORDER, bool VEC>
__global__ __launch_bounds__(kIirBlock) void k_iir(const IirArgs a) {
    using MA = IirMa<ORDER>;
    constexpr int V = VEC ? IirVin<TI>::V : 1;
    using IV = IirVec<TI, V>;
    using WI = typename IV::W;
    using OS = IirStore<TO, V>;
    constexpr int P = kIirP;
    constexpr int NV = P / V;
    const int64_t line = static_cast<int64_t>(blockIdx.x) * kIirBlock + threadIdx.x;
    if (line >= a.nlines) return;
    const int64_t lane = line % a.e;
    int64_t oin, oout;
    outer_offsets(a.outer, line / a.e, &oin, &oout);
    const TI *src = static_cast<const TI *>(a.in) + oin + lane * a.ein;
    TO *dst = static_cast<TO *>(a.out) + oout + lane * a.eout;
    const int64_t ps = a.nlines;
    const int n = static_cast<int>(a.n), B = a.border, L = static_cast<int>(a.L);
    // Causal result for positions B..L-1 only (the anti-causal recursion stops at the
    // first interior sample; the positions before it feed nothing), q = position - B,
    // stored in pairs: element q of this line at pq[(q >> 1) * 2 ps + (q & 1)].
    double *pq = a.p1 + 2 * line;
    const int64_t pstride = 2 * ps;
    auto p1_at = [&](int q) -> double & { return pq[static_cast<int64_t>(q >> 1) * pstride + (q & 1)]; };
    auto p1_pair = [&](int q) { return reinterpret_cast<double2 *>(pq + static_cast<int64_t>(q >> 1) * pstride); };
    const int64_t sin = VEC ? 1 : a.sin, sout = VEC ? 1 : a.sout;
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
    const int nfull = n / P;
    // interior: blocks of P samples from the first; the next block's loads (clamped to
    // the line) are issued before this block's steps
    {
        WI cur[NV], nxt[NV];
        if (nfull > 0) {
#pragma unroll
            for (int v = 0; v < NV; ++v) cur[v] = IV::load(src + static_cast<int64_t>(v * V) * sin);
        }
        for (int blk = 0; blk < nfull; ++blk) {
            const int j0 = blk * P;
            const int jn = blk + 1 < nfull ? j0 + P : j0;
#pragma unroll
            for (int v = 0; v < NV; ++v) nxt[v] = IV::load(src + static_cast<int64_t>(jn + v * V) * sin);
            __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the steps
            double y[P];
            iir_unroll<0, P>([&](auto kc) {
                constexpr int p = decltype(kc)::value;
                y[p] = fstep(IV::template get<p % V>(cur[p / V]));
            });
#pragma unroll
            for (int p = 0; p < P; p += 2) *p1_pair(j0 + p) = make_double2(y[p], y[p + 1]);
#pragma unroll
            for (int v = 0; v < NV; ++v) cur[v] = nxt[v];
        }
        for (int j = nfull * P; j < n; ++j) p1_at(j) = fstep(to_buf<double>(src[static_cast<int64_t>(j) * sin]));
    }
    for (int ii = B + n; ii < L; ++ii) p1_at(ii - B) = fstep(ext(ii));

    // ---- anti-causal recursion: z[i] = c MA(y)[i] - b1 z[i+1] - ... - bnn z[i+nn] ----
    const double c = a.c;
    double g[NN];  // g[k] = z[i + 1 + k]
    double y1 = 0.0, y2 = 0.0;  // y[i + 1], y[i + 2]
    const int qtop = L - 1 - a.bnset - B;
    {
        const double yl = p1_at(L - 1 - B);
        const double yl2 = p1_at(L - 2 - B);
        double init;
        if (a.binit == 2) {
            init = (c * (-yl2 + yl)) / S;
        } else if (a.binit == 1) {
            init = (c * yl) / S;
        } else {
            double v = 0.0;
#pragma unroll
            for (int j = MA::BM4; j <= MA::BM5; ++j) v = v + ba[j] * p1_at(L - 1 - B - MA::BM5 + j);
            init = c * (v / S);
        }
#pragma unroll
        for (int k = 0; k < NN; ++k) g[k] = init;
        if (MA::BM5 >= 1) y1 = p1_at(qtop + 1);
        if (MA::BM5 >= 2) y2 = p1_at(qtop + 2);
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
    // trailing border (state only), then the samples after the last full block
    for (int q = qtop; q >= n; --q) bstep(p1_at(q));
    for (int j = n - 1; j >= nfull * P; --j) dst[static_cast<int64_t>(j) * sout] = static_cast<TO>(bstep(p1_at(j)));
    // full blocks, from the last down, the next (lower) block's pairs in flight
    {
        constexpr int NP = P / 2;
        double2 cur[NP], nxt[NP];
        if (nfull > 0) {
#pragma unroll
            for (int k = 0; k < NP; ++k) cur[k] = *p1_pair((nfull - 1) * P + 2 * k);
        }
        for (int blk = nfull - 1; blk >= 0; --blk) {
            const int j0 = blk * P;
            const int jn = blk > 0 ? j0 - P : j0;
#pragma unroll
            for (int k = 0; k < NP; ++k) nxt[k] = *p1_pair(jn + 2 * k);
            __builtin_amdgcn_sched_barrier(0);
            double z[P];  // z[t]: sample j0 + t
#pragma unroll
            for (int t = P - 1; t >= 0; --t) z[t] = bstep((t & 1) ? cur[t >> 1].y : cur[t >> 1].x);
#pragma unroll
            for (int t = 0; t < P; t += OS::W) OS::store(dst + static_cast<int64_t>(j0 + t) * sout, z + t);
#pragma unroll
            for (int k = 0; k < NP; ++k) cur[k] = nxt[k];
        }
    }
}

}  // namespace dipamd
