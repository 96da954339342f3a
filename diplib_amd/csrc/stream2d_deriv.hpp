// K3D: the derivative family (dip::Gradient, GradientMagnitude, Laplace, Hessian, Dgg,
// LaplacePlusDgg, LaplaceMinusDgg on a 2-D uint8 image, FIR method) in one launch.
//
// Each output of these functions is built from Gaussian derivatives y_b(x_a(in)) with
// a, b in {0, 1, 2} (derivative.cpp:142-169, 348-418, 420-484 of the reference), each
// the separable FIR of the plain path: the x pass into sfloat, then the y pass.  The
// kernel streams the image once, as the register-streaming kernel of single Gaussians
// does (stream2d.hpp: a wave owns a 256-column strip, lane l columns 4l..4l+3, rows
// arrive through a one-period prefetch, the boundary folds along x and y come from the
// same lane and row tables).  Per input row it computes every x derivative order the
// function needs (one conversion of the 12 samples, then 4 outputs per order) into a
// register window per order; per output row it applies each term's y filter to its
// window and combines the terms in registers in the reference's order and type
// (k_elementwise.hip's sequence: products and sums rounded in sfloat, the square root
// correctly rounded), so the results are bit-identical to the separate Gauss calls plus
// the element-wise pass, with one read of the input and one write per output.
//
// Orders 0 and 2 use the even filter form, order 1 the odd one (gauss.cpp:188-199).  The
// radii of the common sigmas (0.75 .. 1: orders 0, 1, 2 of radius 3, 3-4, 4) are
// compile-time constants, so the tap loops are straight-line code; other radii up to
// kStreamRMax come from the arguments, each tap behind a wave-uniform test.
#pragma once

#include "stream2d.hpp"

namespace dipamd {

namespace deriv {

constexpr int kDW = 2 * kStreamRMax + 1;  // window rows = unroll period

template <int MODE>
struct DerivTerms {
    static constexpr bool X1 = MODE != kDerivLaplace;                                 // x order 1 used
    static constexpr bool X2 = MODE != kDerivGradient && MODE != kDerivGradientMagnitude;  // x order 2 used
};

// 4 x-filtered samples of one order from the 12 converted samples of the row
template <int SYM, bool FAST>
__device__ __forceinline__ void deriv_x(const float (&s)[12], const float (&t)[kStreamRMax + 1], int r,
                                        float (&w)[4]) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        float acc = t[0] * s[4 + p];
#pragma unroll
        for (int k = 1; k <= kStreamRMax; ++k) {
            if (k <= r) {
                const float v = SYM == SYM_EVEN ? (s[4 + p + k] + s[4 + p - k]) : (s[4 + p + k] - s[4 + p - k]);
                acc = mac<float, FAST>(acc, t[k], v);
            }
        }
        w[p] = acc;
    }
}

// One output sample of a y filter over a window whose centre row sits in slot cs
template <int SYM, bool FAST>
__device__ __forceinline__ float deriv_y(const float (&w)[kDW][4], int cs, int p, const float (&t)[kStreamRMax + 1],
                                         int r) {
    float acc = t[0] * w[cs][p];
#pragma unroll
    for (int k = 1; k <= kStreamRMax; ++k) {
        if (k <= r) {
            const float hi = w[(cs + k) % kDW][p];
            const float lo = w[(cs - k + kDW) % kDW][p];
            acc = mac<float, FAST>(acc, t[k], SYM == SYM_EVEN ? (hi + lo) : (hi - lo));
        }
    }
    return acc;
}

__device__ __forceinline__ void store4(float *base, int64_t row_stride, int o, int c, const float (&v)[4]) {
    float *d = base + static_cast<int64_t>(o) * row_stride + c;
    *reinterpret_cast<float4 *>(d) = make_float4(v[0], v[1], v[2], v[3]);
}

// R0, R1, R2: the radii of the derivative orders 0, 1, 2 (the same along x and y) as
// compile-time constants, so that every tap loop is straight-line code; 0 takes them
// from the arguments (each tap then behind a wave-uniform test).
template <int MODE, bool FAST, bool XMASK, int R0, int R1, int R2>
__device__ __forceinline__ void deriv_strip(const DerivStreamArgs &a, const uint8_t *__restrict__ src,
                                            const int64_t oout, int c, int z0, int zend) {
    using TT = DerivTerms<MODE>;
    constexpr int RY = kStreamRMax;
    constexpr int W = kDW;
    const StreamArgs &g = a.s;
    const int nx = g.nx, ny = g.ny;
    const int rx0 = R0 ? R0 : a.rx[0], rx1 = R1 ? R1 : a.rx[1], rx2 = R2 ? R2 : a.rx[2];
    const int ry0 = R0 ? R0 : a.ry[0], ry1 = R1 ? R1 : a.ry[1], ry2 = R2 ? R2 : a.ry[2];
    const int E = (zend - z0) + 2 * RY;  // rows streamed: z0-RY .. zend+RY-1
    const StreamLane L = stream_lane(c, nx, g.bcx);
    const bool cok = c < nx;

    float w0[W][4];
    float w1[TT::X1 ? W : 1][4];
    float w2[TT::X2 ? W : 1][4];
    uint32_t pre[W][3];
    int rflag[W];
    auto fetch = [&](int e, uint32_t(&wd)[3], int *flag) {
        const int z = z0 - RY + (e < E ? e : E - 1);  // past the end: reload a valid row, never used
        const bool lo = z < 0, hi = z >= ny;
        const int m = lo ? g.ylo_m : (hi ? g.yhi_m : 1);
        const int cc = lo ? g.ylo_c : (hi ? g.yhi_c : 0);
        *flag = (lo || hi) ? g.yflag : 0;
        const uint8_t *rp = src + static_cast<int64_t>(m * z + cc) * g.syin;
        wd[0] = *reinterpret_cast<const uint32_t *>(rp + L.ol);
        wd[1] = *reinterpret_cast<const uint32_t *>(rp + L.oc);
        wd[2] = *reinterpret_cast<const uint32_t *>(rp + L.orr);
    };
    auto row = [&](const int e, const int u) {
        const uint32_t wl = __builtin_amdgcn_perm(pre[u][1], pre[u][0], L.sell);
        const uint32_t wc = pre[u][1];
        const uint32_t wr = __builtin_amdgcn_perm(pre[u][1], pre[u][2], L.selr);
        const int rf = rflag[u];
        fetch(e + W, pre[u], &rflag[u]);
        float s[12];
        const uint32_t wv[3] = {wl, wc, wr};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s[4 * k + 0] = static_cast<float>(wv[k] & 0xffu);
            s[4 * k + 1] = static_cast<float>((wv[k] >> 8) & 0xffu);
            s[4 * k + 2] = static_cast<float>((wv[k] >> 16) & 0xffu);
            s[4 * k + 3] = static_cast<float>(wv[k] >> 24);
        }
        if constexpr (XMASK) {
#pragma unroll
            for (int i = 0; i < 12; ++i) {
                s[i] = ((L.negbits >> i) & 1u) ? -s[i] : s[i];
                s[i] = ((L.constbits >> i) & 1u) ? g.cvalx : s[i];
            }
        }
        deriv_x<SYM_EVEN, FAST>(s, a.tx[0], rx0, w0[u]);
        if constexpr (TT::X1) deriv_x<SYM_ODD, FAST>(s, a.tx[1], rx1, w1[u]);
        if constexpr (TT::X2) deriv_x<SYM_EVEN, FAST>(s, a.tx[2], rx2, w2[u]);
        if (rf != 0) {  // a row folded across the top or bottom edge (wave-uniform)
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                w0[u][p] = rf == 2 ? g.cvaly : -w0[u][p];
                if constexpr (TT::X1) w1[u][p] = rf == 2 ? g.cvaly : -w1[u][p];
                if constexpr (TT::X2) w2[u][p] = rf == 2 ? g.cvaly : -w2[u][p];
            }
        }
        if (e < 2 * RY) return;
        const int cs = ((u - RY) % W + W) % W;
        const int o = z0 + e - 2 * RY;
        if (!cok) return;
        float r0[4], r1[4], r2[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if constexpr (MODE == kDerivGradient || MODE == kDerivGradientMagnitude) {
                const float d0 = deriv_y<SYM_EVEN, FAST>(w1, cs, p, a.ty[0], ry0);  // d/dx
                const float d1 = deriv_y<SYM_ODD, FAST>(w0, cs, p, a.ty[1], ry1);   // d/dy
                if constexpr (MODE == kDerivGradient) {
                    r0[p] = d0;
                    r1[p] = d1;
                } else {
                    float q = d0 * d0;
                    const float q1 = d1 * d1;
                    q = q + q1;
                    r0[p] = __builtin_sqrtf(q);  // correctly rounded (std::sqrt on sfloat)
                }
            } else if constexpr (MODE == kDerivLaplace) {
                const float h00 = deriv_y<SYM_EVEN, FAST>(w2, cs, p, a.ty[0], ry0);
                const float h11 = deriv_y<SYM_EVEN, FAST>(w0, cs, p, a.ty[2], ry2);
                r0[p] = h00 + h11;
            } else if constexpr (MODE == kDerivHessian) {
                r0[p] = deriv_y<SYM_EVEN, FAST>(w2, cs, p, a.ty[0], ry0);  // d2/dx2
                r1[p] = deriv_y<SYM_EVEN, FAST>(w0, cs, p, a.ty[2], ry2);  // d2/dy2
                r2[p] = deriv_y<SYM_ODD, FAST>(w1, cs, p, a.ty[1], ry1);   // d2/dxdy
            } else {
                // Dgg and its Laplace forms for two dims (k_elementwise.hip, ops 2-4): the
                // mixed term does not enter (the reference's loop over it starts at 1)
                const float g0 = deriv_y<SYM_EVEN, FAST>(w1, cs, p, a.ty[0], ry0);
                const float g1 = deriv_y<SYM_ODD, FAST>(w0, cs, p, a.ty[1], ry1);
                const float h00 = deriv_y<SYM_EVEN, FAST>(w2, cs, p, a.ty[0], ry0);
                const float h11 = deriv_y<SYM_EVEN, FAST>(w0, cs, p, a.ty[2], ry2);
                float gs = g0 * g0;
                float v = gs * h00;
                float t = g1 * g1;
                gs = gs + t;
                t = t * h11;
                v = v + t;
                v = gs == 0.0f ? 0.0f : v / gs;
                if constexpr (MODE != kDerivDgg) {
                    float tr = 0.0f;
                    tr = tr + h00;
                    tr = tr + h11;
                    v = MODE == kDerivLaplacePlusDgg ? tr + v : tr - v;
                }
                r0[p] = v;
            }
        }
        store4(a.out[0] + oout, a.syout[0], o, c, r0);
        if constexpr (MODE == kDerivGradient || MODE == kDerivHessian) store4(a.out[1] + oout, a.syout[1], o, c, r1);
        if constexpr (MODE == kDerivHessian) store4(a.out[2] + oout, a.syout[2], o, c, r2);
    };
#pragma unroll
    for (int u = 0; u < W; ++u) fetch(u, pre[u], &rflag[u]);
    int eb = 0;
    for (; eb + W <= E; eb += W) {  // whole periods
#pragma unroll
        for (int u = 0; u < W; ++u) row(eb + u, u);
    }
#pragma unroll
    for (int u = 0; u < W; ++u)  // the last partial period
        if (eb + u < E) row(eb + u, u);
}

template <int MODE, bool FAST, int R0, int R1, int R2>
__global__ __launch_bounds__(kBlock) void k_stream2d_deriv(const DerivStreamArgs a) {
    const int lane = threadIdx.x & 63;
    const StreamArgs &g = a.s;
    const int64_t gw = xcd_block() * 4 + wave_id();
    if (gw >= g.nwaves) return;
    const int strip = static_cast<int>(gw % g.nstrips);
    const int64_t rest = gw / g.nstrips;
    const int chunk = static_cast<int>(rest % g.nchunks);
    const int64_t plane = rest / g.nchunks;
    int64_t oin, oout;
    outer_offsets(g.outer, plane, &oin, &oout);
    const int c = strip * 256 + 4 * lane;
    const int z0 = chunk * g.chunk;
    const int zend = min(z0 + g.chunk, g.ny);
    const bool xmask = (strip == 0 || strip == g.nstrips - 1) &&
                       (g.bcx == DIPAMD_BC_ASYMMETRIC_MIRROR || g.bcx == DIPAMD_BC_ASYMMETRIC_PERIODIC ||
                        g.bcx == DIPAMD_BC_ADD_ZEROS || g.bcx == DIPAMD_BC_ADD_MAX_VALUE ||
                        g.bcx == DIPAMD_BC_ADD_MIN_VALUE);
    if (xmask)
        deriv_strip<MODE, FAST, true, R0, R1, R2>(a, g.in + oin, oout, c, z0, zend);
    else
        deriv_strip<MODE, FAST, false, R0, R1, R2>(a, g.in + oin, oout, c, z0, zend);
}

template <int MODE, int R0, int R1, int R2>
void deriv_launch(const DerivStreamArgs &a, bool fast, hipStream_t st) {
    const dim3 grid(static_cast<unsigned>((a.s.nwaves + 3) / 4));
    if (fast)
        hipLaunchKernelGGL((k_stream2d_deriv<MODE, true, R0, R1, R2>), grid, dim3(kBlock), 0, st, a);
    else
        hipLaunchKernelGGL((k_stream2d_deriv<MODE, false, R0, R1, R2>), grid, dim3(kBlock), 0, st, a);
}

// Every function of the family for one radius triple (0, 0, 0: from the arguments).
template <int R0, int R1, int R2>
hipError_t deriv_dispatch(int mode, const DerivStreamArgs &a, bool fast, hipStream_t stream) {
    switch (mode) {
        case kDerivGradient: deriv_launch<kDerivGradient, R0, R1, R2>(a, fast, stream); break;
        case kDerivGradientMagnitude: deriv_launch<kDerivGradientMagnitude, R0, R1, R2>(a, fast, stream); break;
        case kDerivLaplace: deriv_launch<kDerivLaplace, R0, R1, R2>(a, fast, stream); break;
        case kDerivHessian: deriv_launch<kDerivHessian, R0, R1, R2>(a, fast, stream); break;
        case kDerivDgg: deriv_launch<kDerivDgg, R0, R1, R2>(a, fast, stream); break;
        case kDerivLaplacePlusDgg: deriv_launch<kDerivLaplacePlusDgg, R0, R1, R2>(a, fast, stream); break;
        case kDerivLaplaceMinusDgg: deriv_launch<kDerivLaplaceMinusDgg, R0, R1, R2>(a, fast, stream); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace deriv
}  // namespace dipamd
