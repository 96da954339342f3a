// Backward form of the Gaussian FIR path: the exact adjoint of the boundary-
// extended separable convolution and the derivative with respect to each sigma.
//
// With the passes in forward order i = 1..m (dims d_1..d_m, ascending input stride
// as Framework::Separable runs them, framework_separable.cpp:146-163):
//     v_0 = x,  v_i = K_i v_{i-1}            (forward intermediates)
//     u_m = g,  u_{i-1} = K_i^T u_i          (adjoint chain; u_0 = dL/dx)
//     dL/dsigma_{d_i} = < u_i , K'_i v_{i-1} >
// where K'_i filters with d(taps)/d(sigma) of the normalised discrete half Gaussian
// at the radius the forward pass used (gauss_taps.hpp), and K_i^T is
//   - periodic / asym periodic: the same boundary condition with reversed taps;
//   - mirror / asym mirror / zero order: reversed taps on the zero-padded signal
//     plus the halo terms folded back through the boundary map (k_adjoint_fold);
//   - constant conditions: reversed taps on the zero-padded signal.
// Reversed taps of an even filter are the filter itself, of an odd one its negative.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dipamd.h"
#include "dipamd_internal.hpp"
#include "gauss_taps.hpp"
#include "sep_common.hpp"
#include "sep_kernels.hpp"

namespace dipamd {

hipError_t launch_adjoint_fold(bool is_double, const void *args, int64_t nlines, hipStream_t stream);
hipError_t launch_dot(bool is_double, const void *a, const void *geo, const void *b, double *partials, double *out,
                      hipStream_t stream);
hipError_t launch_fill_doubles(double *p, int n, double v, hipStream_t stream);
hipError_t launch_bwd2d(int symx, int symy, int rc, const BwdArgs &a, int64_t nblocks, double *gs, int gs_mode,
                        int gs_n, hipStream_t stream);
hipError_t launch_sum_doubles(const double *in, int n, double *out, hipStream_t stream);
size_t bwd2d_lds_bytes(int rc);
int bwd2d_class(int rx, int ry);
int bwd2d_tile_columns(int rc);
hipError_t launch_bwd2d_sum(const double *partials, int64_t nblocks, double *gs, int gs_mode, int gs_n,
                            hipStream_t stream);
bool bwd2s_launch_y3(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y5(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y7(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y9(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y11(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);
bool bwd2s_launch_y13(int cx, int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream);

namespace {

bool contiguous(const View &v, int ndim) {
    int64_t s = 1;
    for (int d = 0; d < ndim; ++d) {
        if (v.sizes[d] > 1 && v.strides[d] != s) return false;
        s *= v.sizes[d];
    }
    return true;
}

template <typename TB>
int fold_pass(const View &g, const View &out, int ndim, int d, const InternFilter &FT, int bc, hipStream_t stream) {
    PassArgs<TB> a;
    std::memset(&a, 0, sizeof(a));
    a.in = g.p;
    a.out = out.p;
    a.n = g.sizes[d];
    a.sin = g.strides[d];
    a.sout = out.strides[d];
    a.border = FT.border;
    a.shift = FT.sym == DIPAMD_SYM_ODD ? 1 : 0;
    a.bc = bc;
    a.zlo = INT32_MIN;
    for (size_t k = 0; k < FT.data.size() && k < static_cast<size_t>(kLineCap); ++k) a.taps[k] = static_cast<TB>(FT.data[k]);
    // lines: every other dim
    int dims[DIPAMD_MAX_DIMS];
    int n = 0;
    int64_t nlines = 1;
    for (int k = 0; k < ndim; ++k) {
        if (k == d) continue;
        nlines *= g.sizes[k];
        if (g.sizes[k] > 1) dims[n++] = k;
    }
    a.outer.n = n;
    for (int i = 0; i < kMaxOuter; ++i) {
        a.outer.size[i] = 1;
        a.outer.sin[i] = 0;
        a.outer.sout[i] = 0;
    }
    for (int i = 0; i < n; ++i) {
        a.outer.size[i] = g.sizes[dims[i]];
        a.outer.sin[i] = g.strides[dims[i]];
        a.outer.sout[i] = out.strides[dims[i]];
    }
    a.nlines = nlines;
    TraceScope ts("adjoint_fold", 0, stream);
    hipError_t e = launch_adjoint_fold(sizeof(TB) == 8, &a, nlines, stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("fold launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

// dst = K^T src along dim d (K: filter F with boundary condition bc).
int adjoint_pass(const View &src, const View &dst, int ndim, int d, const InternFilter &F, int bc, bool is_double,
                 bool fast, hipStream_t stream) {
    InternFilter FT = F;
    if (F.sym == DIPAMD_SYM_ODD)
        for (size_t k = 1; k < FT.data.size(); ++k) FT.data[k] = -FT.data[k];
    const bool periodic = bc == DIPAMD_BC_PERIODIC || bc == DIPAMD_BC_ASYMMETRIC_PERIODIC;
    const bool folds = bc == DIPAMD_BC_SYMMETRIC_MIRROR || bc == DIPAMD_BC_ASYMMETRIC_MIRROR ||
                       bc == DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE;
    int st = run_pass(src, dst, ndim, d, FT, periodic ? bc : DIPAMD_BC_ADD_ZEROS, is_double, fast, stream);
    if (st != DIPAMD_OK || !folds) return st;
    return is_double ? fold_pass<double>(src, dst, ndim, d, FT, bc, stream)
                     : fold_pass<float>(src, dst, ndim, d, FT, bc, stream);
}

bool bwd_fused_enabled() {
    static bool on = [] {
        const char *v = std::getenv("DIPAMD_BWD_FUSED");
        return !(v && v[0] == '0');
    }();
    return on;
}

int bwd_mode(int bc) {
    switch (bc) {
        case DIPAMD_BC_PERIODIC:
        case DIPAMD_BC_ASYMMETRIC_PERIODIC: return kBwdPeriodic;
        case DIPAMD_BC_SYMMETRIC_MIRROR:
        case DIPAMD_BC_ASYMMETRIC_MIRROR:
        case DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE: return kBwdFold;
        case DIPAMD_BC_ADD_ZEROS: return kBwdZero;
        default: return -1;  // add max / min: the constant's halo enters dL/dsigma_y
    }
}

// Odd radius class of the streaming backward (3..13), 0 beyond.
int bwd2s_class(int r) {
    if (r > 13) return 0;
    return r <= 3 ? 3 : (r | 1);
}

int bwd2s_forced() {
    static int on = [] {
        const char *v = std::getenv("DIPAMD_BWD_STREAM");
        return v ? std::atoi(v) : -1;
    }();
    return on;
}

// The streaming 2-D backward (k_bwd2s.hpp), under the same conditions as the fused one
// below plus radius classes it has instances for; -100 when it does not apply.
int streaming_backward_2d(const View &x, const View &g, const View &dx, int nd, const InternFilter &F0,
                          const InternFilter &F1, const InternFilter &dF0, const InternFilter &dF1, int bc0, int bc1,
                          double *gs, int gs_mode, int gs_n, hipStream_t stream) {
    const int forced = bwd2s_forced();
    if (forced == 0) return -100;
    if (x.dt != DIPAMD_DT_SFLOAT || g.dt != DIPAMD_DT_SFLOAT) return -100;
    const bool want_dx = dx.p != nullptr;
    if (x.strides[0] != 1 || g.strides[0] != 1 || (want_dx && dx.strides[0] != 1)) return -100;
    const InternFilter *Fs[2] = {&F0, &F1};
    const int bcs[2] = {bc0, bc1};
    for (int d = 0; d < 2; ++d) {
        const InternFilter &f = *Fs[d];
        if ((f.sym != DIPAMD_SYM_EVEN && f.sym != DIPAMD_SYM_ODD) || f.shift != 0) return -100;
        if (bwd_mode(bcs[d]) < 0) return -100;
        if (x.sizes[d] < 2 * f.border + 1 || x.sizes[d] >= (int64_t(1) << 30)) return -100;
    }
    const int cx = bwd2s_class(F0.border), cy = bwd2s_class(F1.border);
    // equal or adjacent classes (a first derivative along one axis)
    if (cx == 0 || !(cx == cy || cx == cy + 2 || cy == cx + 2)) return -100;
    if (static_cast<int>(F0.data.size()) > kB2sCap || static_cast<int>(F1.data.size()) > kB2sCap) return -100;
    if (F1.border + cy > kB2sFold) return -100;  // rows beyond the y extension table
    if (nd - 2 > kMaxOuter) return -100;
    // planes narrower than two strips keep the tiled kernel unless asked otherwise
    if (forced < 0 && x.sizes[0] < 128) return -100;
    auto plane_bytes = [&](const View &v) -> int64_t {
        if (v.strides[1] <= 0) return -1;
        return ((x.sizes[1] - 1) * v.strides[1] + x.sizes[0]) * static_cast<int64_t>(sizeof(float));
    };
    const int64_t kMaxPlane = (int64_t(1) << 31) - 256;
    const int64_t xb = plane_bytes(x), gb = plane_bytes(g), db = want_dx ? plane_bytes(dx) : 0;
    if (xb < 0 || gb < 0 || db < 0 || xb > kMaxPlane || gb > kMaxPlane || db > kMaxPlane) return -100;
    Bwd2sArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = static_cast<const float *>(x.p);
    a.g = static_cast<const float *>(g.p);
    a.dx = want_dx ? static_cast<float *>(dx.p) : nullptr;
    a.want_dx = want_dx ? 1 : 0;
    a.xbytes = static_cast<int>(xb);
    a.gbytes = static_cast<int>(gb);
    a.dxbytes = static_cast<int>(db);
    a.nx = static_cast<int>(x.sizes[0]);
    a.ny = static_cast<int>(x.sizes[1]);
    a.syx = x.strides[1];
    a.syg = g.strides[1];
    a.sydx = want_dx ? dx.strides[1] : 0;
    a.rx = F0.border;
    a.ry = F1.border;
    a.bcx = bc0;
    a.bcy = bc1;
    a.modex = bwd_mode(bc0);
    a.modey = bwd_mode(bc1);
    int64_t nplanes = 1;
    a.outer.n = 0;
    for (int d = 2; d < nd; ++d) {
        if (x.sizes[d] <= 1) continue;
        const int k = a.outer.n++;
        a.outer.size[k] = x.sizes[d];
        a.outer.sin[k] = x.strides[d];
        a.outer.sout[k] = want_dx ? dx.strides[d] : 0;
        a.gplane[k] = g.strides[d];
        nplanes *= x.sizes[d];
    }
    // adjoint taps (odd filters negated); entries beyond the radius stay zero
    auto put = [](float *dst, const InternFilter &f) {
        for (size_t k = 0; k < f.data.size(); ++k)
            dst[k] = static_cast<float>((f.sym == DIPAMD_SYM_ODD && k > 0) ? -f.data[k] : f.data[k]);
    };
    put(a.txa, F0);
    put(a.tdxa, dF0);
    put(a.tya, F1);
    put(a.tdya, dF1);
    // zero-order x fold: suffix sums of the adjoint taps, H[m] = sum over k = m+1..rx
    auto suffix = [](float *dst, const InternFilter &f) {
        const int r = static_cast<int>(f.data.size()) - 1;
        for (int m = 0; m < r; ++m) {
            double t = 0.0;
            for (int k = m + 1; k <= r; ++k) t += (f.sym == DIPAMD_SYM_ODD) ? -f.data[k] : f.data[k];
            dst[m] = static_cast<float>(t);
        }
    };
    suffix(a.hzv, F0);
    suffix(a.hzw, dF0);
    // the y extension of the rows just outside the plane (source row and sign, or a zero
    // row for the constant conditions)
    for (int i = 0; i < kB2sFold; ++i) {
        for (int side = 0; side < 2; ++side) {
            const int64_t q = side == 0 ? -1 - i : a.ny + i;
            int mul = 1;
            bool is_const = false;
            const int64_t src = bc_map(q, a.ny, bc1, &mul, &is_const);
            a.yfold[side * kB2sFold + i] = is_const ? -1 : static_cast<int>((src << 1) | (mul < 0 ? 1 : 0));
        }
    }
    a.nstrips = (a.nx + 63) / 64;
    // chunks: one round of waves (the kernel holds three waves per SIMD: 12 per CU;
    // 32 x 1024^2 sigma 3 periodic 0.153 -> 0.118 ms against a third more waves).  Under
    // an x fold the edge strips' rows cost about twice as much (the correlations at the
    // halo positions), so they take chunks half as long, dispatched first, and the
    // interior chunks are sized so that both together still make one round (mirror
    // order 0: 0.154 -> 0.136 ms, order 1: 0.163 -> 0.142 ms).  At least 64 rows (and
    // more than ry, so the rows folded onto stay in the chunk that folds them); the last
    // chunk takes the remainder.  DIPAMD_BWD_STREAM_CHUNK / _ECHUNK set the interior /
    // edge rows per chunk for experiments.
    static const int ncu = [] {
        int dev = 0, n = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        return n > 0 ? n : 256;
    }();
    static const int force_chunk = [] {
        const char *v = std::getenv("DIPAMD_BWD_STREAM_CHUNK");
        return v ? std::atoi(v) : 0;
    }();
    static const int force_echunk = [] {
        const char *v = std::getenv("DIPAMD_BWD_STREAM_ECHUNK");
        return v ? std::atoi(v) : 0;
    }();
    const bool split_edges = a.modex == kBwdFold && a.nstrips > 2;
    // (a stack the caches hold whole may take short chunks: its halo re-reads hit the
    // caches and the waves are what it lacks; 128^3 sigma 3: 48 -> 43 us against the
    // tiled kernel)
    const bool cached = static_cast<int64_t>(a.nx) * a.ny * nplanes * 12 <= (int64_t(64) << 20);
    const int minrows = std::max(cached ? 16 : 64, a.ry + 1);
    const int64_t units = static_cast<int64_t>(split_edges ? a.nstrips + 2 : a.nstrips) * nplanes;
    int64_t nchunks = std::max<int64_t>(1, (12 * static_cast<int64_t>(ncu)) / units);
    if (force_chunk > 0) nchunks = std::max<int64_t>(1, a.ny / std::max(force_chunk, a.ry + 1));
    else if (a.ny / nchunks < minrows) nchunks = std::max<int64_t>(1, a.ny / minrows);
    const int64_t chunk = a.ny / nchunks;
    a.chunk = static_cast<int>(chunk);
    a.nchunks = static_cast<int>(nchunks);
    int64_t nchunks_e = split_edges ? std::min<int64_t>(2 * nchunks, a.ny / (a.ry + 1)) : nchunks;
    if (force_echunk > 0) nchunks_e = std::max<int64_t>(1, a.ny / std::max(force_echunk, a.ry + 1));
    nchunks_e = std::max<int64_t>(1, nchunks_e);
    a.nchunks_e = static_cast<int>(nchunks_e);
    a.chunk_e = static_cast<int>(a.ny / nchunks_e);
    a.nplanes = nplanes;
    const int ne = a.nstrips > 1 ? 2 : 1;
    a.nwaves = (static_cast<int64_t>(ne) * nchunks_e + static_cast<int64_t>(a.nstrips - ne) * nchunks) * nplanes;
    const int64_t nblocks = (a.nwaves + 3) / 4;
    if (nblocks >= (int64_t(1) << 31)) return -100;
    double *partials = nullptr;
    hipError_t e = scratch_alloc(reinterpret_cast<void **>(&partials), static_cast<size_t>(2 * nblocks) * sizeof(double),
                                 stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
    a.partials = partials;
    bool ok = false;
    {
        TraceScope ts("bwd2s", static_cast<int64_t>(a.nx) * a.ny * nplanes * (want_dx ? 12 : 8), stream);
        switch (cy) {
            case 3: ok = bwd2s_launch_y3(cx, F0.sym, F1.sym, a, nblocks, stream); break;
            case 5: ok = bwd2s_launch_y5(cx, F0.sym, F1.sym, a, nblocks, stream); break;
            case 7: ok = bwd2s_launch_y7(cx, F0.sym, F1.sym, a, nblocks, stream); break;
            case 9: ok = bwd2s_launch_y9(cx, F0.sym, F1.sym, a, nblocks, stream); break;
            case 11: ok = bwd2s_launch_y11(cx, F0.sym, F1.sym, a, nblocks, stream); break;
            case 13: ok = bwd2s_launch_y13(cx, F0.sym, F1.sym, a, nblocks, stream); break;
        }
        e = ok ? hipGetLastError() : hipSuccess;
    }
    if (ok && e == hipSuccess && gs) e = launch_bwd2d_sum(partials, nblocks, gs, gs_mode, gs_n, stream);
    scratch_free(partials, stream);
    if (!ok) return -100;
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("backward launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

// The fused 2-D backward (k_bwd2d.hip) for the passes along dims 0 and 1, on views:
// x (the forward input), g (dL/d(output of those two passes)), dx (dL/dx, or p == null
// for none), in float with unit x strides, symmetric filters of radius <= 16, no line
// shorter than its filter and a condition it supports; dims >= 2 are stacked planes.
// dL/dsigma along dims 0 and 1 goes to gs (device doubles, nullable) by gs_mode:
// kBwdSumAdd gs[0..1] +=, kBwdSumWrite gs[0..1] = and gs[2..gs_n-1] = 0, kBwdSumTotal
// gs[0] = their sum.
// Returns -100 when it does not apply.
int fused_backward_2d(const View &x, const View &g, const View &dx, int nd, const InternFilter &F0,
                      const InternFilter &F1, const InternFilter &dF0, const InternFilter &dF1, int bc0, int bc1,
                      double *gs, int gs_mode, int gs_n, hipStream_t stream) {
    {
        const int st = streaming_backward_2d(x, g, dx, nd, F0, F1, dF0, dF1, bc0, bc1, gs, gs_mode, gs_n, stream);
        if (st != -100) return st;
    }
    if (!bwd_fused_enabled()) return -100;
    if (x.dt != DIPAMD_DT_SFLOAT || g.dt != DIPAMD_DT_SFLOAT) return -100;
    const bool want_dx = dx.p != nullptr;
    if (x.strides[0] != 1 || g.strides[0] != 1 || (want_dx && dx.strides[0] != 1)) return -100;
    const InternFilter *Fs[2] = {&F0, &F1};
    const int bcs[2] = {bc0, bc1};
    for (int d = 0; d < 2; ++d) {
        const InternFilter &f = *Fs[d];
        if ((f.sym != DIPAMD_SYM_EVEN && f.sym != DIPAMD_SYM_ODD) || f.shift != 0 || f.border > 16) return -100;
        if (bwd_mode(bcs[d]) < 0) return -100;
        if (x.sizes[d] < 2 * f.border + 1 || x.sizes[d] >= (int64_t(1) << 30)) return -100;
    }
    if (nd - 2 > kMaxOuter) return -100;
    // the kernel addresses a plane through 32-bit byte offsets from its first sample
    auto plane_bytes = [&](const View &v) -> int64_t {
        if (v.strides[1] <= 0) return -1;
        return ((x.sizes[1] - 1) * v.strides[1] + x.sizes[0]) * static_cast<int64_t>(sizeof(float));
    };
    const int64_t kMaxPlane = (int64_t(1) << 31) - 256;
    const int64_t xb = plane_bytes(x), gb = plane_bytes(g), db = want_dx ? plane_bytes(dx) : 0;
    if (xb < 0 || gb < 0 || db < 0 || xb > kMaxPlane || gb > kMaxPlane || db > kMaxPlane) return -100;
    BwdArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = static_cast<const float *>(x.p);
    a.g = static_cast<const float *>(g.p);
    a.dx = want_dx ? static_cast<float *>(dx.p) : nullptr;
    a.want_dx = want_dx ? 1 : 0;
    a.xbytes = static_cast<int>(xb);
    a.gbytes = static_cast<int>(gb);
    a.dxbytes = static_cast<int>(db);
    a.nx = static_cast<int>(x.sizes[0]);
    a.ny = static_cast<int>(x.sizes[1]);
    a.syx = x.strides[1];
    a.syg = g.strides[1];
    a.sydx = want_dx ? dx.strides[1] : 0;
    a.rx = F0.border;
    a.ry = F1.border;
    a.bcx = bc0;
    a.bcy = bc1;
    a.modex = bwd_mode(bc0);
    a.modey = bwd_mode(bc1);
    a.cvalx = 0.f;  // add zeros is the only constant condition taken here
    int64_t nplanes = 1;
    a.outer.n = 0;
    for (int d = 2; d < nd; ++d) {
        if (x.sizes[d] <= 1) continue;
        const int k = a.outer.n++;
        a.outer.size[k] = x.sizes[d];
        a.outer.sin[k] = x.strides[d];
        a.outer.sout[k] = want_dx ? dx.strides[d] : 0;
        a.gplane[k] = g.strides[d];
        nplanes *= x.sizes[d];
    }
    auto put = [](float *dst, const InternFilter &f, bool adjoint) {
        for (size_t k = 0; k < f.data.size(); ++k) {
            const double v = (adjoint && f.sym == DIPAMD_SYM_ODD && k > 0) ? -f.data[k] : f.data[k];
            dst[k] = static_cast<float>(v);
        }
    };
    put(a.tx, F0, false);
    put(a.tdx, dF0, false);
    put(a.txa, F0, true);
    put(a.tya, F1, true);
    put(a.tdya, dF1, true);
    const int rc = bwd2d_class(a.rx, a.ry);  // compile-time windows; taps stay zero beyond r
    if (rc == 0) return -100;
    const int txc = bwd2d_tile_columns(rc);
    a.ntiles = (a.nx + txc - 1) / txc;
    const int64_t base = static_cast<int64_t>(a.ntiles) * nplanes;
    // With a folding y condition the rows that receive folded terms get chunks of their
    // own (whole steps from each edge), so that every other chunk runs without folds.
    a.ytop = 0;
    a.ybot = a.ny;
    if (a.modey == kBwdFold) {
        const int e = kBwdStep * ((a.ry + kBwdStep) / kBwdStep);  // S ceil((ry + 1) / S)
        if (2 * e < a.ny) {
            a.ytop = e;
            a.ybot = a.ny - e;
        }
    }
    const int64_t mid = a.ybot - a.ytop;
    int64_t want = std::max<int64_t>(1, (2048 + base - 1) / base);
    int64_t chunk = (mid + want - 1) / want;
    chunk = std::max<int64_t>(kBwdStep, ((chunk + kBwdStep - 1) / kBwdStep) * kBwdStep);
    a.chunk = static_cast<int>(chunk);
    a.nchunks = static_cast<int>((mid + chunk - 1) / chunk) + (a.ytop > 0 ? 2 : 0);
    const int64_t nblocks = base * a.nchunks;
    if (nblocks >= (int64_t(1) << 31)) return -100;
    if (bwd2d_lds_bytes(rc) > lds_limit_bytes()) return -100;
    double *partials = nullptr;
    hipError_t e = scratch_alloc(reinterpret_cast<void **>(&partials), static_cast<size_t>(2 * nblocks) * sizeof(double),
                                 stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
    a.partials = partials;
    {
        TraceScope ts("bwd2d", static_cast<int64_t>(a.nx) * a.ny * nplanes * (want_dx ? 12 : 8), stream);
        e = launch_bwd2d(F0.sym, F1.sym, rc, a, nblocks, gs, gs_mode, gs_n, stream);
    }
    scratch_free(partials, stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("backward launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

}  // namespace
}  // namespace dipamd

using namespace dipamd;

static int gauss_fir_backward_impl(const dipamd_image *x, const dipamd_image *grad_out, const dipamd_image *grad_in,
                                   double *grad_sigma, const double *sigmas, int32_t nsigmas, const int32_t *orders,
                                   int32_t norders, const int32_t *bc, int32_t nbc, double truncation, uint32_t flags,
                                   const dipamd_image *kept, void *stream_) {
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    int st;
    if ((st = check_image(x, "x")) != DIPAMD_OK) return st;
    if ((st = require_real(x)) != DIPAMD_OK || (st = require_real(grad_out)) != DIPAMD_OK) return st;
    if ((st = check_image(grad_out, "grad_out")) != DIPAMD_OK) return st;
    const int nd = x->ndim;
    const bool want_dx = grad_in && grad_in->origin;
    const bool want_ds = grad_sigma != nullptr;
    if (want_dx && (st = check_image(grad_in, "grad_in")) != DIPAMD_OK) return st;
    if (grad_out->ndim != nd || (want_dx && grad_in->ndim != nd))
        return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    for (int d = 0; d < nd; ++d) {
        if (grad_out->sizes[d] != x->sizes[d] || (want_dx && grad_in->sizes[d] != x->sizes[d]))
            return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    }
    if (want_dx && (views_overlap(to_view(grad_in), to_view(grad_out), nd) || views_overlap(to_view(grad_in), to_view(x), nd)))
        return set_error(DIPAMD_E_PARAMETER, "grad_in must not overlap grad_out or x");
    const int buf_dt = suggest_flex(x->dtype);
    const bool is_double = buf_dt == DIPAMD_DT_DFLOAT;
    if (grad_out->dtype != buf_dt || (want_dx && grad_in->dtype != buf_dt))
        return set_error(DIPAMD_E_PARAMETER, "gradients must have the output type of the forward pass");
    if (nsigmas != 1 && nsigmas != nd) return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    if (norders != 0 && norders != 1 && norders != nd)
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    int bcs[DIPAMD_MAX_DIMS];
    if ((st = broadcast_bc(bc, nbc, nd, bcs)) != DIPAMD_OK) return st;
    const bool fast = (flags & DIPAMD_FAST) != 0;
    const bool total = (flags & DIPAMD_GRAD_SIGMA_TOTAL) != 0;
    const bool compute_double = x->dtype == DIPAMD_DT_DFLOAT;

    // filters (GaussFIR, gauss.cpp:238-286) and their sigma derivatives
    std::vector<InternFilter> F(static_cast<size_t>(nd)), dF(static_cast<size_t>(nd));
    std::vector<int> order;
    for (int d = 0; d < nd; ++d) {
        const double s = nsigmas == 1 ? sigmas[0] : sigmas[d];
        const int o = norders == 0 ? 0 : (norders == 1 ? orders[0] : orders[d]);
        if (!(s > 0.0) || x->sizes[d] <= 1) continue;
        if (o < 0 || o > 3) return set_error(DIPAMD_E_ORDER, "Gaussian FIR filter not implemented for order > 3");
        if (!bc_supported(bcs[d])) return set_error(DIPAMD_E_UNSUPPORTED, "boundary condition not supported");
        std::vector<double> h = half_gaussian(s, o, truncation, compute_double);
        if (static_cast<int>(h.size()) - 1 > DIPAMD_MAX_RADIUS)
            return set_error(DIPAMD_E_UNSUPPORTED, "Gaussian radius above DIPAMD_MAX_RADIUS");
        const int sym = (o & 1) ? DIPAMD_SYM_ODD : DIPAMD_SYM_EVEN;
        if ((st = intern_filter(h.data(), static_cast<int>(h.size()), sym, -1, F[d])) != DIPAMD_OK) return st;
        std::vector<double> dh = half_gaussian_dsigma(s, o, static_cast<int>(h.size()));
        if ((st = intern_filter(dh.data(), static_cast<int>(dh.size()), sym, -1, dF[d])) != DIPAMD_OK) return st;
        if (meaningless(F[d], is_double)) continue;
        order.push_back(d);
    }
    const View xv = to_view(x);
    for (size_t i = 1; i < order.size(); ++i) {  // stable sort by input stride
        const int e = order[i];
        size_t j = i;
        while (j > 0 && xv.strides[order[j - 1]] > xv.strides[e]) {
            order[j] = order[j - 1];
            --j;
        }
        order[j] = e;
    }
    const int m = static_cast<int>(order.size());
    // dims 0 and 1 of a float image through the fused 2-D kernel: the whole backward
    // when only they are filtered, else after the adjoint passes of the other dims
    const bool fuse01 = m >= 2 && order[0] == 0 && order[1] == 1;
    if (fuse01 && m == 2) {
        // the partial sums' kernel writes every output entry: no separate fill
        View dxv;
        std::memset(&dxv, 0, sizeof(dxv));
        if (want_dx) dxv = to_view(grad_in);
        st = fused_backward_2d(to_view(x), to_view(grad_out), dxv, nd, F[0], F[1], dF[0], dF[1], bcs[0], bcs[1],
                               want_ds ? grad_sigma : nullptr, total ? kBwdSumTotal : kBwdSumWrite, nd, stream);
        if (st != -100) return st;
    }

    // the general path: dL/dsigma per dim into gsig (grad_sigma, or a scratch array
    // whose entries are summed into grad_sigma[0] afterwards)
    auto general = [&](double *gsig) -> int {
        std::vector<void *> scratch;
        auto alloc = [&](View *v) -> int {
            int64_t n = 1;
            for (int d = 0; d < nd; ++d) n *= x->sizes[d];
            void *p = nullptr;
            hipError_t e = scratch_alloc(&p, static_cast<size_t>(n) * dt_size(buf_dt), stream);
            if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
            scratch.push_back(p);
            *v = contiguous_view(p, buf_dt, nd, xv.sizes);
            return DIPAMD_OK;
        };
        auto finish = [&](int status) {
            for (void *p : scratch) scratch_free(p, stream);
            return status;
        };

        if (want_ds) {
            hipError_t e = launch_fill_doubles(gsig, nd, 0.0, stream);
            if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("fill: ") + hipGetErrorString(e));
        }
        const View gv = to_view(grad_out);
        if (m == 0) {
            // identity forward: dL/dx = g
            if (want_dx) return cast_from_buffer(gv, to_view(grad_in), nd, stream);
            return DIPAMD_OK;
        }

        // the input as a kernel source (uint8 or the buffer type); others are converted
        View x0 = xv;
        if (!(x0.dt == DIPAMD_DT_UINT8 && !is_double) && x0.dt != buf_dt) {
            View t;
            if ((st = alloc(&t)) != DIPAMD_OK) return finish(st);
            if ((st = cast_to_buffer(x0, t, nd, is_double, stream)) != DIPAMD_OK) return finish(st);
            x0 = t;
        }

        // forward intermediates v_1..v_{m-1} (only needed for the sigma derivatives)
        std::vector<View> v(static_cast<size_t>(m));
        v[0] = x0;
        // With dims 0 and 1 first and fused below, only v_2 (after both) is needed: one
        // fused forward over them gives it; v_1 is made only if the fused backward
        // turns out not to apply.
        bool v1_pending = false;
        auto make_v = [&](int i) -> int {
            int s2 = alloc(&v[static_cast<size_t>(i)]);
            if (s2 != DIPAMD_OK) return s2;
            const int d = order[static_cast<size_t>(i - 1)];
            return run_pass(v[static_cast<size_t>(i - 1)], v[static_cast<size_t>(i)], nd, d, F[d], bcs[d], is_double,
                            fast, stream);
        };
        if (want_ds && fuse01 && m > 2) {
            if (kept) {
                v[2] = to_view(kept);  // the forward's own intermediate after dims 0 and 1
            } else {
                if ((st = alloc(&v[2])) != DIPAMD_OK) return finish(st);
                const InternFilter *fp[DIPAMD_MAX_DIMS];
                bool proc[DIPAMD_MAX_DIMS];
                for (int d = 0; d < nd; ++d) {
                    fp[d] = &F[d];
                    proc[d] = d == 0 || d == 1;
                }
                if ((st = separable_device(x0, v[2], nd, fp, bcs, proc, fast ? DIPAMD_FAST : 0u, stream)) !=
                    DIPAMD_OK)
                    return finish(st);
            }
            v1_pending = true;
            for (int i = 3; i < m; ++i)
                if ((st = make_v(i)) != DIPAMD_OK) return finish(st);
        } else if (want_ds) {
            for (int i = 1; i < m; ++i) {
                if ((st = alloc(&v[static_cast<size_t>(i)])) != DIPAMD_OK) return finish(st);
                const int d = order[static_cast<size_t>(i - 1)];
                if ((st = run_pass(v[static_cast<size_t>(i - 1)], v[static_cast<size_t>(i)], nd, d, F[d], bcs[d],
                                   is_double, fast, stream)) != DIPAMD_OK)
                    return finish(st);
            }
        }

        // adjoint chain u_m = g, u_{i-1} = K_i^T u_i
        std::vector<View> u(static_cast<size_t>(m + 1));
        u[static_cast<size_t>(m)] = gv;
        View w;
        bool have_w = false;
        double *partials = nullptr;
        if (want_ds) {
            hipError_t e = scratch_alloc(reinterpret_cast<void **>(&partials), 1024 * sizeof(double), stream);
            if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, "scratch allocation failed"));
            scratch.push_back(partials);
        }
        for (int i = m; i >= 1; --i) {
            const int d = order[static_cast<size_t>(i - 1)];
            if (i == 2 && fuse01 && m > 2) {
                // the remaining passes (dims 0 and 1) and their sigma derivatives in one kernel
                View dxv;
                std::memset(&dxv, 0, sizeof(dxv));
                if (want_dx) dxv = to_view(grad_in);
                st = fused_backward_2d(xv, u[2], dxv, nd, F[0], F[1], dF[0], dF[1], bcs[0], bcs[1],
                                       want_ds ? gsig : nullptr, kBwdSumAdd, nd, stream);
                if (st != -100) return finish(st);
                if (v1_pending && (st = make_v(1)) != DIPAMD_OK) return finish(st);
                v1_pending = false;
            }
            if (want_ds) {
                // dL/dsigma_d = < u_i , K'_i v_{i-1} >: in one sweep where it applies
                st = run_pass_dot(v[static_cast<size_t>(i - 1)], u[static_cast<size_t>(i)], nd, d, dF[d], bcs[d],
                                  is_double, fast, gsig + d, stream);
                if (st != -100 && st != DIPAMD_OK) return finish(st);
            }
            if (want_ds && st == -100) {
                if (!have_w) {
                    if ((st = alloc(&w)) != DIPAMD_OK) return finish(st);
                    have_w = true;
                }
                if ((st = run_pass(v[static_cast<size_t>(i - 1)], w, nd, d, dF[d], bcs[d], is_double, fast, stream)) !=
                    DIPAMD_OK)
                    return finish(st);
                const View &ui = u[static_cast<size_t>(i)];
                CastGeo geo;
                std::memset(&geo, 0, sizeof(geo));
                geo.total = 1;
                for (int k = 0; k < nd; ++k) geo.total *= ui.sizes[k];
                if (!contiguous(ui, nd)) {
                    geo.ndim = nd;
                    for (int k = 0; k < nd; ++k) {
                        geo.size[k] = ui.sizes[k];
                        geo.sin[k] = ui.strides[k];
                    }
                }
                TraceScope ts("dot", geo.total * static_cast<int64_t>(2 * dt_size(buf_dt)), stream);
                hipError_t e = launch_dot(is_double, ui.p, &geo, w.p, partials, gsig + d, stream);
                if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, std::string("dot: ") + hipGetErrorString(e)));
            }
            if (i == 1 && !want_dx) break;
            View dst;
            if (i == 1) {
                dst = to_view(grad_in);
            } else if ((st = alloc(&dst)) != DIPAMD_OK) {
                return finish(st);
            }
            if ((st = adjoint_pass(u[static_cast<size_t>(i)], dst, nd, d, F[d], bcs[d], is_double, fast, stream)) !=
                DIPAMD_OK)
                return finish(st);
            u[static_cast<size_t>(i - 1)] = dst;
        }
        return finish(DIPAMD_OK);
    };
    if (!(total && want_ds)) return general(grad_sigma);
    double *gsd = nullptr;
    hipError_t e = scratch_alloc(reinterpret_cast<void **>(&gsd), static_cast<size_t>(nd) * sizeof(double), stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
    st = general(gsd);
    if (st == DIPAMD_OK) {
        e = launch_sum_doubles(gsd, nd, grad_sigma, stream);
        if (e != hipSuccess) st = set_error(DIPAMD_E_HIP, std::string("sum: ") + hipGetErrorString(e));
    }
    scratch_free(gsd, stream);
    return st;
}

extern "C" int dipamd_gauss_fir_backward(const dipamd_image *x, const dipamd_image *grad_out,
                                         const dipamd_image *grad_in, double *grad_sigma, const double *sigmas,
                                         int32_t nsigmas, const int32_t *orders, int32_t norders, const int32_t *bc,
                                         int32_t nbc, double truncation, uint32_t flags, void *stream) {
    return gauss_fir_backward_impl(x, grad_out, grad_in, grad_sigma, sigmas, nsigmas, orders, norders, bc, nbc,
                                   truncation, flags, nullptr, stream);
}

extern "C" int dipamd_gauss_fir_backward_kept(const dipamd_image *x, const dipamd_image *grad_out,
                                              const dipamd_image *grad_in, double *grad_sigma, const double *sigmas,
                                              int32_t nsigmas, const int32_t *orders, int32_t norders,
                                              const int32_t *bc, int32_t nbc, double truncation, uint32_t flags,
                                              const dipamd_image *kept, void *stream) {
    int st;
    if (kept) {
        if ((st = check_image(kept, "kept")) != DIPAMD_OK) return st;
        if ((st = check_image(x, "x")) != DIPAMD_OK) return st;
        if (kept->ndim != x->ndim) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
        for (int d = 0; d < x->ndim; ++d)
            if (kept->sizes[d] != x->sizes[d]) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
        if (kept->dtype != suggest_flex(x->dtype))
            return set_error(DIPAMD_E_PARAMETER, "kept must have the buffer type of x");
    }
    return gauss_fir_backward_impl(x, grad_out, grad_in, grad_sigma, sigmas, nsigmas, orders, norders, bc, nbc,
                                   truncation, flags, kept, stream);
}
