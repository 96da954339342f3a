// K2R: sweep along a strided axis for short symmetric filters (radius R <= 12, a
// compile-time constant).  One thread per position of the lane axis, as K2, but
// the window is a register array that slides down the chunk: each step keeps the
// 2R samples it shares with the next step, so every input sample is loaded once
// per chunk, and the next step's P samples are loaded before this step's
// arithmetic, so the loads are in flight while it runs.  Every window index is a
// compile-time constant.  The per-output operation order is the reference's
// (convolution.cpp:214-241): centre product first, then the tap pairs k = 1..R.
// k_zsweep_dot runs the same sweep into an FP64 dot product with a second image
// instead of writing the outputs (the sigma derivatives of the backward form).
#include "sep_kernels_impl.hpp"

#include "sep_config.hpp"

namespace dipamd {

namespace {

constexpr int kZP = 8;  // outputs per thread and step

// One column's outputs z0..zend-1 of the sweep, P at a time, handed to sink(zb, acc)
// (outputs zb..zb+P-1, those at or past zend to be ignored).
template <typename TI, typename TB, int R, int SYM, bool FAST, typename Sink>
__device__ __forceinline__ void zsweep_column(const PassArgs<TB> &a, const TI *src, int z0, int zend, Sink &&sink) {
    constexpr int P = kZP;
    constexpr int NW = P + 2 * R;
    const int n = static_cast<int>(a.n);
    const int64_t sin = a.sin;
    TB f[R + 1];
#pragma unroll
    for (int k = 0; k <= R; ++k) f[k] = a.taps[k];

    // sample j of the line, boundary-extended where the block touches an end
    auto load = [&](int j, bool inside) -> TB {
        return inside ? to_buf<TB>(src[static_cast<int64_t>(j) * sin]) : fetch_ext<TI, TB>(src, sin, j, n, a.bc, a.cval, a.zlo);
    };
    TB w[NW];
    {
        const bool inside = z0 - R >= 0 && z0 + P + R <= n;
#pragma unroll
        for (int q = 0; q < NW; ++q) w[q] = load(z0 - R + q, inside);
    }
    for (int zb = z0; zb < zend; zb += P) {
        // next step's new samples: rows zb + P + R .. zb + 2P + R - 1
        TB nx[P];
        const int jn = zb + P + R;
        const bool more = zb + P < zend;
        if (more) {
            const bool inside = jn + P <= n;
#pragma unroll
            for (int p = 0; p < P; ++p) nx[p] = load(jn + p, inside);
        }
        TB acc[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc[p] = f[0] * w[R + p];
#pragma unroll
            for (int k = 1; k <= R; ++k) {
                const TB t = (SYM == SYM_EVEN) ? (w[R + p + k] + w[R + p - k]) : (w[R + p + k] - w[R + p - k]);
                acc[p] = mac<TB, FAST>(acc[p], f[k], t);
            }
        }
        sink(zb, acc);
        if (more) {
#pragma unroll
            for (int q = 0; q < 2 * R; ++q) w[q] = w[q + P];
#pragma unroll
            for (int p = 0; p < P; ++p) w[2 * R + p] = nx[p];
        }
    }
}

template <typename TI, typename TB, int R, int SYM, bool FAST>
__global__ __launch_bounds__(kBlock) void k_zsweep(const PassArgs<TB> a, int64_t ntiles, int64_t nchunks,
                                                    const TI *__restrict__ in, TB *__restrict__ out) {
    constexpr int P = kZP;
    // tile fastest in hardware order, as K2: neighbouring chunks of a column are
    // dispatched later and find their shared rows in L2
    const int64_t bid = blockIdx.x;
    const int64_t tile = bid % ntiles;
    const int64_t rest = bid / ntiles;
    const int64_t chunk = rest % nchunks;
    const int64_t outer = rest / nchunks;
    const int64_t c = tile * blockDim.x + threadIdx.x;
    if (c >= a.e) return;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    const TI *src = in + oin + c * a.ein;
    TB *dst = out + oout + c * a.eout;
    const int z0 = static_cast<int>(chunk * a.chunk);
    const int zend = static_cast<int>(min(chunk * a.chunk + a.chunk, a.n));
    zsweep_column<TI, TB, R, SYM, FAST>(a, src, z0, zend, [&](int zb, const TB (&acc)[P]) {
        TB *d = dst + static_cast<int64_t>(zb) * a.sout;
        if (zb + P <= zend) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                *d = acc[p];
                d += a.sout;
            }
        } else {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                if (zb + p < zend) *d = acc[p];
                d += a.sout;
            }
        }
    });
}

// <u, K v> in FP64 without writing K v: each thread sums its column's products in
// order, each workgroup its threads' sums in a fixed tree order into partials[block];
// k_sum_partials adds those in a fixed order.  u has the output geometry of the pass.
template <typename TB, int R, int SYM, bool FAST>
__global__ __launch_bounds__(kBlock) void k_zsweep_dot(const PassArgs<TB> a, int64_t ntiles, int64_t nchunks,
                                                        const TB *__restrict__ in, const TB *__restrict__ u,
                                                        double *__restrict__ partials) {
    constexpr int P = kZP;
    __shared__ double red[kBlock];
    const int64_t bid = blockIdx.x;
    const int64_t tile = bid % ntiles;
    const int64_t rest = bid / ntiles;
    const int64_t chunk = rest % nchunks;
    const int64_t outer = rest / nchunks;
    const int64_t c = tile * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (c < a.e) {
        int64_t oin, oout;
        outer_offsets(a.outer, outer, &oin, &oout);
        const TB *src = in + oin + c * a.ein;
        const TB *up = u + oout + c * a.eout;
        const int z0 = static_cast<int>(chunk * a.chunk);
        const int zend = static_cast<int>(min(chunk * a.chunk + a.chunk, a.n));
        zsweep_column<TB, TB, R, SYM, FAST>(a, src, z0, zend, [&](int zb, const TB (&acc)[P]) {
            const TB *q = up + static_cast<int64_t>(zb) * a.sout;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                if (zb + p < zend) s += static_cast<double>(q[0]) * static_cast<double>(acc[p]);
                q += a.sout;
            }
        });
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// *out = the n partials added in a fixed order (strided per thread, then a tree).
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *partials, int64_t n, double *out) {
    __shared__ double red[kBlock];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

template <typename TB, int R, int SYM, bool FAST>
hipError_t zsweep_dot_t(const PassArgs<TB> &a, int64_t e, int64_t nchunks, int64_t nouter, const TB *u,
                        double *partials, double *out, hipStream_t st) {
    const int bs = e >= 256 ? 256 : 64;
    const int64_t ntiles = (e + bs - 1) / bs;
    const int64_t nblk = ntiles * nchunks * nouter;
    hipLaunchKernelGGL((k_zsweep_dot<TB, R, SYM, FAST>), dim3(static_cast<unsigned>(nblk)), dim3(bs), 0, st, a,
                       ntiles, nchunks, static_cast<const TB *>(a.in), u, partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, st, partials, nblk, out);
    return hipGetLastError();
}

template <typename TB, int R>
hipError_t zsweep_dot_sym(int sym, bool fast, const PassArgs<TB> &a, int64_t e, int64_t nc, int64_t no, const TB *u,
                          double *partials, double *out, hipStream_t st) {
    if (sym == SYM_EVEN)
        return fast ? zsweep_dot_t<TB, R, SYM_EVEN, true>(a, e, nc, no, u, partials, out, st)
                    : zsweep_dot_t<TB, R, SYM_EVEN, false>(a, e, nc, no, u, partials, out, st);
    return fast ? zsweep_dot_t<TB, R, SYM_ODD, true>(a, e, nc, no, u, partials, out, st)
                : zsweep_dot_t<TB, R, SYM_ODD, false>(a, e, nc, no, u, partials, out, st);
}

template <typename TB>
hipError_t zsweep_dot_r(int r, int sym, bool fast, const PassArgs<TB> &a, int64_t e, int64_t nc, int64_t no,
                        const TB *u, double *partials, double *out, hipStream_t st) {
    switch (r) {
        case 1: return zsweep_dot_sym<TB, 1>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 2: return zsweep_dot_sym<TB, 2>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 3: return zsweep_dot_sym<TB, 3>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 4: return zsweep_dot_sym<TB, 4>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 5: return zsweep_dot_sym<TB, 5>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 6: return zsweep_dot_sym<TB, 6>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 7: return zsweep_dot_sym<TB, 7>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 8: return zsweep_dot_sym<TB, 8>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 9: return zsweep_dot_sym<TB, 9>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 10: return zsweep_dot_sym<TB, 10>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 11: return zsweep_dot_sym<TB, 11>(sym, fast, a, e, nc, no, u, partials, out, st);
        case 12: return zsweep_dot_sym<TB, 12>(sym, fast, a, e, nc, no, u, partials, out, st);
    }
    return hipErrorInvalidValue;
}

template <typename TI, typename TB, int R, int SYM, bool FAST>
hipError_t zsweep_t(const PassArgs<TB> &a, int64_t e, int64_t nchunks, int64_t nouter, hipStream_t st) {
    const int bs = e >= 256 ? 256 : 64;
    const int64_t ntiles = (e + bs - 1) / bs;
    const int64_t nblk = ntiles * nchunks * nouter;
    hipLaunchKernelGGL((k_zsweep<TI, TB, R, SYM, FAST>), dim3(static_cast<unsigned>(nblk)), dim3(bs), 0, st, a,
                       ntiles, nchunks, static_cast<const TI *>(a.in), static_cast<TB *>(a.out));
    return hipGetLastError();
}

template <typename TI, typename TB, int R>
hipError_t zsweep_sym(int sym, bool fast, const PassArgs<TB> &a, int64_t e, int64_t nc, int64_t no, hipStream_t st) {
    if (sym == SYM_EVEN)
        return fast ? zsweep_t<TI, TB, R, SYM_EVEN, true>(a, e, nc, no, st)
                    : zsweep_t<TI, TB, R, SYM_EVEN, false>(a, e, nc, no, st);
    return fast ? zsweep_t<TI, TB, R, SYM_ODD, true>(a, e, nc, no, st)
                : zsweep_t<TI, TB, R, SYM_ODD, false>(a, e, nc, no, st);
}

template <typename TI, typename TB>
hipError_t zsweep_r(int r, int sym, bool fast, const PassArgs<TB> &a, int64_t e, int64_t nc, int64_t no,
                    hipStream_t st) {
    switch (r) {
        case 1: return zsweep_sym<TI, TB, 1>(sym, fast, a, e, nc, no, st);
        case 2: return zsweep_sym<TI, TB, 2>(sym, fast, a, e, nc, no, st);
        case 3: return zsweep_sym<TI, TB, 3>(sym, fast, a, e, nc, no, st);
        case 4: return zsweep_sym<TI, TB, 4>(sym, fast, a, e, nc, no, st);
        case 5: return zsweep_sym<TI, TB, 5>(sym, fast, a, e, nc, no, st);
        case 6: return zsweep_sym<TI, TB, 6>(sym, fast, a, e, nc, no, st);
        case 7: return zsweep_sym<TI, TB, 7>(sym, fast, a, e, nc, no, st);
        case 8: return zsweep_sym<TI, TB, 8>(sym, fast, a, e, nc, no, st);
        case 9: return zsweep_sym<TI, TB, 9>(sym, fast, a, e, nc, no, st);
        case 10: return zsweep_sym<TI, TB, 10>(sym, fast, a, e, nc, no, st);
        case 11: return zsweep_sym<TI, TB, 11>(sym, fast, a, e, nc, no, st);
        case 12: return zsweep_sym<TI, TB, 12>(sym, fast, a, e, nc, no, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_zsweep_dot(bool is_double, int r, int sym, bool fast, const void *args, int64_t e, int64_t nchunks,
                             int64_t nouter, const void *u, double *partials, double *out, hipStream_t stream) {
    if (r < 1 || r > kZSweepMaxR || (sym != SYM_EVEN && sym != SYM_ODD)) return hipErrorInvalidValue;
    if (is_double)
        return zsweep_dot_r<double>(r, sym, fast, *static_cast<const PassArgs<double> *>(args), e, nchunks, nouter,
                                    static_cast<const double *>(u), partials, out, stream);
    return zsweep_dot_r<float>(r, sym, fast, *static_cast<const PassArgs<float> *>(args), e, nchunks, nouter,
                               static_cast<const float *>(u), partials, out, stream);
}

hipError_t launch_zsweep(InType ti, bool is_double, int r, int sym, bool fast, const void *args, int64_t e,
                         int64_t nchunks, int64_t nouter, hipStream_t stream) {
    if (r < 1 || r > kZSweepMaxR || (sym != SYM_EVEN && sym != SYM_ODD)) return hipErrorInvalidValue;
    if (is_double)
        return zsweep_r<double, double>(r, sym, fast, *static_cast<const PassArgs<double> *>(args), e, nchunks,
                                        nouter, stream);
    const auto &a = *static_cast<const PassArgs<float> *>(args);
    return ti == IN_U8 ? zsweep_r<uint8_t, float>(r, sym, fast, a, e, nchunks, nouter, stream)
                       : zsweep_r<float, float>(r, sym, fast, a, e, nchunks, nouter, stream);
}

}  // namespace dipamd
