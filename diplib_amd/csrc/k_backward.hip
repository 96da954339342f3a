// Kernels of the backward form: the boundary fold-back of an adjoint pass and a
// deterministic two-stage FP64 dot product.
//
// Adjoint of one filter pass along an axis.  The forward pass is
//     out[i] = sum_{k=-B..B} c_k * ext(x)[i+k],   ext(x)[j] = s(j) x[fold(j)]  (or a constant)
// so its adjoint is  xbar[m] = sum_{j in [-B, N+B): fold(j) = m} s(j) h[j]  with
//     h[j] = sum_k c_{-k} * g[j+k]    (g zero outside [0, N)).
// The interior term h[m] is a forward pass with reversed taps on zero-padded g
// (launched by the host); this kernel adds the halo terms s(j) h[j] for the 2B
// positions j outside the line, one thread per line, in a fixed order.
#include "sep_kernels_impl.hpp"

namespace dipamd {

constexpr int kFoldFast = 16;  // borders up to this take the register path

template <typename TB>
__global__ __launch_bounds__(kBlock) void k_adjoint_fold(const PassArgs<TB> a) {
    const int64_t line = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    // the boundary map of the 2B halo positions is the same for every line: computed
    // once per workgroup (packed as in bc_fold_packed: source << 1 | negate, -1 constant)
    __shared__ int fmap[2][kFoldFast + 1];
    if (a.border <= kFoldFast) {
        const int i = threadIdx.x;
        if (i < 2 * a.border) {
            const int side = i / a.border, t = i % a.border + 1;
            const int n0 = static_cast<int>(a.n);
            fmap[side][t] = bc_fold_packed(side == 0 ? -t : n0 - 1 + t, n0, a.bc);
        }
        __syncthreads();
    }
    if (line >= a.nlines) return;
    int64_t oin, oout;
    outer_offsets(a.outer, line, &oin, &oout);
    const TB *g = static_cast<const TB *>(a.in) + oin;
    TB *out = static_cast<TB *>(a.out) + oout;
    const int n = static_cast<int>(a.n);
    const int B = a.border;
    const bool odd = a.shift != 0;  // the host passes 1 for an odd filter
    // adjoint correlation taps: cT_k = c_{-k}; taps[] holds |k| magnitudes of cT
    if (B <= kFoldFast) {
        // the B + 1 samples at each end of the line are loaded once, all at a time; the
        // sums run over them in the same order as below (k ascending)
        for (int side = 0; side < 2; ++side) {
            TB e[kFoldFast + 1];
#pragma unroll
            for (int q = 0; q <= kFoldFast; ++q) {
                const int qq = q <= B && q < n ? q : 0;
                e[q] = g[static_cast<int64_t>(side == 0 ? qq : n - 1 - qq) * a.sin];
            }
            TB hs[kFoldFast + 1];
#pragma unroll
            for (int t = 1; t <= kFoldFast; ++t) {
                TB h = TB(0);
                if (side == 0) {  // k = t..B reads sample k - t
#pragma unroll
                    for (int k = 1; k <= kFoldFast; ++k)
                        if (k >= t && k <= B && k - t < n) h += a.taps[k] * e[k - t];
                } else {  // k = -B..-t reads sample n - 1 - q, q = -t - k
#pragma unroll
                    for (int q = kFoldFast - 1; q >= 0; --q) {
                        if (q <= B - t && q < n) {
                            const TB c = a.taps[t + q];
                            h += (odd ? -c : c) * e[q];
                        }
                    }
                }
                hs[t] = h;
            }
#pragma unroll
            for (int t = 1; t <= kFoldFast; ++t) {
                if (t > B) break;
                const int pk = fmap[side][t];
                if (pk < 0) continue;  // constant condition: no dependence on the input
                const int m = pk >> 1, mul = (pk > 0 && (pk & 1)) ? -1 : 1;
                TB *o = out + static_cast<int64_t>(m) * a.sout;
                *o = mul < 0 ? *o - hs[t] : *o + hs[t];
            }
        }
        return;
    }
    for (int side = 0; side < 2; ++side) {
        for (int t = 1; t <= B; ++t) {
            const int j = side == 0 ? -t : n - 1 + t;
            int mul;
            const int m = bc_fold(j, n, a.bc, &mul);
            if (m < 0) continue;  // constant condition: no dependence on the input
            TB h = TB(0);
            for (int k = -B; k <= B; ++k) {
                const int idx = j + k;
                if (idx < 0 || idx >= n) continue;
                const int ak = k < 0 ? -k : k;
                TB c = a.taps[ak];
                if (odd && k < 0) c = -c;
                h += c * g[static_cast<int64_t>(idx) * a.sin];
            }
            TB *o = out + static_cast<int64_t>(m) * a.sout;
            *o = mul < 0 ? *o - h : *o + h;
        }
    }
}

template __global__ void k_adjoint_fold<float>(const PassArgs<float>);
template __global__ void k_adjoint_fold<double>(const PassArgs<double>);

// The same halo terms for the mirror conditions on lines of n >= 2B + 2 samples, one
// thread per (line, side, t): there every halo position j = -t (j = n - 1 + t) folds
// onto its own target (t, or n - 1 - t), so the 2B read-add-writes of a line touch
// distinct samples and run in parallel (B times the threads of the kernel above, whose
// one thread per line leaves a short line stack with few waves).  Each sum runs over
// the same terms in the same order as above.
template <typename TB>
__global__ __launch_bounds__(kBlock) void k_adjoint_fold_pt(const PassArgs<TB> a) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int B = a.border;
    if (idx >= a.nlines * 2 * B) return;
    const int64_t line = idx % a.nlines;  // consecutive threads: consecutive lines
    const int r = static_cast<int>(idx / a.nlines);
    const int side = r / B, t = r % B + 1;
    int64_t oin, oout;
    outer_offsets(a.outer, line, &oin, &oout);
    const TB *g = static_cast<const TB *>(a.in) + oin;
    TB *out = static_cast<TB *>(a.out) + oout;
    const int n = static_cast<int>(a.n);
    const bool odd = a.shift != 0;
    TB h = TB(0);
    if (side == 0) {  // k = t..B reads sample k - t
        for (int k = t; k <= B; ++k) h += a.taps[k] * g[static_cast<int64_t>(k - t) * a.sin];
    } else {  // q = B - t..0 reads sample n - 1 - q
        for (int q = B - t; q >= 0; --q) {
            const TB c = a.taps[t + q];
            h += (odd ? -c : c) * g[static_cast<int64_t>(n - 1 - q) * a.sin];
        }
    }
    int mul;
    const int m = bc_fold(side == 0 ? -t : n - 1 + t, n, a.bc, &mul);
    TB *o = out + static_cast<int64_t>(m) * a.sout;
    *o = mul < 0 ? *o - h : *o + h;
}
template __global__ void k_adjoint_fold_pt<float>(const PassArgs<float>);
template __global__ void k_adjoint_fold_pt<double>(const PassArgs<double>);

hipError_t launch_adjoint_fold(bool is_double, const void *args, int64_t nlines, hipStream_t stream) {
    // (the fields read here precede the first one whose type depends on TB)
    const PassArgs<float> &af = *static_cast<const PassArgs<float> *>(args);
    const int B = af.border;
    const bool mirror = af.bc == DIPAMD_BC_SYMMETRIC_MIRROR || af.bc == DIPAMD_BC_ASYMMETRIC_MIRROR;
    if (mirror && B >= 1 && af.n >= 2 * static_cast<int64_t>(B) + 2) {
        const int64_t nblk = (nlines * 2 * B + kBlock - 1) / kBlock;
        if (nblk < (int64_t(1) << 31)) {
            if (is_double)
                hipLaunchKernelGGL(k_adjoint_fold_pt<double>, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0,
                                   stream, *static_cast<const PassArgs<double> *>(args));
            else
                hipLaunchKernelGGL(k_adjoint_fold_pt<float>, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0,
                                   stream, af);
            return hipGetLastError();
        }
    }
    const int64_t nblk = (nlines + kBlock - 1) / kBlock;
    if (is_double) {
        hipLaunchKernelGGL(k_adjoint_fold<double>, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, stream,
                           *static_cast<const PassArgs<double> *>(args));
    } else {
        hipLaunchKernelGGL(k_adjoint_fold<float>, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, stream, af);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// <a, b> in FP64: stage 1 gives each of `nparts` blocks a fixed, contiguous index
// range (reduced in a fixed tree order), stage 2 adds the partials in order.
// `a` is a strided view (geometry `g`; g.ndim == 0 means contiguous), `b` is
// contiguous in the same index order.
// ---------------------------------------------------------------------------
constexpr int kDotParts = 1024;

template <typename TB>
__global__ __launch_bounds__(kBlock) void k_dot_partial(const TB *a, const CastGeo g, const TB *b, double *partials) {
    __shared__ double red[kBlock];
    const int64_t per = (g.total + kDotParts - 1) / kDotParts;
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * per;
    const int64_t hi = min(lo + per, g.total);
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        int64_t oa = i;
        if (g.ndim > 0) {
            int64_t r = i;
            oa = 0;
            for (int d = 0; d < g.ndim; ++d) {
                const int64_t c = r % g.size[d];
                r /= g.size[d];
                oa += c * g.sin[d];
            }
        }
        s += static_cast<double>(a[oa]) * static_cast<double>(b[i]);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void k_dot_final(const double *partials, double *out) {
    __shared__ double red[kBlock];
    double s = 0.0;
    for (int i = threadIdx.x; i < kDotParts; i += blockDim.x) s += partials[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_fill_doubles(double *p, int n, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_dot(bool is_double, const void *a, const void *geo, const void *b, double *partials, double *out,
                      hipStream_t stream) {
    const CastGeo &g = *static_cast<const CastGeo *>(geo);
    if (is_double) {
        hipLaunchKernelGGL(k_dot_partial<double>, dim3(kDotParts), dim3(kBlock), 0, stream,
                           static_cast<const double *>(a), g, static_cast<const double *>(b), partials);
    } else {
        hipLaunchKernelGGL(k_dot_partial<float>, dim3(kDotParts), dim3(kBlock), 0, stream,
                           static_cast<const float *>(a), g, static_cast<const float *>(b), partials);
    }
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kBlock), 0, stream, partials, out);
    return hipGetLastError();
}

// out[0] = in[0] + in[1] + ... + in[n-1], in that order.
__global__ void k_sum_doubles(const double *in, int n, double *out) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += in[i];
    *out = s;
}

hipError_t launch_sum_doubles(const double *in, int n, double *out, hipStream_t stream) {
    hipLaunchKernelGGL(k_sum_doubles, dim3(1), dim3(1), 0, stream, in, n, out);
    return hipGetLastError();
}

hipError_t launch_fill_doubles(double *p, int n, double v, hipStream_t stream) {
    hipLaunchKernelGGL(k_fill_doubles, dim3((n + 63) / 64), dim3(64), 0, stream, p, n, v);
    return hipGetLastError();
}

}  // namespace dipamd
