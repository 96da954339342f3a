// Complex separable convolution (dip::SeparableConvolution with a complex buffer type).
//
// One pass along one axis of an image of interleaved (re, im) samples, one thread per
// output sample.  The arithmetic is the reference line filter's for complex buffers
// (convolution.cpp:184-297 with TPI = scomplex / dcomplex): a real filter multiplies
// both parts (std::complex<T> times T), a complex filter multiplies as std::complex
// does -- re = a.re*b.re - a.im*b.im, im = a.re*b.im + a.im*b.re, each product rounded,
// no contraction, with C's Annex G infinity recovery where both parts come out NaN --
// and the CONJ forms add f*x[i+k] and conj(f)*x[i-k] before adding to the sum.
// Positions outside the line take the boundary condition component by component
// (copy_buffer.cpp:802-808 expands the real and imaginary parts as two real lines).
#include "sep_kernels_impl.hpp"
#include "complex_kernels.hpp"

namespace dipamd {

namespace {

template <typename TS>
struct cx {
    TS re, im;
};

template <typename TS>
__device__ __forceinline__ cx<TS> cadd(cx<TS> a, cx<TS> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename TS>
__device__ __forceinline__ cx<TS> csub(cx<TS> a, cx<TS> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename TS>
__device__ __forceinline__ TS csign(TS mag, TS sgn) {
    if constexpr (sizeof(TS) == 4) {
        return __builtin_copysignf(mag, sgn);
    } else {
        return __builtin_copysign(mag, sgn);
    }
}

// The infinity recovery of C11 Annex G.5.1 for (a + ib)(c + id) when the plain formula
// gave NaN + iNaN: an infinite factor is boxed to +-1 / +-0 per part with NaN parts of
// the other factor set to signed zero, overflowed products likewise, and the product
// recomputed times infinity.  The compiled std::complex product does this (GCC lowers it
// to the plain formula plus a call of __mulsc3 / __muldc3 when both parts are NaN).
template <typename TS>
__device__ __attribute__((noinline)) cx<TS> cmul_recover(TS a, TS b, TS c, TS d, cx<TS> r) {
    const TS one = TS(1), zero = TS(0);
    bool recalc = false;
    if (__builtin_isinf(a) || __builtin_isinf(b)) {
        a = csign(__builtin_isinf(a) ? one : zero, a);
        b = csign(__builtin_isinf(b) ? one : zero, b);
        if (__builtin_isnan(c)) c = csign(zero, c);
        if (__builtin_isnan(d)) d = csign(zero, d);
        recalc = true;
    }
    if (__builtin_isinf(c) || __builtin_isinf(d)) {
        c = csign(__builtin_isinf(c) ? one : zero, c);
        d = csign(__builtin_isinf(d) ? one : zero, d);
        if (__builtin_isnan(a)) a = csign(zero, a);
        if (__builtin_isnan(b)) b = csign(zero, b);
        recalc = true;
    }
    if (!recalc && (__builtin_isinf(a * c) || __builtin_isinf(b * d) || __builtin_isinf(a * d) ||
                    __builtin_isinf(b * c))) {
        if (__builtin_isnan(a)) a = csign(zero, a);
        if (__builtin_isnan(b)) b = csign(zero, b);
        if (__builtin_isnan(c)) c = csign(zero, c);
        if (__builtin_isnan(d)) d = csign(zero, d);
        recalc = true;
    }
    if (recalc) {
        const TS inf = __builtin_huge_valf();
        r.re = inf * (a * c - b * d);
        r.im = inf * (a * d + b * c);
    }
    return r;
}

// f * x for a complex f (the first operand of std::complex's product)
template <typename TS>
__device__ __forceinline__ cx<TS> cmul(cx<TS> f, cx<TS> x) {
    const TS rr = f.re * x.re, ii = f.im * x.im, ri = f.re * x.im, ir = f.im * x.re;
    cx<TS> r{rr - ii, ri + ir};
    if (__builtin_expect(__builtin_isnan(r.re) && __builtin_isnan(r.im), 0)) r = cmul_recover(f.re, f.im, x.re, x.im, r);
    return r;
}
// f * x for a real f
template <typename TS>
__device__ __forceinline__ cx<TS> rmul(TS f, cx<TS> x) {
    return {f * x.re, f * x.im};
}

template <typename TS>
__device__ __forceinline__ cx<TS> fetch(const CxPassArgs<TS> &a, const TS *line, int j) {
    if (static_cast<unsigned>(j) < static_cast<unsigned>(a.n)) {
        const TS *p = line + 2 * static_cast<int64_t>(j) * a.sin;
        return {p[0], p[1]};
    }
    if (j < a.zlo) return {TS(0), TS(0)};
    const int v = bc_fold_packed(j, static_cast<int>(a.n), a.bc);
    if (v < 0) return {a.cval, a.cval};
    const TS *p = line + 2 * static_cast<int64_t>(v >> 1) * a.sin;
    return (v & 1) ? cx<TS>{-p[0], -p[1]} : cx<TS>{p[0], p[1]};
}

// sum += f_k * t (FC: complex taps, else real taps)
template <typename TS, bool FC>
__device__ __forceinline__ cx<TS> tap(const CxPassArgs<TS> &a, int k, cx<TS> t) {
    if constexpr (FC) {
        return cmul(cx<TS>{a.taps[2 * k], a.taps[2 * k + 1]}, t);
    } else {
        return rmul(a.taps[k], t);
    }
}
template <typename TS>
__device__ __forceinline__ cx<TS> conj_tap(const CxPassArgs<TS> &a, int k, cx<TS> t) {
    return cmul(cx<TS>{a.taps[2 * k], -a.taps[2 * k + 1]}, t);
}

template <typename TS, bool FC, int SYM>
__global__ __launch_bounds__(kBlock) void k_cpass(const CxPassArgs<TS> a) {
    const int64_t total = a.n * a.nlines;
    for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int i = static_cast<int>(t % a.n);
        const int64_t line = t / a.n;
        int64_t oin, oout;
        outer_offsets(a.outer, line, &oin, &oout);
        const TS *src = a.in + 2 * oin;
        const int ds = a.ntaps;
        cx<TS> sum;
        if constexpr (SYM == SYM_GENERAL) {
            sum = {TS(0), TS(0)};
            const int b = i - a.origin;
            for (int k = 0; k < ds; ++k) sum = cadd(sum, tap<TS, FC>(a, k, fetch(a, src, b + k)));
        } else if constexpr (SYM == SYM_EVEN || SYM == SYM_ODD || SYM == kSymConj) {
            const int c = i - a.origin + ds - 1;
            sum = tap<TS, FC>(a, 0, fetch(a, src, c));
            for (int k = 1; k < ds; ++k) {
                const cx<TS> r = fetch(a, src, c + k), l = fetch(a, src, c - k);
                if constexpr (SYM == kSymConj) {
                    sum = cadd(sum, cadd(tap<TS, true>(a, k, r), conj_tap(a, k, l)));
                } else {
                    sum = cadd(sum, tap<TS, FC>(a, k, SYM == SYM_EVEN ? cadd(r, l) : csub(r, l)));
                }
            }
        } else {  // D_EVEN, D_ODD, D_CONJ
            const int c = i - a.origin + ds - 1;
            sum = {TS(0), TS(0)};
            for (int k = 0; k < ds; ++k) {
                const cx<TS> r = fetch(a, src, c + k), l = fetch(a, src, c - 1 - k);
                if constexpr (SYM == kSymDConj) {
                    sum = cadd(sum, cadd(tap<TS, true>(a, k, r), conj_tap(a, k, l)));
                } else {
                    sum = cadd(sum, tap<TS, FC>(a, k, SYM == SYM_D_EVEN ? cadd(r, l) : csub(r, l)));
                }
            }
        }
        TS *d = a.out + 2 * (oout + static_cast<int64_t>(i) * a.sout);
        d[0] = sum.re;
        d[1] = sum.im;
    }
}

// Conversion into the complex buffer type (the reference's CopyBuffer casts): a real
// sample v becomes (TS(v), 0), a complex one keeps both parts in TS.
template <typename TI, typename TS, bool SRC_COMPLEX>
__global__ __launch_bounds__(kBlock) void k_to_complex(const TI *in, TS *out, const CastGeo g) {
    for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < g.total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t r = t, oi = 0, oo = 0;
        for (int e = 0; e < g.ndim; ++e) {
            const int64_t c = r % g.size[e];
            r /= g.size[e];
            oi += c * g.sin[e];
            oo += c * g.sout[e];
        }
        if constexpr (SRC_COMPLEX) {
            out[2 * oo] = static_cast<TS>(in[2 * oi]);
            out[2 * oo + 1] = static_cast<TS>(in[2 * oi + 1]);
        } else {
            out[2 * oo] = to_buf<TS>(in[oi]);
            out[2 * oo + 1] = TS(0);
        }
    }
}

template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void k_complex_copy(const TS *in, TD *out, const CastGeo g) {
    for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < g.total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t r = t, oi = 0, oo = 0;
        for (int e = 0; e < g.ndim; ++e) {
            const int64_t c = r % g.size[e];
            r /= g.size[e];
            oi += c * g.sin[e];
            oo += c * g.sout[e];
        }
        out[2 * oo] = static_cast<TD>(in[2 * oi]);
        out[2 * oo + 1] = static_cast<TD>(in[2 * oi + 1]);
    }
}

unsigned grid_of(int64_t total) {
    int64_t nblk = (total + kBlock - 1) / kBlock;
    if (nblk > 65536) nblk = 65536;
    if (nblk < 1) nblk = 1;
    return static_cast<unsigned>(nblk);
}

template <typename TS, bool FC>
hipError_t cpass_sym(int sym, const CxPassArgs<TS> &a, hipStream_t st) {
    const unsigned g = grid_of(a.n * a.nlines);
    switch (sym) {
        case SYM_GENERAL: hipLaunchKernelGGL((k_cpass<TS, FC, SYM_GENERAL>), dim3(g), dim3(kBlock), 0, st, a); break;
        case SYM_EVEN: hipLaunchKernelGGL((k_cpass<TS, FC, SYM_EVEN>), dim3(g), dim3(kBlock), 0, st, a); break;
        case SYM_ODD: hipLaunchKernelGGL((k_cpass<TS, FC, SYM_ODD>), dim3(g), dim3(kBlock), 0, st, a); break;
        case SYM_D_EVEN: hipLaunchKernelGGL((k_cpass<TS, FC, SYM_D_EVEN>), dim3(g), dim3(kBlock), 0, st, a); break;
        case SYM_D_ODD: hipLaunchKernelGGL((k_cpass<TS, FC, SYM_D_ODD>), dim3(g), dim3(kBlock), 0, st, a); break;
        case kSymConj:
            if constexpr (!FC) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_cpass<TS, true, kSymConj>), dim3(g), dim3(kBlock), 0, st, a);
            break;
        case kSymDConj:
            if constexpr (!FC) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_cpass<TS, true, kSymDConj>), dim3(g), dim3(kBlock), 0, st, a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename TS>
hipError_t to_complex_dt(int dt, const void *in, TS *out, const CastGeo &g, hipStream_t st) {
    const unsigned n = grid_of(g.total);
#define DIPAMD_TOCX(CODE, T, CPLX)                                                                          \
    case CODE:                                                                                              \
        hipLaunchKernelGGL((k_to_complex<T, TS, CPLX>), dim3(n), dim3(kBlock), 0, st, static_cast<const T *>(in), \
                           out, g);                                                                         \
        break;
    switch (dt) {
        DIPAMD_TOCX(DIPAMD_DT_UINT8, uint8_t, false)
        DIPAMD_TOCX(DIPAMD_DT_UINT16, uint16_t, false)
        DIPAMD_TOCX(DIPAMD_DT_UINT32, uint32_t, false)
        DIPAMD_TOCX(DIPAMD_DT_SINT8, int8_t, false)
        DIPAMD_TOCX(DIPAMD_DT_SINT16, int16_t, false)
        DIPAMD_TOCX(DIPAMD_DT_SINT32, int32_t, false)
        DIPAMD_TOCX(DIPAMD_DT_SFLOAT, float, false)
        DIPAMD_TOCX(DIPAMD_DT_DFLOAT, double, false)
        DIPAMD_TOCX(DIPAMD_DT_BIN, bin_t, false)
        DIPAMD_TOCX(DIPAMD_DT_SCOMPLEX, float, true)
        DIPAMD_TOCX(DIPAMD_DT_DCOMPLEX, double, true)
        default: return hipErrorInvalidValue;
    }
#undef DIPAMD_TOCX
    return hipGetLastError();
}

}  // namespace

hipError_t launch_cpass(bool is_double, bool complex_filter, int sym, const void *args, hipStream_t stream) {
    if (is_double) {
        const auto &a = *static_cast<const CxPassArgs<double> *>(args);
        return complex_filter ? cpass_sym<double, true>(sym, a, stream) : cpass_sym<double, false>(sym, a, stream);
    }
    const auto &a = *static_cast<const CxPassArgs<float> *>(args);
    return complex_filter ? cpass_sym<float, true>(sym, a, stream) : cpass_sym<float, false>(sym, a, stream);
}

hipError_t launch_to_complex(int dt_in, bool to_double, const void *in, void *out, const void *geo,
                             hipStream_t stream) {
    const CastGeo &g = *static_cast<const CastGeo *>(geo);
    return to_double ? to_complex_dt<double>(dt_in, in, static_cast<double *>(out), g, stream)
                     : to_complex_dt<float>(dt_in, in, static_cast<float *>(out), g, stream);
}

hipError_t launch_complex_copy(bool from_double, bool to_double, const void *in, void *out, const void *geo,
                               hipStream_t stream) {
    const CastGeo &g = *static_cast<const CastGeo *>(geo);
    const unsigned n = grid_of(g.total);
    if (from_double && to_double)
        hipLaunchKernelGGL((k_complex_copy<double, double>), dim3(n), dim3(kBlock), 0, stream,
                           static_cast<const double *>(in), static_cast<double *>(out), g);
    else if (from_double)
        hipLaunchKernelGGL((k_complex_copy<double, float>), dim3(n), dim3(kBlock), 0, stream,
                           static_cast<const double *>(in), static_cast<float *>(out), g);
    else if (to_double)
        hipLaunchKernelGGL((k_complex_copy<float, double>), dim3(n), dim3(kBlock), 0, stream,
                           static_cast<const float *>(in), static_cast<double *>(out), g);
    else
        hipLaunchKernelGGL((k_complex_copy<float, float>), dim3(n), dim3(kBlock), 0, stream,
                           static_cast<const float *>(in), static_cast<float *>(out), g);
    return hipGetLastError();
}

}  // namespace dipamd
