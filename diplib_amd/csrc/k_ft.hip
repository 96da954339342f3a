// Element-wise kernels of the Fourier-domain Gaussian (dip::GaussFT): the boundary
// extension of the image into the padded transform buffer (ExpandInput /
// ExtendImageToSize, gauss.cpp:380-393, with the image in the centred window of
// image_manip.cpp:561-564), the product of the half spectrum with the separable
// frequency response (GaussFTLineFilter::Filter, gauss.cpp:360-376), and the crop of
// the centred window with the inverse transform's 1/N scale.
#include "ft_common.hpp"

namespace dipamd {

namespace {

constexpr int kFtBlock = 256;

// Coordinates of element t of the padded buffer: transformed coordinates i[] and the
// batch index.
__device__ __forceinline__ int64_t ft_split(const FtGeo &g, int64_t t, int64_t *i) {
#pragma unroll
    for (int k = 0; k < kFtMaxT; ++k) {
        i[k] = t % g.P[k];
        t /= g.P[k];
    }
    return t;
}

// Input offset of batch index b (over the batch dims, in increasing dim order).
__device__ __forceinline__ int64_t ft_batch_offset(const FtGeo &g, int64_t b, const int64_t *stride) {
    int64_t o = 0;
    for (int k = 0; k < g.nb; ++k) {
        const int d = g.bdim[k];
        o += (b % g.size[d]) * stride[d];
        b /= g.size[d];
    }
    return o;
}

template <typename TI, typename TB>
__global__ __launch_bounds__(kFtBlock) void k_ft_expand(const TI *in, TB *pad, const FtGeo g) {
    for (int64_t t = blockIdx.x * static_cast<int64_t>(kFtBlock) + threadIdx.x; t < g.nreal;
         t += static_cast<int64_t>(gridDim.x) * kFtBlock) {
        int64_t i[kFtMaxT];
        const int64_t b = ft_split(g, t, i);
        int64_t src = ft_batch_offset(g, b, g.sin);
        // ExtendRegion extends one dim after the other (boundary.cpp:297-319), each over
        // the lines the earlier dims already extended: the last dim's condition applies
        // outermost, and a constant condition replaces whatever lower dims would give
        bool is_const = false;
        TB cval = TB(0);
        int sign = 1;
        for (int kk = 0; kk < g.nt; ++kk) {
            const int k = g.ext_order[kk];
            const int d = g.tdim[k];
            int64_t j = i[k] - g.off[k];
            const int64_t n = g.size[d];
            if (j < 0 || j >= n) {
                int mul;
                bool c;
                j = bc_map(j, n, g.bc[k], &mul, &c);
                if (c) {
                    cval = static_cast<TB>(g.cval[k]);
                    is_const = true;
                    break;
                }
                sign *= mul;
            }
            src += j * g.sin[d];
        }
        const TB v = is_const ? cval : to_buf<TB>(in[src]);
        pad[t] = sign < 0 ? -v : v;
    }
}

template <typename TB>
struct Cplx {
    TB x, y;
};

template <typename TB>
__device__ __forceinline__ Cplx<TB> cmul(Cplx<TB> a, Cplx<TB> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// spec: the half spectrum, (P0 / 2 + 1) x P1 x P2 x batch; luts: each transformed
// dim's response in FFT order (frequency k at index k).
template <typename TB>
__global__ __launch_bounds__(kFtBlock) void k_ft_multiply(Cplx<TB> *spec, const Cplx<TB> *luts, const FtGeo g) {
    const int64_t h0 = g.P[0] / 2 + 1;
    for (int64_t t = blockIdx.x * static_cast<int64_t>(kFtBlock) + threadIdx.x; t < g.nhalf;
         t += static_cast<int64_t>(gridDim.x) * kFtBlock) {
        int64_t r = t;
        const int64_t k0 = r % h0;
        r /= h0;
        const int64_t k1 = r % g.P[1];
        r /= g.P[1];
        const int64_t k2 = r % g.P[2];
        // the other dims' factors first, in dim order, then this one's
        Cplx<TB> w = {TB(1), TB(0)};
        if (g.nt > 1) w = cmul(w, luts[g.lut_off[1] + k1]);
        if (g.nt > 2) w = cmul(w, luts[g.lut_off[2] + k2]);
        Cplx<TB> v = cmul(spec[t], w);
        spec[t] = cmul(v, luts[g.lut_off[0] + k0]);
    }
}

template <typename TB, typename TO>
__global__ __launch_bounds__(kFtBlock) void k_ft_crop(const TB *pad, TO *out, TB scale, const FtGeo g,
                                                     int64_t npx) {
    for (int64_t t = blockIdx.x * static_cast<int64_t>(kFtBlock) + threadIdx.x; t < npx;
         t += static_cast<int64_t>(gridDim.x) * kFtBlock) {
        // image coordinates, dim 0 fastest
        int64_t r = t, o = 0, p = 0, bidx = 0, bmul = 1;
        int64_t c[DIPAMD_MAX_DIMS];
        for (int d = 0; d < g.ndim; ++d) {
            c[d] = r % g.size[d];
            r /= g.size[d];
            o += c[d] * g.sout[d];
        }
        for (int k = 0; k < g.nb; ++k) {
            const int d = g.bdim[k];
            bidx += c[d] * bmul;
            bmul *= g.size[d];
        }
        p = bidx;
        for (int k = kFtMaxT - 1; k >= 0; --k) p = p * g.P[k] + (k < g.nt ? c[g.tdim[k]] + g.off[k] : 0);
        out[o] = static_cast<TO>(pad[p] * scale);
    }
}

unsigned grid_for(int64_t n) {
    int64_t b = (n + kFtBlock - 1) / kFtBlock;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return static_cast<unsigned>(b);
}

}  // namespace

hipError_t launch_ft_expand(int in_dt, bool is_double, const void *in, void *pad, const FtGeo &g, hipStream_t st) {
    const unsigned nb = grid_for(g.nreal);
    if (is_double) {
        if (in_dt != DIPAMD_DT_DFLOAT) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_ft_expand<double, double>), dim3(nb), dim3(kFtBlock), 0, st,
                           static_cast<const double *>(in), static_cast<double *>(pad), g);
    } else if (in_dt == DIPAMD_DT_UINT8) {
        hipLaunchKernelGGL((k_ft_expand<uint8_t, float>), dim3(nb), dim3(kFtBlock), 0, st,
                           static_cast<const uint8_t *>(in), static_cast<float *>(pad), g);
    } else if (in_dt == DIPAMD_DT_SFLOAT) {
        hipLaunchKernelGGL((k_ft_expand<float, float>), dim3(nb), dim3(kFtBlock), 0, st,
                           static_cast<const float *>(in), static_cast<float *>(pad), g);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ft_multiply(bool is_double, void *spec, const void *luts, const FtGeo &g, hipStream_t st) {
    const unsigned nb = grid_for(g.nhalf);
    if (is_double)
        hipLaunchKernelGGL((k_ft_multiply<double>), dim3(nb), dim3(kFtBlock), 0, st, static_cast<Cplx<double> *>(spec),
                           static_cast<const Cplx<double> *>(luts), g);
    else
        hipLaunchKernelGGL((k_ft_multiply<float>), dim3(nb), dim3(kFtBlock), 0, st, static_cast<Cplx<float> *>(spec),
                           static_cast<const Cplx<float> *>(luts), g);
    return hipGetLastError();
}

hipError_t launch_ft_crop(bool is_double, const void *pad, void *out, int out_dt, double scale, const FtGeo &g,
                          hipStream_t st) {
    int64_t npx = 1;
    for (int d = 0; d < g.ndim; ++d) npx *= g.size[d];
    const unsigned nb = grid_for(npx);
    if (is_double) {
        if (out_dt != DIPAMD_DT_DFLOAT) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_ft_crop<double, double>), dim3(nb), dim3(kFtBlock), 0, st,
                           static_cast<const double *>(pad), static_cast<double *>(out), scale, g, npx);
    } else {
        if (out_dt != DIPAMD_DT_SFLOAT) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_ft_crop<float, float>), dim3(nb), dim3(kFtBlock), 0, st, static_cast<const float *>(pad),
                           static_cast<float *>(out), static_cast<float>(scale), g, npx);
    }
    return hipGetLastError();
}

namespace {

template <typename TB>
__global__ __launch_bounds__(kFtBlock) void k_ft_cshift(const Cplx<TB> *in, Cplx<TB> *out, const FtCGeo g) {
    for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < g.total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t r = t, o = 0, st = 1;
        for (int d = 0; d < g.ndim; ++d) {
            const int64_t n = g.size[d];
            const int64_t i = r % n;
            r /= n;
            int64_t j = i + g.shift[d];
            j = j >= n ? j - n : j;
            o += j * st;
            st *= n;
        }
        out[o] = in[t];
    }
}

// the centred response products in dim order, then the scale (the reference multiplies
// the spectrum by each dim's response; the inverse scales by 1/N)
template <typename TB>
__global__ __launch_bounds__(kFtBlock) void k_ft_cmultiply(Cplx<TB> *data, const Cplx<TB> *luts, TB scale,
                                                           const FtCGeo g) {
    for (int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; t < g.total;
         t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t r = t;
        Cplx<TB> v = data[t];
        for (int d = 0; d < g.ndim; ++d) {
            const int64_t i = r % g.size[d];
            r /= g.size[d];
            if (g.lut_off[d] >= 0) {
                const Cplx<TB> h = luts[g.lut_off[d] + i];
                v = Cplx<TB>{v.x * h.x - v.y * h.y, v.x * h.y + v.y * h.x};
            }
        }
        if (scale != TB(1)) v = Cplx<TB>{v.x * scale, v.y * scale};
        data[t] = v;
    }
}

unsigned cgrid(int64_t total) {
    int64_t n = (total + kFtBlock - 1) / kFtBlock;
    return static_cast<unsigned>(n > 65536 ? 65536 : (n < 1 ? 1 : n));
}

}  // namespace

hipError_t launch_ft_cshift(bool is_double, const void *in, void *out, const FtCGeo &g, hipStream_t st) {
    if (is_double)
        hipLaunchKernelGGL(k_ft_cshift<double>, dim3(cgrid(g.total)), dim3(kFtBlock), 0, st,
                           static_cast<const Cplx<double> *>(in), static_cast<Cplx<double> *>(out), g);
    else
        hipLaunchKernelGGL(k_ft_cshift<float>, dim3(cgrid(g.total)), dim3(kFtBlock), 0, st,
                           static_cast<const Cplx<float> *>(in), static_cast<Cplx<float> *>(out), g);
    return hipGetLastError();
}

hipError_t launch_ft_cmultiply(bool is_double, void *data, const void *luts, double scale, const FtCGeo &g,
                               hipStream_t st) {
    if (is_double)
        hipLaunchKernelGGL(k_ft_cmultiply<double>, dim3(cgrid(g.total)), dim3(kFtBlock), 0, st,
                           static_cast<Cplx<double> *>(data), static_cast<const Cplx<double> *>(luts), scale, g);
    else
        hipLaunchKernelGGL(k_ft_cmultiply<float>, dim3(cgrid(g.total)), dim3(kFtBlock), 0, st,
                           static_cast<Cplx<float> *>(data), static_cast<const Cplx<float> *>(luts),
                           static_cast<float>(scale), g);
    return hipGetLastError();
}

}  // namespace dipamd
