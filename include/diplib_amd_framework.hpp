/*
 * diplib_amd_framework.hpp -- the plug-in operator API of the separable framework on
 * the device: dip::Framework::Separable with a user line filter
 * (include/diplib/framework.h:589-669 SeparableBuffer, SeparableLineFilterParameters,
 * SeparableLineFilter; :760-770 Separable; framework_separable.cpp:35-431).
 *
 * Include it from a HIP translation unit (hipcc); it builds on diplib_amd.hpp.  The
 * line filter is a type with a member
 *
 *     __device__ void Filter(dip::Framework::SeparableLineFilterParameters const &params) const;
 *
 * (the device counterpart of SeparableLineFilter::Filter, framework.h:654-658; the
 * reference's virtual class becomes a template argument, since a device function is
 * bound at compile time).  Separable calls it once per image line of each processed
 * dim, one GPU thread per line, with the line's samples in an input buffer of the
 * buffer type extended by `border` samples on each side by the boundary condition, and
 * an output buffer of the buffer type whose samples are then cast to the output type.
 * Both buffers hold all lines interleaved: sample i of a line is at buffer[i * stride]
 * (stride = the number of lines, so the threads of a wave touch consecutive
 * addresses); the input buffer's sample -k (k <= border) is at buffer[-k * stride].
 *
 * As the reference, the dims are processed in order of increasing input stride, each
 * pass reading the previous pass's result (held in the buffer type); the first pass
 * reads the input (any real type, cast to the buffer type), the last writes `out`
 * (reforged to the input's sizes and outType unless protected).  Dims with process
 * false or size 1 are skipped; border and boundaryCondition broadcast from one element.
 * Differences: the line position (params.position) is given as the line index
 * (params.line); buffers are always used (the UseInputBuffer / UseOutputBuffer
 * behaviour); images are scalar; the buffer type is sfloat or dfloat.  A filter that
 * must round as the reference's compiled loop does is compiled without contraction
 * (hipcc -ffp-contract=off), as this library is.
 */
#ifndef DIPLIB_AMD_FRAMEWORK_HPP
#define DIPLIB_AMD_FRAMEWORK_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "diplib_amd.hpp"

namespace dip {
namespace Framework {

/* framework.h:589-598 */
struct SeparableBuffer {
    void *buffer;
    dip::uint length;  // samples of the line
    dip::uint border;  // extension samples on each side (input buffer)
    dip::sint stride;  // elements between consecutive samples
};

/* framework.h:601-611 */
struct SeparableLineFilterParameters {
    SeparableBuffer const &inBuffer;
    SeparableBuffer &outBuffer;
    dip::uint dimension;  // the dim being filtered
    dip::uint pass;       // 0 .. nPasses - 1
    dip::uint nPasses;
    dip::uint line;       // index of the line (the other dims, dim 0 fastest)
};

namespace detail {

template <typename LineFilter>
__global__ void SeparableLineKernel(LineFilter filter, void *in, void *out, dip::uint length, dip::uint border,
                                    dip::uint nlines, dip::uint sampleSize, dip::uint dimension, dip::uint pass,
                                    dip::uint nPasses) {
    const dip::uint l = static_cast<dip::uint>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    char *ib = static_cast<char *>(in) + (border * nlines + l) * sampleSize;
    char *ob = static_cast<char *>(out) + l * sampleSize;
    const SeparableBuffer inBuffer{ib, length, border, static_cast<dip::sint>(nlines)};
    SeparableBuffer outBuffer{ob, length, 0, static_cast<dip::sint>(nlines)};
    const SeparableLineFilterParameters params{inBuffer, outBuffer, dimension, pass, nPasses, l};
    filter.Filter(params);
}

/* Device scratch of the stream, released when the stream reaches it. */
struct StreamBuffer {
    void *p = nullptr;
    hipStream_t s = nullptr;
    StreamBuffer(std::size_t bytes, hipStream_t stream) : s(stream) {
        if (hipMallocAsync(&p, bytes ? bytes : 1, stream) != hipSuccess) throw RunTimeError("hipMallocAsync failed");
    }
    ~StreamBuffer() { (void)hipFreeAsync(p, s); }
    StreamBuffer(StreamBuffer const &) = delete;
    StreamBuffer &operator=(StreamBuffer const &) = delete;
};

}  // namespace detail

/* framework.h:760-770; framework_separable.cpp:35-431 */
template <typename LineFilter>
void Separable(Image const &c_in, Image &c_out, DataType bufferType, DataType outType, BooleanArray process,
               UnsignedArray border, StringArray const &boundaryCondition, LineFilter const &lineFilter) {
    if (!c_in.IsForged()) throw ParameterError("Image is not forged");
    if (bufferType != DT_SFLOAT && bufferType != DT_DFLOAT) throw ParameterError("Data type not supported");
    const dip::uint nDims = c_in.Dimensionality();
    if (nDims < 1) throw ParameterError("Dimensionality not supported");
    if (process.empty()) process.assign(nDims, true);
    if (process.size() != nDims) throw ParameterError("Array parameter has the wrong number of elements");
    if (border.empty()) border.assign(nDims, 0);
    if (border.size() == 1) border.assign(nDims, border[0]);
    if (border.size() != nDims) throw ParameterError("Array parameter has the wrong number of elements");
    std::vector<int32_t> bc = dip::detail::BoundaryCodes(boundaryCondition);
    if (bc.empty()) bc.push_back(DIPAMD_BC_SYMMETRIC_MIRROR);
    if (bc.size() == 1) bc.assign(nDims, bc[0]);
    if (bc.size() != nDims) throw ParameterError("Array parameter has the wrong number of elements");

    // processing order: increasing input stride (framework_separable.cpp:146-163)
    std::vector<dip::uint> order;
    for (dip::uint d = 0; d < nDims; ++d)
        if (process[d] && c_in.Size(d) > 1) order.push_back(d);
    std::stable_sort(order.begin(), order.end(),
                     [&](dip::uint a, dip::uint b) { return c_in.Stride(a) < c_in.Stride(b); });

    Image in = c_in;  // keeps the input's data alive if `out` is `in`
    dip::detail::PrepareOutput(in, c_out, outType);
    const hipStream_t s = dipamd::CurrentStream();
    const dip::uint nPasses = order.size();
    if (nPasses == 0) {
        // out = in, cast (a Gaussian of sigma 0 processes no dim: the same copy)
        dipamd_image a = in.View(), b = c_out.View();
        const double zero = 0.0;
        dip::detail::Check(dipamd_gauss_fir(&a, &b, &zero, 1, nullptr, 0, nullptr, 0, 3.0, DIPAMD_EXACT, s));
        return;
    }
    const std::size_t ss = bufferType.SizeOf();
    Image cur = in;
    Image tmp;  // the intermediate, in the buffer type
    for (dip::uint k = 0; k < nPasses; ++k) {
        const dip::uint d = order[k];
        const dip::uint n = in.Size(d);
        dipamd_image src = cur.View();
        const int64_t nl = dipamd_separable_lines_count(&src, static_cast<int32_t>(d));
        if (nl < 0) dip::detail::Check(static_cast<int>(nl));
        const dip::uint nlines = static_cast<dip::uint>(nl);
        detail::StreamBuffer ib(nlines * (n + 2 * border[d]) * ss, s);
        detail::StreamBuffer ob(nlines * n * ss, s);
        dip::detail::Check(dipamd_separable_lines_gather(&src, static_cast<int32_t>(d), static_cast<int32_t>(border[d]),
                                                         bc[d], bufferType.Code(), ib.p, s));
        const unsigned threads = 256;
        const unsigned blocks = static_cast<unsigned>((nlines + threads - 1) / threads);
        hipLaunchKernelGGL(detail::SeparableLineKernel<LineFilter>, dim3(blocks), dim3(threads), 0, s, lineFilter, ib.p,
                           ob.p, n, border[d], nlines, ss, d, k, nPasses);
        if (hipGetLastError() != hipSuccess) throw RunTimeError("line filter launch failed");
        Image *dst = &c_out;
        if (k + 1 < nPasses) {
            if (!tmp.IsForged()) tmp = Image(in.Sizes(), 1, bufferType);
            dst = &tmp;
        }
        dipamd_image o = dst->View();
        dip::detail::Check(dipamd_separable_lines_scatter(ob.p, bufferType.Code(), &o, static_cast<int32_t>(d), s));
        cur = *dst;
    }
}

}  // namespace Framework
}  // namespace dip

#endif  // DIPLIB_AMD_FRAMEWORK_HPP
