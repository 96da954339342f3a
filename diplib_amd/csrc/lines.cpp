// Line buffers of the separable framework for user line filters (the plug-in operator
// API of Framework::Separable, framework_separable.cpp:146-331): the gather of every
// image line along one dim into a buffer of the buffer type with its boundary
// extension (copy_buffer.cpp:284-350 CopyBuffer with the cast, :751-813 ExpandBuffer,
// every boundary condition), and the scatter of filtered lines back into an image
// with the cast to its type (CopyBuffer, clamp_cast).  The buffers hold the lines
// interleaved -- sample i of line l at i * nlines + l -- so that a kernel with one thread
// per line reads and writes whole cache lines across a wave; include/diplib_amd_framework.hpp
// runs the user's filter between the two.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "dipamd.h"
#include "dipamd_internal.hpp"
#include "sep_kernels.hpp"

using namespace dipamd;

namespace {

int lines_of(const dipamd_image *im, int dim, int64_t *nlines) {
    int64_t n = 1;
    for (int d = 0; d < im->ndim; ++d)
        if (d != dim) n *= im->sizes[d];
    *nlines = n;
    return DIPAMD_OK;
}

// Strides of the interleaved layout: dim `dim` steps by nlines, the other dims number
// the lines (dim 0 fastest).
void interleaved_strides(const dipamd_image *im, int dim, int64_t nlines, int64_t *st) {
    int64_t s = 1;
    for (int d = 0; d < im->ndim; ++d) {
        if (d == dim) {
            st[d] = nlines;
        } else {
            st[d] = s;
            s *= im->sizes[d];
        }
    }
}

int64_t npx_of_image(const dipamd_image *im) {
    int64_t n = 1;
    for (int d = 0; d < im->ndim; ++d) n *= im->sizes[d];
    return n;
}

int check_lines(const dipamd_image *im, int32_t dim, int32_t buffer_dt, const void *buffer) {
    int st;
    if ((st = check_image(im, "image")) != DIPAMD_OK) return st;
    if ((st = require_real(im)) != DIPAMD_OK) return st;
    if (dim < 0 || dim >= im->ndim) return set_error(DIPAMD_E_PARAMETER, "dimension out of range");
    if (buffer_dt != DIPAMD_DT_SFLOAT && buffer_dt != DIPAMD_DT_DFLOAT)
        return set_error(DIPAMD_E_UNSUPPORTED, "line buffers are sfloat or dfloat");
    if (!buffer) return set_error(DIPAMD_E_PARAMETER, "null line buffer");
    return DIPAMD_OK;
}

}  // namespace

extern "C" {

int64_t dipamd_separable_lines_count(const dipamd_image *im, int32_t dim) {
    if (!im || dim < 0 || dim >= im->ndim) return set_error(DIPAMD_E_PARAMETER, "dimension out of range");
    int64_t n = 0;
    lines_of(im, dim, &n);
    return n;
}

int dipamd_separable_lines_gather(const dipamd_image *in, int32_t dim, int32_t border, int32_t bc, int32_t buffer_dt,
                                  void *buffer, void *stream) {
    int st;
    if ((st = check_lines(in, dim, buffer_dt, buffer)) != DIPAMD_OK) return st;
    if (border < 0) return set_error(DIPAMD_E_PARAMETER, "negative border");
    if (bc < 0 || bc > 12) return set_error(DIPAMD_E_PARAMETER, "Boundary condition not recognized");
    if (bc == DIPAMD_BC_ALREADY_EXPANDED)
        return set_error(DIPAMD_E_UNSUPPORTED, "Functionality has not (yet) been implemented");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const bool dbl = buffer_dt == DIPAMD_DT_DFLOAT;
    View src = to_view(in);
    void *tmp = nullptr;
    // the extension kernel reads uint8 (into sfloat) or the buffer type; other types
    // are converted to the buffer type first
    if (!(src.dt == DIPAMD_DT_UINT8 && !dbl) && src.dt != buffer_dt) {
        const hipError_t e = scratch_alloc(&tmp, static_cast<size_t>(npx_of_image(in)) * dt_size(buffer_dt), s);
        if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
        const View t = contiguous_view(tmp, buffer_dt, in->ndim, src.sizes);
        if ((st = cast_to_buffer(src, t, in->ndim, dbl, s)) != DIPAMD_OK) {
            scratch_free(tmp, s);
            return st;
        }
        src = t;
    }
    int64_t nlines = 0;
    lines_of(in, dim, &nlines);
    CastGeo g;
    std::memset(&g, 0, sizeof(g));
    g.ndim = in->ndim;
    g.total = 1;
    int64_t ost[DIPAMD_MAX_DIMS];
    interleaved_strides(in, dim, nlines, ost);
    for (int d = 0; d < in->ndim; ++d) {
        g.size[d] = d == dim ? in->sizes[d] + 2 * static_cast<int64_t>(border) : in->sizes[d];
        g.sin[d] = src.strides[d];
        g.sout[d] = ost[d];
        g.total *= g.size[d];
    }
    const hipError_t e = launch_expand(src.dt == DIPAMD_DT_UINT8 ? IN_U8 : IN_BUF, dbl, src.p, buffer, &g, dim,
                                       static_cast<int>(in->sizes[dim]), border, 0, bc, s);
    if (tmp) scratch_free(tmp, s);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

int dipamd_separable_lines_scatter(const void *buffer, int32_t buffer_dt, const dipamd_image *out, int32_t dim,
                                   void *stream) {
    int st;
    if ((st = check_lines(out, dim, buffer_dt, buffer)) != DIPAMD_OK) return st;
    int64_t nlines = 0;
    lines_of(out, dim, &nlines);
    View src;
    std::memset(&src, 0, sizeof(src));
    src.p = const_cast<void *>(buffer);
    src.dt = buffer_dt;
    int64_t ist[DIPAMD_MAX_DIMS];
    interleaved_strides(out, dim, nlines, ist);
    for (int d = 0; d < DIPAMD_MAX_DIMS; ++d) {
        src.sizes[d] = d < out->ndim ? out->sizes[d] : 1;
        src.strides[d] = d < out->ndim ? ist[d] : 0;
    }
    return cast_from_buffer(src, to_view(out), out->ndim, static_cast<hipStream_t>(stream));
}

}  // extern "C"
