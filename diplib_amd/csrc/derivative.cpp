// The derivative family on the Gaussian FIR path (derivative.cpp:120-418 of the
// reference): dip::Gradient, dip::GradientMagnitude, dip::Hessian and dip::Laplace
// with the FIR method.  Each Gaussian derivative is the plain drop-in call
// (dipamd_gauss_fir); GradientMagnitude and Laplace combine them in one element-wise
// pass (k_elementwise.hip) with the reference's operation order and type.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dipamd.h"
#include "dipamd_internal.hpp"
#include "sep_kernels.hpp"

namespace dipamd {
hipError_t launch_combine(bool is_double, int op, const void *const *d, int nd, void *out, int64_t n,
                          hipStream_t stream);
}  // namespace dipamd

using namespace dipamd;

namespace {

// FindGradientDimensions (derivative.cpp:120-137), the Gaussian form: dims with
// process set, size > 1 and sigma > 0.
std::vector<int> gradient_dims(const dipamd_image *in, const double *sig, const uint8_t *proc) {
    std::vector<int> dims;
    for (int d = 0; d < in->ndim; ++d)
        if (proc[d] && in->sizes[d] > 1 && sig[d] > 0.0) dims.push_back(d);
    return dims;
}

bool contiguous(const dipamd_image *im) {
    int64_t st = 1;
    for (int d = 0; d < im->ndim; ++d) {
        if (im->sizes[d] > 1 && im->strides[d] != st) return false;
        st *= im->sizes[d];
    }
    return true;
}

}  // namespace

extern "C" {

int dipamd_derivative_outputs(int32_t function, int32_t ndiff) {
    if (ndiff < 1 || ndiff > DIPAMD_MAX_DIMS) return set_error(DIPAMD_E_PARAMETER, "Dimensionality not supported");
    switch (function) {
        case DIPAMD_FN_GRADIENT: return ndiff;
        case DIPAMD_FN_HESSIAN: return ndiff * (ndiff + 1) / 2;
        case DIPAMD_FN_GRADIENT_MAGNITUDE:
        case DIPAMD_FN_LAPLACE: return 1;
    }
    return set_error(DIPAMD_E_PARAMETER, "unknown derivative function");
}

int dipamd_derivative(int32_t function, const dipamd_image *in, const dipamd_image *outs, int32_t nouts,
                      const double *sigmas, int32_t nsigmas, const int32_t *bc, int32_t nbc, const uint8_t *process,
                      int32_t nprocess, double truncation, uint32_t flags, void *stream) {
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    const int nd = in->ndim;
    // ArrayUseParameter: sigmas default 1, process default all (derivative.cpp:127-129)
    double sig[DIPAMD_MAX_DIMS];
    uint8_t proc[DIPAMD_MAX_DIMS];
    if (nsigmas != 0 && nsigmas != 1 && nsigmas != nd)
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    if (nprocess != 0 && nprocess != 1 && nprocess != nd)
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    for (int d = 0; d < nd; ++d) {
        sig[d] = nsigmas == 0 ? 1.0 : (nsigmas == 1 ? sigmas[0] : sigmas[d]);
        proc[d] = nprocess == 0 ? 1 : (nprocess == 1 ? process[0] : process[d]);
    }
    const std::vector<int> dims = gradient_dims(in, sig, proc);
    const int n = static_cast<int>(dims.size());
    if (n < 1) return set_error(DIPAMD_E_PARAMETER, "Dimensionality not supported");
    const int want = dipamd_derivative_outputs(function, n);
    if (want < 0) return want;
    if (nouts != want) return set_error(DIPAMD_E_ARRAY_LENGTH, "wrong number of output images for this function");
    for (int k = 0; k < nouts; ++k) {
        if ((st = check_image(&outs[k], "out")) != DIPAMD_OK) return st;
        if (outs[k].ndim != nd) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
        for (int d = 0; d < nd; ++d)
            if (outs[k].sizes[d] != in->sizes[d]) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
        // the reference strips an output that aliases the input (derivative.cpp:156-158);
        // here the caller owns the memory, so an aliasing output is an error
        if (function != DIPAMD_FN_GRADIENT_MAGNITUDE && function != DIPAMD_FN_LAPLACE &&
            views_overlap(to_view(in), to_view(&outs[k]), nd))
            return set_error(DIPAMD_E_PARAMETER, "an output image overlaps the input");
    }
    int32_t order[DIPAMD_MAX_DIMS] = {0};
    auto derivative = [&](const dipamd_image *out) {
        return dipamd_gauss_fir(in, out, sig, nd, order, nd, bc, nbc, truncation, flags, stream);
    };
    if (function == DIPAMD_FN_GRADIENT) {  // derivative.cpp:160-169
        for (int i = 0; i < n; ++i) {
            order[dims[i]] = 1;
            if ((st = derivative(&outs[i])) != DIPAMD_OK) return st;
            order[dims[i]] = 0;
        }
        return DIPAMD_OK;
    }
    if (function == DIPAMD_FN_HESSIAN) {  // derivative.cpp:348-369
        int k = 0;
        for (int i = 0; i < n; ++i) {  // the diagonal first
            order[dims[i]] = 2;
            if ((st = derivative(&outs[k++])) != DIPAMD_OK) return st;
            order[dims[i]] = 0;
        }
        for (int j = 1; j < n; ++j) {  // then above the diagonal, column by column
            for (int i = 0; i < j; ++i) {
                order[dims[i]] = 1;
                order[dims[j]] = 1;
                if ((st = derivative(&outs[k++])) != DIPAMD_OK) return st;
                order[dims[i]] = 0;
                order[dims[j]] = 0;
            }
        }
        return DIPAMD_OK;
    }
    // GradientMagnitude (derivative.cpp:174-210) and Laplace (:374-418): the terms in
    // buffer-type scratch images, combined in one element-wise pass
    const int buf_dt = suggest_flex(in->dtype);
    const bool is_double = buf_dt == DIPAMD_DT_DFLOAT;
    if (outs[0].dtype != buf_dt)
        return set_error(DIPAMD_E_UNSUPPORTED, "the output must have the floating-point type of the input");
    const hipStream_t hs = static_cast<hipStream_t>(stream);
    int64_t npx = 1;
    for (int d = 0; d < nd; ++d) npx *= in->sizes[d];
    const size_t bytes = static_cast<size_t>(npx) * (is_double ? 8 : 4);
    std::vector<void *> bufs;
    auto release = [&]() {
        for (void *p : bufs) scratch_free(p, hs);
    };
    const void *terms[DIPAMD_MAX_DIMS];
    for (int i = 0; i < n; ++i) {
        void *p = nullptr;
        hipError_t e = scratch_alloc(&p, bytes, hs);
        if (e != hipSuccess) {
            release();
            return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
        }
        bufs.push_back(p);
        dipamd_image t;
        t.origin = p;
        t.dtype = buf_dt;
        t.ndim = nd;
        int64_t stride = 1;
        for (int d = 0; d < DIPAMD_MAX_DIMS; ++d) {
            t.sizes[d] = d < nd ? in->sizes[d] : 1;
            t.strides[d] = d < nd ? stride : 0;
            if (d < nd) stride *= in->sizes[d];
        }
        order[dims[i]] = function == DIPAMD_FN_LAPLACE ? 2 : 1;
        st = derivative(&t);
        order[dims[i]] = 0;
        if (st != DIPAMD_OK) {
            release();
            return st;
        }
        terms[i] = p;
    }
    const int op = function == DIPAMD_FN_LAPLACE ? 0 : 1;
    hipError_t e;
    if (contiguous(&outs[0])) {
        e = launch_combine(is_double, op, terms, n, outs[0].origin, npx, hs);
    } else {  // combine into the first term, then copy into the strided output
        e = launch_combine(is_double, op, terms, n, bufs[0], npx, hs);
        if (e == hipSuccess) {
            int64_t sz[DIPAMD_MAX_DIMS];
            for (int d = 0; d < nd; ++d) sz[d] = in->sizes[d];
            st = cast_from_buffer(contiguous_view(bufs[0], buf_dt, nd, sz), to_view(&outs[0]), nd, hs);
            if (st != DIPAMD_OK) {
                release();
                return st;
            }
        }
    }
    release();
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

}  // extern "C"
