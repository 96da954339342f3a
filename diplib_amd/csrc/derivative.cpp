// The derivative family on the Gaussian FIR path (derivative.cpp:120-521 of the
// reference) and the structure tensor (structure_tensor.cpp:32-57): dip::Gradient,
// GradientMagnitude, Hessian, Laplace, Dgg, LaplacePlusDgg, LaplaceMinusDgg and
// StructureTensor with the FIR method.  A 2-D uint8 image whose Gaussian derivatives
// all have radius <= 4 takes the one-launch form (k_stream2d_deriv.hip: one read of the
// input, every output written once, bit-identical to the separate calls).  Otherwise
// each Gaussian derivative is the plain drop-in call (dipamd_gauss_fir, or
// dipamd_gauss_iir with DIPAMD_GAUSS_IIR) and the combined forms add one element-wise
// pass (k_elementwise.hip) with the reference's operation order and type.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>
#include <vector>

#include "dipamd.h"
#include "dipamd_internal.hpp"
#include "sep_common.hpp"
#include "sep_kernels.hpp"

namespace dipamd {
hipError_t launch_combine(bool is_double, int op, const void *const *d, int nd, void *const *out, int64_t n,
                          hipStream_t stream);
hipError_t launch_masked(bool is_double, int op, const void *x, const void *m, void *a, const void *den, const void *dw,
                         const void *dm, void *out, int64_t n, hipStream_t stream);
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

int array_param(const double *v, int32_t nv, int nd, double dflt, double *out) {
    if (nv != 0 && nv != 1 && nv != nd)
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    for (int d = 0; d < nd; ++d) out[d] = nv == 0 ? dflt : (nv == 1 ? v[0] : v[d]);
    return DIPAMD_OK;
}

int check_outputs(int function, const dipamd_image *in, const dipamd_image *outs, int32_t nouts, int want) {
    int st;
    if (nouts != want) return set_error(DIPAMD_E_ARRAY_LENGTH, "wrong number of output images for this function");
    for (int k = 0; k < nouts; ++k) {
        if ((st = check_image(&outs[k], "out")) != DIPAMD_OK) return st;
        if (outs[k].ndim != in->ndim) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
        for (int d = 0; d < in->ndim; ++d)
            if (outs[k].sizes[d] != in->sizes[d]) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
        // the reference strips an output that aliases the input (derivative.cpp:156-158);
        // here the caller owns the memory, so an aliasing output is an error
        if ((function == DIPAMD_FN_GRADIENT || function == DIPAMD_FN_HESSIAN ||
             function == DIPAMD_FN_STRUCTURE_TENSOR) &&
            views_overlap(to_view(in), to_view(&outs[k]), in->ndim))
            return set_error(DIPAMD_E_PARAMETER, "an output image overlaps the input");
    }
    return DIPAMD_OK;
}

// Scratch images of the buffer type, contiguous, released together.
struct Scratch {
    hipStream_t stream;
    std::vector<void *> bufs;
    ~Scratch() {
        for (void *p : bufs) scratch_free(p, stream);
    }
    int make(const dipamd_image *in, int dt, dipamd_image *t) {
        int64_t npx = 1;
        for (int d = 0; d < in->ndim; ++d) npx *= in->sizes[d];
        void *p = nullptr;
        const hipError_t e = scratch_alloc(&p, static_cast<size_t>(npx) * (dt == DIPAMD_DT_DFLOAT ? 8 : 4), stream);
        if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
        bufs.push_back(p);
        t->origin = p;
        t->dtype = dt;
        t->ndim = in->ndim;
        int64_t stride = 1;
        for (int d = 0; d < DIPAMD_MAX_DIMS; ++d) {
            t->sizes[d] = d < in->ndim ? in->sizes[d] : 1;
            t->strides[d] = d < in->ndim ? stride : 0;
            if (d < in->ndim) stride *= in->sizes[d];
        }
        return DIPAMD_OK;
    }
};

// The Gaussian derivatives of `in` with the given per-dim orders (one list of nd
// orders per image), each into `dst[k]`.
// One dip::Gauss call with the method the flags select: 0 FIR, 1 IIR, 2 FT.
int gauss(int method, const dipamd_image *in, const dipamd_image *out, const double *sig, const int32_t *orders,
          int32_t norders, const int32_t *bc, int32_t nbc, double truncation, uint32_t flags, void *stream) {
    if (method == 1)
        return dipamd_gauss_iir(in, out, sig, in->ndim, orders, norders, bc, nbc, nullptr, 0,
                                DIPAMD_IIR_DISCRETE_TIME_FIT, truncation, 0, stream);
    if (method == 2) {
        const int32_t mirror = DIPAMD_BC_SYMMETRIC_MIRROR;  // dip::Gauss's FT default (derivative.cpp:74-77)
        return nbc == 0 ? dipamd_gauss_ft(in, out, sig, in->ndim, orders, norders, &mirror, 1, truncation, 0, stream)
                        : dipamd_gauss_ft(in, out, sig, in->ndim, orders, norders, bc, nbc, truncation, 0, stream);
    }
    return dipamd_gauss_fir(in, out, sig, in->ndim, orders, norders, bc, nbc, truncation,
                            flags & static_cast<uint32_t>(DIPAMD_FAST | DIPAMD_NO_FUSION | DIPAMD_FORCE_FUSION),
                            stream);
}

// DIPAMD_DERIV_FUSION=0 turns the one-launch form off (experiments).
bool deriv_fusion_enabled() {
    static const bool on = [] {
        const char *e = std::getenv("DIPAMD_DERIV_FUSION");
        return !(e && e[0] == '0');
    }();
    return on;
}

// The one-launch form (run_stream2d_deriv) for a 2-D uint8 image, or -100.
int try_stream_derivative(int32_t function, const dipamd_image *in, const dipamd_image *outs, int32_t nouts,
                          const double *sig, const int32_t *bc, int32_t nbc, double truncation, uint32_t flags,
                          void *stream) {
    if (!deriv_fusion_enabled() || (flags & DIPAMD_NO_FUSION) || (flags & (DIPAMD_GAUSS_IIR | DIPAMD_GAUSS_FT)))
        return -100;
    if (in->ndim != 2 || in->dtype != DIPAMD_DT_UINT8) return -100;
    int mode, maxorder;
    switch (function) {
        case DIPAMD_FN_GRADIENT: mode = kDerivGradient, maxorder = 1; break;
        case DIPAMD_FN_GRADIENT_MAGNITUDE: mode = kDerivGradientMagnitude, maxorder = 1; break;
        case DIPAMD_FN_LAPLACE: mode = kDerivLaplace, maxorder = 2; break;
        case DIPAMD_FN_HESSIAN: mode = kDerivHessian, maxorder = 2; break;
        case DIPAMD_FN_DGG: mode = kDerivDgg, maxorder = 2; break;
        case DIPAMD_FN_LAPLACE_PLUS_DGG: mode = kDerivLaplacePlusDgg, maxorder = 2; break;
        case DIPAMD_FN_LAPLACE_MINUS_DGG: mode = kDerivLaplaceMinusDgg, maxorder = 2; break;
        default: return -100;
    }
    int bcs[2];
    if (nbc != 0 && nbc != 1 && nbc != 2) return -100;
    for (int d = 0; d < 2; ++d) bcs[d] = nbc == 0 ? DIPAMD_BC_SYMMETRIC_MIRROR : (nbc == 1 ? bc[0] : bc[d]);
    // the orders each function combines (Laplace: 0 and 2 only)
    InternFilter F[2][3];
    const InternFilter *fx[3] = {nullptr, nullptr, nullptr}, *fy[3] = {nullptr, nullptr, nullptr};
    for (int o = 0; o <= maxorder; ++o) {
        if (function == DIPAMD_FN_LAPLACE && o == 1) continue;
        for (int d = 0; d < 2; ++d) {
            if (gauss_radius(sig[d], o, truncation, false) > kStreamRMaxHost) return -100;
            const std::vector<double> t = half_gaussian(sig[d], o, truncation, false);
            if (intern_filter(t.data(), static_cast<int>(t.size()), (o & 1) ? DIPAMD_SYM_ODD : DIPAMD_SYM_EVEN, -1,
                              F[d][o]) != DIPAMD_OK)
                return -100;
        }
        fx[o] = &F[0][o];
        fy[o] = &F[1][o];
    }
    View v[3];
    for (int k = 0; k < nouts && k < 3; ++k) v[k] = to_view(&outs[k]);
    return run_stream2d_deriv(mode, to_view(in), v, nouts, fx, fy, bcs[0], bcs[1], (flags & DIPAMD_FAST) != 0,
                              static_cast<hipStream_t>(stream));
}

int derivative_method(uint32_t flags) {
    return (flags & DIPAMD_GAUSS_FT) ? 2 : ((flags & DIPAMD_GAUSS_IIR) ? 1 : 0);
}
int tensor_method(uint32_t flags) {
    return (flags & DIPAMD_TENSOR_FT) ? 2 : ((flags & DIPAMD_TENSOR_IIR) ? 1 : 0);
}

int derivatives(const dipamd_image *in, const std::vector<std::vector<int32_t>> &orders, const dipamd_image *dst,
                const double *sig, const int32_t *bc, int32_t nbc, double truncation, uint32_t flags, void *stream) {
    const int method = derivative_method(flags);
    for (size_t k = 0; k < orders.size(); ++k) {
        const int st = gauss(method, in, &dst[k], sig, orders[k].data(), in->ndim, bc, nbc, truncation, flags,
                             stream);
        if (st != DIPAMD_OK) return st;
    }
    return DIPAMD_OK;
}

// Orders of the gradient (first derivative along each of `dims`) and of the Hessian
// (second derivatives on the diagonal, then the mixed ones above the diagonal column
// by column: derivative.cpp:348-369).
std::vector<std::vector<int32_t>> gradient_orders(int nd, const std::vector<int> &dims) {
    std::vector<std::vector<int32_t>> o;
    for (int d : dims) {
        o.emplace_back(nd, 0);
        o.back()[d] = 1;
    }
    return o;
}
std::vector<std::vector<int32_t>> hessian_orders(int nd, const std::vector<int> &dims) {
    std::vector<std::vector<int32_t>> o;
    for (int d : dims) {
        o.emplace_back(nd, 0);
        o.back()[d] = 2;
    }
    for (size_t j = 1; j < dims.size(); ++j)
        for (size_t i = 0; i < j; ++i) {
            o.emplace_back(nd, 0);
            o.back()[dims[i]] = 1;
            o.back()[dims[j]] = 1;
        }
    return o;
}

// Combine `terms` (contiguous scratch images) into the outputs with element-wise
// operation `op`; outputs that are not contiguous are written through a scratch copy.
int combine(int op, bool is_double, const std::vector<dipamd_image> &terms, int n, const dipamd_image *outs,
            int nouts, Scratch &scr, const dipamd_image *in) {
    std::vector<const void *> d;
    for (const auto &t : terms) d.push_back(t.origin);
    std::vector<void *> o(static_cast<size_t>(nouts));
    std::vector<dipamd_image> staged(static_cast<size_t>(nouts));
    int st;
    for (int k = 0; k < nouts; ++k) {
        if (contiguous(&outs[k])) {
            o[k] = outs[k].origin;
        } else {
            if ((st = scr.make(in, outs[k].dtype, &staged[k])) != DIPAMD_OK) return st;
            o[k] = staged[k].origin;
        }
    }
    int64_t npx = 1;
    for (int dd = 0; dd < in->ndim; ++dd) npx *= in->sizes[dd];
    const hipError_t e = launch_combine(is_double, op, d.data(), n, o.data(), npx, scr.stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    for (int k = 0; k < nouts; ++k) {
        if (o[k] == outs[k].origin) continue;
        st = cast_from_buffer(to_view(&staged[k]), to_view(&outs[k]), in->ndim, scr.stream);
        if (st != DIPAMD_OK) return st;
    }
    return DIPAMD_OK;
}

// The masked forms (derivative.cpp:522-600): the mask's Gaussian (the denominator) and
// the weighted input, in buffer-type scratch images.
struct MaskedTerms {
    dipamd_image den, w, m;
};

int64_t npx_of_image(const dipamd_image *im) {
    int64_t n = 1;
    for (int d = 0; d < im->ndim; ++d) n *= im->sizes[d];
    return n;
}

int masked_terms(const dipamd_image *in, const dipamd_image *mask, int buf_dt, const double *sig, const int32_t *bc,
                 int32_t nbc, double truncation, uint32_t flags, Scratch &scr, MaskedTerms *t) {
    int st;
    if ((st = check_image(mask, "mask")) != DIPAMD_OK) return st;
    if ((st = require_real(mask)) != DIPAMD_OK) return st;
    if (mask->ndim != in->ndim) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    for (int d = 0; d < in->ndim; ++d)
        if (mask->sizes[d] != in->sizes[d]) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    const int method = derivative_method(flags);
    const bool dbl = buf_dt == DIPAMD_DT_DFLOAT;
    dipamd_image xc;
    if ((st = scr.make(in, buf_dt, &t->den)) != DIPAMD_OK || (st = scr.make(in, buf_dt, &t->w)) != DIPAMD_OK ||
        (st = scr.make(in, buf_dt, &t->m)) != DIPAMD_OK || (st = scr.make(in, buf_dt, &xc)) != DIPAMD_OK)
        return st;
    // denominator = Gauss( mask ) (in the buffer type: SafeDivide's computation type)
    if ((st = gauss(method, mask, &t->den, sig, nullptr, 0, bc, nbc, truncation, flags, scr.stream)) != DIPAMD_OK)
        return st;
    // weighted = the input masked (binary mask) or times the mask, in the buffer type
    if ((st = cast_to_buffer(to_view(in), to_view(&xc), in->ndim, dbl, scr.stream)) != DIPAMD_OK) return st;
    if ((st = cast_to_buffer(to_view(mask), to_view(&t->m), in->ndim, dbl, scr.stream)) != DIPAMD_OK) return st;
    const hipError_t e = launch_masked(dbl, mask->dtype == DIPAMD_DT_BIN ? 0 : 1, xc.origin, t->m.origin, t->w.origin,
                                       nullptr, nullptr, nullptr, nullptr, npx_of_image(in), scr.stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

// NormalizedDifferentialConvolution along `dim` into the contiguous buffer-type `out`.
int ndc_into(const dipamd_image *mask, const MaskedTerms &t, int dim, int buf_dt, const double *sig, const int32_t *bc,
             int32_t nbc, double truncation, uint32_t flags, Scratch &scr, const dipamd_image *out) {
    int st;
    const int method = derivative_method(flags);
    const bool dbl = buf_dt == DIPAMD_DT_DFLOAT;
    const int nd = t.w.ndim;
    dipamd_image nc, dm, dw;
    if ((st = scr.make(&t.w, buf_dt, &nc)) != DIPAMD_OK || (st = scr.make(&t.w, buf_dt, &dm)) != DIPAMD_OK ||
        (st = scr.make(&t.w, buf_dt, &dw)) != DIPAMD_OK)
        return st;
    std::vector<int32_t> ord(static_cast<size_t>(nd), 0);
    ord[static_cast<size_t>(dim)] = 1;
    const int64_t n = npx_of_image(&t.w);
    // NC = SafeDivide( Gauss( weighted ), denominator )
    if ((st = gauss(method, &t.w, &nc, sig, nullptr, 0, bc, nbc, truncation, flags, scr.stream)) != DIPAMD_OK)
        return st;
    hipError_t e = launch_masked(dbl, 2, nullptr, nullptr, nc.origin, t.den.origin, nullptr, nullptr, nullptr, n,
                                 scr.stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    // the derivatives of the mask and of the weighted input along `dim`
    if ((st = gauss(method, mask, &dm, sig, ord.data(), nd, bc, nbc, truncation, flags, scr.stream)) != DIPAMD_OK)
        return st;
    if ((st = gauss(method, &t.w, &dw, sig, ord.data(), nd, bc, nbc, truncation, flags, scr.stream)) != DIPAMD_OK)
        return st;
    // out = SafeDivide( dw - NC * dm, denominator )
    e = launch_masked(dbl, 3, nullptr, nullptr, nc.origin, t.den.origin, dw.origin, dm.origin, out->origin, n,
                      scr.stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

// Writes the contiguous buffer-type `res` into `out` (a copy when out is not contiguous).
int deliver(const dipamd_image *res, const dipamd_image *out, hipStream_t stream) {
    if (res->origin == out->origin) return DIPAMD_OK;
    return cast_from_buffer(to_view(res), to_view(out), out->ndim, stream);
}

}  // namespace

extern "C" {

int dipamd_normalized_convolution(const dipamd_image *in, const dipamd_image *mask, const dipamd_image *out,
                                  const double *sigmas, int32_t nsigmas, const int32_t *bc, int32_t nbc,
                                  double truncation, uint32_t flags, void *stream) {
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if ((st = require_real(in)) != DIPAMD_OK) return st;
    if ((st = check_outputs(DIPAMD_FN_LAPLACE, in, out, 1, 1)) != DIPAMD_OK) return st;
    const int buf_dt = suggest_flex(in->dtype);
    if (out->dtype != buf_dt)
        return set_error(DIPAMD_E_UNSUPPORTED, "the output must have the floating-point type of the input");
    double sig[DIPAMD_MAX_DIMS];
    if ((st = array_param(sigmas, nsigmas, in->ndim, 1.0, sig)) != DIPAMD_OK) return st;
    Scratch scr{static_cast<hipStream_t>(stream), {}};
    MaskedTerms t;
    if ((st = masked_terms(in, mask, buf_dt, sig, bc, nbc, truncation, flags, scr, &t)) != DIPAMD_OK) return st;
    dipamd_image g;
    if ((st = scr.make(in, buf_dt, &g)) != DIPAMD_OK) return st;
    // out = SafeDivide( Gauss( weighted ), Gauss( mask ) ) (derivative.cpp:522-542)
    if ((st = gauss(derivative_method(flags), &t.w, &g, sig, nullptr, 0, bc, nbc, truncation, flags, stream)) != DIPAMD_OK)
        return st;
    const hipError_t e = launch_masked(buf_dt == DIPAMD_DT_DFLOAT, 2, nullptr, nullptr, g.origin, t.den.origin, nullptr,
                                       nullptr, nullptr, npx_of_image(in), scr.stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return deliver(&g, out, scr.stream);
}

int dipamd_normalized_differential_convolution(const dipamd_image *in, const dipamd_image *mask,
                                               const dipamd_image *out, int32_t dimension, const double *sigmas,
                                               int32_t nsigmas, const int32_t *bc, int32_t nbc, double truncation,
                                               uint32_t flags, void *stream) {
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if ((st = require_real(in)) != DIPAMD_OK) return st;
    if ((st = check_outputs(DIPAMD_FN_LAPLACE, in, out, 1, 1)) != DIPAMD_OK) return st;
    if (dimension < 0 || dimension >= in->ndim) return set_error(DIPAMD_E_PARAMETER, "Parameter has invalid value");
    const int buf_dt = suggest_flex(in->dtype);
    if (out->dtype != buf_dt)
        return set_error(DIPAMD_E_UNSUPPORTED, "the output must have the floating-point type of the input");
    double sig[DIPAMD_MAX_DIMS];
    if ((st = array_param(sigmas, nsigmas, in->ndim, 1.0, sig)) != DIPAMD_OK) return st;
    Scratch scr{static_cast<hipStream_t>(stream), {}};
    MaskedTerms t;
    if ((st = masked_terms(in, mask, buf_dt, sig, bc, nbc, truncation, flags, scr, &t)) != DIPAMD_OK) return st;
    dipamd_image r;
    if ((st = scr.make(in, buf_dt, &r)) != DIPAMD_OK) return st;
    if ((st = ndc_into(mask, t, dimension, buf_dt, sig, bc, nbc, truncation, flags, scr, &r)) != DIPAMD_OK) return st;
    return deliver(&r, out, scr.stream);
}

int dipamd_structure_tensor_masked(const dipamd_image *in, const dipamd_image *mask, const dipamd_image *outs,
                                   int32_t nouts, const double *gradient_sigmas, int32_t ngs,
                                   const double *tensor_sigmas, int32_t nts, const int32_t *bc, int32_t nbc,
                                   double truncation, uint32_t flags, void *stream) {
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if ((st = require_real(in)) != DIPAMD_OK) return st;
    const int nd = in->ndim;
    // every dim gets its normalized derivative (structure_tensor.cpp:46-52)
    const int want = nd * (nd + 1) / 2;
    if ((st = check_outputs(DIPAMD_FN_STRUCTURE_TENSOR, in, outs, nouts, want)) != DIPAMD_OK) return st;
    const int buf_dt = suggest_flex(in->dtype);
    for (int k = 0; k < nouts; ++k)
        if (outs[k].dtype != buf_dt)
            return set_error(DIPAMD_E_UNSUPPORTED, "the output must have the floating-point type of the input");
    double gsig[DIPAMD_MAX_DIMS], tsig[DIPAMD_MAX_DIMS];
    if ((st = array_param(gradient_sigmas, ngs, nd, 1.0, gsig)) != DIPAMD_OK) return st;
    if ((st = array_param(tensor_sigmas, nts, nd, 5.0, tsig)) != DIPAMD_OK) return st;
    Scratch scr{static_cast<hipStream_t>(stream), {}};
    MaskedTerms t;
    if ((st = masked_terms(in, mask, buf_dt, gsig, bc, nbc, truncation, flags, scr, &t)) != DIPAMD_OK) return st;
    std::vector<dipamd_image> g(static_cast<size_t>(nd));
    for (int d = 0; d < nd; ++d) {
        if ((st = scr.make(in, buf_dt, &g[static_cast<size_t>(d)])) != DIPAMD_OK) return st;
        if ((st = ndc_into(mask, t, d, buf_dt, gsig, bc, nbc, truncation, flags, scr, &g[static_cast<size_t>(d)])) !=
            DIPAMD_OK)
            return st;
    }
    // g g^T, then each element smoothed with the tensor sigmas (structure_tensor.cpp:55-56)
    if ((st = combine(5, buf_dt == DIPAMD_DT_DFLOAT, g, nd, outs, nouts, scr, in)) != DIPAMD_OK) return st;
    for (int k = 0; k < nouts; ++k) {
        st = gauss(tensor_method(flags), &outs[k], &outs[k], tsig, nullptr, 0, bc, nbc, truncation, flags, stream);
        if (st != DIPAMD_OK) return st;
    }
    return DIPAMD_OK;
}

int dipamd_derivative_outputs(int32_t function, int32_t ndiff) {
    if (ndiff < 1 || ndiff > DIPAMD_MAX_DIMS) return set_error(DIPAMD_E_PARAMETER, "Dimensionality not supported");
    switch (function) {
        case DIPAMD_FN_GRADIENT: return ndiff;
        case DIPAMD_FN_HESSIAN:
        case DIPAMD_FN_STRUCTURE_TENSOR: return ndiff * (ndiff + 1) / 2;
        case DIPAMD_FN_GRADIENT_MAGNITUDE:
        case DIPAMD_FN_LAPLACE:
        case DIPAMD_FN_DGG:
        case DIPAMD_FN_LAPLACE_PLUS_DGG:
        case DIPAMD_FN_LAPLACE_MINUS_DGG: return 1;
    }
    return set_error(DIPAMD_E_PARAMETER, "unknown derivative function");
}

int dipamd_derivative(int32_t function, const dipamd_image *in, const dipamd_image *outs, int32_t nouts,
                      const double *sigmas, int32_t nsigmas, const int32_t *bc, int32_t nbc, const uint8_t *process,
                      int32_t nprocess, double truncation, uint32_t flags, void *stream) {
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if ((st = require_real(in)) != DIPAMD_OK) return st;
    if (function == DIPAMD_FN_STRUCTURE_TENSOR)
        return set_error(DIPAMD_E_PARAMETER, "the structure tensor takes dipamd_structure_tensor");
    const int nd = in->ndim;
    // ArrayUseParameter: sigmas default 1, process default all (derivative.cpp:127-129)
    double sig[DIPAMD_MAX_DIMS];
    uint8_t proc[DIPAMD_MAX_DIMS];
    if ((st = array_param(sigmas, nsigmas, nd, 1.0, sig)) != DIPAMD_OK) return st;
    if (nprocess != 0 && nprocess != 1 && nprocess != nd)
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    for (int d = 0; d < nd; ++d) proc[d] = nprocess == 0 ? 1 : (nprocess == 1 ? process[0] : process[d]);
    const std::vector<int> dims = gradient_dims(in, sig, proc);
    const int n = static_cast<int>(dims.size());
    if (n < 1) return set_error(DIPAMD_E_PARAMETER, "Dimensionality not supported");
    const int want = dipamd_derivative_outputs(function, n);
    if (want < 0) return want;
    if ((st = check_outputs(function, in, outs, nouts, want)) != DIPAMD_OK) return st;
    if (n == 2 && (st = try_stream_derivative(function, in, outs, nouts, sig, bc, nbc, truncation, flags, stream)) !=
                      -100)
        return st;
    if (function == DIPAMD_FN_GRADIENT)  // derivative.cpp:160-169
        return derivatives(in, gradient_orders(nd, dims), outs, sig, bc, nbc, truncation, flags, stream);
    if (function == DIPAMD_FN_HESSIAN)  // derivative.cpp:348-369
        return derivatives(in, hessian_orders(nd, dims), outs, sig, bc, nbc, truncation, flags, stream);
    // GradientMagnitude (derivative.cpp:174-210), Laplace (:374-418) and the Dgg family
    // (:420-521): the terms in buffer-type scratch images, combined in one pass
    const int buf_dt = suggest_flex(in->dtype);
    if (outs[0].dtype != buf_dt)
        return set_error(DIPAMD_E_UNSUPPORTED, "the output must have the floating-point type of the input");
    std::vector<std::vector<int32_t>> orders;
    int op = 0;
    if (function == DIPAMD_FN_LAPLACE) {
        for (int d : dims) {
            orders.emplace_back(nd, 0);
            orders.back()[d] = 2;
        }
        op = 0;
    } else if (function == DIPAMD_FN_GRADIENT_MAGNITUDE) {
        orders = gradient_orders(nd, dims);
        op = 1;
    } else {
        orders = gradient_orders(nd, dims);
        for (auto &o : hessian_orders(nd, dims)) orders.push_back(o);
        op = function == DIPAMD_FN_DGG ? 2 : (function == DIPAMD_FN_LAPLACE_PLUS_DGG ? 3 : 4);
    }
    Scratch scr{static_cast<hipStream_t>(stream), {}};
    std::vector<dipamd_image> terms(orders.size());
    for (auto &t : terms)
        if ((st = scr.make(in, buf_dt, &t)) != DIPAMD_OK) return st;
    if ((st = derivatives(in, orders, terms.data(), sig, bc, nbc, truncation, flags, stream)) != DIPAMD_OK) return st;
    return combine(op, buf_dt == DIPAMD_DT_DFLOAT, terms, n, outs, 1, scr, in);
}

int dipamd_structure_tensor(const dipamd_image *in, const dipamd_image *outs, int32_t nouts,
                            const double *gradient_sigmas, int32_t ngs, const double *tensor_sigmas, int32_t nts,
                            const int32_t *bc, int32_t nbc, double truncation, uint32_t flags, void *stream) {
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if ((st = require_real(in)) != DIPAMD_OK) return st;
    const int nd = in->ndim;
    double gsig[DIPAMD_MAX_DIMS], tsig[DIPAMD_MAX_DIMS];
    if ((st = array_param(gradient_sigmas, ngs, nd, 1.0, gsig)) != DIPAMD_OK) return st;
    if ((st = array_param(tensor_sigmas, nts, nd, 5.0, tsig)) != DIPAMD_OK) return st;
    uint8_t proc[DIPAMD_MAX_DIMS];
    for (int d = 0; d < nd; ++d) proc[d] = 1;
    const std::vector<int> dims = gradient_dims(in, gsig, proc);
    const int n = static_cast<int>(dims.size());
    if (n < 1) return set_error(DIPAMD_E_PARAMETER, "Dimensionality not supported");
    const int want = dipamd_derivative_outputs(DIPAMD_FN_STRUCTURE_TENSOR, n);
    if ((st = check_outputs(DIPAMD_FN_STRUCTURE_TENSOR, in, outs, nouts, want)) != DIPAMD_OK) return st;
    const int buf_dt = suggest_flex(in->dtype);
    for (int k = 0; k < nouts; ++k)
        if (outs[k].dtype != buf_dt)
            return set_error(DIPAMD_E_UNSUPPORTED, "the output must have the floating-point type of the input");
    // gradient, products g g^T, then each element smoothed with the tensor sigmas
    // (structure_tensor.cpp:53-56)
    Scratch scr{static_cast<hipStream_t>(stream), {}};
    std::vector<dipamd_image> g(static_cast<size_t>(n));
    for (auto &t : g)
        if ((st = scr.make(in, buf_dt, &t)) != DIPAMD_OK) return st;
    if ((st = derivatives(in, gradient_orders(nd, dims), g.data(), gsig, bc, nbc, truncation, flags, stream)) !=
        DIPAMD_OK)
        return st;
    if ((st = combine(5, buf_dt == DIPAMD_DT_DFLOAT, g, n, outs, nouts, scr, in)) != DIPAMD_OK) return st;
    for (int k = 0; k < nouts; ++k) {
        st = gauss(tensor_method(flags), &outs[k], &outs[k], tsig, nullptr, 0, bc, nbc, truncation, flags, stream);
        if (st != DIPAMD_OK) return st;
    }
    return DIPAMD_OK;
}

}  // extern "C"
