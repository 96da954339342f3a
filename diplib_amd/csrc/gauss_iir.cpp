// dip::GaussIIR on the device (src/linear/gaussiir.cpp:741-806): per-axis filter
// design on the host, then one k_iir pass per processed axis in the order of
// Framework::Separable (ascending input stride, framework_separable.cpp:161), with a
// double buffer type: when the output is not double the passes before the last go
// through a double intermediate image (framework_separable.cpp:190-206), and the last
// pass casts its result to the output type.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "dipamd_internal.hpp"
#include "iir_common.hpp"

namespace dipamd {

namespace {

template <typename T>
T const_value_d(int bc);
template <>
double const_value_d<double>(int bc) {
    switch (bc) {
        case DIPAMD_BC_ADD_MAX_VALUE: return std::numeric_limits<double>::max();
        case DIPAMD_BC_ADD_MIN_VALUE: return std::numeric_limits<double>::lowest();
        default: return 0.0;
    }
}

// Kernel arguments of one axis (the initial-value forms of gaussiir.cpp:485-738: the
// unrolled loops for 3..5 poles start from different expressions than the general
// loop; which one runs depends on the pole count and the derivative order).
void fill_filter_args(IirArgs &a, const IirFilter &f, int order) {
    for (int k = 0; k < 6; ++k) a.b[k] = k <= f.nn ? f.b[k] : 0.0;
    a.b[0] = 0.0;
    for (int k = 0; k < 3; ++k) {
        a.a1[k] = f.a1[k];
        a.a2[k] = f.a2[k];
    }
    a.S = 1.0 + a.b[1] + a.b[2] + a.b[3] + a.b[4] + a.b[5];
    a.c = f.cc;
    a.m1 = f.m1;
    a.m2 = f.m2;
    a.m4 = f.m4;
    a.m5 = f.m5;
    const int nn = f.nn;
    const int order1 = std::max(nn, f.m2);
    const int order2 = std::max(nn, f.m5);
    const bool copy_f = f.m2 == 0 && f.a1[0] == 1.0;
    const bool copy_b = f.m5 == 0 && f.a2[0] == 1.0;
    // causal part
    if (copy_f && order1 >= 3 && order1 <= 5) {
        a.finit_diff = 2;  // x[0] / S, recursion from position 0
        a.fnset = 0;
    } else if (order1 == 4 && f.a1[0] == 0.5 && f.a1[1] == 0.0 && f.a1[2] == -0.5 && f.a1[3] == 0.0) {
        a.finit_diff = 1;  // first derivative, 4 poles
        a.fnset = 2;
    } else if (order1 == 5 && f.a1[0] == 1.0 && f.a1[1] == -1.0 && f.a1[2] == 0.0 && f.a1[3] == 0.0) {
        a.finit_diff = 1;  // second derivative, 5 poles
        a.fnset = 1;
    } else {
        a.finit_diff = 0;
        a.fnset = order1;
    }
    // anti-causal part
    if (copy_b && order2 >= 3 && order2 <= 5) {
        a.binit = 1;
        a.bnset = 0;
    } else if (order2 == 4 && f.a2[0] == 0.0 && f.a2[1] == 1.0 && f.a2[2] == 0.0 && f.a2[3] == 0.0) {
        a.binit = 1;
        a.bnset = 1;
    } else if (order2 == 5 && f.a2[0] == -1.0 && f.a2[1] == 1.0 && f.a2[2] == 0.0 && f.a2[3] == 0.0) {
        a.binit = 2;
        a.bnset = 1;
    } else {
        a.binit = 0;
        a.bnset = order2;
    }
    (void)order;
}

int64_t npx(const View &v, int ndim) {
    int64_t p = 1;
    for (int d = 0; d < ndim; ++d) p *= v.sizes[d];
    return p;
}

int run_iir_pass(const View &src, const View &dst, int ndim, int d, const IirFilter &f, int order, int bc,
                 double *p1, hipStream_t stream) {
    IirArgs a;
    std::memset(&a, 0, sizeof(a));
    a.in = src.p;
    a.out = dst.p;
    a.p1 = p1;
    a.n = src.sizes[d];
    a.border = f.border;
    a.L = a.n + 2 * static_cast<int64_t>(f.border);
    a.sin = src.strides[d];
    a.sout = dst.strides[d];
    a.bc = bc;
    a.cval = const_value_d<double>(bc);
    a.out_double = dst.dt == DIPAMD_DT_DFLOAT;
    fill_filter_args(a, f, order);
    int lane = -1;
    for (int k = 0; k < ndim; ++k) {
        if (k == d || src.sizes[k] <= 1) continue;
        if (lane < 0 || std::llabs(src.strides[k]) < std::llabs(src.strides[lane])) lane = k;
    }
    if (lane >= 0) {
        a.e = src.sizes[lane];
        a.ein = src.strides[lane];
        a.eout = dst.strides[lane];
    } else {
        a.e = 1;
    }
    fill_outer(a.outer, src, dst, ndim, d, lane);
    a.nlines = npx(src, ndim) / a.n;
    if (a.L > (int64_t(1) << 31) - 1 || a.nlines >= (int64_t(1) << 31))
        return set_error(DIPAMD_E_UNSUPPORTED, "image too large for the IIR kernel");
    const int64_t px = npx(src, ndim);
    TraceScope ts("iir", px * static_cast<int64_t>(dt_size(src.dt) + dt_size(dst.dt)), stream,
                  a.nlines * a.L * static_cast<int64_t>(2 * (2 * f.nn + 3)));
    // vector loads and stores along a contiguous filter axis when every line starts
    // aligned: V input samples of dt_size bytes, W output samples per access
    const int64_t vin = src.dt == DIPAMD_DT_DFLOAT ? 2 : 4;
    const int64_t wout = dst.dt == DIPAMD_DT_DFLOAT ? 2 : 4;
    bool vec = src.strides[d] == 1 && dst.strides[d] == 1 &&
               reinterpret_cast<uintptr_t>(src.p) % static_cast<uintptr_t>(vin * dt_size(src.dt)) == 0 &&
               reinterpret_cast<uintptr_t>(dst.p) % static_cast<uintptr_t>(wout * dt_size(dst.dt)) == 0;
    for (int k = 0; k < ndim && vec; ++k) {
        if (k == d || src.sizes[k] <= 1) continue;
        if (src.strides[k] % vin != 0 || dst.strides[k] % wout != 0) vec = false;
    }
    hipError_t e;
    switch (src.dt) {
        case DIPAMD_DT_UINT8:
            e = vec ? launch_iir_u8_vec(f.nn, order, a, stream) : launch_iir_u8(f.nn, order, a, stream);
            break;
        case DIPAMD_DT_SFLOAT:
            e = vec ? launch_iir_f32_vec(f.nn, order, a, stream) : launch_iir_f32(f.nn, order, a, stream);
            break;
        case DIPAMD_DT_DFLOAT:
            e = vec ? launch_iir_f64_vec(f.nn, order, a, stream) : launch_iir_f64(f.nn, order, a, stream);
            break;
        default: return set_error(DIPAMD_E_PARAMETER, "Data type not supported");
    }
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

}  // namespace

int gauss_iir_device(const View &in, const View &out, int ndim, const IirFilter *filters, const int *orders,
                     const bool *process, const int *bc, hipStream_t stream) {
    std::vector<int> order;
    for (int d = 0; d < ndim; ++d)
        if (process[d]) order.push_back(d);
    for (size_t i = 1; i < order.size(); ++i) {
        const int e = order[i];
        size_t j = i;
        while (j > 0 && in.strides[order[j - 1]] > in.strides[e]) {
            order[j] = order[j - 1];
            --j;
        }
        order[j] = e;
    }
    for (int d : order) {
        if (bc[d] == DIPAMD_BC_ALREADY_EXPANDED)
            return set_error(DIPAMD_E_UNSUPPORTED, "Functionality has not (yet) been implemented");
        if (!bc_supported(bc[d]) && !bc_value_based(bc[d]))
            return set_error(DIPAMD_E_UNSUPPORTED, "boundary condition not supported");
    }
    if (order.empty()) {
        // output.Copy( input )
        const InternFilter *none[DIPAMD_MAX_DIMS] = {};
        InternFilter id;
        id.data = {1.0};
        id.size = 1;
        for (int d = 0; d < ndim; ++d) none[d] = &id;
        bool proc[DIPAMD_MAX_DIMS] = {};
        return separable_device(in, out, ndim, none, bc, proc, 0, stream);
    }
    std::vector<void *> scratch;
    auto alloc = [&](size_t bytes, void **p) -> int {
        hipError_t e = scratch_alloc(p, bytes, stream);
        if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
        scratch.push_back(*p);
        return DIPAMD_OK;
    };
    auto finish = [&](int st) {
        for (void *p : scratch) scratch_free(p, stream);
        return st;
    };
    const int64_t px = npx(in, ndim);
    int st;
    // causal scratch (positions border .. n + 2 border - 1 of every line), sized for the
    // longest pass
    size_t p1_bytes = 0;
    for (int d : order) {
        const int64_t kept = (in.sizes[d] + static_cast<int64_t>(filters[d].border) + 1) / 2 * 2;  // in pairs
        p1_bytes = std::max(p1_bytes, static_cast<size_t>(px / in.sizes[d] * kept) * sizeof(double));
    }
    void *p1 = nullptr;
    if ((st = alloc(p1_bytes, &p1)) != DIPAMD_OK) return finish(st);

    // source of the first pass: inputs without a kernel go through a double copy
    View src = in;
    if (in.dt != DIPAMD_DT_UINT8 && in.dt != DIPAMD_DT_SFLOAT && in.dt != DIPAMD_DT_DFLOAT) {
        void *t;
        if ((st = alloc(static_cast<size_t>(px) * sizeof(double), &t)) != DIPAMD_OK) return finish(st);
        View tv = contiguous_view(t, DIPAMD_DT_DFLOAT, ndim, in.sizes);
        if ((st = cast_to_buffer(in, tv, ndim, true, stream)) != DIPAMD_OK) return finish(st);
        src = tv;
    }
    // each thread reads its whole line before writing it, so a pass may write the
    // view it reads; a destination that overlaps the source any other way is not used
    const bool direct_out = (out.dt == DIPAMD_DT_DFLOAT || out.dt == DIPAMD_DT_SFLOAT) &&
                            !(views_overlap(src, out, ndim) &&
                              !(src.p == out.p && std::memcmp(src.strides, out.strides, sizeof(src.strides)) == 0));
    View inter;
    bool have_inter = false;
    const bool need_inter = order.size() > 1 || !direct_out;
    if (need_inter) {
        if (out.dt == DIPAMD_DT_DFLOAT && direct_out) {
            inter = out;  // in -> out -> out -> ... (the output holds the buffer type)
        } else {
            void *t;
            if ((st = alloc(static_cast<size_t>(px) * sizeof(double), &t)) != DIPAMD_OK) return finish(st);
            inter = contiguous_view(t, DIPAMD_DT_DFLOAT, ndim, in.sizes);
            have_inter = true;
        }
    }
    for (size_t k = 0; k < order.size(); ++k) {
        const int d = order[k];
        const bool last = k + 1 == order.size();
        const View dst = (last && direct_out) ? out : inter;
        if ((st = run_iir_pass(src, dst, ndim, d, filters[d], orders[d], bc[d], static_cast<double *>(p1),
                               stream)) != DIPAMD_OK)
            return finish(st);
        src = dst;
    }
    if (!direct_out) st = cast_from_buffer(inter, out, ndim, stream);
    (void)have_inter;
    return finish(st);
}

}  // namespace dipamd

using namespace dipamd;

extern "C" int dipamd_gauss_iir(const dipamd_image *in, const dipamd_image *out, const double *sigmas,
                                int32_t nsigmas, const int32_t *orders, int32_t norders, const int32_t *bc,
                                int32_t nbc, const int32_t *filter_orders, int32_t nfilter_orders,
                                int32_t design_method, double truncation, uint32_t flags, void *stream) {
    (void)flags;
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if ((st = require_real(in)) != DIPAMD_OK) return st;
    if ((st = check_image(out, "out")) != DIPAMD_OK) return st;
    const int nd = in->ndim;
    if (out->ndim != nd) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    for (int d = 0; d < nd; ++d)
        if (out->sizes[d] != in->sizes[d]) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    auto bad_len = [](int n, int nd) { return n != 0 && n != 1 && n != nd; };
    if (bad_len(nsigmas, nd) || bad_len(norders, nd) || bad_len(nfilter_orders, nd))
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    if (design_method != DIPAMD_IIR_FORWARD_BACKWARD && design_method != DIPAMD_IIR_DISCRETE_TIME_FIT)
        return set_error(DIPAMD_E_PARAMETER, "Invalid flag");
    int bcs[DIPAMD_MAX_DIMS];
    if ((st = broadcast_bc(bc, nbc, nd, bcs)) != DIPAMD_OK) return st;
    double s[DIPAMD_MAX_DIMS];
    int o[DIPAMD_MAX_DIMS], fo[DIPAMD_MAX_DIMS];
    bool proc[DIPAMD_MAX_DIMS];
    IirFilter filt[DIPAMD_MAX_DIMS];
    for (int d = 0; d < nd; ++d) {
        s[d] = nsigmas == 0 ? 1.0 : (nsigmas == 1 ? sigmas[0] : sigmas[d]);
        o[d] = norders == 0 ? 0 : (norders == 1 ? orders[0] : orders[d]);
        if (o[d] < 0) return set_error(DIPAMD_E_PARAMETER, "derivative order must be non-negative");
        if (nfilter_orders == 0) {
            fo[d] = o[d] > 2 ? 5 : o[d] + 3;
        } else {
            fo[d] = nfilter_orders == 1 ? filter_orders[0] : filter_orders[d];
            if (fo[d] < 0) return set_error(DIPAMD_E_PARAMETER, "filter order must be non-negative");
        }
        proc[d] = s[d] > 0.0 && in->sizes[d] > 1;
        if (proc[d] && (st = design_gauss_iir(s[d], o[d], fo[d], design_method, truncation, &filt[d])) != DIPAMD_OK)
            return st;
    }
    return gauss_iir_device(to_view(in), to_view(out), nd, filt, o, proc, bcs, static_cast<hipStream_t>(stream));
}
