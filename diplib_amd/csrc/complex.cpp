// dip::SeparableConvolution over a complex buffer type: with a complex filter in the
// array (buffer SuggestComplex(in)) or a complex input (buffer SuggestFlex(in), its own
// complex type) -- convolution.cpp:319-421, with the framework of
// framework_separable.cpp:35-431 (passes in order of increasing input stride).
//
// The input is first converted into a contiguous complex scratch (the reference's
// CopyBuffer casts: a real sample becomes (v, 0)); each pass then filters one axis
// (k_complex.hip), the passes ping-pong between scratch images, and the last one writes
// the output when it holds the buffer's complex type (else a copy converts the
// precision).  Conditions computed from the edge values expand the real and imaginary
// parts of the line separately (copy_buffer.cpp:802-808) into an extended scratch first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "complex_kernels.hpp"
#include "dipamd.h"
#include "dipamd_internal.hpp"
#include "sep_common.hpp"
#include "sep_kernels.hpp"

namespace dipamd {

namespace {

// InternOneDimensionalFilter (convolution.cpp:67-162) for a complex buffer: taps
// reversed into correlation order, complex when any filter of the array is.
struct CxFilter {
    std::vector<double> re, im;
    int size = 0, origin = 0, border = 0, sym = DIPAMD_SYM_GENERAL, zpad = 0;
};

int intern_cx(const dipamd_filter &f, bool use_complex, CxFilter &F) {
    const int n = f.ntaps;
    int sym = f.symmetry;
    if (sym < 0 || sym > DIPAMD_SYM_D_CONJ) return set_error(DIPAMD_E_PARAMETER, "Symmetry string not recognized");
    if (sym == DIPAMD_SYM_CONJ && !use_complex) sym = DIPAMD_SYM_EVEN;
    if (sym == DIPAMD_SYM_D_CONJ && !use_complex) sym = DIPAMD_SYM_D_EVEN;
    F.sym = sym;
    F.re.resize(static_cast<size_t>(n));
    F.im.assign(static_cast<size_t>(n), 0.0);
    for (int k = 0; k < n; ++k) {  // reversed
        const int s = n - 1 - k;
        F.re[static_cast<size_t>(k)] = f.is_complex ? f.taps[2 * s] : f.taps[s];
        if (f.is_complex) F.im[static_cast<size_t>(k)] = f.taps[2 * s + 1];
    }
    F.size = n;
    F.origin = F.border = F.zpad = 0;
    if (n == 0) return DIPAMD_OK;
    if (sym == DIPAMD_SYM_EVEN || sym == DIPAMD_SYM_ODD || sym == DIPAMD_SYM_CONJ) F.size = 2 * n - 1;
    if (sym == DIPAMD_SYM_D_EVEN || sym == DIPAMD_SYM_D_ODD || sym == DIPAMD_SYM_D_CONJ) F.size = 2 * n;
    const int origin = f.origin < 0 ? F.size / 2 : f.origin;
    if (origin >= F.size) return set_error(DIPAMD_E_PARAMETER, "Origin outside of filter");
    F.origin = F.size - origin - 1;
    F.border = std::max(F.origin, F.size - F.origin - 1);
    if ((sym == DIPAMD_SYM_D_EVEN || sym == DIPAMD_SYM_D_ODD || sym == DIPAMD_SYM_D_CONJ) && F.origin >= n) {
        // the d-* filter reads one sample past the border (see intern_filter): read as 0
        F.border = F.origin + 1;
        F.zpad = 1;
    }
    return DIPAMD_OK;
}

// IsMeaninglessFilter (convolution.cpp:304-314) in the buffer's precision.
bool meaningless_cx(const CxFilter &F, bool use_double) {
    if (F.size == 0) return true;
    if (F.size != 1) return false;
    if (use_double) return F.re[0] == 1.0 && F.im[0] == 0.0;
    return static_cast<float>(F.re[0]) == 1.0f && static_cast<float>(F.im[0]) == 0.0f;
}

bool is_complex_dt(int dt) { return dt == DIPAMD_DT_SCOMPLEX || dt == DIPAMD_DT_DCOMPLEX; }

// SuggestComplex (datatype.cpp:95-106)
int suggest_complex(int dt) {
    return (dt == DIPAMD_DT_UINT32 || dt == DIPAMD_DT_SINT32 || dt == DIPAMD_DT_DFLOAT || dt == DIPAMD_DT_DCOMPLEX)
               ? DIPAMD_DT_DCOMPLEX
               : DIPAMD_DT_SCOMPLEX;
}

CastGeo geo_of(const View &src, const View &dst, int ndim, int64_t in_scale, int64_t out_scale) {
    CastGeo g;
    std::memset(&g, 0, sizeof(g));
    g.ndim = ndim;
    g.total = 1;
    for (int d = 0; d < ndim; ++d) {
        g.size[d] = src.sizes[d];
        g.sin[d] = src.strides[d] * in_scale;
        g.sout[d] = dst.strides[d] * out_scale;
        g.total *= src.sizes[d];
    }
    return g;
}

template <typename TS>
int run_cpass(const View &src, const View &dst, int ndim, int d, const CxFilter &F, const TS *taps, bool fc, int bc,
              hipStream_t stream) {
    CxPassArgs<TS> a;
    std::memset(&a, 0, sizeof(a));
    a.in = static_cast<const TS *>(src.p);
    a.out = static_cast<TS *>(dst.p);
    a.n = src.sizes[d];
    a.sin = src.strides[d];
    a.sout = dst.strides[d];
    a.nlines = 1;
    for (int k = 0; k < ndim; ++k)
        if (k != d) a.nlines *= src.sizes[k];
    fill_outer(a.outer, src, dst, ndim, d, -1);
    a.ntaps = static_cast<int32_t>(F.re.size());
    a.origin = F.origin;
    a.bc = bc;
    a.zlo = F.zpad ? -(F.border - F.zpad) : INT32_MIN;
    a.cval = bc == DIPAMD_BC_ADD_MAX_VALUE ? std::numeric_limits<TS>::max()
                                           : (bc == DIPAMD_BC_ADD_MIN_VALUE ? std::numeric_limits<TS>::lowest() : TS(0));
    a.taps = taps;
    int64_t npx = 1;
    for (int k = 0; k < ndim; ++k) npx *= src.sizes[k];
    TraceScope ts("complex_pass", npx * static_cast<int64_t>(4 * sizeof(TS)), stream);
    hipError_t e = launch_cpass(sizeof(TS) == 8, fc, F.sym, &a, stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("complex pass: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

}  // namespace

int separable_complex(const View &in, const View &out, int ndim, const dipamd_filter *filters, int nfilters,
                      const int *bc, const uint8_t *process, bool any_complex_filter, hipStream_t stream) {
    const int buf_dt = any_complex_filter ? suggest_complex(in.dt)
                                          : (in.dt == DIPAMD_DT_DCOMPLEX ? DIPAMD_DT_DCOMPLEX : DIPAMD_DT_SCOMPLEX);
    const bool dbl = buf_dt == DIPAMD_DT_DCOMPLEX;
    if (!is_complex_dt(out.dt))
        return set_error(DIPAMD_E_UNSUPPORTED, "a complex result needs a complex output image");
    std::vector<CxFilter> F(static_cast<size_t>(nfilters));
    int st;
    for (int i = 0; i < nfilters; ++i) {
        if (filters[i].ntaps < 0 || (filters[i].ntaps > 0 && !filters[i].taps))
            return set_error(DIPAMD_E_PARAMETER, "bad filter");
        if ((st = intern_cx(filters[i], any_complex_filter, F[static_cast<size_t>(i)])) != DIPAMD_OK) return st;
        if (F[static_cast<size_t>(i)].border > DIPAMD_MAX_RADIUS + 1)
            return set_error(DIPAMD_E_UNSUPPORTED, "filter longer than the kernels' tap capacity");
    }
    bool proc[DIPAMD_MAX_DIMS];
    for (int d = 0; d < ndim; ++d) {
        const CxFilter &f = F[nfilters == 1 ? 0 : static_cast<size_t>(d)];
        proc[d] = process ? process[d] != 0 : true;
        if (nfilters == 1) {
            if (meaningless_cx(F[0], dbl)) proc[d] = false;
        } else if (in.sizes[d] <= 1 || meaningless_cx(f, dbl)) {
            proc[d] = false;
        }
        if (in.sizes[d] == 1) proc[d] = false;
    }
    std::vector<int> order;
    for (int d = 0; d < ndim; ++d)
        if (proc[d]) order.push_back(d);
    for (size_t i = 1; i < order.size(); ++i) {  // stable sort by input stride
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
    const bool out_dbl = out.dt == DIPAMD_DT_DCOMPLEX;
    if (order.empty()) {  // output.Copy( input ) with conversion
        CastGeo g = geo_of(in, out, ndim, 1, 1);
        hipError_t e = launch_to_complex(in.dt, out_dbl, in.p, out.p, &g, stream);
        if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("complex copy: ") + hipGetErrorString(e));
        return DIPAMD_OK;
    }
    const size_t es = dbl ? 16 : 8;
    int64_t npx = 1;
    for (int d = 0; d < ndim; ++d) npx *= in.sizes[d];
    std::vector<void *> scratch;
    auto release = [&](int status) {
        for (void *p : scratch) scratch_free(p, stream);
        return status;
    };
    auto alloc = [&](size_t bytes, void **p) -> int {
        hipError_t e = scratch_alloc(p, bytes, stream);
        if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
        scratch.push_back(*p);
        return DIPAMD_OK;
    };
    // taps of every filter, in the buffer's precision (complex pairs when any filter is)
    const size_t per = any_complex_filter ? 2 : 1;
    std::vector<size_t> toff(F.size());
    size_t ntot = 0;
    for (size_t i = 0; i < F.size(); ++i) {
        toff[i] = ntot;
        ntot += per * std::max<size_t>(1, F[i].re.size());
    }
    std::vector<double> hd(ntot, 0.0);
    std::vector<float> hf(ntot, 0.f);
    for (size_t i = 0; i < F.size(); ++i)
        for (size_t k = 0; k < F[i].re.size(); ++k) {
            const size_t o = toff[i] + per * k;
            hd[o] = F[i].re[k];
            hf[o] = static_cast<float>(F[i].re[k]);
            if (per == 2) {
                hd[o + 1] = F[i].im[k];
                hf[o + 1] = static_cast<float>(F[i].im[k]);
            }
        }
    void *dtaps = nullptr;
    if ((st = alloc(ntot * (dbl ? 8 : 4), &dtaps)) != DIPAMD_OK) return release(st);
    hipError_t e = hipMemcpyAsync(dtaps, dbl ? static_cast<const void *>(hd.data()) : static_cast<const void *>(hf.data()),
                                  ntot * (dbl ? 8 : 4), hipMemcpyHostToDevice, stream);  // pageable: synchronous
    if (e != hipSuccess) return release(set_error(DIPAMD_E_HIP, std::string("taps upload: ") + hipGetErrorString(e)));

    // the input in the buffer type
    void *p0 = nullptr;
    if ((st = alloc(static_cast<size_t>(npx) * es, &p0)) != DIPAMD_OK) return release(st);
    View src = contiguous_view(p0, buf_dt, ndim, in.sizes);
    {
        CastGeo g = geo_of(in, src, ndim, 1, 1);
        e = launch_to_complex(in.dt, dbl, in.p, src.p, &g, stream);
        if (e != hipSuccess) return release(set_error(DIPAMD_E_HIP, std::string("complex cast: ") + hipGetErrorString(e)));
    }
    void *p1 = nullptr, *p2 = nullptr;
    const bool out_is_buf = out.dt == buf_dt;
    for (size_t k = 0; k < order.size(); ++k) {
        const int d = order[k];
        const size_t fi = nfilters == 1 ? 0 : static_cast<size_t>(d);
        const CxFilter &f = F[fi];
        const bool last = k + 1 == order.size();
        View dst;
        if (last && out_is_buf && !views_overlap(out, src, ndim)) {
            dst = out;
        } else {
            void *&q = (src.p == p1) ? p2 : p1;
            if (!q && (st = alloc(static_cast<size_t>(npx) * es, &q)) != DIPAMD_OK) return release(st);
            dst = contiguous_view(q, buf_dt, ndim, in.sizes);
        }
        if (bc_value_based(bc[d])) {
            // the line extended by the border, real and imaginary parts separately
            const int B = std::max(1, f.border);
            int64_t esz[DIPAMD_MAX_DIMS];
            int64_t etot = 1;
            for (int q = 0; q < ndim; ++q) {
                esz[q] = in.sizes[q] + (q == d ? 2 * static_cast<int64_t>(B) : 0);
                etot *= esz[q];
            }
            void *pe = nullptr, *pf = nullptr;
            if ((st = alloc(static_cast<size_t>(etot) * es, &pe)) != DIPAMD_OK) return release(st);
            if ((st = alloc(static_cast<size_t>(etot) * es, &pf)) != DIPAMD_OK) return release(st);
            const View ext = contiguous_view(pe, buf_dt, ndim, esz);
            const View filt = contiguous_view(pf, buf_dt, ndim, esz);
            CastGeo g;
            std::memset(&g, 0, sizeof(g));
            g.ndim = ndim;
            g.total = etot;
            for (int q = 0; q < ndim; ++q) {
                g.size[q] = esz[q];
                g.sin[q] = 2 * src.strides[q];
                g.sout[q] = 2 * ext.strides[q];
            }
            const size_t ss = dbl ? 8 : 4;
            for (int part = 0; part < 2 && e == hipSuccess; ++part)
                e = launch_expand(IN_BUF, dbl, static_cast<const char *>(src.p) + part * ss,
                                  static_cast<char *>(ext.p) + part * ss, &g, d, static_cast<int>(in.sizes[d]),
                                  B - f.zpad, f.zpad, bc[d], stream);
            if (e != hipSuccess)
                return release(set_error(DIPAMD_E_HIP, std::string("expand: ") + hipGetErrorString(e)));
            st = dbl ? run_cpass<double>(ext, filt, ndim, d, f, static_cast<const double *>(dtaps) + toff[fi],
                                         any_complex_filter, DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE, stream)
                     : run_cpass<float>(ext, filt, ndim, d, f, static_cast<const float *>(dtaps) + toff[fi],
                                        any_complex_filter, DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE, stream);
            if (st != DIPAMD_OK) return release(st);
            View inner = filt;
            inner.p = static_cast<char *>(filt.p) + static_cast<int64_t>(B) * filt.strides[d] * static_cast<int64_t>(es);
            for (int q = 0; q < ndim; ++q) inner.sizes[q] = in.sizes[q];
            CastGeo gc = geo_of(inner, dst, ndim, 1, 1);
            e = launch_complex_copy(dbl, dst.dt == DIPAMD_DT_DCOMPLEX, inner.p, dst.p, &gc, stream);
            if (e != hipSuccess) return release(set_error(DIPAMD_E_HIP, std::string("copy: ") + hipGetErrorString(e)));
        } else {
            st = dbl ? run_cpass<double>(src, dst, ndim, d, f, static_cast<const double *>(dtaps) + toff[fi],
                                         any_complex_filter, bc[d], stream)
                     : run_cpass<float>(src, dst, ndim, d, f, static_cast<const float *>(dtaps) + toff[fi],
                                        any_complex_filter, bc[d], stream);
            if (st != DIPAMD_OK) return release(st);
        }
        src = dst;
    }
    if (src.p != out.p) {  // the output type differs from the buffer's, or it overlapped
        CastGeo g = geo_of(src, out, ndim, 1, 1);
        e = launch_complex_copy(dbl, out_dbl, src.p, out.p, &g, stream);
        if (e != hipSuccess) return release(set_error(DIPAMD_E_HIP, std::string("copy: ") + hipGetErrorString(e)));
    }
    return release(DIPAMD_OK);
}

}  // namespace dipamd
