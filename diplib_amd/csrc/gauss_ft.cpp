// dip::GaussFT on the device (src/linear/gauss.cpp:397-487).  Real images in the
// spatial domain: the image is extended by the boundary condition to a 5-smooth size
// that holds its Gaussian border (ExpandInput, gauss.cpp:380-393; without a condition
// it is transformed as it is), transformed with rocFFT (real to half-complex), the
// spectrum is multiplied by the separable frequency response of each filtered dim
// (GaussFTLineFilter, gauss.cpp:290-378: the response computed on the host in the
// transform's precision, as the reference computes it), transformed back, scaled by
// 1/N and cropped to the centred window.  Dims that are neither smoothed nor
// differentiated are not transformed (their response is 1), except the dim of the
// reference's real-to-complex pass.  Single precision for 8-bit and sfloat images,
// double for dfloat, as DIPlib's DFT.  Where that pass's padded length is odd the
// reference's output is not shift invariant (a defect of its odd-length real DFT path);
// this implementation keeps the exact convolution there.
// A frequency-domain input or output (gauss_ft_centred) takes the complex path: the
// centred DFT of every dim (rocFFT complex-to-complex between cyclic shifts), the
// product with the centred responses, and the centred inverse where asked.
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "dipamd_internal.hpp"
#include "ft_common.hpp"
#include "complex_kernels.hpp"

namespace dipamd {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Smallest 2-3-5-smooth number >= n (GetOptimalDFTSize(n, true, 5),
// dft_optimal_size.cpp:63-108: the size the reference pads a real image to).
int64_t next_smooth5(int64_t n) {
    if (n <= 6) return n;
    for (int64_t m = n;; ++m) {
        int64_t r = m;
        for (int f : {2, 3, 5})
            while (r % f == 0) r /= f;
        if (r == 1) return m;
    }
}

// Frequency response of one dim (GaussFTLineFilter's constructor, gauss.cpp:294-345) in
// the centred order of DIPlib's DFT (frequency 0 at size / 2).
template <typename TF>
std::vector<std::complex<TF>> ft_response_centred(int64_t size, double sigma, int order, double truncation) {
    using C = std::complex<TF>;
    std::vector<C> lut(static_cast<size_t>(size), C(0));
    const int64_t origin = size / 2;
    TF b = static_cast<TF>(2.0 * kPi * sigma) / static_cast<TF>(size);
    b = -TF(0.5) * b * b;
    const int64_t N =
        b == 0 ? size
               : gauss_radius(static_cast<double>(size) / (2.0 * kPi * sigma), order, truncation, sizeof(TF) == 8);
    const int64_t begin = std::max<int64_t>(0, origin - N);
    const int64_t end = std::min<int64_t>(size, origin + N + 1);
    if (order > 0) {
        const TF o = static_cast<TF>(order);
        C a = {0, static_cast<TF>(2.0 * kPi) / static_cast<TF>(size)};
        a = std::pow(a, o);
        for (int64_t jj = begin; jj < end; ++jj) {
            const TF x = static_cast<TF>(jj - origin);
            lut[static_cast<size_t>(jj)] = b != 0 ? a * std::pow(x, o) * std::exp(b * x * x) : a * std::pow(x, o);
        }
    } else if (b != 0) {
        for (int64_t jj = begin; jj < end; ++jj) {
            const TF x = static_cast<TF>(jj - origin);
            lut[static_cast<size_t>(jj)] = std::exp(b * x * x);
        }
    } else {
        std::fill(lut.begin(), lut.end(), C(1));
    }
    return lut;
}

// The same rotated to FFT order (frequency k at k).
template <typename TF>
std::vector<std::complex<TF>> ft_response(int64_t size, double sigma, int order, double truncation) {
    using C = std::complex<TF>;
    const std::vector<C> lut = ft_response_centred<TF>(size, sigma, order, truncation);
    const int64_t origin = size / 2;
    std::vector<C> out(lut.size());
    for (int64_t k = 0; k < size; ++k) out[static_cast<size_t>(k)] = lut[static_cast<size_t>((k + origin) % size)];
    return out;
}

// rocFFT plans, kept for the process (keyed by lengths, batch, precision, direction).
std::mutex g_fft_mu;
std::once_flag g_fft_once;
std::map<std::tuple<int64_t, int64_t, int64_t, int, int64_t, bool, bool>, rocfft_plan> g_plans;

int get_plan(const FtGeo &g, bool is_double, bool inverse, rocfft_plan *plan) {
    std::call_once(g_fft_once, []() { rocfft_setup(); });
    const auto key = std::make_tuple(g.P[0], g.P[1], g.P[2], g.nt, g.nbatch, is_double, inverse);
    std::lock_guard<std::mutex> lk(g_fft_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
        *plan = it->second;
        return DIPAMD_OK;
    }
    const size_t lengths[3] = {static_cast<size_t>(g.P[0]), static_cast<size_t>(g.P[1]), static_cast<size_t>(g.P[2])};
    rocfft_plan p = nullptr;
    const rocfft_status s = rocfft_plan_create(
        &p, rocfft_placement_notinplace, inverse ? rocfft_transform_type_real_inverse : rocfft_transform_type_real_forward,
        is_double ? rocfft_precision_double : rocfft_precision_single, static_cast<size_t>(g.nt), lengths,
        static_cast<size_t>(g.nbatch), nullptr);
    if (s != rocfft_status_success) return set_error(DIPAMD_E_HIP, "rocFFT plan creation failed");
    g_plans[key] = p;
    *plan = p;
    return DIPAMD_OK;
}

int run_fft(rocfft_plan plan, void *in, void *out, hipStream_t stream, std::vector<void *> &scratch) {
    size_t wbytes = 0;
    if (rocfft_plan_get_work_buffer_size(plan, &wbytes) != rocfft_status_success)
        return set_error(DIPAMD_E_HIP, "rocFFT work buffer query failed");
    rocfft_execution_info info = nullptr;
    if (rocfft_execution_info_create(&info) != rocfft_status_success)
        return set_error(DIPAMD_E_HIP, "rocFFT execution info failed");
    void *work = nullptr;
    if (wbytes > 0) {
        if (scratch_alloc(&work, wbytes, stream) != hipSuccess) {
            rocfft_execution_info_destroy(info);
            return set_error(DIPAMD_E_HIP, "scratch allocation failed");
        }
        scratch.push_back(work);
        rocfft_execution_info_set_work_buffer(info, work, wbytes);
    }
    rocfft_execution_info_set_stream(info, stream);
    void *ib[1] = {in};
    void *ob[1] = {out};
    const rocfft_status s = rocfft_execute(plan, ib, ob, info);
    rocfft_execution_info_destroy(info);
    if (s != rocfft_status_success) return set_error(DIPAMD_E_HIP, "rocFFT execution failed");
    return DIPAMD_OK;
}

// Complex-to-complex plans of the centred path (keyed by lengths, precision, direction).
std::map<std::tuple<int64_t, int64_t, int64_t, int, bool, bool>, rocfft_plan> g_cplans;

int get_cplan(const int64_t *len, int rank, bool is_double, bool inverse, rocfft_plan *plan) {
    std::call_once(g_fft_once, []() { rocfft_setup(); });
    const auto key = std::make_tuple(len[0], rank > 1 ? len[1] : 1, rank > 2 ? len[2] : 1, rank, is_double, inverse);
    std::lock_guard<std::mutex> lk(g_fft_mu);
    auto it = g_cplans.find(key);
    if (it != g_cplans.end()) {
        *plan = it->second;
        return DIPAMD_OK;
    }
    size_t lengths[3] = {1, 1, 1};
    for (int k = 0; k < rank; ++k) lengths[k] = static_cast<size_t>(len[k]);
    rocfft_plan p = nullptr;
    const rocfft_status s = rocfft_plan_create(
        &p, rocfft_placement_inplace,
        inverse ? rocfft_transform_type_complex_inverse : rocfft_transform_type_complex_forward,
        is_double ? rocfft_precision_double : rocfft_precision_single, static_cast<size_t>(rank), lengths, 1, nullptr);
    if (s != rocfft_status_success) return set_error(DIPAMD_E_HIP, "rocFFT plan creation failed");
    g_cplans[key] = p;
    *plan = p;
    return DIPAMD_OK;
}

int64_t npx_of(const View &v, int ndim) {
    int64_t n = 1;
    for (int d = 0; d < ndim; ++d) n *= v.sizes[d];
    return n;
}

}  // namespace

int gauss_ft_device(const View &in, const View &out, int ndim, const double *sig, const int *ord, const int *bc,
                    bool expand, double truncation, hipStream_t stream) {
    const int buf_dt = suggest_flex(in.dt);
    const bool is_double = buf_dt == DIPAMD_DT_DFLOAT;
    FtGeo g;
    std::memset(&g, 0, sizeof(g));
    g.ndim = ndim;
    g.expand = expand ? 1 : 0;
    for (int d = 0; d < ndim; ++d) {
        g.size[d] = in.sizes[d];
        g.sin[d] = in.strides[d];
    }
    bool filtered[DIPAMD_MAX_DIMS], any = false;
    for (int d = 0; d < ndim; ++d) {
        filtered[d] = sig[d] > 0.0 || ord[d] > 0;
        any = any || filtered[d];
    }
    if (any) {
        // Padded sizes of every dim (ExpandInput pads them all) decide the dim of the
        // real-to-complex pass (FourierTransform, fourier.cpp:548-562: the first of the
        // smallest |stride| among sizes > 2; the padded image has normal strides).  That
        // dim is transformed even when it is not filtered: the inverse pass along it
        // shapes the result at unpaired Nyquist bins.  Other unfiltered dims are a batch.
        int64_t Pall[DIPAMD_MAX_DIMS], str[DIPAMD_MAX_DIMS];
        int64_t acc = 1;
        for (int d = 0; d < ndim; ++d) {
            Pall[d] = expand ? next_smooth5(in.sizes[d] + 2 * static_cast<int64_t>(gauss_radius(
                                                               filtered[d] ? sig[d] : 0.0, ord[d], truncation,
                                                               in.dt == DIPAMD_DT_DFLOAT)))
                             : in.sizes[d];
            str[d] = expand ? acc : in.strides[d];
            acc *= Pall[d];
        }
        int opt = 0;
        for (int d = 1; d < ndim; ++d)
            if (Pall[d] > 2 && (Pall[opt] <= 2 || std::llabs(str[d]) < std::llabs(str[opt]))) opt = d;
        g.tdim[g.nt++] = opt;
        for (int d = 0; d < ndim; ++d) {
            if (d == opt) continue;
            if (filtered[d]) {
                if (g.nt == kFtMaxT)
                    return set_error(DIPAMD_E_UNSUPPORTED, "the FT method transforms at most 3 dimensions at once");
                g.tdim[g.nt++] = d;
            } else {
                g.bdim[g.nb++] = d;
            }
        }
        // extension order: the highest dim outermost (ExtendRegion runs dim 0 first)
        for (int k = 0; k < g.nt; ++k) g.ext_order[k] = k;
        std::sort(g.ext_order, g.ext_order + g.nt, [&](int a, int b) { return g.tdim[a] > g.tdim[b]; });
    }
    if (g.nt == 0) {
        // no dim filtered: a copy with conversion (the reference's `out = in`)
        InternFilter id;
        id.data = {1.0};
        id.size = 1;
        const InternFilter *none[DIPAMD_MAX_DIMS];
        for (int d = 0; d < ndim; ++d) none[d] = &id;
        bool proc[DIPAMD_MAX_DIMS] = {};
        int bcs[DIPAMD_MAX_DIMS] = {};
        return separable_device(in, out, ndim, none, bcs, proc, 0, stream);
    }
    g.nbatch = 1;
    for (int k = 0; k < g.nb; ++k) g.nbatch *= in.sizes[g.bdim[k]];
    for (int k = 0; k < kFtMaxT; ++k) {
        g.P[k] = 1;
        g.off[k] = 0;
    }
    for (int k = 0; k < g.nt; ++k) {
        const int d = g.tdim[k];
        const int64_t n = in.sizes[d];
        int64_t P = n;
        if (expand) {
            // HalfGaussianSize with the input's type (gauss.cpp:385-387)
            P = next_smooth5(n + 2 * static_cast<int64_t>(gauss_radius(sig[d] > 0.0 ? sig[d] : 0.0, ord[d],
                                                                       truncation, in.dt == DIPAMD_DT_DFLOAT)));
            const int b = bc[d];
            if (b == DIPAMD_BC_ALREADY_EXPANDED)
                return set_error(DIPAMD_E_UNSUPPORTED, "Functionality has not (yet) been implemented");
            if (!bc_supported(b))
                return set_error(DIPAMD_E_UNSUPPORTED,
                                 "the FT method takes the boundary conditions that copy samples (not extrapolation)");
            g.bc[k] = b;
            g.cval[k] = b == DIPAMD_BC_ADD_MAX_VALUE
                            ? (is_double ? std::numeric_limits<double>::max() : std::numeric_limits<float>::max())
                        : b == DIPAMD_BC_ADD_MIN_VALUE
                            ? (is_double ? std::numeric_limits<double>::lowest()
                                         : static_cast<double>(std::numeric_limits<float>::lowest()))
                            : 0.0;
        }
        g.P[k] = P;
        // the image's window in the padded line (CropWindow CENTER, image_manip.cpp:561-564)
        g.off[k] = (P - n) / 2 + ((P % 2 == 0) && (n % 2 == 1) ? 1 : 0);
    }
    g.nreal = g.P[0] * g.P[1] * g.P[2] * g.nbatch;
    g.nhalf = (g.P[0] / 2 + 1) * g.P[1] * g.P[2] * g.nbatch;
    // frequency responses (complex, in the transform's precision)
    std::vector<double> lutd;
    std::vector<float> lutf;
    int64_t lut_len = 0;
    for (int k = 0; k < g.nt; ++k) {
        const int d = g.tdim[k];
        g.lut_off[k] = lut_len;
        // conj(H(-k)) in place of H(k): the reference's inverse reads the non-positive half
        // of the real-to-complex dim's centred spectrum (fourier.cpp:596-616), which is the
        // standard half-complex inverse of the spectrum times conj(H(-k)) -- the same as
        // H(k) except at the unpaired Nyquist bin of odd derivative orders
        const int64_t P = g.P[k];
        if (is_double) {
            auto r = ft_response<double>(P, sig[d] > 0 ? sig[d] : 0.0, ord[d], truncation);
            for (int64_t q = 0; q < P; ++q) {
                const auto c = std::conj(r[static_cast<size_t>((P - q) % P)]);
                lutd.push_back(c.real());
                lutd.push_back(c.imag());
            }
        } else {
            auto r = ft_response<float>(P, sig[d] > 0 ? sig[d] : 0.0, ord[d], truncation);
            for (int64_t q = 0; q < P; ++q) {
                const auto c = std::conj(r[static_cast<size_t>((P - q) % P)]);
                lutf.push_back(c.real());
                lutf.push_back(c.imag());
            }
        }
        lut_len += g.P[k];
    }
    const size_t es = is_double ? 8 : 4;
    std::vector<void *> scratch;
    auto finish = [&](int st) {
        for (void *p : scratch) scratch_free(p, stream);
        return st;
    };
    auto alloc = [&](size_t bytes, void **p) -> int {
        if (scratch_alloc(p, bytes, stream) != hipSuccess) return set_error(DIPAMD_E_HIP, "scratch allocation failed");
        scratch.push_back(*p);
        return DIPAMD_OK;
    };
    void *pad = nullptr, *spec = nullptr, *luts = nullptr, *src = nullptr;
    int st;
    if ((st = alloc(static_cast<size_t>(g.nreal) * es, &pad)) != DIPAMD_OK) return finish(st);
    if ((st = alloc(static_cast<size_t>(g.nhalf) * 2 * es, &spec)) != DIPAMD_OK) return finish(st);
    if ((st = alloc(static_cast<size_t>(lut_len) * 2 * es, &luts)) != DIPAMD_OK) return finish(st);
    if (hipMemcpyAsync(luts, is_double ? static_cast<const void *>(lutd.data()) : static_cast<const void *>(lutf.data()),
                       static_cast<size_t>(lut_len) * 2 * es, hipMemcpyHostToDevice, stream) != hipSuccess)
        return finish(set_error(DIPAMD_E_HIP, "copy of the frequency response failed"));
    // inputs without an expand kernel: a copy in the buffer type first
    View sv = in;
    if (!(in.dt == DIPAMD_DT_UINT8 && !is_double) && in.dt != buf_dt) {
        if ((st = alloc(static_cast<size_t>(npx_of(in, ndim)) * es, &src)) != DIPAMD_OK) return finish(st);
        sv = contiguous_view(src, buf_dt, ndim, in.sizes);
        if ((st = cast_to_buffer(in, sv, ndim, is_double, stream)) != DIPAMD_OK) return finish(st);
        for (int d = 0; d < ndim; ++d) g.sin[d] = sv.strides[d];
    }
    hipError_t e;
    {
        TraceScope ts("ft_expand", g.nreal * static_cast<int64_t>(es) + npx_of(in, ndim) * dt_size(sv.dt), stream);
        e = launch_ft_expand(sv.dt, is_double, sv.p, pad, g, stream);
    }
    if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e)));
    rocfft_plan fwd, inv;
    if ((st = get_plan(g, is_double, false, &fwd)) != DIPAMD_OK) return finish(st);
    if ((st = get_plan(g, is_double, true, &inv)) != DIPAMD_OK) return finish(st);
    {
        TraceScope ts("fft_forward", g.nreal * static_cast<int64_t>(es) + g.nhalf * 2 * static_cast<int64_t>(es), stream);
        if ((st = run_fft(fwd, pad, spec, stream, scratch)) != DIPAMD_OK) return finish(st);
    }
    {
        TraceScope ts("ft_multiply", 2 * g.nhalf * 2 * static_cast<int64_t>(es), stream);
        e = launch_ft_multiply(is_double, spec, luts, g, stream);
    }
    if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e)));
    {
        TraceScope ts("fft_inverse", g.nreal * static_cast<int64_t>(es) + g.nhalf * 2 * static_cast<int64_t>(es), stream);
        if ((st = run_fft(inv, spec, pad, stream, scratch)) != DIPAMD_OK) return finish(st);
    }
    // crop (and 1/N) into the output, or into a buffer-type image cast afterwards
    const double scale = 1.0 / static_cast<double>(g.P[0] * g.P[1] * g.P[2]);
    const bool direct = out.dt == buf_dt;
    View dst = out;
    void *tmp = nullptr;
    if (!direct) {
        if ((st = alloc(static_cast<size_t>(npx_of(in, ndim)) * es, &tmp)) != DIPAMD_OK) return finish(st);
        dst = contiguous_view(tmp, buf_dt, ndim, in.sizes);
    }
    for (int d = 0; d < ndim; ++d) g.sout[d] = dst.strides[d];
    {
        TraceScope ts("ft_crop", npx_of(in, ndim) * static_cast<int64_t>(2 * es), stream);
        e = launch_ft_crop(is_double, pad, dst.p, dst.dt, scale, g, stream);
    }
    if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e)));
    if (!direct && (st = cast_from_buffer(dst, out, ndim, stream)) != DIPAMD_OK) return finish(st);
    return finish(DIPAMD_OK);
}

// GaussFT with a frequency-domain input or output (gauss.cpp:397-487 without
// ExpandInput): the centred DFT of the input (fourier.cpp:116-150: ifftshift, DFT,
// fftshift; unscaled) unless it is a spectrum, times each filtered dim's centred
// response, and the centred inverse (scaled by 1/N) unless a spectrum is asked for.
int gauss_ft_centred(const View &in, const View &out, int ndim, const double *sig, const int *ord, bool in_freq,
                     bool out_freq, double truncation, hipStream_t stream) {
    const bool is_double = in.dt == DIPAMD_DT_DFLOAT || in.dt == DIPAMD_DT_DCOMPLEX || in.dt == DIPAMD_DT_SINT32 ||
                           in.dt == DIPAMD_DT_UINT32;
    if (out.dt != DIPAMD_DT_SCOMPLEX && out.dt != DIPAMD_DT_DCOMPLEX)
        return set_error(DIPAMD_E_UNSUPPORTED, "a frequency-domain GaussFT gives a complex image");
    if (in_freq && in.dt != DIPAMD_DT_SCOMPLEX && in.dt != DIPAMD_DT_DCOMPLEX)
        return set_error(DIPAMD_E_PARAMETER, "Data type not supported");
    if (ndim > 3) return set_error(DIPAMD_E_UNSUPPORTED, "the FT method transforms at most 3 dimensions at once");
    const int64_t n = npx_of(in, ndim);
    const size_t ce = is_double ? 16 : 8;
    std::vector<void *> scratch;
    auto finish = [&](int st) {
        for (void *p : scratch) scratch_free(p, stream);
        return st;
    };
    auto alloc = [&](size_t bytes, void **p) -> int {
        if (scratch_alloc(p, bytes ? bytes : 1, stream) != hipSuccess)
            return set_error(DIPAMD_E_HIP, "scratch allocation failed");
        scratch.push_back(*p);
        return DIPAMD_OK;
    };
    void *a = nullptr, *b = nullptr, *luts = nullptr;
    int st;
    if ((st = alloc(static_cast<size_t>(n) * ce, &a)) != DIPAMD_OK) return finish(st);
    if ((st = alloc(static_cast<size_t>(n) * ce, &b)) != DIPAMD_OK) return finish(st);
    CastGeo cg;
    std::memset(&cg, 0, sizeof(cg));
    cg.ndim = ndim;
    cg.total = n;
    int64_t s = 1;
    for (int d = 0; d < ndim; ++d) {
        cg.size[d] = in.sizes[d];
        cg.sin[d] = in.strides[d];
        cg.sout[d] = s;
        s *= in.sizes[d];
    }
    hipError_t e = (in.dt == DIPAMD_DT_SCOMPLEX || in.dt == DIPAMD_DT_DCOMPLEX)
                       ? launch_complex_copy(in.dt == DIPAMD_DT_DCOMPLEX, is_double, in.p, a, &cg, stream)
                       : launch_to_complex(in.dt, is_double, in.p, a, &cg, stream);
    if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e)));
    FtCGeo g;
    std::memset(&g, 0, sizeof(g));
    g.ndim = ndim;
    g.total = n;
    int64_t len[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) {
        g.size[d] = in.sizes[d];
        g.lut_off[d] = -1;
        len[d] = in.sizes[d];
    }
    auto shift = [&](const void *src, void *dst, bool to_corner) -> int {
        for (int d = 0; d < ndim; ++d) g.shift[d] = to_corner ? (g.size[d] - g.size[d] / 2) % g.size[d] : g.size[d] / 2;
        const hipError_t err = launch_ft_cshift(is_double, src, dst, g, stream);
        return err == hipSuccess ? DIPAMD_OK : set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(err));
    };
    auto transform = [&](bool inverse) -> int {
        // a -> b (origin to the corner), DFT in place, b -> a (origin to the middle)
        int r;
        if ((r = shift(a, b, true)) != DIPAMD_OK) return r;
        rocfft_plan plan;
        if ((r = get_cplan(len, ndim, is_double, inverse, &plan)) != DIPAMD_OK) return r;
        if ((r = run_fft(plan, b, b, stream, scratch)) != DIPAMD_OK) return r;
        return shift(b, a, false);
    };
    if (!in_freq && (st = transform(false)) != DIPAMD_OK) return finish(st);
    // the responses of the filtered dims, centred, in the transform's precision
    std::vector<double> lutd;
    std::vector<float> lutf;
    int64_t lut_len = 0;
    for (int d = 0; d < ndim; ++d) {
        if (!(sig[d] > 0.0 || ord[d] > 0)) continue;
        g.lut_off[d] = lut_len;
        if (is_double) {
            for (const auto &c : ft_response_centred<double>(in.sizes[d], sig[d], ord[d], truncation)) {
                lutd.push_back(c.real());
                lutd.push_back(c.imag());
            }
        } else {
            for (const auto &c : ft_response_centred<float>(in.sizes[d], sig[d], ord[d], truncation)) {
                lutf.push_back(c.real());
                lutf.push_back(c.imag());
            }
        }
        lut_len += in.sizes[d];
    }
    if (lut_len > 0) {
        if ((st = alloc(static_cast<size_t>(lut_len) * ce, &luts)) != DIPAMD_OK) return finish(st);
        if (hipMemcpyAsync(luts, is_double ? static_cast<const void *>(lutd.data()) : static_cast<const void *>(lutf.data()),
                           static_cast<size_t>(lut_len) * ce, hipMemcpyHostToDevice, stream) != hipSuccess)
            return finish(set_error(DIPAMD_E_HIP, "copy of the frequency response failed"));
        e = launch_ft_cmultiply(is_double, a, luts, 1.0, g, stream);
        if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e)));
    }
    if (!out_freq) {
        if ((st = transform(true)) != DIPAMD_OK) return finish(st);
        for (int d = 0; d < ndim; ++d) g.lut_off[d] = -1;
        e = launch_ft_cmultiply(is_double, a, nullptr, 1.0 / static_cast<double>(n), g, stream);
        if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e)));
    }
    for (int d = 0; d < ndim; ++d) {
        cg.sin[d] = cg.sout[d];
        cg.sout[d] = out.strides[d];
    }
    e = launch_complex_copy(is_double, out.dt == DIPAMD_DT_DCOMPLEX, a, out.p, &cg, stream);
    if (e != hipSuccess) return finish(set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e)));
    return finish(DIPAMD_OK);
}

}  // namespace dipamd

using namespace dipamd;

extern "C" int dipamd_gauss_ft(const dipamd_image *in, const dipamd_image *out, const double *sigmas, int32_t nsigmas,
                               const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc,
                               double truncation, uint32_t flags, void *stream) {
    int st;
    const bool in_freq = (flags & DIPAMD_FT_IN_FREQUENCY) != 0, out_freq = (flags & DIPAMD_FT_OUT_FREQUENCY) != 0;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if (!in_freq && !out_freq && (st = require_real(in)) != DIPAMD_OK) return st;
    if ((st = check_image(out, "out")) != DIPAMD_OK) return st;
    const int nd = in->ndim;
    if (out->ndim != nd) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    for (int d = 0; d < nd; ++d)
        if (out->sizes[d] != in->sizes[d]) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    if ((nsigmas != 0 && nsigmas != 1 && nsigmas != nd) || (norders != 0 && norders != 1 && norders != nd))
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    if (in->dtype == DIPAMD_DT_BIN) return set_error(DIPAMD_E_PARAMETER, "Data type not supported");
    double s[DIPAMD_MAX_DIMS];
    int o[DIPAMD_MAX_DIMS];
    for (int d = 0; d < nd; ++d) {
        s[d] = nsigmas == 0 ? 1.0 : (nsigmas == 1 ? sigmas[0] : sigmas[d]);
        o[d] = norders == 0 ? 0 : (norders == 1 ? orders[0] : orders[d]);
        if (o[d] < 0) return set_error(DIPAMD_E_PARAMETER, "derivative order must be non-negative");
        if (in->sizes[d] == 1) {  // gauss.cpp:419-427
            s[d] = 0;
            o[d] = 0;
        } else if (s[d] < 0) {
            s[d] = 0;
        }
    }
    if (in_freq || out_freq) {
        if (nbc > 0)
            return set_error(DIPAMD_E_UNSUPPORTED,
                             "a frequency-domain GaussFT takes no boundary condition (the image is not extended)");
        return gauss_ft_centred(to_view(in), to_view(out), nd, s, o, in_freq, out_freq, truncation,
                                static_cast<hipStream_t>(stream));
    }
    int bcs[DIPAMD_MAX_DIMS] = {};
    const bool expand = nbc > 0;
    if (expand && (st = broadcast_bc(bc, nbc, nd, bcs)) != DIPAMD_OK) return st;
    return gauss_ft_device(to_view(in), to_view(out), nd, s, o, bcs, expand, truncation,
                           static_cast<hipStream_t>(stream));
}
