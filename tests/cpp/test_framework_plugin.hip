// A user line filter plugged into the device separable framework
// (include/diplib_amd_framework.hpp): the reference's GENERAL convolution loop
// (convolution.cpp:202-212: sum = 0, sum += f * x over the reversed taps from
// x[i - origin]) written as a plug-in, run through dip::Framework::Separable and
// compared, bit for bit, with dip::SeparableConvolution of the same filter (which the
// reference's fixtures pin), over boundary conditions (index-map and value-based),
// input types, 1-D to 3-D, per-dim process flags, strided input and in-place use; and a
// non-linear plug-in (a running maximum over the extended line) against a host loop.
// Prints "PASSED" when every check holds.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <iostream>
#include <random>
#include <vector>

#include "diplib_amd.hpp"
#include "diplib_amd_framework.hpp"

namespace {

constexpr int kMaxTaps = 16;

// taps reversed (correlation order), internal origin = size - origin - 1
template <typename TB>
struct GeneralLine {
    TB taps[kMaxTaps];
    int ntaps, origin;
    __device__ void Filter(dip::Framework::SeparableLineFilterParameters const &p) const {
        const TB *in = static_cast<const TB *>(p.inBuffer.buffer);
        TB *out = static_cast<TB *>(p.outBuffer.buffer);
        const dip::sint si = p.inBuffer.stride, so = p.outBuffer.stride;
        for (dip::uint i = 0; i < p.inBuffer.length; ++i) {
            const TB *x = in + (static_cast<dip::sint>(i) - origin) * si;
            TB sum = 0;
            for (int k = 0; k < ntaps; ++k) sum += taps[k] * x[k * si];
            out[static_cast<dip::sint>(i) * so] = sum;
        }
    }
};

// out[i] = max(x[i - r] .. x[i + r]) over the extended line
struct RunningMax {
    int r;
    __device__ void Filter(dip::Framework::SeparableLineFilterParameters const &p) const {
        const float *in = static_cast<const float *>(p.inBuffer.buffer);
        float *out = static_cast<float *>(p.outBuffer.buffer);
        const dip::sint si = p.inBuffer.stride;
        for (dip::uint i = 0; i < p.inBuffer.length; ++i) {
            float m = in[(static_cast<dip::sint>(i) - r) * si];
            for (int k = -r + 1; k <= r; ++k) m = fmaxf(m, in[(static_cast<dip::sint>(i) + k) * si]);
            out[static_cast<dip::sint>(i) * p.outBuffer.stride] = m;
        }
    }
};

int fails = 0;
void check(bool ok, std::string const &what) {
    std::cout << (ok ? "ok: " : "FAIL: ") << what << std::endl;
    if (!ok) ++fails;
}

template <typename T>
std::vector<T> host(dip::Image const &im) {
    std::vector<T> v(im.NumberOfPixels());
    (void)hipDeviceSynchronize();
    im.CopyToHost(v.data());
    return v;
}

template <typename T>
dip::Image device_image(dip::UnsignedArray sizes, dip::DataType dt, unsigned seed) {
    dip::Image im(sizes, 1, dt);
    std::vector<T> v(im.NumberOfPixels());
    std::mt19937 g(seed);
    for (auto &e : v) e = static_cast<T>(std::is_floating_point<T>::value ? (g() % 100000) / 1000.0 : g() % 256);
    im.CopyFromHost(v.data());
    return im;
}

// the plug-in against dip::SeparableConvolution with a "general" filter
template <typename TB>
void general_case(std::string const &name, dip::Image const &in, std::vector<double> const &taps, int origin,
                  dip::StringArray const &bc, dip::BooleanArray const &process) {
    GeneralLine<TB> f{};
    const int n = static_cast<int>(taps.size());
    for (int k = 0; k < n; ++k) f.taps[k] = static_cast<TB>(taps[static_cast<size_t>(n - 1 - k)]);
    f.ntaps = n;
    const int o = origin < 0 ? n / 2 : origin;
    f.origin = n - o - 1;
    const dip::uint border = static_cast<dip::uint>(std::max(f.origin, n - f.origin - 1));
    const dip::DataType bt = dip::DataType::SuggestFlex(in.DataType());
    dip::Image a, b;
    dip::Framework::Separable(in, a, bt, bt, process, {border}, bc, f);
    dip::SeparableConvolution(in, b, {dip::OneDimensionalFilter{taps, origin, "general"}}, bc, process);
    const auto ha = host<TB>(a), hb = host<TB>(b);
    check(a.DataType() == b.DataType() && ha.size() == hb.size() &&
              std::memcmp(ha.data(), hb.data(), ha.size() * sizeof(TB)) == 0,
          name + ": plug-in GENERAL filter bit-identical to SeparableConvolution");
}

}  // namespace

int main() {
    try {
        const std::vector<double> t5 = {0.1, -0.25, 0.6, 0.3, -0.05};
        const std::vector<double> t4 = {0.2, 0.5, -0.3, 0.35};
        general_case<float>("f32 2-D mirror", device_image<float>({97, 61}, dip::DT_SFLOAT, 1), t5, -1, {"mirror"}, {});
        general_case<float>("f32 2-D periodic origin 0", device_image<float>({64, 40}, dip::DT_SFLOAT, 2), t4, 0,
                            {"periodic"}, {});
        general_case<float>("u8 2-D first order", device_image<uint8_t>({80, 33}, dip::DT_UINT8, 3), t5, 3,
                            {"first order"}, {});
        general_case<float>("f32 2-D third order / add zeros", device_image<float>({40, 30}, dip::DT_SFLOAT, 4), t4,
                            -1, {"third order", "add zeros"}, {});
        general_case<float>("u16 2-D y only", device_image<uint16_t>({50, 44}, dip::DT_UINT16, 5), t5, -1,
                            {"asym mirror"}, {false, true});
        general_case<float>("f32 3-D", device_image<float>({21, 17, 9}, dip::DT_SFLOAT, 6), t4, 1,
                            {"mirror", "periodic", "zero order"}, {});
        general_case<double>("f64 2-D antisym reflect", device_image<double>({45, 33}, dip::DT_DFLOAT, 7), t5, -1,
                             {"antisym reflect"}, {});
        general_case<float>("f32 1-D short line", device_image<float>({3}, dip::DT_SFLOAT, 8), t5, -1, {"mirror"},
                            {});
        // a transposed view: the passes follow the strides
        {
            dip::Image base = device_image<float>({48, 36}, dip::DT_SFLOAT, 9);
            dip::Image view(dip::NonOwnedRefToDataSegment(base.Origin()), base.Origin(), dip::DT_SFLOAT, {36, 48}, {48, 1});
            general_case<float>("f32 transposed view", view, t5, 2, {"mirror"}, {});
        }
        // in place
        {
            dip::Image x = device_image<float>({60, 50}, dip::DT_SFLOAT, 10);
            dip::Image ref;
            dip::SeparableConvolution(x, ref, {dip::OneDimensionalFilter{t5, -1, "general"}}, {"mirror"}, {});
            GeneralLine<float> f{};
            for (int k = 0; k < 5; ++k) f.taps[k] = static_cast<float>(t5[static_cast<size_t>(4 - k)]);
            f.ntaps = 5;
            f.origin = 2;
            dip::Framework::Separable(x, x, dip::DT_SFLOAT, dip::DT_SFLOAT, {}, {2}, {"mirror"}, f);
            const auto hx = host<float>(x), hr = host<float>(ref);
            check(std::memcmp(hx.data(), hr.data(), hx.size() * sizeof(float)) == 0, "in place");
        }
        // a non-linear plug-in: running maximum, mirror, against a host loop
        {
            const dip::uint W = 70, H = 25;
            dip::Image x = device_image<float>({W, H}, dip::DT_SFLOAT, 11);
            const auto hx = host<float>(x);
            dip::Image y;
            dip::Framework::Separable(x, y, dip::DT_SFLOAT, dip::DT_SFLOAT, {true, false}, {2}, {"mirror"},
                                      RunningMax{2});
            const auto hy = host<float>(y);
            auto mir = [&](long i) { return i < 0 ? -i - 1 : (i >= static_cast<long>(W) ? 2 * W - i - 1 : i); };
            bool ok = true;
            for (dip::uint yy = 0; yy < H; ++yy)
                for (long i = 0; i < static_cast<long>(W); ++i) {
                    float m = hx[yy * W + mir(i - 2)];
                    for (long k = -1; k <= 2; ++k) m = std::max(m, hx[yy * W + mir(i + k)]);
                    ok = ok && m == hy[yy * W + static_cast<dip::uint>(i)];
                }
            check(ok, "running maximum plug-in (x only) against a host loop");
        }
    } catch (std::exception const &e) {
        std::cout << "FAIL: exception " << e.what() << std::endl;
        ++fails;
    }
    std::cout << (fails == 0 ? "PASSED" : "FAILED") << std::endl;
    return fails == 0 ? 0 : 1;
}
