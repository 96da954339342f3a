// Exercises the C++ drop-in (include/diplib_amd.hpp) the way a DIPlib caller
// would: dip::Gauss / GaussFIR / SeparableConvolution on device images, value-
// returning and in/out forms, a protected output, in-place, and the error paths.
// Usage: test_dip_shim <in.f32> <W> <H> <expected.f32> <sigma>
//   in.f32 / expected.f32: raw little-endian float32, W*H samples (dip order: x fastest).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "diplib_amd.hpp"

static std::vector<float> read_f32(const char *path, size_t n) {
    std::vector<float> v(n);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(n * sizeof(float)));
    if (!f) throw std::runtime_error(std::string("cannot read ") + path);
    return v;
}

static int fails = 0;
#define CHECK(cond, what)                                              \
    do {                                                               \
        if (!(cond)) {                                                 \
            std::cout << "FAIL: " << what << std::endl;                \
            ++fails;                                                   \
        } else {                                                       \
            std::cout << "ok: " << what << std::endl;                  \
        }                                                              \
    } while (0)

int main(int argc, char **argv) {
    if (argc < 6) {
        std::cerr << "usage: test_dip_shim in.f32 W H expected.f32 sigma\n";
        return 2;
    }
    const size_t W = std::stoul(argv[2]), H = std::stoul(argv[3]);
    const double sigma = std::stod(argv[5]);
    std::vector<float> in = read_f32(argv[1], W * H);
    std::vector<float> exp = read_f32(argv[4], W * H);
    std::vector<float> got(W * H);
    dip::SetNumberOfThreads(1);
    CHECK(dip::GetNumberOfThreads() == 1, "SetNumberOfThreads(1) kept");
    dip::SetNumberOfThreads(0);
    CHECK(dip::GetNumberOfThreads() >= 1, "SetNumberOfThreads(0) restores the default");
    try {
        dip::Image img({W, H}, 1, dip::DT_SFLOAT);
        img.CopyFromHost(in.data());
        // value-returning form
        dip::Image out = dip::Gauss(img, {sigma}, {0}, "best", {"mirror"}, 3);
        (void)hipDeviceSynchronize();
        out.CopyToHost(got.data());
        CHECK(out.DataType() == dip::DT_SFLOAT && out.Sizes() == img.Sizes(), "output forged as sfloat, same sizes");
        CHECK(std::memcmp(got.data(), exp.data(), W * H * sizeof(float)) == 0, "dip::Gauss bit-identical to the reference");
        // in/out form into an existing output, GaussFIR
        dip::Image out2;
        dip::GaussFIR(img, out2, {sigma}, {0}, {"mirror"}, 3);
        (void)hipDeviceSynchronize();
        out2.CopyToHost(got.data());
        CHECK(std::memcmp(got.data(), exp.data(), W * H * sizeof(float)) == 0, "dip::GaussFIR bit-identical");
        // SeparableConvolution with the half Gaussian, even symmetry
        std::vector<double> h = dip::MakeHalfGaussian(sigma, 0, 3, dip::DT_SFLOAT);
        dip::OneDimensionalFilterArray fa(1);
        fa[0].filter = h;
        fa[0].symmetry = "even";
        dip::Image out3 = dip::SeparableConvolution(img, fa, {"mirror"});
        (void)hipDeviceSynchronize();
        out3.CopyToHost(got.data());
        CHECK(std::memcmp(got.data(), exp.data(), W * H * sizeof(float)) == 0,
              "dip::SeparableConvolution(MakeHalfGaussian) bit-identical");
        // in-place
        dip::Image inplace({W, H}, 1, dip::DT_SFLOAT);
        inplace.CopyFromHost(in.data());
        dip::Gauss(inplace, inplace, {sigma}, {0}, "FIR", {"mirror"}, 3);
        (void)hipDeviceSynchronize();
        inplace.CopyToHost(got.data());
        CHECK(std::memcmp(got.data(), exp.data(), W * H * sizeof(float)) == 0, "in-place bit-identical");
        // the derivative family: elements equal the Gauss derivatives; GradientMagnitude and
        // Laplace combine them in float, in the reference's order
        {
            auto host = [&](dip::Image const &im) {
                std::vector<float> v(W * H);
                (void)hipDeviceSynchronize();
                im.CopyToHost(v.data());
                return v;
            };
            std::vector<float> dx = host(dip::GaussFIR(img, {sigma}, {1, 0}, {"mirror"}, 3));
            std::vector<float> dy = host(dip::GaussFIR(img, {sigma}, {0, 1}, {"mirror"}, 3));
            std::vector<float> dxx = host(dip::GaussFIR(img, {sigma}, {2, 0}, {"mirror"}, 3));
            std::vector<float> dyy = host(dip::GaussFIR(img, {sigma}, {0, 2}, {"mirror"}, 3));
            std::vector<float> dxy = host(dip::GaussFIR(img, {sigma}, {1, 1}, {"mirror"}, 3));
            dip::ImageArray g, hs;
            dip::Gradient(img, g, {sigma}, "FIR", {"mirror"});
            dip::Hessian(img, hs, {sigma}, "best", {"mirror"});
            CHECK(g.size() == 2 && host(g[0]) == dx && host(g[1]) == dy, "dip::Gradient elements are the derivatives");
            CHECK(hs.size() == 3 && host(hs[0]) == dxx && host(hs[1]) == dyy && host(hs[2]) == dxy,
                  "dip::Hessian elements in dip's order");
            std::vector<float> lap = host(dip::Laplace(img, {sigma}, "FIR", {"mirror"}));
            std::vector<float> mag = host(dip::GradientMagnitude(img, {sigma}, "FIR", {"mirror"}));
            bool lap_ok = true, mag_ok = true;
            for (size_t i = 0; i < W * H; ++i) {
                const float l = dxx[i] + dyy[i];
                const float a = dx[i] * dx[i];
                const float b = dy[i] * dy[i];
                const float m = std::sqrt(a + b);
                lap_ok &= std::memcmp(&l, &lap[i], 4) == 0;
                mag_ok &= std::memcmp(&m, &mag[i], 4) == 0;
            }
            CHECK(lap_ok, "dip::Laplace = dxx + dyy in float");
            CHECK(mag_ok, "dip::GradientMagnitude = sqrt(dx*dx + dy*dy) in float");
        }
        // the IIR method: "best" picks it for sigma > 10, and "IIR" is dip::GaussIIR with
        // the discrete-time-fit design
        {
            auto host = [&](dip::Image const &im) {
                std::vector<float> v(W * H);
                (void)hipDeviceSynchronize();
                im.CopyToHost(v.data());
                return v;
            };
            std::vector<float> iir = host(dip::GaussIIR(img, {12.0}, {1, 0}, {"mirror"}));
            CHECK(host(dip::Gauss(img, {12.0}, {1, 0}, "best", {"mirror"})) == iir, "dip::Gauss best -> GaussIIR");
            CHECK(host(dip::Gauss(img, {12.0}, {1, 0}, "IIR", {"mirror"})) == iir, "dip::Gauss IIR -> GaussIIR");
            CHECK(host(dip::GaussIIR(img, {12.0}, {1, 0}, {"mirror"}, {4, 3}, "discrete time fit")) == iir,
                  "default filter orders: 3 + derivative order");
            dip::ImageArray gi;
            dip::Gradient(img, gi, {12.0}, "best", {"mirror"});
            CHECK(gi.size() == 2 && host(gi[0]) == iir, "dip::Gradient best -> the IIR derivatives");
        }
        // tensor images: every element filtered as a scalar image (AsScalarImage); the
        // Gradient and Hessian as tensor images; a tensor element is a view
        {
            auto host = [&](dip::Image const &im) {
                std::vector<float> v(im.NumberOfSamples());
                (void)hipDeviceSynchronize();
                im.CopyToHost(v.data());
                return v;
            };
            dip::Image t({W, H}, 2, dip::DT_SFLOAT);
            std::vector<float> planes(2 * W * H);
            for (size_t i = 0; i < W * H; ++i) {
                planes[i] = in[i];
                planes[W * H + i] = in[W * H - 1 - i];
            }
            t.CopyFromHost(planes.data());
            dip::Image tg = dip::Gauss(t, {sigma}, {0}, "FIR", {"mirror"}, 3);
            std::vector<float> rev(in.rbegin(), in.rend());
            dip::Image r({W, H}, 1, dip::DT_SFLOAT);
            r.CopyFromHost(rev.data());
            std::vector<float> e1 = host(dip::Gauss(r, {sigma}, {0}, "FIR", {"mirror"}, 3));
            std::vector<float> tgh = host(tg);
            CHECK(tg.TensorElements() == 2 && std::memcmp(tgh.data(), exp.data(), W * H * 4) == 0 &&
                      std::memcmp(tgh.data() + W * H, e1.data(), W * H * 4) == 0,
                  "tensor image: each element filtered as a scalar image");
            dip::Image tsep = dip::SeparableConvolution(t, fa, {"mirror"});
            CHECK(host(tsep) == tgh, "tensor image through SeparableConvolution");
            dip::Image gt = dip::Gradient(img, {sigma}, "FIR", {"mirror"});
            dip::ImageArray ga;
            dip::Gradient(img, ga, {sigma}, "FIR", {"mirror"});
            CHECK(gt.TensorElements() == 2 && host(gt.TensorElement(0)) == host(ga[0]) &&
                      host(gt.TensorElement(1)) == host(ga[1]),
                  "dip::Gradient as a tensor image");
            dip::Image ht;
            dip::Hessian(img, ht, {sigma}, "FIR", {"mirror"});
            CHECK(ht.TensorElements() == 3, "dip::Hessian as a tensor image of 3 elements");
            bool threw_t = false;
            try {
                dip::Laplace(t, {sigma});
            } catch (dip::ParameterError const &) {
                threw_t = true;
            }
            CHECK(threw_t, "the derivative family takes scalar images");
        }
        // host images: staged through the device, the result in host memory
        {
            dip::Image hin({W, H}, 1, dip::DT_SFLOAT, dip::MemoryLocation::HOST);
            std::memcpy(hin.Origin(), in.data(), W * H * sizeof(float));
            dip::Image hout = dip::Gauss(hin, {sigma}, {0}, "FIR", {"mirror"}, 3);
            CHECK(hout.IsOnHost() && std::memcmp(hout.Origin(), exp.data(), W * H * 4) == 0,
                  "host image in, host image out, bit-identical");
            std::vector<float> user(in);  // the caller's own host buffer around an image
            dip::Image wrap(dip::NonOwnedRefToDataSegment(user.data()), user.data(), dip::DT_SFLOAT, {W, H}, {}, 1, 0,
                            dip::MemoryLocation::HOST);
            dip::Gauss(wrap, wrap, {sigma}, {0}, "FIR", {"mirror"}, 3);
            CHECK(std::memcmp(user.data(), exp.data(), W * H * 4) == 0, "in place on a host buffer");
            dip::Image dev_from_host({W, H}, 1, dip::DT_SFLOAT);
            dip::GaussFIR(hin, dev_from_host, {sigma}, {0}, {"mirror"}, 3);
            std::vector<float> dv(W * H);
            (void)hipDeviceSynchronize();
            dev_from_host.CopyToHost(dv.data());
            CHECK(!dev_from_host.IsOnHost() && dv == exp, "host input into a device output");
        }
        // protected output of another type keeps its type (clamp_cast)
        dip::Image prot({W, H}, 1, dip::DT_DFLOAT);
        prot.Protect();
        dip::Gauss(img, prot, {sigma}, {0}, "FIR", {"mirror"}, 3);
        (void)hipDeviceSynchronize();
        std::vector<double> gotd(W * H);
        prot.CopyToHost(gotd.data());
        bool same = true;
        for (size_t i = 0; i < W * H; ++i) same &= gotd[i] == static_cast<double>(exp[i]);
        CHECK(prot.DataType() == dip::DT_DFLOAT && same, "protected dfloat output keeps its type");
        // protected output of the wrong size throws
        dip::Image small({W / 2 + 1, H}, 1, dip::DT_SFLOAT);
        small.Protect();
        bool threw = false;
        try {
            dip::Gauss(img, small, {sigma});
        } catch (dip::ParameterError const &e) {
            threw = std::string(e.what()).find("protected") != std::string::npos;
        }
        CHECK(threw, "protected output of the wrong size throws \"Image is protected\"");
        // bad boundary condition and order > 3 throw ParameterError
        threw = false;
        try {
            dip::Gauss(img, {sigma}, {0}, "FIR", {"nonsense"});
        } catch (dip::ParameterError const &) {
            threw = true;
        }
        CHECK(threw, "unknown boundary condition throws");
        threw = false;
        try {
            dip::GaussFIR(img, {sigma}, {4});
        } catch (dip::ParameterError const &) {
            threw = true;
        }
        CHECK(threw, "order > 3 throws");
        threw = false;
        try {
            dip::Image unforged;
            dip::Gauss(unforged, {sigma});
        } catch (dip::ParameterError const &) {
            threw = true;
        }
        CHECK(threw, "unforged input throws");
    } catch (std::exception const &e) {
        std::cout << "FAIL: exception " << e.what() << std::endl;
        return 1;
    }
    std::cout << (fails ? "FAILED" : "PASSED") << std::endl;
    return fails ? 1 : 0;
}
