// Half-Gaussian tap generator (MakeHalfGaussianInternal, src/linear/gauss.cpp:53-146)
// written once for any scalar type: double for the taps themselves, and
// std::complex<double> for their derivative with respect to sigma by the complex
// step  df/dsigma = Im f(sigma + i*h) / h  (h = 1e-30: no cancellation, exact to
// rounding, the radius held fixed).
#pragma once

#include <cmath>
#include <complex>
#include <vector>

namespace dipamd {

template <typename T>
std::vector<T> half_gaussian_t(T sigma, int order, int n) {
    using std::exp;
    using std::sqrt;
    std::vector<T> f(static_cast<size_t>(n), T(0));
    const int r0 = n - 1;
    const T sigma2 = sigma * sigma;
    const double pi = 3.14159265358979323846264338327950288;
    switch (order) {
        case 0: {
            const T factor = -0.5 / sigma2;
            T normalization = 0;
            f[r0] = 1.0;
            for (int rr = 1; rr < n; rr++) {
                const double rad = rr;
                const T g = exp(factor * (rad * rad));
                f[r0 - rr] = g;
                normalization += g;
            }
            normalization = 1.0 / (normalization * 2.0 + 1.0);
            for (int rr = 0; rr < n; rr++) f[rr] *= normalization;
            break;
        }
        case 1: {
            const T factor = -0.5 / sigma2;
            T moment = 0.0;
            f[r0] = 0.0;
            for (int rr = 1; rr < n; rr++) {
                const double rad = rr;
                const T g = rad * exp(factor * (rad * rad));
                f[r0 - rr] = g;
                moment += rad * g;
            }
            const T normalization = 1.0 / (2.0 * moment);
            for (int rr = 0; rr < n - 1; rr++) f[rr] *= normalization;
            break;
        }
        case 2: {
            const T norm = 1.0 / (sqrt(2.0 * pi) * sigma * sigma2);
            T mean = 0.0;
            f[r0] = -norm;
            for (int rr = 1; rr < n; rr++) {
                const double rad = rr;
                const T sr2 = rad * rad / sigma2;
                const T g = (sr2 - 1.0) * norm * exp(-0.5 * sr2);
                f[r0 - rr] = g;
                mean += g;
            }
            mean = (mean * 2.0 + f[r0]) / (static_cast<double>(r0) * 2.0 + 1.0);
            f[r0] -= mean;
            T moment = 0.0;
            for (int rr = 1; rr < n; rr++) {
                const double rad = rr;
                f[r0 - rr] -= mean;
                moment += rad * rad * f[r0 - rr];
            }
            const T normalization = 1.0 / moment;
            for (int rr = 0; rr < n; rr++) f[rr] *= normalization;
            break;
        }
        case 3: {
            const T norm = 1.0 / (sqrt(2.0 * pi) * sigma * sigma2 * sigma2);
            f[r0] = 0.0;
            T moment = 0.0;
            for (int rr = 1; rr < n; rr++) {
                const double rad = rr;
                const double rr2 = rad * rad;
                const T sr2 = rr2 / sigma2;
                const T g = norm * exp(-0.5 * sr2) * (rad * (3.0 - sr2));
                f[r0 - rr] = g;
                moment += g * rr2 * rad;
            }
            const T normalization = 3.0 / moment;
            for (int rr = 0; rr < n; rr++) f[rr] *= normalization;
            break;
        }
        default: break;
    }
    return f;
}

// d(taps)/d(sigma) at a fixed half size n (same layout: x = 0 last).
inline std::vector<double> half_gaussian_dsigma(double sigma, int order, int n) {
    const double h = 1e-30;
    std::vector<std::complex<double>> fc = half_gaussian_t<std::complex<double>>({sigma, h}, order, n);
    std::vector<double> d(fc.size());
    for (size_t i = 0; i < fc.size(); ++i) d[i] = fc[i].imag() / h;
    return d;
}

}  // namespace dipamd
