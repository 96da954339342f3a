// Host design of the recursive Gaussian (src/linear/gaussiir.cpp:54-425): pole
// tables, the scale q that gives the requested sigma, the pole scaling, the
// recursion coefficients from the scaled poles, the gain and the moving-average
// parts of the derivative orders.  All arithmetic is double, in the reference's
// order, so the coefficients carry the same bits.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "dipamd_internal.hpp"
#include "iir_common.hpp"

namespace dipamd {

namespace {

struct Cx {
    double re, im;
};
inline Cx cmul(Cx x, Cx y) { return {x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re}; }
inline Cx cadd(Cx x, Cx y) { return {x.re + y.re, x.im + y.im}; }

// Normalised poles in the z-domain (index 1..nn; conjugate pairs first, a real pole
// last).  Discrete-time fit: van Vliet, Young and Verbeek (1998), per derivative
// order 0..2 (gaussiir.cpp:90-175).  Forward-backward: Young and van Vliet (1995)
// (gaussiir.cpp:176-205).
struct PoleSet {
    int nn;
    double p[5][2];
};
const PoleSet kDtf[6][3] = {
    {},
    {{1, {{2.00000, 0.0}}}, {1, {{2.00000, 0.0}}}, {1, {{2.00000, 0.0}}}},
    {{2, {{1.69580, 0.59955}, {1.69580, -0.59955}}},
     {2, {{1.69580, 0.59955}, {1.69580, -0.59955}}},
     {2, {{1.69580, 0.59955}, {1.69580, -0.59955}}}},
    {{3, {{1.41650, 1.00829}, {1.41650, -1.00829}, {1.86131, 0.0}}},
     {3, {{1.32094, 0.97057}, {1.32094, -0.97057}, {1.77635, 0.0}}},
     {3, {{1.21969, 0.91724}, {1.21969, -0.91724}, {1.69485, 0.0}}}},
    {{4, {{1.13231, 1.28122}, {1.13231, -1.28122}, {1.78532, 0.46766}, {1.78532, -0.46766}}},
     {4, {{1.04198, 1.25046}, {1.04198, -1.25046}, {1.69337, 0.45006}, {1.69337, -0.45006}}},
     {4, {{0.94576, 1.21364}, {0.94576, -1.21364}, {1.59892, 0.42668}, {1.59892, -0.42668}}}},
    {{5, {{0.85991, 1.45235}, {0.85991, -1.45235}, {1.60953, 0.83009}, {1.60953, -0.83009}, {1.87040, 0.0}}},
     {5, {{0.70237, 1.38717}, {0.70237, -1.38717}, {1.43280, 0.77903}, {1.43280, -0.77903}, {1.70346, 0.0}}},
     {5, {{0.70381, 1.38271}, {0.70381, -1.38271}, {1.42239, 0.77978}, {1.42239, -0.77978}, {1.69319, 0.0}}}},
};
const PoleSet kFb3 = {3, {{2.1078345, 1.4058574}, {2.1078345, -1.4058574}, {2.1668048, 0.0}}};
const PoleSet kFb4 = {4, {{2.06954, 1.90416}, {2.06954, -1.90416}, {2.18403, 0.593635}, {2.18403, -0.593635}}};
const PoleSet kFb5 = {
    5, {{2.19406, 1.90251}, {2.19406, -1.90251}, {2.31029, 0.598022}, {2.31029, -0.598022}, {7.64071, 0.0}}};

// Pole p raised to the power 1/q in polar form.
inline void pole_power(double re, double im, double q, double *ore, double *oim) {
    double modulus = std::sqrt(re * re + im * im);
    double phase = std::atan(im / re);
    modulus = std::exp(std::log(modulus) / q);
    phase = phase / q;
    *ore = modulus * std::cos(phase);
    *oim = modulus * std::sin(phase);
}

// Standard deviation of the discrete-time-fit filter scaled by q (gaussiir.cpp:54-80).
double dtf_sigma(int nn, const Cx *pp, double q) {
    double var = 0.0;
    for (int m = 1; m <= nn - (nn & 1); m += 2) {
        double re, im;
        pole_power(pp[m].re, pp[m].im, q, &re, &im);
        double d = 1 - 2 * re + re * re + im * im;
        d *= d;
        var += (4 * (re + (re - 2) * (re * re + im * im))) / d;
    }
    if (nn & 1) {
        double re, im;
        pole_power(pp[nn].re, pp[nn].im, q, &re, &im);
        var += (2 * re) / ((re - 1) * (re - 1));
    }
    return std::sqrt(var);
}

// Adds every product of k of the poles p[1..hi] (indices strictly decreasing, the
// largest chosen first and multiplied first) to *acc, in the order of the
// reference's recursion (gaussiir.cpp:207-231).
void add_products(Cx *acc, Cx prod, int k, int hi, const Cx *p) {
    for (int i = k; i <= hi; ++i) {
        const Cx t = cmul(prod, p[i]);
        if (k == 1) {
            *acc = cadd(t, *acc);
        } else {
            add_products(acc, t, k - 1, i - 1, p);
        }
    }
}

}  // namespace

int design_gauss_iir(double sigma, int order, int filter_order, int method, double truncation, IirFilter *out) {
    IirFilter f;
    if (truncation <= 0.0) truncation = 3;
    const double max_trunc = std::sqrt(-2.0 * std::log(static_cast<double>(1.1920928955078125e-07)));
    truncation = std::min(truncation, max_trunc);
    const double bv = sigma * truncation + 0.5;
    const int64_t bi = bv >= 9.2e18 ? INT64_MAX : static_cast<int64_t>(bv);
    if (bi > (1 << 28)) return set_error(DIPAMD_E_UNSUPPORTED, "IIR border too large");
    f.border = static_cast<int>(bi < 5 ? 5 : bi);

    // poles (FillPoleCoefficients); `method` may switch to forward-backward
    const PoleSet *ps;
    if (method == DIPAMD_IIR_DISCRETE_TIME_FIT) {
        if (filter_order >= 1 && filter_order <= 5) {
            ps = &kDtf[filter_order][(order == 1 || order == 2) ? order : 0];
        } else {
            ps = &kFb3;
            method = DIPAMD_IIR_FORWARD_BACKWARD;
        }
    } else {
        method = DIPAMD_IIR_FORWARD_BACKWARD;
        ps = filter_order == 5 ? &kFb5 : (filter_order == 4 ? &kFb4 : &kFb3);
    }
    const int nn = ps->nn;
    Cx pp[6];
    for (int m = 1; m <= nn; ++m) pp[m] = {ps->p[m - 1][0], ps->p[m - 1][1]};

    // q: the scale that gives the requested sigma
    double q = 0;
    if (method == DIPAMD_IIR_FORWARD_BACKWARD) {
        const double q0term = -sigma * sigma;
        double q1term = 0.0, q2term = 0.0;
        for (int m = 1; m <= nn - (nn & 1); m += 2) {
            const double re = pp[m].re, im = pp[m].im;
            const double t1 = 4 * (-1 + 3 * re - 3 * re * re - im * im + re * (re * re + im * im));
            const double t2 = 4 * (1 - 2 * re + re * re - im * im);
            double d = 1 - 2 * re + re * re + im * im;
            d *= d;
            q1term += t1 / d;
            q2term += t2 / d;
        }
        if (nn & 1) {
            const double d = pp[nn].re - 1.0;
            q1term += 2.0 / d;
            q2term += 2.0 / (d * d);
        }
        const double t = std::sqrt(q1term * q1term - 4.0 * q2term * q0term);
        const double qa = (-q1term + t) / (2 * q2term);
        const double qb = (-q1term - t) / (2 * q2term);
        q = qa > qb ? qa : qb;
        for (int m = 1; m <= nn; ++m) pp[m] = {1.0 + (pp[m].re - 1.0) / q, pp[m].im / q};
    } else {
        // fixed-point iteration on sigma(q) (gaussiir.cpp:299-307)
        q = sigma / 2.0;
        const double q0 = q;
        double qs = dtf_sigma(nn, pp, q);
        for (int it = 0; std::abs(sigma - qs) > 0.000001 && it < 100; ++it) {
            q += q0 - qs / 2.0;
            qs = dtf_sigma(nn, pp, q);
        }
        for (int m = 1; m <= nn; ++m) {
            double re, im;
            pole_power(pp[m].re, pp[m].im, q, &re, &im);
            pp[m] = {re, im};
        }
    }

    // recursion coefficients: bb[nn - m] = sum of the products of m poles
    Cx bb[6];
    for (int m = 0; m <= nn; ++m) {
        if (m == 0) {
            bb[nn] = {1.0, 0.0};
        } else {
            bb[nn - m] = {0.0, 0.0};
            add_products(&bb[nn - m], {1.0, 0.0}, m, nn, pp);
        }
    }
    for (int m = 0; m <= nn; ++m) {
        const double s = std::pow(-1.0, static_cast<double>(nn - m)) / bb[0].re;
        bb[nn - m].re *= s;
        bb[nn - m].im *= s;
        f.b[nn - m] = bb[nn - m].re;
    }
    double cc = 0.0;  // b[0] (the scaled leading coefficient, 1 up to rounding) included
    for (int m = 0; m <= nn; ++m) cc += f.b[m];
    f.cc = cc * cc;
    f.nn = nn;

    // moving-average parts (gaussiir.cpp:351-387)
    switch (order) {
        case 0: f.a1[0] = 1.0; f.a2[0] = 1.0; break;
        case 1: f.a1[0] = 0.5; f.a1[2] = -0.5; f.a2[1] = 1.0; break;
        case 2: f.a1[0] = 1.0; f.a1[1] = -1.0; f.a2[1] = 1.0; f.a2[0] = -1.0; break;
        case 3: f.a1[0] = 1.0; f.a1[1] = -2.0; f.a1[2] = 1.0; f.a2[2] = 0.5; f.a2[0] = -0.5; break;
        case 4: f.a1[0] = 1.0; f.a1[1] = -2.0; f.a1[2] = 1.0; f.a2[2] = 1.0; f.a2[1] = -2.0; f.a2[0] = 1.0; break;
        default: return set_error(DIPAMD_E_ORDER, "Functionality has not (yet) been implemented");
    }
    f.m1 = f.m4 = 6;
    for (int j = 0; j < 6; ++j) {
        if (f.a1[j] != 0.0 && f.m1 > 5) f.m1 = j;
        if (f.a2[j] != 0.0 && f.m4 > 5) f.m4 = j;
    }
    f.m2 = f.m5 = 6;
    for (int j = 5; j >= 0; --j) {
        if (f.a1[j] != 0.0 && f.m2 > 5) f.m2 = j;
        if (f.a2[j] != 0.0 && f.m5 > 5) f.m5 = j;
    }
    *out = f;
    return DIPAMD_OK;
}

}  // namespace dipamd
