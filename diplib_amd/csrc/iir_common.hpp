// The recursive (IIR) Gaussian of src/linear/gaussiir.cpp: filter parameters as the
// host designs them, and the arguments of the line kernel (k_iir.hip).
#pragma once

#include "sep_common.hpp"

namespace dipamd {

// One axis of dip::GaussIIR (FillGaussIIRParams, gaussiir.cpp:224-425): a causal
// recursion of nn poles followed by an anti-causal one, both in double, each with a
// short moving-average part (the derivative order's difference filter).
struct IirFilter {
    int border = 0;      // max(5, (int)(sigma * truncation + 0.5)) samples each side
    int nn = 0;          // poles per direction (1..5)
    double b[6] = {};    // b[1..nn]: feedback coefficients (b[0] = 1 is implied)
    double a1[6] = {};   // forward moving average, a1[j] multiplies x[i - j]
    double a2[6] = {};   // backward moving average, a2[j] multiplies y[i + j]
    double cc = 0;       // gain, applied in the backward recursion
    int m1 = 0, m2 = 0;  // first and last non-zero a1 index
    int m4 = 0, m5 = 0;  // first and last non-zero a2 index
};

// Designs the filter of one axis; returns 0, or a DIPAMD_E_* code for an order the
// reference does not implement.
int design_gauss_iir(double sigma, int order, int filter_order, int method, double truncation, IirFilter *out);

// Kernel arguments of one pass: one thread per line; the causal result of each line
// goes to a double scratch laid out [position pair][line][2], so that both recursions
// read and write it coalesced, 16 bytes per access, whatever the filter axis.
struct IirArgs {
    const void *in;
    void *out;
    double *p1;
    int64_t n;            // samples along the filter axis
    int64_t L;            // n + 2 * border
    int64_t sin, sout;    // strides along the filter axis
    int64_t e, ein, eout; // lane axis: size and strides
    int64_t nlines;       // e * product of the outer sizes
    Outer outer;
    int32_t border;
    int32_t bc;
    double cval;          // constant boundary value
    double b[6];          // b[1..nn]
    double a1[3];         // a1[0..2]
    double a2[3];         // a2[0..2]
    double S;             // 1 + b[1] + ... + b[5], summed in that order
    double c;             // gain
    int32_t m1, m2, m4, m5;
    int32_t fnset;        // forward: positions 0..fnset-1 take the initial value
    int32_t finit_diff;   // forward initial value: 0 MA sum / S, 1 (x[1] - x[0]) / S, 2 x[0] / S
    int32_t bnset;        // backward: positions L-bnset..L-1 take the initial value
    int32_t binit;        // backward initial value: 0 c * (MA sum / S), 1 (c * y[L-1]) / S,
                          // 2 (c * (-y[L-2] + y[L-1])) / S
    int32_t out_double;   // destination holds doubles (else floats)
    int32_t pad_;
};

// Launch one pass (k_iir_<type>.hip; the _vec forms need a contiguous filter axis,
// stride 1 in input and output, with every line aligned to the vector widths).
hipError_t launch_iir_u8(int nn, int order, const IirArgs &a, hipStream_t st);
hipError_t launch_iir_f32(int nn, int order, const IirArgs &a, hipStream_t st);
hipError_t launch_iir_f64(int nn, int order, const IirArgs &a, hipStream_t st);
hipError_t launch_iir_u8_vec(int nn, int order, const IirArgs &a, hipStream_t st);
hipError_t launch_iir_f32_vec(int nn, int order, const IirArgs &a, hipStream_t st);
hipError_t launch_iir_f64_vec(int nn, int order, const IirArgs &a, hipStream_t st);

}  // namespace dipamd
