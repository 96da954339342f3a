// The Fourier-domain Gaussian (dip::GaussFT, src/linear/gauss.cpp:397-487): geometry of
// the padded transform buffer and the element-wise kernels around the FFTs (k_ft.hip).
#pragma once

#include "sep_common.hpp"

namespace dipamd {

constexpr int kFtMaxT = 3;  // dims transformed together (one rocFFT plan)

// The padded buffer holds the transformed dims innermost -- t = 0 (fastest) is the dim
// of the real-to-complex pass, then the others in increasing dim order -- and the other
// dims (none filtered) as a batch, outermost:
// index = ((b * P2 + i2) * P1 + i1) * P0 + i0.  Dims beyond nt have P = 1.
struct FtGeo {
    int32_t ndim;
    int32_t nt;                          // transformed dims
    int32_t tdim[kFtMaxT];               // their image dims
    int32_t nb;                          // batch dims
    int32_t bdim[DIPAMD_MAX_DIMS];       // their image dims
    int64_t size[DIPAMD_MAX_DIMS];       // image sizes
    int64_t sin[DIPAMD_MAX_DIMS];        // input strides (samples)
    int64_t sout[DIPAMD_MAX_DIMS];       // output strides
    int64_t P[kFtMaxT];                  // padded sizes
    int64_t off[kFtMaxT];                // image origin inside the padded line (centred window)
    int32_t bc[kFtMaxT];                 // boundary condition per transformed dim (index maps only)
    int32_t ext_order[kFtMaxT];          // transformed dims (t indices) by decreasing image dim
    int32_t expand;                      // 0: no padding (P = size, off = 0)
    double cval[kFtMaxT];                // constant boundary values
    int64_t nbatch;                      // product of the batch sizes
    int64_t nreal;                       // P0 P1 P2 nbatch
    int64_t nhalf;                       // (P0 / 2 + 1) P1 P2 nbatch
    int64_t lut_off[kFtMaxT];            // offset of each dim's LUT in the LUT array
};

// The centred-spectrum path (GaussFT with a frequency-domain input or output): a
// contiguous complex array of the image's sizes, dim 0 fastest.
struct FtCGeo {
    int32_t ndim;
    int32_t pad_;
    int64_t size[DIPAMD_MAX_DIMS];
    int64_t shift[DIPAMD_MAX_DIMS];     // cyclic shift of each dim (k_ft_cshift)
    int64_t lut_off[DIPAMD_MAX_DIMS];   // LUT of each dim in the LUT array, -1: none
    int64_t total;
};
// out[(i_d + shift_d) mod size_d] = in[i] for every dim (complex, contiguous)
hipError_t launch_ft_cshift(bool is_double, const void *in, void *out, const FtCGeo &g, hipStream_t st);
// data *= the product of each dim's LUT entry (in dim order), then times `scale`
hipError_t launch_ft_cmultiply(bool is_double, void *data, const void *luts, double scale, const FtCGeo &g,
                               hipStream_t st);

hipError_t launch_ft_expand(int in_dt, bool is_double, const void *in, void *pad, const FtGeo &g, hipStream_t st);
hipError_t launch_ft_multiply(bool is_double, void *spec, const void *luts, const FtGeo &g, hipStream_t st);
hipError_t launch_ft_crop(bool is_double, const void *pad, void *out, int out_dt, double scale, const FtGeo &g,
                          hipStream_t st);

}  // namespace dipamd
