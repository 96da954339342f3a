// Shared device/host definitions for the separable FIR kernels (gfx950).
//
// Sample loads convert to the buffer type exactly as dip::detail::CopyBuffer does
// (copy_buffer.cpp:44-121: a static_cast for these types; bin -> 0/1); the
// boundary extension of a line follows copy_buffer.cpp:515-629 in closed form:
// an out-of-range coordinate j maps to a source index plus a sign, or to a
// constant.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dipamd.h"

namespace dipamd {

constexpr int kMaxOuter = DIPAMD_MAX_DIMS;  // outer (non-processed) dims of a pass
constexpr int kLineCap = 256;               // taps per filter in the line / sweep kernels
constexpr int kFusedCap = 65;               // taps per axis in the fused 2-D kernel (radius <= 64)
constexpr int kBlock = 256;                 // threads per workgroup (4 wave64)

// SYM_NONE: no filter along that axis (the fused kernel's x stage copies rows into
// the ring, making it a y-only sweep).
enum Sym : int { SYM_GENERAL = 0, SYM_EVEN = 1, SYM_ODD = 2, SYM_D_EVEN = 4, SYM_D_ODD = 5, SYM_NONE = 7 };

struct bin_t {
    uint8_t v;
};

template <typename T>
__device__ __forceinline__ T dev_load(const T *p) {
    return *p;
}

template <typename TB, typename TI>
__device__ __forceinline__ TB to_buf(TI v) {
    return static_cast<TB>(v);
}
template <>
__device__ __forceinline__ float to_buf<float, bin_t>(bin_t v) {
    return v.v ? 1.0f : 0.0f;
}
template <>
__device__ __forceinline__ double to_buf<double, bin_t>(bin_t v) {
    return v.v ? 1.0 : 0.0;
}

// Boundary map of coordinate j (outside [0, n)) for a line of n samples.
// Returns the source index; *mul is +1 or -1; *is_const is set for the constant
// conditions (value in the pass arguments).
__host__ __device__ __forceinline__ int64_t bc_map(int64_t j, int64_t n, int bc, int *mul, bool *is_const) {
    *mul = 1;
    *is_const = false;
    switch (bc) {
        case DIPAMD_BC_SYMMETRIC_MIRROR:
        case DIPAMD_BC_ASYMMETRIC_MIRROR: {
            if (n == 1) {
                if (bc == DIPAMD_BC_ASYMMETRIC_MIRROR) *mul = -1;
                return 0;
            }
            int64_t P = 2 * (n - 1);
            int64_t p = j % P;
            if (p < 0) p += P;
            int64_t idx = p < n ? p : P - p;
            if (bc == DIPAMD_BC_ASYMMETRIC_MIRROR) {
                int64_t m = j < 0 ? -j : j - (n - 1);
                int64_t c = (m - 1) / (n - 1) + 1;  // edge reflections crossed
                *mul = (c & 1) ? -1 : 1;
            }
            return idx;
        }
        case DIPAMD_BC_PERIODIC:
        case DIPAMD_BC_ASYMMETRIC_PERIODIC: {
            int64_t q = j >= 0 ? j / n : -((-j - 1) / n) - 1;
            if (bc == DIPAMD_BC_ASYMMETRIC_PERIODIC && (q & 1)) *mul = -1;
            if (n == 1 && bc == DIPAMD_BC_ASYMMETRIC_PERIODIC) *mul = -1;
            return j - q * n;
        }
        case DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE:
            return j < 0 ? 0 : n - 1;
        default:
            *is_const = true;
            return 0;
    }
}

// Padded LDS index: one spare word per 32, so lanes reading with a stride of
// P (a power of two <= 32) words hit distinct banks.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int lpad_size(int n) { return n + (n >> 5) + 1; }

// Geometry of the dims that are neither the filter axis nor the lane axis.
struct Outer {
    int32_t n;
    int32_t pad_;
    int64_t size[kMaxOuter];
    int64_t sin[kMaxOuter];
    int64_t sout[kMaxOuter];
};

// Logical block index with each XCD owning a contiguous range: the hardware hands
// consecutive workgroups to the 8 XCDs in turn, and each XCD has its own L2, so
// blocks that share halo rows or columns (neighbours in the logical order) should
// run on the same XCD.  A bijection of [0, gridDim.x).
constexpr int kNumXcd = 8;
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t bid = blockIdx.x, nblk = gridDim.x;
    const int64_t per = nblk / kNumXcd, rem = nblk % kNumXcd;
    const int64_t x = bid % kNumXcd, i = bid / kNumXcd;
    return x * per + (x < rem ? x : rem) + i;
}

// Workgroup barrier ordering LDS accesses only.  __syncthreads() also waits for
// every outstanding global load (its fence covers all address spaces), which would
// drain the prefetched rows at each barrier; these kernels never read back global
// data written by another thread of the block, so LDS ordering is all they need.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Wave index within the block as a wave-uniform (scalar) value.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)); }

__device__ __forceinline__ void outer_offsets(const Outer &o, int64_t idx, int64_t *oin, int64_t *oout) {
    int64_t a = 0, b = 0;
    for (int d = 0; d < o.n; ++d) {
        int64_t c = idx % o.size[d];
        idx /= o.size[d];
        a += c * o.sin[d];
        b += c * o.sout[d];
    }
    *oin = a;
    *oout = b;
}

// One pass along one axis (line kernel and sweep kernel).
template <typename TB>
struct PassArgs {
    const void *in;
    void *out;
    int64_t n;        // samples along the filter axis
    int64_t sin;      // input stride along the filter axis
    int64_t sout;     // output stride along the filter axis
    int64_t e;        // sweep kernel: samples along the lane axis
    int64_t ein;      // lane-axis strides
    int64_t eout;
    int64_t nlines;   // line kernel: number of lines (product of the other sizes)
    int64_t chunk;    // sweep kernel: outputs along the filter axis per workgroup
    Outer outer;
    int32_t border;   // samples of boundary extension on each side
    int32_t ntaps;    // stored taps (data size)
    int32_t shift;    // GENERAL: first tap reads x[i - shift]
    int32_t bc;
    int32_t zlo;      // extended-line coordinates below zlo read as 0 (INT32_MIN: none)
    TB cval;          // constant boundary value
    TB taps[kLineCap];
};

// The fused x+y kernel of a 2-D plane (or each plane of a stack of them).
template <typename TB>
struct FusedArgs {
    const void *in;
    void *out;
    int64_t nx, ny;   // sizes along x (lane / x-filter axis) and y (sweep axis)
    int64_t sxin, syin, sxout, syout;
    int64_t chunk;    // output rows per workgroup
    Outer outer;
    int32_t rx, ry;
    int32_t bcx, bcy;
    TB cvalx, cvaly;
    TB tx[kFusedCap];
    TB ty[kFusedCap];
};

// Register-streaming fused kernel (stream2d.hpp): radii up to kStreamRMax.
constexpr int kStreamRMax = 4;
constexpr int kStreamRMaxHost = kStreamRMax;

struct StreamArgs {
    const uint8_t *in;
    float *out;
    int nx, ny;
    int64_t syin, syout;  // row strides in samples (x strides are 1)
    int chunk;            // output rows per wave
    int nstrips, nchunks;
    int rx, ry, bcx, bcy;
    // rows outside [0, ny) fold branch-free: z < 0 -> ylo_m*z + ylo_c, z >= ny ->
    // yhi_m*z + yhi_c, with flag yflag (0 as is, 1 negate, 2 constant) -- one reflection
    int ylo_m, ylo_c, yhi_m, yhi_c, yflag;
    float cvalx, cvaly;
    int aligned_out;      // output rows 16-byte aligned: vector stores
    int64_t nwaves;       // nstrips * nchunks * nplanes
    int64_t nplanes;      // product of the outer sizes
    float tx[kStreamRMax + 1];
    float ty[kStreamRMax + 1];
    Outer outer;          // stack of planes
};

// The derivative family's fused streaming form (stream2d_deriv.hpp): every Gaussian
// derivative of orders 0..2 per axis that the function needs, from one read of the
// input, combined in registers.  `s` carries the geometry, strips and folds (its out,
// rx, ry, tx, ty are unused); filters are indexed by derivative order.
enum : int32_t {
    kDerivGradient = 0,
    kDerivGradientMagnitude = 1,
    kDerivLaplace = 2,
    kDerivHessian = 3,
    kDerivDgg = 4,
    kDerivLaplacePlusDgg = 5,
    kDerivLaplaceMinusDgg = 6,
};
struct DerivStreamArgs {
    StreamArgs s;
    float *out[3];
    int64_t syout[3];
    int32_t rx[3], ry[3];
    float tx[3][kStreamRMax + 1];
    float ty[3][kStreamRMax + 1];
};

// Batched form: per-image parameters, work items and kernel arguments.
template <typename TB>
struct BatchImage {
    const void *in;
    TB *out;
    int64_t nx, ny, syin, syout;
    int32_t rx, ry;
    int32_t tap_offset;  // into the taps table: rx+1 x taps then ry+1 y taps
    int32_t vec_ok;      // input rows allow the radius-class kernel's vector loads
};

struct BatchItem {
    int32_t image;
    int32_t tile;
    int32_t z0;
    int32_t z1;
};

template <typename TB>
struct BatchArgs {
    const BatchImage<TB> *images;
    const BatchItem *items;
    const TB *taps;
    int32_t *counter;
    int32_t nitems;
    int32_t bcx, bcy;
    TB cvalx, cvaly;
};

// Geometry of an element-wise conversion pass.
struct CastGeo {
    int32_t ndim;
    int32_t pad_;
    int64_t size[DIPAMD_MAX_DIMS];
    int64_t sin[DIPAMD_MAX_DIMS];
    int64_t sout[DIPAMD_MAX_DIMS];
    int64_t total;
};

// Fused backward pass over stacks of 2-D planes (k_bwd2d.hip): x, dL/dy and dL/dx are
// float with unit x stride; `outer` holds the plane sizes with the x (sin) and dL/dx
// (sout) strides, gplane the dL/dy strides.  Taps are half filters, centre first.
struct BwdArgs {
    const float *x;
    const float *g;
    float *dx;
    double *partials;  // two per block: <h, Ax' x> (sigma x), <h', Ax x> (sigma y)
    int nx, ny;
    int64_t syx, syg, sydx;  // row strides
    int rx, ry;
    int bcx, bcy;
    int modex, modey;  // kBwdZero, kBwdPeriodic, kBwdFold
    float cvalx;       // the forward x extension's constant
    int chunk, nchunks, ntiles;
    int ytop, ybot;  // > 0: chunk 0 is rows [0, ytop), the last [ybot, ny), the others split the rest
    int want_dx;
    int xbytes, gbytes, dxbytes;  // byte extent of one plane of x, g, dx (each < 2^31)
    Outer outer;
    int64_t gplane[kMaxOuter];
    float tx[kFusedCap], tdx[kFusedCap];    // forward x filter and its sigma derivative
    float txa[kFusedCap];                   // adjoint x filter (odd: negated)
    float tya[kFusedCap], tdya[kFusedCap];  // adjoint y filter and adjoint y derivative
};
// How an adjoint pass treats positions outside the line: dL/dy zero-padded (constant
// conditions), dL/dy periodically extended (periodic conditions: self-adjoint with
// reversed taps), or zero-padded plus the halo outputs folded back onto the pixels the
// forward extension copied them from (mirror conditions, zero order).
constexpr int kBwdZero = 0, kBwdPeriodic = 1, kBwdFold = 2;
constexpr int kBwdStep = 16;   // rows per step
constexpr int kBwdSumAdd = 0, kBwdSumWrite = 1, kBwdSumTotal = 2;  // how dL/dsigma reaches its output

// Arguments of the streaming backward kernel (k_bwd2s.hpp).
constexpr int kB2sCap = 16;
constexpr int kB2sFold = 32;  // rows of the y extension table on each side  // taps per filter (radius <= 15)

struct Bwd2sArgs {
    const float *x;
    const float *g;
    float *dx;
    double *partials;  // two per workgroup: dL/dsigma_x, dL/dsigma_y
    int nx, ny;
    int64_t syx, syg, sydx;  // row strides (samples)
    int rx, ry;
    int bcx, bcy;
    int modex, modey;  // kBwdZero, kBwdPeriodic, kBwdFold
    int nstrips, nchunks, chunk;  // rows per chunk; the last chunk runs to ny
    int nchunks_e, chunk_e;       // the same for the edge strips
    int64_t nplanes, nwaves;
    int want_dx;
    int xbytes, gbytes, dxbytes;  // byte extent of one plane of x, g, dx (each < 2^31)
    Outer outer;                  // sin: x, sout: dx
    int64_t gplane[kMaxOuter];
    float txa[kB2sCap], tdxa[kB2sCap];  // adjoint x filter and adjoint x derivative
    float hzv[kB2sCap], hzw[kB2sCap];   // zero-order x fold: sums of txa / tdxa over m+1..rx
    int yfold[2 * kB2sFold];            // y extension of rows -1-i and ny+i: source << 1 | negate, -1 zero
    float tya[kB2sCap], tdya[kB2sCap];  // adjoint y filter and adjoint y derivative
};


}  // namespace dipamd
