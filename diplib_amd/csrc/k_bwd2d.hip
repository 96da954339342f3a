// Instances of the fused 2-D backward kernel (k_bwd2d.hpp) with equal x and y radius
// classes, the fixed-order sum of its dL/dsigma partials, and the launch dispatch.
#include "k_bwd2d.hpp"

namespace dipamd {

// Sum of the per-block partials in a fixed order (sx: even entries, sy: odd), then by
// mode: kBwdSumAdd out[0] += sx, out[1] += sy; kBwdSumWrite out[0] = sx, out[1] = sy and
// out[2 .. nout-1] = 0; kBwdSumTotal out[0] = sx + sy.
__global__ __launch_bounds__(kBlock) void k_bwd2d_sum(const double *partials, int64_t nblocks, double *out, int mode,
                                                      int nout) {
    __shared__ double red[2][kBlock];
    double sx = 0.0, sy = 0.0;
#pragma unroll 8
    for (int64_t b = threadIdx.x; b < nblocks; b += kBlock) {
        sx += partials[2 * b];
        sy += partials[2 * b + 1];
    }
    red[0][threadIdx.x] = sx;
    red[1][threadIdx.x] = sy;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode == kBwdSumAdd) {
            out[0] += red[0][0];
            out[1] += red[1][0];
        } else if (mode == kBwdSumWrite) {
            out[0] = red[0][0];
            out[1] = red[1][0];
        } else {
            out[0] = red[0][0] + red[1][0];
        }
    }
    if (mode == kBwdSumWrite)
        for (int i = 2 + static_cast<int>(threadIdx.x); i < nout; i += kBlock) out[i] = 0.0;
}

// The fixed-order sum of per-block partials (for the streaming backward, k_bwd2s).
hipError_t launch_bwd2d_sum(const double *partials, int64_t nblocks, double *gs, int gs_mode, int gs_n,
                            hipStream_t stream) {
    hipLaunchKernelGGL(k_bwd2d_sum, dim3(1), dim3(kBlock), 0, stream, partials, nblocks, gs, gs_mode, gs_n);
    return hipGetLastError();
}

namespace {

constexpr int kBwdClasses[] = {3, 5, 7, 9, 11, 13, 16};  // 20 would exceed 64 KB of LDS (67200 B)
constexpr int kNumBwdClasses = sizeof(kBwdClasses) / sizeof(kBwdClasses[0]);

int class_index(int r) {
    for (int i = 0; i < kNumBwdClasses; ++i)
        if (r <= kBwdClasses[i]) return i;
    return -1;
}

}  // namespace

// Radius classes (compile-time windows) for radii rx, ry, as key 32 cx + cy: each the
// smallest class >= its radius when the two are the same or adjacent classes (a first
// derivative along one axis: radius ceil(3.5 sigma) against ceil(3 sigma)), else the
// larger for both; 0 when a radius exceeds every class.
int bwd2d_class(int rx, int ry) {
    const int ix = class_index(rx), iy = class_index(ry);
    if (ix < 0 || iy < 0) return 0;
    int cx = kBwdClasses[ix], cy = kBwdClasses[iy];
    if (ix - iy > 1 || iy - ix > 1) cx = cy = cx > cy ? cx : cy;
    return 32 * cx + cy;
}

int bwd2d_tile_columns(int key) { return 128 - 2 * (key / 32); }

size_t bwd2d_lds_bytes(int key) {
    const int cx = key / 32, cy = key % 32;
    if (cx != cy) return bwd2d_lds_xy(cx, cy);
    switch (cx) {
        case 3: return bwd2d_lds_of<3, 3>();
        case 5: return bwd2d_lds_of<5, 5>();
        case 7: return bwd2d_lds_of<7, 7>();
        case 9: return bwd2d_lds_of<9, 9>();
        case 11: return bwd2d_lds_of<11, 11>();
        case 13: return bwd2d_lds_of<13, 13>();
        case 16: return bwd2d_lds_of<16, 16>();
        default: return 0;
    }
}

hipError_t launch_bwd2d(int symx, int symy, int key, const BwdArgs &a, int64_t nblocks, double *gs, int gs_mode,
                        int gs_n, hipStream_t stream) {
    const int cx = key / 32, cy = key % 32;
    bool ok = false;
    if (cx != cy) {
        ok = bwd2d_launch_xy(cx, cy, symx, symy, a, nblocks, stream);
    } else {
        switch (cx) {
            case 3: ok = bwd2d_launch_pair<3, 3>(symx, symy, a, nblocks, stream); break;
            case 5: ok = bwd2d_launch_pair<5, 5>(symx, symy, a, nblocks, stream); break;
            case 7: ok = bwd2d_launch_pair<7, 7>(symx, symy, a, nblocks, stream); break;
            case 9: ok = bwd2d_launch_pair<9, 9>(symx, symy, a, nblocks, stream); break;
            case 11: ok = bwd2d_launch_pair<11, 11>(symx, symy, a, nblocks, stream); break;
            case 13: ok = bwd2d_launch_pair<13, 13>(symx, symy, a, nblocks, stream); break;
            case 16: ok = bwd2d_launch_pair<16, 16>(symx, symy, a, nblocks, stream); break;
            default: ok = false;
        }
    }
    if (!ok) return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (!gs) return hipSuccess;
    hipLaunchKernelGGL(k_bwd2d_sum, dim3(1), dim3(kBlock), 0, stream, a.partials, nblocks, gs, gs_mode, gs_n);
    return hipGetLastError();
}

}  // namespace dipamd
