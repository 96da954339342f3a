// K3S: register-streaming fused x+y pass for short filters (radius <= 4) on uint8
// planes -- the HBM-bound regime.
//
// A wave owns a strip of 256 columns: lane l holds columns 4l..4l+3 of every row as
// one 32-bit word and also loads the two neighbouring words (its x halo; those
// loads hit L1).  It converts the 12 samples, computes its 4 x-filtered samples in
// registers in the reference's order (convolution.cpp:214-241), and pushes them
// into a register window of W = 2*RMAX+1 rows.  Output row o is produced when row
// o+RMAX has arrived, from window slots whose indices are compile-time constants
// (the row loop is unrolled by W); the actual radii (<= RMAX) are wave-uniform
// guards on the tap loops.  Input words for the next W rows are prefetched one
// period ahead: the loads are unconditional and call-free so they stay in flight.
// No LDS, no barriers; one 16-byte store per lane and output row.
//
// Strips that touch the left/right image edge run the same loop with every sample
// loaded through the boundary map; rows outside the image are folded per row.
// Both folds assume one reflection or wrap reaches the image (the host only picks
// this kernel for planes of at least 16 x 16).
#pragma once

#include "sep_kernels_impl.hpp"

namespace dipamd {

// Coordinate x outside [0, n) folded once: index, and *flag = 0 (as is), 1 (negate)
// or 2 (constant condition).  Valid for overshoots below n - 1.
__device__ __forceinline__ int fold_once(int x, int n, int bc, int *flag) {
    *flag = 0;
    if (x >= 0 && x < n) return x;
    switch (bc) {
        case DIPAMD_BC_ASYMMETRIC_MIRROR:
            *flag = 1;
            [[fallthrough]];
        case DIPAMD_BC_SYMMETRIC_MIRROR:
            return x < 0 ? -x : 2 * (n - 1) - x;
        case DIPAMD_BC_ASYMMETRIC_PERIODIC:
            *flag = 1;
            [[fallthrough]];
        case DIPAMD_BC_PERIODIC:
            return x < 0 ? x + n : x - n;
        case DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE:
            return x < 0 ? 0 : n - 1;
        default:
            *flag = 2;
            return 0;
    }
}

template <int SX, int SY, bool FAST, bool EDGE>
__device__ __forceinline__ void stream_strip(const StreamArgs &a, const uint8_t *__restrict__ src,
                                             float *__restrict__ dst, int c, int z0, int zend) {
    // (__restrict__: input and output never overlap, so loads need not wait for stores)
    constexpr int RMAX = kStreamRMax;
    constexpr int W = 2 * RMAX + 1;  // window rows = unroll period
    constexpr int NS = 12;           // samples per lane per row: 4 own + 4 each side
    const int nx = a.nx, ny = a.ny, rx = a.rx, ry = a.ry;
    const int E = (zend - z0) + 2 * RMAX;  // rows streamed: z0-RMAX .. zend+RMAX-1

    // edge strips: folded column of each sample, and negate / constant bits
    int cofs[EDGE ? NS : 1];
    uint32_t negbits = 0, constbits = 0;
    if constexpr (EDGE) {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            int flag;
            cofs[i] = fold_once(c - 4 + i, nx, a.bcx, &flag);
            if (c - 4 + i >= nx + 4 || c - 4 + i < -4) cofs[i] = 0;  // lanes beyond the last strip's end
            negbits |= (flag == 1 ? 1u : 0u) << i;
            constbits |= (flag == 2 ? 1u : 0u) << i;
        }
    }

    float win[W][4];
    uint32_t pre[W][3];
    int rflag[W];  // row fold flag of the prefetched row
    auto fetch = [&](int e, uint32_t (&wd)[3], int *flag) {
        const int eclamp = e < E ? e : E - 1;  // past the end: reload a valid row, never used
        const int j = fold_once(z0 - RMAX + eclamp, ny, a.bcy, flag);
        const uint8_t *rp = src + static_cast<int64_t>(j) * a.syin;
        if constexpr (!EDGE) {
            const uint32_t *wp = reinterpret_cast<const uint32_t *>(rp + c);
            wd[0] = wp[-1];
            wd[1] = wp[0];
            wd[2] = wp[1];
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                uint32_t v = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) v |= static_cast<uint32_t>(rp[cofs[4 * k + b]]) << (8 * b);
                wd[k] = v;
            }
        }
    };
#pragma unroll
    for (int u = 0; u < W; ++u) fetch(u, pre[u], &rflag[u]);

    const bool cok = c < nx;
    for (int eb = 0; eb < E; eb += W) {
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const int e = eb + u;
            if (e >= E) break;  // wave-uniform
            // --- x stage: 4 outputs of row e from 12 samples ---
            float s[NS];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const uint32_t v = pre[u][k];
                // byte -> float (exact); the compiler emits v_cvt_f32_ubyte0..3
                s[4 * k + 0] = static_cast<float>(v & 0xffu);
                s[4 * k + 1] = static_cast<float>((v >> 8) & 0xffu);
                s[4 * k + 2] = static_cast<float>((v >> 16) & 0xffu);
                s[4 * k + 3] = static_cast<float>(v >> 24);
            }
            if constexpr (EDGE) {
#pragma unroll
                for (int i = 0; i < NS; ++i) {
                    s[i] = ((negbits >> i) & 1u) ? -s[i] : s[i];
                    s[i] = ((constbits >> i) & 1u) ? a.cvalx : s[i];
                }
            }
            const int rf = rflag[u];
            fetch(e + W, pre[u], &rflag[u]);  // prefetch one period ahead into this slot
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                float acc = a.tx[0] * s[4 + p];
#pragma unroll
                for (int k = 1; k <= RMAX; ++k) {
                    if (k <= rx) {
                        const float t = (SX == SYM_EVEN) ? (s[4 + p + k] + s[4 + p - k]) : (s[4 + p + k] - s[4 + p - k]);
                        acc = mac<float, FAST>(acc, a.tx[k], t);
                    }
                }
                win[u][p] = rf == 2 ? a.cvaly : (rf == 1 ? -acc : acc);
            }
            // --- y stage: output row z0 + e - 2*RMAX, centre slot (u - RMAX) mod W ---
            if (e >= 2 * RMAX) {
                const int cs = ((u - RMAX) % W + W) % W;
                float out[4];
#pragma unroll
                for (int p = 0; p < 4; ++p) {
                    float acc = a.ty[0] * win[cs][p];
#pragma unroll
                    for (int k = 1; k <= RMAX; ++k) {
                        if (k <= ry) {
                            const float hi = win[(cs + k) % W][p];
                            const float lo = win[(cs - k + W) % W][p];
                            const float t = (SY == SYM_EVEN) ? (hi + lo) : (hi - lo);
                            acc = mac<float, FAST>(acc, a.ty[k], t);
                        }
                    }
                    out[p] = acc;
                }
                const int o = z0 + e - 2 * RMAX;
                float *d = dst + static_cast<int64_t>(o) * a.syout + c;
                if (a.aligned_out && c + 4 <= nx) {
                    *reinterpret_cast<float4 *>(d) = make_float4(out[0], out[1], out[2], out[3]);
                } else if (cok) {
#pragma unroll
                    for (int p = 0; p < 4; ++p)
                        if (c + p < nx) d[p] = out[p];
                }
            }
        }
    }
}

// EDGE = false: the interior strips 1 .. nstrips-2 (strip index = 1 + wave's strip
// slot); EDGE = true: the first and the last strip (slots 0 and 1).  A strip is an
// edge strip when some lane's halo leaves [0, nx).
template <int SX, int SY, bool FAST, bool EDGE>
__global__ __launch_bounds__(kBlock) void k_stream2d(const StreamArgs a) {
    const int lane = threadIdx.x & 63;
    const int64_t gw = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int nslots = EDGE ? (a.nstrips > 1 ? 2 : 1) : a.nstrips - 2;
    if (nslots <= 0) return;
    const int64_t nw = static_cast<int64_t>(nslots) * a.nchunks *
                       (a.nwaves / (static_cast<int64_t>(a.nstrips) * a.nchunks));
    if (gw >= nw) return;
    const int slot = static_cast<int>(gw % nslots);
    const int strip = EDGE ? (slot == 0 ? 0 : a.nstrips - 1) : 1 + slot;
    const int64_t rest = gw / nslots;
    const int chunk = static_cast<int>(rest % a.nchunks);
    const int64_t outer = rest / a.nchunks;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    const int c = strip * 256 + 4 * lane;
    const int z0 = chunk * a.chunk;
    const int zend = min(z0 + a.chunk, a.ny);
    stream_strip<SX, SY, FAST, EDGE>(a, a.in + oin, a.out + oout, c, z0, zend);
}

}  // namespace dipamd
