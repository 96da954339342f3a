// K3S: register-streaming fused x+y pass for short filters (radius <= 4) on uint8
// planes -- the HBM-bound regime.
//
// A wave owns a strip of 256 columns: lane l holds columns 4l..4l+3 of every row as
// one 32-bit word and also loads the two neighbouring words (its x halo; those
// loads hit L1).  It converts the 12 samples, computes its 4 x-filtered samples in
// registers in the reference's order (convolution.cpp:214-241), and pushes them
// into a register window of W = 2*RY+1 rows.  Output row o is produced when row
// o+RY has arrived, from window slots whose indices are compile-time constants
// (the row loop is unrolled by W, and the radii RX, RY are template parameters so
// every tap loop unrolls completely).  Input words for the next W rows are
// prefetched one period ahead: the loads are unconditional and call-free so they
// stay in flight, and the main loop runs whole periods only (no exit inside the
// unrolled body, so no register rotation that would wait for the loads).  No LDS,
// no barriers; one 16-byte store per lane and output row.
//
// Rows are a multiple of 4 samples and 16-byte aligned in the output (the host
// checks).  The boundary extension along x concerns two lanes only (columns 0..3
// and nx-4..nx-1, radius <= 4): each lane holds a load offset and a byte selector
// per halo word, and the edge lanes load another word of the row and rearrange its
// bytes with v_perm_b32 into the folded halo (mirror: reversed bytes, zero order: a
// repeated byte, periodic: the word at the other end); the sign flips of the
// antisymmetric conditions and the constant conditions are per-sample masks
// applied in the first and last strip only.  Rows outside the image fold with
// wave-uniform arithmetic.  Both folds assume one reflection or wrap reaches the
// image: the host only picks this kernel for planes of at least 16 x 16.
#pragma once

#include "sep_kernels_impl.hpp"

namespace dipamd {

// Per-lane halo layout for a lane at column c (a multiple of 4) of a row of nx
// samples: byte offsets of the left, own and right words, v_perm_b32 selectors
// (source 0 = the own word, source 1 = the loaded halo word; 0x03020100 = the halo
// word as loaded), and per-sample negate / constant masks (bit i = sample c-4+i).
struct StreamLane {
    uint32_t ol, oc, orr, sell, selr, negbits, constbits;
};

__device__ __forceinline__ StreamLane stream_lane(int c, int nx, int bc) {
    StreamLane L;
    const uint32_t ident = 0x03020100u;
    if (c >= nx) {  // beyond the row: loads stay in it, nothing is stored
        L = {0u, 0u, 0u, ident, ident, 0u, 0u};
        return L;
    }
    L.ol = c >= 4 ? static_cast<uint32_t>(c - 4) : 0u;
    L.oc = static_cast<uint32_t>(c);
    L.orr = c + 4 < nx ? static_cast<uint32_t>(c + 4) : 0u;
    L.sell = ident;
    L.selr = ident;
    L.negbits = 0u;
    L.constbits = 0u;
    const bool asym = bc == DIPAMD_BC_ASYMMETRIC_MIRROR || bc == DIPAMD_BC_ASYMMETRIC_PERIODIC;
    if (c == 0) {  // samples -4..-1
        switch (bc) {
            case DIPAMD_BC_SYMMETRIC_MIRROR:
            case DIPAMD_BC_ASYMMETRIC_MIRROR: L.ol = 4u; L.sell = 0x05060700u; break;  // x4 x3 x2 x1
            case DIPAMD_BC_PERIODIC:
            case DIPAMD_BC_ASYMMETRIC_PERIODIC: L.ol = static_cast<uint32_t>(nx - 4); break;
            case DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE: L.sell = 0x04040404u; break;  // x0 x0 x0 x0
            default: L.constbits |= 0x00fu; break;
        }
        if (asym) L.negbits |= 0x00fu;
    }
    if (c + 4 == nx) {  // samples nx..nx+3
        switch (bc) {
            case DIPAMD_BC_SYMMETRIC_MIRROR:
            case DIPAMD_BC_ASYMMETRIC_MIRROR:
                L.orr = static_cast<uint32_t>(nx - 8);
                L.selr = 0x03040506u;  // x[nx-2] x[nx-3] x[nx-4] x[nx-5]
                break;
            case DIPAMD_BC_PERIODIC:
            case DIPAMD_BC_ASYMMETRIC_PERIODIC: L.orr = 0u; break;
            case DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE: L.selr = 0x07070707u; break;  // x[nx-1] x4
            default: L.constbits |= 0xf00u; break;
        }
        if (asym) L.negbits |= 0xf00u;
    }
    return L;
}

template <int SX, int SY, int RX, int RY, bool FAST, bool XMASK>
__device__ __forceinline__ void stream_strip(const StreamArgs &a, const uint8_t *__restrict__ src,
                                             float *__restrict__ dst, int c, int z0, int zend) {
    // (__restrict__: input and output never overlap, so loads need not wait for stores)
    constexpr int W = 2 * RY + 1;  // window rows = unroll period
    constexpr int NS = 12;         // samples per lane per row: 4 own + 4 each side
    const int nx = a.nx, ny = a.ny;
    const int E = (zend - z0) + 2 * RY;  // rows streamed: z0-RY .. zend+RY-1
    const int ylo_m = a.ylo_m, ylo_c = a.ylo_c, yhi_m = a.yhi_m, yhi_c = a.yhi_c, yflag = a.yflag;
    const int64_t syin = a.syin;
    const StreamLane L = stream_lane(c, nx, a.bcx);

    float win[W][4];
    uint32_t pre[W][3];
    int rflag[W];  // row fold flag of the prefetched row: 0 as is, 1 negate, 2 constant
    auto fetch = [&](int e, uint32_t(&wd)[3], int *flag) {
        const int z = z0 - RY + (e < E ? e : E - 1);  // past the end: reload a valid row, never used
        const bool lo = z < 0, hi = z >= ny;
        const int m = lo ? ylo_m : (hi ? yhi_m : 1);
        const int cc = lo ? ylo_c : (hi ? yhi_c : 0);
        *flag = (lo || hi) ? yflag : 0;
        const uint8_t *rp = src + static_cast<int64_t>(m * z + cc) * syin;
        wd[0] = *reinterpret_cast<const uint32_t *>(rp + L.ol);
        wd[1] = *reinterpret_cast<const uint32_t *>(rp + L.oc);
        wd[2] = *reinterpret_cast<const uint32_t *>(rp + L.orr);
    };
    const bool cok = c < nx;
    // one row: x stage into window slot u, then (once 2*RY rows are in) one output row
    auto row = [&](const int e, const int u) {
        const uint32_t wl = __builtin_amdgcn_perm(pre[u][1], pre[u][0], L.sell);
        const uint32_t wc = pre[u][1];
        const uint32_t wr = __builtin_amdgcn_perm(pre[u][1], pre[u][2], L.selr);
        const int rf = rflag[u];
        fetch(e + W, pre[u], &rflag[u]);  // prefetch one period ahead into this slot
        float s[NS];
        const uint32_t wv[3] = {wl, wc, wr};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            // byte -> float (exact)
            s[4 * k + 0] = static_cast<float>(wv[k] & 0xffu);
            s[4 * k + 1] = static_cast<float>((wv[k] >> 8) & 0xffu);
            s[4 * k + 2] = static_cast<float>((wv[k] >> 16) & 0xffu);
            s[4 * k + 3] = static_cast<float>(wv[k] >> 24);
        }
        if constexpr (XMASK) {  // sign flips / constants along x (first and last strip)
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                s[i] = ((L.negbits >> i) & 1u) ? -s[i] : s[i];
                s[i] = ((L.constbits >> i) & 1u) ? a.cvalx : s[i];
            }
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            float acc = a.tx[0] * s[4 + p];
#pragma unroll
            for (int k = 1; k <= RX; ++k) {
                const float t = (SX == SYM_EVEN) ? (s[4 + p + k] + s[4 + p - k]) : (s[4 + p + k] - s[4 + p - k]);
                acc = mac<float, FAST>(acc, a.tx[k], t);
            }
            win[u][p] = acc;
        }
        if (rf != 0) {  // a row folded across the top or bottom edge (wave-uniform)
#pragma unroll
            for (int p = 0; p < 4; ++p) win[u][p] = rf == 2 ? a.cvaly : -win[u][p];
        }
        // y stage: output row z0 + e - 2*RY, centre slot (u - RY) mod W
        if (e >= 2 * RY) {
            const int cs = ((u - RY) % W + W) % W;
            float out[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                float acc = a.ty[0] * win[cs][p];
#pragma unroll
                for (int k = 1; k <= RY; ++k) {
                    const float hi = win[(cs + k) % W][p];
                    const float lo = win[(cs - k + W) % W][p];
                    const float t = (SY == SYM_EVEN) ? (hi + lo) : (hi - lo);
                    acc = mac<float, FAST>(acc, a.ty[k], t);
                }
                out[p] = acc;
            }
            const int o = z0 + e - 2 * RY;
            float *d = dst + static_cast<int64_t>(o) * a.syout + c;
            if (cok) *reinterpret_cast<float4 *>(d) = make_float4(out[0], out[1], out[2], out[3]);
        }
    };
#pragma unroll
    for (int u = 0; u < W; ++u) fetch(u, pre[u], &rflag[u]);
    int eb = 0;
    for (; eb + W <= E; eb += W) {  // whole periods
#pragma unroll
        for (int u = 0; u < W; ++u) row(eb + u, u);
    }
#pragma unroll
    for (int u = 0; u < W; ++u)  // the last partial period
        if (eb + u < E) row(eb + u, u);
}

// Wave gw: strip gw % nstrips (adjacent waves take adjacent strips of the same rows),
// then chunk of rows, then plane.
template <int SX, int SY, int RX, int RY, bool FAST>
__global__ __launch_bounds__(kBlock) void k_stream2d(const StreamArgs a) {
    const int lane = threadIdx.x & 63;
    // the wave index through readfirstlane: everything derived from it (rows, row
    // folds, row pointers, loop bounds) is then known wave-uniform and lives in SGPRs
    const int64_t gw = xcd_block() * 4 + wave_id();
    if (gw >= a.nwaves) return;
    const int strip = static_cast<int>(gw % a.nstrips);
    const int64_t rest = gw / a.nstrips;
    const int chunk = static_cast<int>(rest % a.nchunks);
    const int64_t outer = rest / a.nchunks;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    const int c = strip * 256 + 4 * lane;
    const int z0 = chunk * a.chunk;
    const int zend = min(z0 + a.chunk, a.ny);
    // the first and last strip under an antisymmetric or constant condition along x
    // take the copy of the loop with per-sample masks (the branch is wave-uniform and
    // outside the loop)
    const bool xmask = (strip == 0 || strip == a.nstrips - 1) &&
                       (a.bcx == DIPAMD_BC_ASYMMETRIC_MIRROR || a.bcx == DIPAMD_BC_ASYMMETRIC_PERIODIC ||
                        a.bcx == DIPAMD_BC_ADD_ZEROS || a.bcx == DIPAMD_BC_ADD_MAX_VALUE ||
                        a.bcx == DIPAMD_BC_ADD_MIN_VALUE);
    if (xmask)
        stream_strip<SX, SY, RX, RY, FAST, true>(a, a.in + oin, a.out + oout, c, z0, zend);
    else
        stream_strip<SX, SY, RX, RY, FAST, false>(a, a.in + oin, a.out + oout, c, z0, zend);
}

// Launch of one (x symmetry, y symmetry) pair: the streaming kernel for the radii
// (1..4 each).  Instantiated once per
// pair in its own translation unit (k_stream2d_<sx><sy>.hip).
template <int SX, int SY, int RX, int RY, bool FAST>
void stream2d_main(const StreamArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((k_stream2d<SX, SY, RX, RY, FAST>), dim3(static_cast<unsigned>((a.nwaves + 3) / 4)),
                       dim3(kBlock), 0, st, a);
}

template <int SX, int SY, bool FAST>
hipError_t stream2d_fast(const StreamArgs &a, hipStream_t st) {
    using Fn = void (*)(const StreamArgs &, hipStream_t);
    static const Fn table[4][4] = {
        {stream2d_main<SX, SY, 1, 1, FAST>, stream2d_main<SX, SY, 1, 2, FAST>, stream2d_main<SX, SY, 1, 3, FAST>,
         stream2d_main<SX, SY, 1, 4, FAST>},
        {stream2d_main<SX, SY, 2, 1, FAST>, stream2d_main<SX, SY, 2, 2, FAST>, stream2d_main<SX, SY, 2, 3, FAST>,
         stream2d_main<SX, SY, 2, 4, FAST>},
        {stream2d_main<SX, SY, 3, 1, FAST>, stream2d_main<SX, SY, 3, 2, FAST>, stream2d_main<SX, SY, 3, 3, FAST>,
         stream2d_main<SX, SY, 3, 4, FAST>},
        {stream2d_main<SX, SY, 4, 1, FAST>, stream2d_main<SX, SY, 4, 2, FAST>, stream2d_main<SX, SY, 4, 3, FAST>,
         stream2d_main<SX, SY, 4, 4, FAST>}};
    if (a.rx < 1 || a.rx > kStreamRMax || a.ry < 1 || a.ry > kStreamRMax) return hipErrorInvalidValue;
    table[a.rx - 1][a.ry - 1](a, st);
    return hipGetLastError();
}

template <int SX, int SY>
hipError_t stream2d_launch(const StreamArgs &a, bool fast, hipStream_t st) {
    return fast ? stream2d_fast<SX, SY, true>(a, st) : stream2d_fast<SX, SY, false>(a, st);
}

}  // namespace dipamd
