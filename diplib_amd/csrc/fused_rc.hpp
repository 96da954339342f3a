// K3R: the fused x+y pass with a compile-time x radius class (gfx950).
//
// Same pass structure and arithmetic as K3 (fused_tile in sep_kernels_impl.hpp): rows
// of a column tile are staged in LDS, x-filtered into a ring of rows, and the y filter
// runs out of the ring; every output sample is computed in the reference's order
// (convolution.cpp:214-241), so results are bit-identical to K3 and to the reference.
// What differs is how the x stage reaches its samples:
//   * the staged rows hold the columns [c0 - RXC, c0 + TXC + RXC) for the class radius
//     RXC >= rx (a multiple of 4), unpadded, 16-byte aligned, with a row pitch of 4 words
//     more than a multiple of 32;
//   * an x-stage lane owns P consecutive outputs of one row and loads its whole window
//     (P + 2 RXC samples) with 128-bit LDS reads at compile-time offsets; the tap loop
//     is unrolled to RXC with a uniform exit at the runtime radius rx, so every window
//     index is a constant (no address arithmetic per tap);
//   * the eight lanes of a group are eight rows at one column offset (row pitch 4 words
//     mod 32): their 16-byte accesses cover all 32 banks;
//   * interior tiles load whole 4-sample words (uint8) or 16-byte vectors (float,
//     double) per lane and write converted 16-byte vectors to LDS; tiles at the left or
//     right edge take per-sample boundary-folded loads.
#pragma once

#include "sep_kernels_impl.hpp"

namespace dipamd {

// Row pitch in elements of TB: >= n, 16-byte rows, and 4 words beyond a multiple of 32.
template <typename TB>
__host__ __device__ constexpr int rc_pitch(int n) {
    constexpr int M = sizeof(TB) == 8 ? 16 : 32;  // elements per 32 words
    constexpr int O = sizeof(TB) == 8 ? 2 : 4;    // elements per 4 words
    return n <= O ? O : ((n - O + M - 1) / M) * M + O;
}

// x radius class of a radius: the next multiple of 4 (at least 4); 0 for no x filter.
__host__ __device__ constexpr int rc_class(int r) { return r <= 0 ? 0 : ((r + 3) / 4) * 4; }

template <typename TB>
__host__ __device__ constexpr size_t rc_lds_bytes(int txc, int rxc, int ry) {
    const int S = 16;
    const int A = (2 * ry + S - 1) / S;
    return (static_cast<size_t>(S) * rc_pitch<TB>(txc + 2 * rxc) + static_cast<size_t>((A + 1) * S) * rc_pitch<TB>(txc)) *
           sizeof(TB);
}

// A vector load of the input: N samples of TI in one load.
template <typename TI>
struct RcVec;
template <>
struct RcVec<uint8_t> {
    using V = uint32_t;
    static constexpr int N = 4;
    __device__ static void convert(V v, float (&o)[4]) {
        o[0] = static_cast<float>(v & 0xffu);
        o[1] = static_cast<float>((v >> 8) & 0xffu);
        o[2] = static_cast<float>((v >> 16) & 0xffu);
        o[3] = static_cast<float>(v >> 24);
    }
};
template <>
struct RcVec<float> {
    using V = float4;
    static constexpr int N = 4;
    __device__ static void convert(V v, float (&o)[4]) {
        o[0] = v.x;
        o[1] = v.y;
        o[2] = v.z;
        o[3] = v.w;
    }
};
template <>
struct RcVec<double> {
    using V = double2;
    static constexpr int N = 2;
    __device__ static void convert(V v, double (&o)[2]) {
        o[0] = v.x;
        o[1] = v.y;
    }
};

// 16-byte LDS vector of the buffer type.
template <typename TB>
struct RcLds;
template <>
struct RcLds<float> {
    using V = float4;
    static constexpr int N = 4;
};
template <>
struct RcLds<double> {
    using V = double2;
    static constexpr int N = 2;
};

// One column of the unpadded ring; rows wrap modulo rc.
template <typename TB>
struct RingSrcRc {
    const TB *col;
    int pitch, rc, span, r0, ro, lo;
    __device__ int wrap(int row) const {
        row += row < 0 ? rc : 0;
        row -= row >= rc ? rc : 0;
        return row;
    }
    __device__ TB at(int q) const { return col[wrap(r0 + q) * pitch]; }
    __device__ void rbegin(int q) { ro = wrap(r0 + q) * pitch; }
    __device__ TB rnext() {
        const TB v = col[ro];
        ro += pitch;
        ro = ro == span ? 0 : ro;
        return v;
    }
    __device__ void lbegin(int q) { lo = wrap(r0 + q) * pitch; }
    __device__ TB lnext() {
        const TB v = col[lo];
        lo = lo == 0 ? span - pitch : lo - pitch;
        return v;
    }
};

// The tile.  INTERIOR: every staged column lies inside the plane (vector loads);
// otherwise per-sample folded loads.
template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST, int RXC, bool INTERIOR, typename TapP>
__device__ __forceinline__ void fused_tile_rc_body(const Tile<TB> &t, const TI *__restrict__ src,
                                                   TB *__restrict__ dst, TB *lds, TapP tx, TapP ty) {
    constexpr int S = 16;
    static_assert(TXC / P == 16 && S * 16 == kBlock, "tile shape: 16 rows x 16 lanes of P outputs");
    static_assert(RXC % 4 == 0, "radius class: a multiple of 4");
    constexpr int WX = TXC + 2 * RXC;      // staged columns
    constexpr int PS = rc_pitch<TB>(WX);   // staged row pitch
    constexpr int PR = rc_pitch<TB>(TXC);  // ring row pitch
    constexpr int VN = RcVec<TI>::N;       // samples per vector load
    constexpr int NVEC = WX / VN;          // vector loads per staged row
    constexpr int NV = (NVEC + 63) / 64;   // per lane
    constexpr int NLD = (WX + 63) / 64;    // per-sample loads per lane (edge tiles)
    constexpr int LN = RcLds<TB>::N;       // TB per 16-byte LDS access
    constexpr int WN = P + 2 * RXC;        // window samples of an x-stage lane
    static_assert(WX % VN == 0 && WN % LN == 0 && P % LN == 0, "vector widths");
    using LV = typename RcLds<TB>::V;
    using IV = typename RcVec<TI>::V;

    const int rx = t.rx, By = t.ry;
    const int A = (2 * By + S - 1) / S;
    const int RC = (A + 1) * S;
    TB *stg = lds;
    TB *ring = lds + S * PS;
    const int c0 = t.c0, z0 = t.z0, zend = t.zend;
    const int nsteps = (zend - z0 + S - 1) / S;
    const int M = A + nsteps;
    const int tid = threadIdx.x;
    const int wave = wave_id(), lane = tid & 63;

    // edge tiles: boundary-folded column offsets and signs of this lane's staged samples
    int cofs[INTERIOR ? 1 : NLD];
    int cmul[INTERIOR ? 1 : NLD];
    if constexpr (!INTERIOR) {
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int col = lane + 64 * i;
            int x = c0 - RXC + col;
            int m = 1;
            if (col < WX && (x < 0 || x >= t.nx)) {
                x = bc_fold(x, t.nx, t.bcx, &m);
                if (x < 0) {
                    x = 0;
                    m = 0;
                }
            }
            if (col >= WX) x = 0;
            cofs[i] = x;
            cmul[i] = m;
        }
    }

    // --- staged-row loads: wave w loads rows w, w+4, w+8, w+12 of each step ---
    IV prev[INTERIOR ? 4 : 1][INTERIOR ? NV : 1];
    TI pres[INTERIOR ? 1 : 4][INTERIOR ? 1 : NLD];
    auto load_step = [&](int m) {
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            int j = z0 - By + m * S + wave + 4 * rr;  // extended row (wave-uniform)
            if (j < 0 || j >= t.ny) {
                int mul;
                j = bc_fold(j, t.ny, t.bcy, &mul);  // sign / constant applied after the x stage
                if (j < 0) j = 0;
            }
            const TI *rowp = src + static_cast<int64_t>(j) * t.syin;
            if constexpr (INTERIOR) {
                const IV *rv = reinterpret_cast<const IV *>(rowp + (c0 - RXC));
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    const int v = lane + 64 * i;
                    if (NVEC % 64 == 0 || v < NVEC) prev[rr][i] = rv[v];
                }
            } else {
#pragma unroll
                for (int i = 0; i < NLD; ++i) pres[rr][i] = rowp[cofs[i]];
            }
        }
    };

    load_step(0);
    for (int m = 0; m < M; ++m) {
        lds_barrier();  // previous x stage done with `stg`
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            TB *srow = stg + (wave + 4 * rr) * PS;
            if constexpr (INTERIOR) {
#pragma unroll
                for (int i = 0; i < NV; ++i) {
                    const int v = lane + 64 * i;
                    if (NVEC % 64 == 0 || v < NVEC) {
                        TB o[VN];
                        RcVec<TI>::convert(prev[rr][i], o);
                        if constexpr (VN == LN) {
                            LV w;
                            __builtin_memcpy(&w, o, sizeof(w));
                            reinterpret_cast<LV *>(srow)[v] = w;
                        } else {
#pragma unroll
                            for (int e = 0; e < VN; ++e) srow[v * VN + e] = o[e];
                        }
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < NLD; ++i) {
                    const int col = lane + 64 * i;
                    if (col < WX) {
                        const TB v = to_buf<TB>(pres[rr][i]);
                        srow[col] = cmul[i] > 0 ? v : (cmul[i] < 0 ? -v : t.cvalx);
                    }
                }
            }
        }
        lds_barrier();  // `stg` full; previous y stage done with the ring slots we overwrite
        if (m + 1 < M) load_step(m + 1);

        // --- x stage: row xr, outputs xl*P .. xl*P+P-1; the 8 lanes of a group are 8 rows ---
        {
            const int xr = (lane & 7) + 8 * (wave & 1);
            const int xl = (lane >> 3) + 8 * (wave >> 1);
            const int j = z0 - By + m * S + xr;
            const int jstep = z0 - By + m * S;
            TB acc[P];
            TB w[WN];
            const LV *wv = reinterpret_cast<const LV *>(stg + xr * PS + xl * P);
#pragma unroll
            for (int i = 0; i < WN / LN; ++i) {
                const LV v = wv[i];
                __builtin_memcpy(&w[i * LN], &v, sizeof(v));
            }
            if constexpr (SX == SYM_NONE) {
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = w[RXC + p];
            } else {
                const TB f0 = tx[0];
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = f0 * w[RXC + p];
#pragma unroll
                for (int k = 1; k <= RXC; ++k) {
                    if (k <= rx) {  // uniform: the taps beyond the radius are not there
                        const TB f = tx[k];
#pragma unroll
                        for (int p = 0; p < P; ++p) {
                            const TB a = w[RXC + p + k];
                            const TB b = w[RXC + p - k];
                            acc[p] = mac<TB, FAST>(acc[p], f, SX == SYM_EVEN ? a + b : a - b);
                        }
                    }
                }
            }
            // rows outside the plane only in the first and last steps (a uniform test):
            // the boundary condition's sign or constant along y
            if (jstep < 0 || jstep + S > t.ny) {
                int mul = 1;
                bool isc = false;
                if (j < 0 || j >= t.ny) isc = bc_fold(j, t.ny, t.bcy, &mul) < 0;
                if (isc) {
#pragma unroll
                    for (int p = 0; p < P; ++p) acc[p] = t.cvaly;
                } else if (mul < 0) {
#pragma unroll
                    for (int p = 0; p < P; ++p) acc[p] = -acc[p];
                }
            }
            // ring slot of row m*S + xr: RC is a whole number of steps
            const int slot = (m % (A + 1)) * S + xr;
            LV *rrow = reinterpret_cast<LV *>(ring + slot * PR + xl * P);
#pragma unroll
            for (int i = 0; i < P / LN; ++i) {
                LV v;
                __builtin_memcpy(&v, &acc[i * LN], sizeof(v));
                rrow[i] = v;
            }
        }
        lds_barrier();  // ring rows of this step written

        // --- y stage: step n = m - A ---
        const int n = m - A;
        if (n >= 0) {
            const int col = tid % TXC, grp = (wave * 64) / TXC;
            const int c = c0 + col;
            const int rel0 = n * S + grp * P + By;  // ring-relative row of output p = 0
            RingSrcRc<TB> s{ring + col, PR, RC, RC * PR, rel0 % RC, 0, 0};
            TB acc[P];
            fir_window<TB, P, SY, FAST>(ty, By + 1, 0, s, acc);
            const int zs = z0 + n * S + grp * P;
            if (c < t.nx) {
                TB *d = dst + static_cast<int64_t>(zs) * t.syout + static_cast<int64_t>(c) * t.sxout;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    if (zs + p < zend) *d = acc[p];
                    d += t.syout;
                }
            }
        }
    }
}

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST, int RXC, typename TapP>
__device__ __forceinline__ void fused_tile_rc(const Tile<TB> &t, const TI *__restrict__ src, TB *__restrict__ dst,
                                              TB *lds, TapP tx, TapP ty) {
    if (t.vec_ok && t.c0 - RXC >= 0 && t.c0 + TXC + RXC <= t.nx)
        fused_tile_rc_body<TI, TB, TXC, P, SX, SY, FAST, RXC, true>(t, src, dst, lds, tx, ty);
    else
        fused_tile_rc_body<TI, TB, TXC, P, SX, SY, FAST, RXC, false>(t, src, dst, lds, tx, ty);
}

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST, int RXC>
__global__ __launch_bounds__(kBlock) void k_fused2d_rc(const FusedArgs<TB> a, int64_t ntiles, int64_t nchunks) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TB *lds = reinterpret_cast<TB *>(smem_raw);
    const int64_t bid = xcd_block();
    const int64_t tile = bid % ntiles;
    const int64_t rest = bid / ntiles;
    const int64_t chunk = rest % nchunks;
    const int64_t outer = rest / nchunks;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    Tile<TB> t;
    t.in = static_cast<const TI *>(a.in) + oin;
    t.out = static_cast<TB *>(a.out) + oout;
    t.nx = static_cast<int>(a.nx);
    t.ny = static_cast<int>(a.ny);
    t.sxin = 1;
    t.syin = a.syin;
    t.sxout = a.sxout;
    t.syout = a.syout;
    t.c0 = static_cast<int>(tile * TXC);
    t.z0 = static_cast<int>(chunk * a.chunk);
    t.zend = static_cast<int>(min(chunk * a.chunk + a.chunk, a.ny));
    t.rx = a.rx;
    t.ry = a.ry;
    t.bcx = a.bcx;
    t.bcy = a.bcy;
    t.cvalx = a.cvalx;
    t.cvaly = a.cvaly;
    t.vec_ok = 1;  // checked by the host
    fused_tile_rc<TI, TB, TXC, P, SX, SY, FAST, RXC>(t, static_cast<const TI *>(t.in), t.out, lds, a.tx, a.ty);
}

}  // namespace dipamd

namespace dipamd {

// K4R: the batched persistent form on the radius-class tile.  One launch covers the
// work items of one radius class (x radius up to RXC): each workgroup takes the next
// item from the group's counter and runs the tile with that image's taps.
// The loop ends when the counter passes the item count, so every wave exits.
template <typename TI, typename TB, int TXC, int P, bool FAST, int RXC>
__global__ __launch_bounds__(kBlock) void k_batched_rc(const BatchArgs<TB> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_item;
    TB *lds = reinterpret_cast<TB *>(smem_raw);
    // the taps table is written before the launch and only read here: through the
    // constant address space its reads are scalar loads, as the plain kernel's taps
    using CT = const __attribute__((address_space(4))) TB;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(a.counter, 1);
        lds_barrier();
        const int item = s_item;
        if (item >= a.nitems) break;  // uniform: every wave of the workgroup leaves
        const BatchItem it = a.items[item];
        const BatchImage<TB> im = a.images[it.image];
        const int ntx = im.rx + 1;
        CT *taps = (CT *)(a.taps + __builtin_amdgcn_readfirstlane(im.tap_offset));
        Tile<TB> t;
        t.in = im.in;
        t.out = im.out;
        t.nx = static_cast<int>(im.nx);
        t.ny = static_cast<int>(im.ny);
        t.sxin = 1;
        t.syin = im.syin;
        t.sxout = 1;
        t.syout = im.syout;
        t.c0 = it.tile * TXC;
        t.z0 = it.z0;
        t.zend = it.z1;
        t.rx = im.rx;
        t.ry = im.ry;
        t.bcx = a.bcx;
        t.bcy = a.bcy;
        t.cvalx = a.cvalx;
        t.cvaly = a.cvaly;
        t.vec_ok = im.vec_ok;
        fused_tile_rc<TI, TB, TXC, P, SYM_EVEN, SYM_EVEN, FAST, RXC>(t, static_cast<const TI *>(t.in), t.out, lds,
                                                                     taps, taps + ntx);
        lds_barrier();  // LDS (ring) free and s_item read by all before the next item
    }
}

}  // namespace dipamd
