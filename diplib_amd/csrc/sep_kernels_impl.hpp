// Separable FIR kernels for gfx950 (MI355X): the line pass along the fastest axis
// (LDS-staged rows with boundary folding), the strided-axis sweep (one thread per
// column, a register window sliding along the axis, no LDS) and the fused 2-D
// pass (x filter into an LDS ring of rows, y filter out of it, one HBM read and
// one HBM write per pixel).
//
// Every output sample is computed in the order of the reference line filter
// (convolution.cpp:214-241):
//     EVEN:  s = f0*x[i];  s += f_k*(x[i+k] + x[i-k])  for k = 1..r
//     ODD:   s = f0*x[i];  s += f_k*(x[i+k] - x[i-k])  for k = 1..r
//     GENERAL (convolution.cpp:202-213), D_EVEN / D_ODD (:256-283) likewise.
// With DIPAMD_EXACT the code is compiled without contraction (-ffp-contract=off)
// and rounds like the reference; with DIPAMD_FAST the products are fused (fma).
//
// The window: a lane computes P consecutive outputs.  Right-hand samples x[i+p+k]
// live in a P-slot register ring indexed (p+k) mod P, left-hand ones in a second
// ring indexed (p-k) mod P; the tap loop is unrolled by P so every slot index is a
// compile-time constant, and each tap step loads just one new sample per side.
#pragma once

#include "sep_common.hpp"
#include "sep_kernels.hpp"

namespace dipamd {

template <typename TB, bool FAST>
__device__ __forceinline__ TB mac(TB acc, TB f, TB t) {
    if constexpr (FAST) {
        return __builtin_fma(f, t, acc);
    } else {
        return acc + f * t;
    }
}
template <>
__device__ __forceinline__ float mac<float, true>(float acc, float f, float t) {
    return __builtin_fmaf(f, t, acc);
}

// P outputs from a sample accessor get(q): q = 0 is output p = 0's centre sample.
// `shift` moves the window: get(q) is asked for sample i + q - shift (GENERAL: the
// reversed origin; the symmetric forms: origin - (ntaps - 1), zero for a centred filter).
template <typename TB, int P, int SYM, bool FAST, typename Get0>
__device__ __forceinline__ void fir_window(const TB *taps, int ntaps, int shift, Get0 get0, TB (&acc)[P]) {
    auto get = [&](int q) { return get0(q - shift); };
    if constexpr (SYM == SYM_EVEN || SYM == SYM_ODD) {
        const int r = ntaps - 1;
        TB rw[P], lw[P];
        TB w[P + 2];
#pragma unroll
        for (int q = 0; q < P + 2; ++q) w[q] = get(q - 1);
        const TB f0 = taps[0];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc[p] = f0 * w[p + 1];
            rw[(p + 1) % P] = w[p + 2];
            lw[(p - 1 + P) % P] = w[p];
        }
        for (int kb = 1; kb <= r; kb += P) {
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const int k = kb + u;
                if (k <= r) {
                    const TB f = taps[k];
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        const TB a = rw[(p + 1 + u) % P];
                        const TB b = lw[((p - 1 - u) % P + P) % P];
                        const TB t = (SYM == SYM_EVEN) ? (a + b) : (a - b);
                        acc[p] = mac<TB, FAST>(acc[p], f, t);
                    }
                    if (k < r) {
                        rw[(1 + u) % P] = get(k + P);
                        lw[((P - 2 - u) % P + P) % P] = get(-k - 1);
                    }
                }
            }
        }
    } else if constexpr (SYM == SYM_GENERAL) {
        TB win[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc[p] = TB(0);
            win[p] = get(p);
        }
        for (int kb = 0; kb < ntaps; kb += P) {
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const int k = kb + u;
                if (k < ntaps) {
                    const TB f = taps[k];
#pragma unroll
                    for (int p = 0; p < P; ++p) acc[p] = mac<TB, FAST>(acc[p], f, win[(p + u) % P]);
                    if (k + 1 < ntaps) win[u % P] = get(k + P);
                }
            }
        }
    } else {  // D_EVEN / D_ODD: s += f_k*(x[i+k] +- x[i-k-1]), k = 0..ntaps-1
        TB rw[P], lw[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc[p] = TB(0);
            rw[p] = get(p);
            lw[(p - 1 + P) % P] = get(p - 1);
        }
        for (int kb = 0; kb < ntaps; kb += P) {
#pragma unroll
            for (int u = 0; u < P; ++u) {
                const int k = kb + u;
                if (k < ntaps) {
                    const TB f = taps[k];
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        const TB a = rw[(p + u) % P];
                        const TB b = lw[((p - 1 - u) % P + P) % P];
                        const TB t = (SYM == SYM_D_EVEN) ? (a + b) : (a - b);
                        acc[p] = mac<TB, FAST>(acc[p], f, t);
                    }
                    if (k + 1 < ntaps) {
                        rw[u % P] = get(k + P);
                        lw[((P - 2 - u) % P + P) % P] = get(-k - 2);
                    }
                }
            }
        }
    }
}

// Sample j of a line of n samples starting at `src` with stride `st`, extended by
// the boundary condition.
template <typename TI, typename TB>
__device__ __forceinline__ TB load_ext(const TI *src, int64_t st, int64_t j, int64_t n, int bc, TB cval) {
    if (j >= 0 && j < n) return to_buf<TB>(src[j * st]);
    int mul;
    bool isc;
    int64_t s = bc_map(j, n, bc, &mul, &isc);
    if (isc) return cval;
    TB v = to_buf<TB>(src[s * st]);
    return mul < 0 ? -v : v;
}

// ---------------------------------------------------------------------------
// K1: line pass along the fastest axis.  One wave per line segment of 64*P
// outputs, 4 lines per workgroup; the segment plus `border` samples each side
// is staged in LDS (converted to the buffer type, boundary folded).
// ---------------------------------------------------------------------------
template <typename TI, typename TB, int P, int SYM, bool FAST>
__global__ __launch_bounds__(kBlock) void k_line(const PassArgs<TB> a, int64_t nseg) {
    constexpr int TX = 64 * P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TB *lds = reinterpret_cast<TB *>(smem_raw);
    const int B = a.border;
    const int W = TX + 2 * B;
    const int pitch = lpad_size(W);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t seg = blockIdx.x % nseg;
    const int64_t line = (blockIdx.x / nseg) * 4 + wave;
    const int64_t x0 = seg * TX;
    TB *row = lds + wave * pitch;
    int64_t oin = 0, oout = 0;
    const bool active = line < a.nlines;
    if (active) {
        outer_offsets(a.outer, line, &oin, &oout);
        const TI *src = static_cast<const TI *>(a.in) + oin;
        const bool interior = (x0 - B >= 0) && (x0 + TX + B <= a.n);
        if (interior) {
            const TI *s0 = src + (x0 - B) * a.sin;
            for (int i = lane; i < W; i += 64) row[lpad(i)] = to_buf<TB>(s0[i * a.sin]);
        } else {
            for (int i = lane; i < W; i += 64) row[lpad(i)] = load_ext<TI, TB>(src, a.sin, x0 - B + i, a.n, a.bc, a.cval);
        }
    }
    __syncthreads();
    if (!active) return;
    const int base = B + lane * P;
    TB acc[P];
    fir_window<TB, P, SYM, FAST>(a.taps, a.ntaps, a.shift, [&](int q) { return row[lpad(base + q)]; }, acc);
    TB *dst = static_cast<TB *>(a.out) + oout;
    const int64_t xs = x0 + lane * P;
    if (a.sout == 1 && xs + P <= a.n) {
#pragma unroll
        for (int p = 0; p < P; ++p) dst[xs + p] = acc[p];
    } else {
#pragma unroll
        for (int p = 0; p < P; ++p)
            if (xs + p < a.n) dst[(xs + p) * a.sout] = acc[p];
    }
}

// ---------------------------------------------------------------------------
// K2: sweep along a strided axis.  One thread per position of the lane axis
// (coalesced across the wave), each running down a chunk of the filter axis P
// outputs at a time with the register window reading global memory (the window
// rows are re-read from L1/L2, HBM sees each row once).  No LDS, no barriers.
// ---------------------------------------------------------------------------
template <typename TI, typename TB, int P, int SYM, bool FAST>
__global__ __launch_bounds__(kBlock) void k_sweep(const PassArgs<TB> a, int64_t ntiles, int64_t nchunks) {
    const int64_t bid = blockIdx.x;
    const int64_t tile = bid % ntiles;
    const int64_t rest = bid / ntiles;
    const int64_t chunk = rest % nchunks;
    const int64_t outer = rest / nchunks;
    const int64_t c = tile * blockDim.x + threadIdx.x;
    if (c >= a.e) return;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    const TI *src = static_cast<const TI *>(a.in) + oin + c * a.ein;
    TB *dst = static_cast<TB *>(a.out) + oout + c * a.eout;
    const int64_t z0 = chunk * a.chunk;
    const int64_t zend = min(z0 + a.chunk, a.n);
    const int64_t n = a.n;
    const int64_t sin = a.sin;
    for (int64_t zb = z0; zb < zend; zb += P) {
        TB acc[P];
        const bool interior = (zb - a.border >= 0) && (zb + P + a.border <= n);
        if (interior) {
            const TI *s0 = src + zb * sin;
            fir_window<TB, P, SYM, FAST>(a.taps, a.ntaps, a.shift,
                                         [&](int q) { return to_buf<TB>(s0[q * sin]); }, acc);
        } else {
            fir_window<TB, P, SYM, FAST>(a.taps, a.ntaps, a.shift,
                                         [&](int q) { return load_ext<TI, TB>(src, sin, zb + q, n, a.bc, a.cval); },
                                         acc);
        }
#pragma unroll
        for (int p = 0; p < P; ++p)
            if (zb + p < zend) dst[(zb + p) * a.sout] = acc[p];
    }
}

// ---------------------------------------------------------------------------
// K3: fused x+y pass over 2-D planes.  A workgroup owns TXC columns and a chunk
// of rows; it streams input rows (with x and y halos, both boundary-folded)
// through an LDS staging buffer, x-filters them into an LDS ring of
// RC = (A+1)*S rows (A = ceil(2*ry/S)), and y-filters each step's S output rows
// out of the ring straight to HBM.  The intermediate never leaves LDS and stays
// in the buffer type, so the result equals the reference's two passes.
//   threads: x stage = 16 rows x (TXC/P) lanes, P outputs each;
//            y stage = TXC columns x (256/TXC) groups, P rows each;  S = 16.
// Next step's input rows are prefetched into registers before the compute.
// ---------------------------------------------------------------------------
// One (column tile, row chunk) of a 2-D plane: the body of the fused kernel,
// shared by the plain launch (one workgroup per tile) and the batched persistent
// kernel (tiles pulled from a work queue).  `lds` holds S*pitchS staging words
// followed by RC*pitchR ring words; `tx`/`ty` point at rx+1 / ry+1 taps.
template <typename TB>
struct Tile {
    const void *in;   // plane origin
    TB *out;
    int64_t nx, ny;
    int64_t sxin, syin, sxout, syout;
    int64_t c0, z0, zend;  // first column, first and end output rows
    int rx, ry, bcx, bcy;
    TB cvalx, cvaly;
};

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST>
__device__ __forceinline__ void fused_tile(const Tile<TB> &t, TB *lds, const TB *tx, const TB *ty) {
    constexpr int S = 16;
    constexpr int XT = TXC / P;            // x-stage lanes per row (16)
    static_assert(XT * S == kBlock, "tile shape");
    constexpr int NLD = (TXC + 2 * (kFusedCap - 1) + 63) / 64;  // loads per staged row per lane (max)
    const int Bx = t.rx, By = t.ry;
    const int Wx = TXC + 2 * Bx;
    const int pitchS = lpad_size(Wx);
    const int pitchR = lpad_size(TXC);
    const int A = (2 * By + S - 1) / S;
    const int RC = (A + 1) * S;
    TB *stg = lds;
    TB *ring = lds + S * pitchS;
    const TI *src = static_cast<const TI *>(t.in);
    TB *dst = t.out;
    const int64_t c0 = t.c0, z0 = t.z0, zend = t.zend;
    const int nsteps = static_cast<int>((zend - z0 + S - 1) / S);
    const int M = A + nsteps;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const bool xinterior = (c0 - Bx >= 0) && (c0 + TXC + Bx <= t.nx);
    const int nld = (Wx + 63) / 64;

    // --- staged-row loads: wave w loads rows w, w+4, w+8, w+12 of each step ---
    TB pre[4][NLD];
    auto load_step = [&](int m) {
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int64_t j = z0 - By + static_cast<int64_t>(m) * S + wave + 4 * rr;  // extended row
            int mul = 1;
            bool isc = false;
            int64_t sr = j;
            if (j < 0 || j >= t.ny) sr = bc_map(j, t.ny, t.bcy, &mul, &isc);
            const TI *srow = src + sr * t.syin;
            if (isc) continue;  // constant row: filled after the x stage
            if (xinterior) {
                const TI *s0 = srow + (c0 - Bx) * t.sxin;
#pragma unroll
                for (int i = 0; i < NLD; ++i) {
                    const int col = lane + 64 * i;
                    if (i < nld && col < Wx) pre[rr][i] = to_buf<TB>(s0[col * t.sxin]);
                }
            } else {
#pragma unroll
                for (int i = 0; i < NLD; ++i) {
                    const int col = lane + 64 * i;
                    if (i < nld && col < Wx)
                        pre[rr][i] = load_ext<TI, TB>(srow, t.sxin, c0 - Bx + col, t.nx, t.bcx, t.cvalx);
                }
            }
        }
    };

    load_step(0);
    for (int m = 0; m < M; ++m) {
        __syncthreads();  // previous x stage done with `stg`
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            TB *srow = stg + (wave + 4 * rr) * pitchS;
#pragma unroll
            for (int i = 0; i < NLD; ++i) {
                const int col = lane + 64 * i;
                if (i < nld && col < Wx) srow[lpad(col)] = pre[rr][i];
            }
        }
        __syncthreads();  // `stg` full; previous y stage done with the ring slots we overwrite
        if (m + 1 < M) load_step(m + 1);

        // --- x stage: row xr of this step, outputs xl*P .. xl*P+P-1 ---
        {
            const int xr = tid / XT, xl = tid % XT;
            const int64_t j = z0 - By + static_cast<int64_t>(m) * S + xr;
            int mul = 1;
            bool isc = false;
            if (j < 0 || j >= t.ny) bc_map(j, t.ny, t.bcy, &mul, &isc);
            TB *rrow = ring + ((m * S + xr) % RC) * pitchR;
            if (isc) {
#pragma unroll
                for (int p = 0; p < P; ++p) rrow[lpad(xl * P + p)] = t.cvaly;
            } else {
                const TB *srow = stg + xr * pitchS;
                const int base = Bx + xl * P;
                TB acc[P];
                fir_window<TB, P, SX, FAST>(tx, Bx + 1, 0, [&](int q) { return srow[lpad(base + q)]; }, acc);
#pragma unroll
                for (int p = 0; p < P; ++p) rrow[lpad(xl * P + p)] = mul < 0 ? -acc[p] : acc[p];
            }
        }
        __syncthreads();  // ring rows of this step written

        // --- y stage: step n = m - A ---
        const int n = m - A;
        if (n >= 0) {
            const int col = tid % TXC, grp = tid / TXC;
            const int64_t c = c0 + col;
            const int rel0 = n * S + grp * P + By;  // ring-relative row of output p = 0
            TB acc[P];
            fir_window<TB, P, SY, FAST>(ty, By + 1, 0,
                                        [&](int q) { return ring[((rel0 + q) % RC) * pitchR + lpad(col)]; }, acc);
            const int64_t zs = z0 + static_cast<int64_t>(n) * S + grp * P;
            if (c < t.nx) {
#pragma unroll
                for (int p = 0; p < P; ++p)
                    if (zs + p < zend) dst[(zs + p) * t.syout + c * t.sxout] = acc[p];
            }
        }
    }
}

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST>
__global__ __launch_bounds__(kBlock) void k_fused2d(const FusedArgs<TB> a, int64_t ntiles, int64_t nchunks) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TB *lds = reinterpret_cast<TB *>(smem_raw);
    const int64_t bid = blockIdx.x;
    const int64_t tile = bid % ntiles;
    const int64_t rest = bid / ntiles;
    const int64_t chunk = rest % nchunks;
    const int64_t outer = rest / nchunks;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    Tile<TB> t;
    t.in = static_cast<const TI *>(a.in) + oin;
    t.out = static_cast<TB *>(a.out) + oout;
    t.nx = a.nx;
    t.ny = a.ny;
    t.sxin = a.sxin;
    t.syin = a.syin;
    t.sxout = a.sxout;
    t.syout = a.syout;
    t.c0 = tile * TXC;
    t.z0 = chunk * a.chunk;
    t.zend = min(t.z0 + a.chunk, a.ny);
    t.rx = a.rx;
    t.ry = a.ry;
    t.bcx = a.bcx;
    t.bcy = a.bcy;
    t.cvalx = a.cvalx;
    t.cvaly = a.cvaly;
    fused_tile<TI, TB, TXC, P, SX, SY, FAST>(t, lds, a.tx, a.ty);
}

// ---------------------------------------------------------------------------
// K4: batched persistent form.  One launch covers a list of 2-D images of
// different sizes and sigmas: each workgroup repeatedly takes the next work item
// (image, column tile, row chunk) from a global counter, copies that image's taps
// to LDS and runs the fused tile.  Items are ordered largest-cost first.  The
// loop ends when the counter passes the item count, so every wave exits.
// ---------------------------------------------------------------------------
template <typename TI, typename TB, int TXC, int P, bool FAST>
__global__ __launch_bounds__(kBlock) void k_batched(const BatchArgs<TB> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_item;
    __shared__ TB s_taps[2 * kFusedCap];
    TB *lds = reinterpret_cast<TB *>(smem_raw);
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(a.counter, 1);
        __syncthreads();
        const int item = s_item;
        if (item >= a.nitems) break;  // uniform: every wave of the workgroup leaves
        const BatchItem it = a.items[item];
        const BatchImage<TB> im = a.images[it.image];
        const int ntx = im.rx + 1, nty = im.ry + 1;
        for (int k = threadIdx.x; k < ntx + nty; k += blockDim.x) s_taps[k] = a.taps[im.tap_offset + k];
        __syncthreads();  // taps visible; s_item read by all before it is rewritten
        Tile<TB> t;
        t.in = im.in;
        t.out = im.out;
        t.nx = im.nx;
        t.ny = im.ny;
        t.sxin = 1;
        t.syin = im.syin;
        t.sxout = 1;
        t.syout = im.syout;
        t.c0 = static_cast<int64_t>(it.tile) * TXC;
        t.z0 = it.z0;
        t.zend = it.z1;
        t.rx = im.rx;
        t.ry = im.ry;
        t.bcx = a.bcx;
        t.bcy = a.bcy;
        t.cvalx = a.cvalx;
        t.cvaly = a.cvaly;
        fused_tile<TI, TB, TXC, P, SYM_EVEN, SYM_EVEN, FAST>(t, lds, s_taps, s_taps + ntx);
        __syncthreads();  // LDS (ring, taps) free before the next item
    }
}


// ---------------------------------------------------------------------------
// Element-wise conversions: input -> buffer type (contiguous), buffer type ->
// output type with clamp_cast (clamp_cast.h:210-235: clamp, then truncate).
// ---------------------------------------------------------------------------


template <typename T>
__device__ __forceinline__ T clamp_to(double v, double lo, double hi) {
    v = v < lo ? lo : v;
    v = v > hi ? hi : v;
    return static_cast<T>(v);
}

template <typename TS, typename TD>
__device__ __forceinline__ TD convert(TS v);

#define DIPAMD_CONV_FLOAT(TD)                                        \
    template <>                                                      \
    __device__ __forceinline__ TD convert<float, TD>(float v) {      \
        return static_cast<TD>(v);                                   \
    }                                                                \
    template <>                                                      \
    __device__ __forceinline__ TD convert<double, TD>(double v) {    \
        return static_cast<TD>(v);                                   \
    }
DIPAMD_CONV_FLOAT(float)
DIPAMD_CONV_FLOAT(double)

// float sources clamp against limits cast to float (as clamp_cast< T >( sfloat ) does)
#define DIPAMD_CONV_INT(TD, LO, HI)                                                                          \
    template <>                                                                                              \
    __device__ __forceinline__ TD convert<float, TD>(float v) {                                             \
        const float lo = static_cast<float>(LO), hi = static_cast<float>(HI);                               \
        float c = v < lo ? lo : (v > hi ? hi : v);                                                           \
        if (static_cast<double>(c) >= static_cast<double>(HI)) return static_cast<TD>(HI);                  \
        return static_cast<TD>(c);                                                                           \
    }                                                                                                        \
    template <>                                                                                              \
    __device__ __forceinline__ TD convert<double, TD>(double v) {                                           \
        const double lo = static_cast<double>(LO), hi = static_cast<double>(HI);                            \
        double c = v < lo ? lo : (v > hi ? hi : v);                                                          \
        return static_cast<TD>(c);                                                                           \
    }
DIPAMD_CONV_INT(uint8_t, 0, 255)
DIPAMD_CONV_INT(uint16_t, 0, 65535)
DIPAMD_CONV_INT(uint32_t, 0u, 4294967295u)
DIPAMD_CONV_INT(int8_t, -128, 127)
DIPAMD_CONV_INT(int16_t, -32768, 32767)
DIPAMD_CONV_INT(int32_t, (-2147483647 - 1), 2147483647)
template <>
__device__ __forceinline__ bin_t convert<float, bin_t>(float v) {
    return bin_t{static_cast<uint8_t>(v != 0.0f)};
}
template <>
__device__ __forceinline__ bin_t convert<double, bin_t>(double v) {
    return bin_t{static_cast<uint8_t>(v != 0.0)};
}

template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void k_cast_out(const TS *in, TD *out, const CastGeo g) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < g.total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t r = i, oi = 0, oo = 0;
        for (int d = 0; d < g.ndim; ++d) {
            int64_t c = r % g.size[d];
            r /= g.size[d];
            oi += c * g.sin[d];
            oo += c * g.sout[d];
        }
        out[oo] = convert<TS, TD>(in[oi]);
    }
}

template <typename TS, typename TB>
__global__ __launch_bounds__(kBlock) void k_cast_in(const TS *in, TB *out, const CastGeo g) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < g.total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t r = i, oi = 0, oo = 0;
        for (int d = 0; d < g.ndim; ++d) {
            int64_t c = r % g.size[d];
            r /= g.size[d];
            oi += c * g.sin[d];
            oo += c * g.sout[d];
        }
        out[oo] = to_buf<TB>(in[oi]);
    }
}

}  // namespace dipamd

