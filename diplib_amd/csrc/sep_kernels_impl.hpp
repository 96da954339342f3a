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
// compile-time constant, and each tap step reads one new sample per side through
// a cursor of the sample source (LDS row, LDS ring column, or global column).
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

// Boundary map of coordinate j outside [0, n), packed as (source index << 1) | negate,
// or -1 for a constant condition.  32-bit and out of line: it runs for edge samples
// only, and returning by value keeps the call free of stack memory.
static __device__ __noinline__ int bc_fold_packed(int j, int n, int bc) {
    switch (bc) {
        case DIPAMD_BC_SYMMETRIC_MIRROR:
        case DIPAMD_BC_ASYMMETRIC_MIRROR: {
            if (n == 1) return bc == DIPAMD_BC_ASYMMETRIC_MIRROR ? 1 : 0;
            const int P = 2 * (n - 1);
            int p = j % P;
            if (p < 0) p += P;
            const int idx = p < n ? p : P - p;
            int neg = 0;
            if (bc == DIPAMD_BC_ASYMMETRIC_MIRROR) {
                const int m = j < 0 ? -j : j - (n - 1);
                neg = ((m - 1) / (n - 1) + 1) & 1;  // odd number of edge reflections crossed
            }
            return (idx << 1) | neg;
        }
        case DIPAMD_BC_PERIODIC:
        case DIPAMD_BC_ASYMMETRIC_PERIODIC: {
            const int q = j >= 0 ? j / n : -((-j - 1) / n) - 1;
            const int neg = (bc == DIPAMD_BC_ASYMMETRIC_PERIODIC && ((q & 1) || n == 1)) ? 1 : 0;
            return ((j - q * n) << 1) | neg;
        }
        case DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE:
            return (j < 0 ? 0 : n - 1) << 1;
        default:
            return -1;
    }
}

// Source index (-1: constant) and sign (*mul = +1 / -1) of coordinate j.
__device__ __forceinline__ int bc_fold(int j, int n, int bc, int *mul) {
    const int v = bc_fold_packed(j, n, bc);
    *mul = (v > 0 && (v & 1)) ? -1 : 1;
    return v < 0 ? -1 : (v >> 1);
}

// Sample j of a line of n samples at `base` with stride `st`, boundary-extended;
// coordinates below `zlo` (past the reference's border) read as 0.
template <typename TI, typename TB>
__device__ __forceinline__ TB fetch_ext(const TI *base, int64_t st, int j, int n, int bc, TB cval,
                                        int zlo = INT32_MIN) {
    if (static_cast<unsigned>(j) < static_cast<unsigned>(n)) return to_buf<TB>(base[static_cast<int64_t>(j) * st]);
    if (j < zlo) return TB(0);
    int mul;
    const int s = bc_fold(j, n, bc, &mul);
    if (s < 0) return cval;
    const TB v = to_buf<TB>(base[static_cast<int64_t>(s) * st]);
    return mul < 0 ? -v : v;
}

// ---------------------------------------------------------------------------
// Window sources: at(q) reads sample q (q = 0: output p = 0's centre);
// rbegin(q)/rnext() stream q, q+1, ...; lbegin(q)/lnext() stream q, q-1, ...
// ---------------------------------------------------------------------------
template <typename TB>
struct LdsRowSrc {  // a staged row in LDS, padded index
    const TB *row;
    int base, ri, li;
    __device__ TB at(int q) const { return row[lpad(base + q)]; }
    __device__ void rbegin(int q) { ri = base + q; }
    __device__ TB rnext() { return row[lpad(ri++)]; }
    __device__ void lbegin(int q) { li = base + q; }
    __device__ TB lnext() { return row[lpad(li--)]; }
};

template <typename TB>
struct RingSrc {  // one column of the LDS ring; rows wrap modulo rc
    const TB *col;  // ring + lpad(column)
    int pitch, rc, span, r0, ro, lo;  // span = rc * pitch; r0 = row of q = 0
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

template <typename TI, typename TB>
struct GColSrc {  // a column in global memory, all samples in range
    const TI *p;
    int64_t st;
    const TI *rp, *lp;
    __device__ TB at(int q) const { return to_buf<TB>(p[q * st]); }
    __device__ void rbegin(int q) { rp = p + q * st; }
    __device__ TB rnext() {
        const TB v = to_buf<TB>(*rp);
        rp += st;
        return v;
    }
    __device__ void lbegin(int q) { lp = p + q * st; }
    __device__ TB lnext() {
        const TB v = to_buf<TB>(*lp);
        lp -= st;
        return v;
    }
};

template <typename TI, typename TB>
struct GColExtSrc {  // a column in global memory near its ends: boundary-extended
    const TI *line;
    int64_t st;
    int z, n, bc;
    TB cval;
    int ri, li;
    int zlo;
    __device__ TB at(int q) const { return fetch_ext<TI, TB>(line, st, z + q, n, bc, cval, zlo); }
    __device__ void rbegin(int q) { ri = z + q; }
    __device__ TB rnext() { return fetch_ext<TI, TB>(line, st, ri++, n, bc, cval, zlo); }
    __device__ void lbegin(int q) { li = z + q; }
    __device__ TB lnext() { return fetch_ext<TI, TB>(line, st, li--, n, bc, cval, zlo); }
};

// P outputs from a window source.  `shift` moves the window: sample q is read as
// q - shift (GENERAL: the reversed origin; the symmetric forms: origin - (ntaps-1),
// zero for a centred filter).
// `taps`: any pointer to the taps (generic, or the constant address space for taps that
// are to be read with scalar loads).
template <typename TB, int P, int SYM, bool FAST, typename Src, typename TapP>
__device__ __forceinline__ void fir_window(TapP taps, int ntaps, int shift, Src &s, TB (&acc)[P]) {
    if constexpr (SYM == SYM_EVEN || SYM == SYM_ODD) {
        const int r = ntaps - 1;
        TB rw[P], lw[P];
        TB w[P + 2];
        // the outer two samples feed the k = 1 pairs: a single-tap filter (r = 0, border
        // 0) must not read them, they lie outside its window
#pragma unroll
        for (int q = 0; q < P + 2; ++q) w[q] = (q > 0 && q < P + 1) || r > 0 ? s.at(q - 1 - shift) : TB(0);
        const TB f0 = taps[0];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc[p] = f0 * w[p + 1];
            rw[(p + 1) % P] = w[p + 2];
            lw[(p - 1 + P) % P] = w[p];
        }
        s.rbegin(P + 1 - shift);
        s.lbegin(-2 - shift);
        // tap step k (slot indices fold to constants once the u loops unroll)
        auto step = [&](const int u, const int k) {
            const TB f = taps[k];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const TB a = rw[(p + 1 + u) % P];
                const TB b = lw[((p - 1 - u) % P + P) % P];
                const TB t = (SYM == SYM_EVEN) ? (a + b) : (a - b);
                acc[p] = mac<TB, FAST>(acc[p], f, t);
            }
        };
        auto advance = [&](const int u) {
            rw[(1 + u) % P] = s.rnext();
            lw[((P - 2 - u) % P + P) % P] = s.lnext();
        };
        int kb = 1;
        for (; kb + P <= r; kb += P) {  // whole blocks: every step also reads the next pair
#pragma unroll
            for (int u = 0; u < P; ++u) {
                step(u, kb + u);
                advance(u);
            }
        }
#pragma unroll
        for (int u = 0; u < P; ++u) {  // the last (partial) block
            const int k = kb + u;
            if (k <= r) {
                step(u, k);
                if (k < r) advance(u);
            }
        }
    } else if constexpr (SYM == SYM_GENERAL) {
        TB win[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc[p] = TB(0);
            win[p] = s.at(p - shift);
        }
        s.rbegin(P - shift);
        auto step = [&](const int u, const int k) {
            const TB f = taps[k];
#pragma unroll
            for (int p = 0; p < P; ++p) acc[p] = mac<TB, FAST>(acc[p], f, win[(p + u) % P]);
        };
        int kb = 0;
        for (; kb + P < ntaps; kb += P) {
#pragma unroll
            for (int u = 0; u < P; ++u) {
                step(u, kb + u);
                win[u % P] = s.rnext();
            }
        }
#pragma unroll
        for (int u = 0; u < P; ++u) {
            const int k = kb + u;
            if (k < ntaps) {
                step(u, k);
                if (k + 1 < ntaps) win[u % P] = s.rnext();
            }
        }
    } else {  // D_EVEN / D_ODD: s += f_k*(x[i+k] +- x[i-k-1]), k = 0..ntaps-1
        TB rw[P], lw[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            acc[p] = TB(0);
            rw[p] = s.at(p - shift);
            lw[(p - 1 + P) % P] = s.at(p - 1 - shift);
        }
        s.rbegin(P - shift);
        s.lbegin(-2 - shift);
        auto step = [&](const int u, const int k) {
            const TB f = taps[k];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const TB a = rw[(p + u) % P];
                const TB b = lw[((p - 1 - u) % P + P) % P];
                const TB t = (SYM == SYM_D_EVEN) ? (a + b) : (a - b);
                acc[p] = mac<TB, FAST>(acc[p], f, t);
            }
        };
        auto advance = [&](const int u) {
            rw[u % P] = s.rnext();
            lw[((P - 2 - u) % P + P) % P] = s.lnext();
        };
        int kb = 0;
        for (; kb + P < ntaps; kb += P) {
#pragma unroll
            for (int u = 0; u < P; ++u) {
                step(u, kb + u);
                advance(u);
            }
        }
#pragma unroll
        for (int u = 0; u < P; ++u) {
            const int k = kb + u;
            if (k < ntaps) {
                step(u, k);
                if (k + 1 < ntaps) advance(u);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// K1: line pass along the fastest axis.  One wave per line segment of 64*P
// outputs, 4 lines per workgroup; the segment plus `border` samples each side
// is staged in LDS (converted to the buffer type, boundary folded).
// ---------------------------------------------------------------------------
template <typename TI, typename TB, int P, int SYM, bool FAST>
__global__ __launch_bounds__(kBlock) void k_line(const PassArgs<TB> a, int64_t nseg, const TI *__restrict__ in,
                                                  TB *__restrict__ out) {
    constexpr int TX = 64 * P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TB *lds = reinterpret_cast<TB *>(smem_raw);
    const int B = a.border;
    const int W = TX + 2 * B;
    const int pitch = lpad_size(W);
    const int wave = wave_id(), lane = threadIdx.x & 63;
    const int64_t bid = xcd_block();
    const int64_t seg = bid % nseg;
    const int64_t line = (bid / nseg) * 4 + wave;
    const int x0 = static_cast<int>(seg) * TX;
    const int n = static_cast<int>(a.n);
    TB *row = lds + wave * pitch;
    int64_t oin = 0, oout = 0;
    const bool active = line < a.nlines;
    if (active) {
        outer_offsets(a.outer, line, &oin, &oout);
        const TI *src = in + oin;
        const bool interior = (x0 - B >= 0) && (x0 + TX + B <= n);
        if (interior && a.sin == 1) {
            const TI *s0 = src + (x0 - B);
            for (int i = lane; i < W; i += 64) row[lpad(i)] = to_buf<TB>(s0[i]);
        } else {
            for (int i = lane; i < W; i += 64) row[lpad(i)] = fetch_ext<TI, TB>(src, a.sin, x0 - B + i, n, a.bc, a.cval, a.zlo);
        }
    }
    lds_barrier();
    if (!active) return;
    LdsRowSrc<TB> s{row, B + lane * P, 0, 0};
    TB acc[P];
    fir_window<TB, P, SYM, FAST>(a.taps, a.ntaps, a.shift, s, acc);
    TB *dst = out + oout;
    const int xs = x0 + lane * P;
    if (a.sout == 1 && xs + P <= n) {
        TB *d = dst + xs;
#pragma unroll
        for (int p = 0; p < P; ++p) d[p] = acc[p];
    } else {
        TB *d = dst + static_cast<int64_t>(xs) * a.sout;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (xs + p < n) *d = acc[p];
            d += a.sout;
        }
    }
}

// ---------------------------------------------------------------------------
// K2: sweep along a strided axis.  One thread per position of the lane axis
// (coalesced across the wave), each running down a chunk of the filter axis P
// outputs at a time with the register window streaming from global memory (the
// window rows are re-read from L1/L2; HBM sees each row once).  No LDS, no
// barriers.
// ---------------------------------------------------------------------------
template <typename TI, typename TB, int P, int SYM, bool FAST>
__global__ __launch_bounds__(kBlock) void k_sweep(const PassArgs<TB> a, int64_t ntiles, int64_t nchunks,
                                                   const TI *__restrict__ in, TB *__restrict__ out) {
    // tile fastest in hardware order: the blocks of neighbouring chunks of a column are
    // ntiles apart, dispatched later (their shared halo rows are then L2 hits, not two
    // concurrent misses) and, for ntiles a multiple of 8, on the same XCD
    const int64_t bid = blockIdx.x;
    const int64_t tile = bid % ntiles;
    const int64_t rest = bid / ntiles;
    const int64_t chunk = rest % nchunks;
    const int64_t outer = rest / nchunks;
    const int64_t c = tile * blockDim.x + threadIdx.x;
    if (c >= a.e) return;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    const TI *src = in + oin + c * a.ein;
    TB *dst = out + oout + c * a.eout;
    const int n = static_cast<int>(a.n);
    const int z0 = static_cast<int>(chunk * a.chunk);
    const int zend = static_cast<int>(min(chunk * a.chunk + a.chunk, a.n));
    const int64_t sin = a.sin;
    for (int zb = z0; zb < zend; zb += P) {
        TB acc[P];
        if (zb - a.border >= 0 && zb + P + a.border <= n) {
            GColSrc<TI, TB> s{src + zb * sin, sin, nullptr, nullptr};
            fir_window<TB, P, SYM, FAST>(a.taps, a.ntaps, a.shift, s, acc);
        } else {
            GColExtSrc<TI, TB> s{src, sin, zb, n, a.bc, a.cval, 0, 0, a.zlo};
            fir_window<TB, P, SYM, FAST>(a.taps, a.ntaps, a.shift, s, acc);
        }
        TB *d = dst + static_cast<int64_t>(zb) * a.sout;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (zb + p < zend) *d = acc[p];
            d += a.sout;
        }
    }
}

// ---------------------------------------------------------------------------
// K2T: column-tile sweep along a strided axis.  A workgroup stages a block of 64
// lane-axis positions x (R + 2*border) filter-axis samples in LDS, column-major
// with an odd pitch (lanes = columns hit distinct banks), then each thread runs
// the window down its column for R/4 outputs, P at a time, reading LDS at
// compile-time offsets.  One barrier per block; stores are coalesced along the
// lane axis.  Used for long filters, where the ring of the fused kernel is large.
// ---------------------------------------------------------------------------
template <typename TB>
struct LdsColSrc {  // one staged column in LDS, unit stride
    const TB *col;
    int base, ri, li;
    __device__ TB at(int q) const { return col[base + q]; }
    __device__ void rbegin(int q) { ri = base + q; }
    __device__ TB rnext() { return col[ri++]; }
    __device__ void lbegin(int q) { li = base + q; }
    __device__ TB lnext() { return col[li--]; }
};

template <typename TI, typename TB, int P, int SYM, bool FAST, int NW>
__global__ __launch_bounds__(64 * NW) void k_colsweep(const PassArgs<TB> a, int64_t ntiles, int64_t nchunks,
                                                       const TI *__restrict__ in, TB *__restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TB *lds = reinterpret_cast<TB *>(smem_raw);
    // tile fastest in hardware order: the blocks of neighbouring chunks of a column are
    // ntiles apart, dispatched later (their shared halo rows are then L2 hits, not two
    // concurrent misses) and, for ntiles a multiple of 8, on the same XCD
    const int64_t bid = blockIdx.x;
    const int64_t tile = bid % ntiles;
    const int64_t rest = bid / ntiles;
    const int64_t chunk = rest % nchunks;
    const int64_t outer = rest / nchunks;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    const int B = a.border;
    const int R = static_cast<int>(a.chunk);  // outputs per column in this block (multiple of P)
    const int H = R + 2 * B;                   // staged samples per column
    const int cpitch = H | 1;
    const int n = static_cast<int>(a.n);
    const int z0 = static_cast<int>(chunk) * R;
    const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int64_t c = tile * 64 + lane;
    const bool cok = c < a.e;
    const TI *src = in + oin + (cok ? c : 0) * a.ein;
    TB *col = lds + lane * cpitch;
    // stage: wave `grp` loads samples grp, grp+NW, ... of every column (rows are
    // coalesced across the lanes), eight loads in flight before their LDS stores
    const bool interior = z0 - B >= 0 && z0 + R + B <= n;
    for (int h0 = grp; h0 < H; h0 += 8 * NW) {
        TB v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int h = h0 + NW * i;
            if (h < H) {
                const int j = z0 - B + h;
                v[i] = interior ? to_buf<TB>(src[static_cast<int64_t>(j) * a.sin])
                                : fetch_ext<TI, TB>(src, a.sin, j, n, a.bc, a.cval, a.zlo);
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int h = h0 + NW * i;
            if (h < H) col[h] = v[i];
        }
    }
    lds_barrier();
    if (!cok) return;
    TB *dst = out + oout + c * a.eout;
    const int zend = min(z0 + R, n);
    for (int blk = grp * P; blk < R; blk += NW * P) {  // wave `grp` takes blocks of P rows grp, grp+NW, ...
        const int zs = z0 + blk;
        if (zs >= zend) break;
        LdsColSrc<TB> s{col, B + blk, 0, 0};
        TB acc[P];
        fir_window<TB, P, SYM, FAST>(a.taps, a.ntaps, a.shift, s, acc);
        TB *d = dst + static_cast<int64_t>(zs) * a.sout;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (zs + p < zend) *d = acc[p];
            d += a.sout;
        }
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
// Next step's input rows are prefetched into registers during the compute.
// Column offsets of the staged row (boundary-folded) are computed once per tile.
// ---------------------------------------------------------------------------
template <typename TB>
struct Tile {
    const void *in;   // plane origin
    TB *out;
    int nx, ny;
    int64_t sxin, syin, sxout, syout;
    int c0, z0, zend;  // first column, first and end output rows
    int rx, ry, bcx, bcy;
    TB cvalx, cvaly;
    int vec_ok;  // rows allow vector loads (aligned origin and row stride; fused_rc.hpp)
};

// LDS row pitch (words) of a staged or ring row: padded, and congruent to P/2
// modulo P.  Sixteen lanes reading with a stride of P words cover four runs of
// P/2... banks; a row offset of P/2 (mod P) moves the next row onto the runs in between,
// so the two rows a half-wave reads hit disjoint banks.
__host__ __device__ constexpr int fused_pitch(int n, int P) {
    return ((lpad_size(n) - P / 2 + P - 1) / P) * P + P / 2;
}
__host__ __device__ constexpr int fused_ring_rows(int ry) { return ((2 * ry + 15) / 16 + 1) * 16; }

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST>
__device__ __forceinline__ void fused_tile(const Tile<TB> &t, const TI *__restrict__ src, TB *__restrict__ dst,
                                           TB *lds, const TB *tx, const TB *ty) {
    constexpr int S = 16;
    constexpr int XT = TXC / P;            // x-stage lanes per row (16)
    static_assert(XT * S == kBlock, "tile shape");
    constexpr int NLD = (TXC + 2 * (kFusedCap - 1) + 63) / 64;  // loads per staged row per lane (max)
    const int Bx = t.rx, By = t.ry;
    const int Wx = TXC + 2 * Bx;
    const int pitchS = fused_pitch(Wx, P);
    const int pitchR = fused_pitch(TXC, P);
    const int A = (2 * By + S - 1) / S;
    const int RC = (A + 1) * S;
    TB *stg = lds;
    TB *ring = lds + S * pitchS;
    const int c0 = t.c0, z0 = t.z0, zend = t.zend;
    const int nsteps = (zend - z0 + S - 1) / S;
    const int M = A + nsteps;
    const int tid = threadIdx.x;
    const int wave = wave_id(), lane = tid & 63;
    const int nld = (Wx + 63) / 64;

    // boundary-folded column offsets of this lane's staged samples (fixed per tile)
    int cofs[NLD];
    int cmul[NLD];  // +1, -1, or 0 for the constant condition
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
        const int col = lane + 64 * i;
        int x = c0 - Bx + col;
        int m = 1;
        if (col < Wx && (x < 0 || x >= t.nx)) {
            x = bc_fold(x, t.nx, t.bcx, &m);
            if (x < 0) {
                x = 0;
                m = 0;
            }
        }
        if (col >= Wx) x = 0;
        cofs[i] = x * static_cast<int>(t.sxin);
        cmul[i] = m;
    }

    // --- staged-row loads: wave w loads rows w, w+4, w+8, w+12 of each step ---
    // Raw samples, loaded unconditionally (offsets stay in the row) and converted only
    // when they are written to LDS one step later, so the loads stay in flight.
    TI pre[4][NLD];
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
#pragma unroll
            for (int i = 0; i < NLD; ++i) pre[rr][i] = rowp[cofs[i]];
        }
    };

    load_step(0);
    for (int m = 0; m < M; ++m) {
        lds_barrier();  // previous x stage done with `stg`
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            TB *srow = stg + (wave + 4 * rr) * pitchS;
#pragma unroll
            for (int i = 0; i < NLD; ++i) {
                const int col = lane + 64 * i;
                if (i < nld && col < Wx) {
                    const TB v = to_buf<TB>(pre[rr][i]);
                    srow[lpad(col)] = cmul[i] > 0 ? v : (cmul[i] < 0 ? -v : t.cvalx);
                }
            }
        }
        lds_barrier();  // `stg` full; previous y stage done with the ring slots we overwrite
        if (m + 1 < M) load_step(m + 1);

        // --- x stage: row xr of this step, outputs xl*P .. xl*P+P-1 ---
        {
            const int xr = tid / XT, xl = tid % XT;
            const int j = z0 - By + m * S + xr;
            int mul = 1;
            bool isc = false;
            // rows outside the plane only in the first and last steps (a uniform test)
            const int jstep = z0 - By + m * S;
            if (jstep < 0 || jstep + S > t.ny) {
                if (j < 0 || j >= t.ny) isc = bc_fold(j, t.ny, t.bcy, &mul) < 0;
            }
            TB *rrow = ring + ((m * S + xr) % RC) * pitchR;
            if (isc) {
#pragma unroll
                for (int p = 0; p < P; ++p) rrow[lpad(xl * P + p)] = t.cvaly;
            } else {
                LdsRowSrc<TB> s{stg + xr * pitchS, Bx + xl * P, 0, 0};
                TB acc[P];
                if constexpr (SX == SYM_NONE) {
#pragma unroll
                    for (int p = 0; p < P; ++p) acc[p] = s.at(p);
                } else {
                    fir_window<TB, P, SX, FAST>(tx, Bx + 1, 0, s, acc);
                }
#pragma unroll
                for (int p = 0; p < P; ++p) rrow[lpad(xl * P + p)] = mul < 0 ? -acc[p] : acc[p];
            }
        }
        lds_barrier();  // ring rows of this step written

        // --- y stage: step n = m - A ---
        const int n = m - A;
        if (n >= 0) {
            // a wave lies within one row group (TXC >= 64): taken from the wave index, the
            // ring row offsets are wave-uniform and their wrap-around is scalar work
            const int col = tid % TXC, grp = (wave * 64) / TXC;
            const int c = c0 + col;
            const int rel0 = n * S + grp * P + By;  // ring-relative row of output p = 0
            RingSrc<TB> s{ring + lpad(col), pitchR, RC, RC * pitchR, rel0 % RC, 0, 0};
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

template <typename TI, typename TB, int TXC, int P, int SX, int SY, bool FAST>
__global__ __launch_bounds__(kBlock) void k_fused2d(const FusedArgs<TB> a, int64_t ntiles, int64_t nchunks) {
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
    t.sxin = a.sxin;
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
    t.vec_ok = 0;
    fused_tile<TI, TB, TXC, P, SX, SY, FAST>(t, static_cast<const TI *>(t.in), t.out, lds, a.tx, a.ty);
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
        lds_barrier();
        const int item = s_item;
        if (item >= a.nitems) break;  // uniform: every wave of the workgroup leaves
        const BatchItem it = a.items[item];
        const BatchImage<TB> im = a.images[it.image];
        const int ntx = im.rx + 1, nty = im.ry + 1;
        for (int k = threadIdx.x; k < ntx + nty; k += blockDim.x) s_taps[k] = a.taps[im.tap_offset + k];
        lds_barrier();  // taps visible; s_item read by all before it is rewritten
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
        t.vec_ok = 0;
        fused_tile<TI, TB, TXC, P, SYM_EVEN, SYM_EVEN, FAST>(t, static_cast<const TI *>(t.in), t.out, lds, s_taps,
                                                             s_taps + ntx);
        lds_barrier();  // LDS (ring, taps) free before the next item
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

