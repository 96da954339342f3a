// Fused backward pass of the separable Gaussian over stacks of 2-D planes.
//
// The forward pass is y = A_y A_x x with A_d = C_d E_d (E_d: boundary extension
// along axis d, C_d: the symmetric FIR on the extended line).  For a loss L with
// g = dL/dy, one pass over x and g gives
//     dL/dx         = A_x^T h,            h  = A_y^T g
//     dL/dsigma_x   = < h , A'_x x >       (A'_d: the filter with d(taps)/d(sigma))
//     dL/dsigma_y   = < h', A_x x >,      h' = A'_y^T g
// where the adjoint of an extended filter is the correlation with reversed taps over
// the zero-padded line plus, for every halo position j, its output added to the pixel
// the extension copied into j (with the extension's sign); for periodic conditions
// it is the same filter with reversed taps on the periodically extended line.
//
// A workgroup owns 128 staged columns (a tile of 128 - 2R output columns plus an x
// halo of R each side) and a chunk of rows, walked in steps of 16 rows.  Per step it
// stages g (with a y halo of R rows, column-major) and x in LDS; the y stage gives
// each thread one column and 8 rows: it loads the 8 + 2R samples of its window into
// registers once and runs both adjoint y filters (h, h'), adding the folded halo
// terms near the top and bottom of the plane; the x stage gives each thread 8
// consecutive columns of one row and runs the adjoint x filter along h (dL/dx, with
// folded halo terms near the left and right edges) and the forward x filter and its
// derivative along x (for the dot products) from two register windows.
//
// RX, RY are compile-time radius classes of the x and y filters (taps zero-padded up
// to them), so every window offset is an immediate.  Rows of h, h' and x are stored P-interleaved: sample i at
// (i % 8) * LP + i / 8, so the eight window samples of consecutive lanes are
// consecutive words (no bank conflicts) at offsets that depend on the tap only.
// Arithmetic is fused multiply-add: gradients are checked against FP64 references.
#pragma once

#include <cstddef>
#include <type_traits>

#include "sep_kernels_impl.hpp"

namespace dipamd {

namespace {

// Folded halo terms at target c of a line of n >= 2r+1 samples (no bounce): the sum,
// over the halo positions l the extension copies from c, of sign(l) * T[l], where
// T[l] is the adjoint correlation (reversed taps ct, symmetric form SYM) at l over
// the zero-padded line.  Mirror: l = -c (1 <= c <= r) and l = 2(n-1) - c
// (n-1-r <= c <= n-2), sign -1 for the antisymmetric mirror; zero order: every
// l < 0 onto c = 0 and every l > n-1 onto c = n-1.  At l = -t only samples k - t
// (k >= t) are inside the line, at l = n-1+t only n-1+t-k (negated for odd filters).
// The mirror sums run over k = 1..RMAX (the compile-time radius class) with the terms
// outside c..r selected away, so the reads are independent of each other and of the
// sums (all in flight at once) instead of a chain of dependent loads.
template <int SYM, int RMAX, typename F>
__device__ __forceinline__ float fold_sum(const float *ct, int r, int bc, int n, int c, F src) {
    float acc = 0.f;
    if (bc == DIPAMD_BC_SYMMETRIC_MIRROR || bc == DIPAMD_BC_ASYMMETRIC_MIRROR) {
        const float sg = bc == DIPAMD_BC_ASYMMETRIC_MIRROR ? -1.f : 1.f;
        if (c >= 1 && c <= r) {
            float t = 0.f;
#pragma unroll
            for (int k = 1; k <= RMAX; ++k) {
                const bool in = k >= c && k <= r;
                const float v = src(in ? k - c : 0);
                t = __builtin_fmaf(ct[k], in ? v : 0.f, t);
            }
            acc = __builtin_fmaf(sg, t, acc);
        }
        if (c >= n - 1 - r && c <= n - 2) {
            const int tt = n - 1 - c;
            float t = 0.f;
#pragma unroll
            for (int k = 1; k <= RMAX; ++k) {
                const bool in = k >= tt && k <= r;
                const float v = src(in ? n - 1 + tt - k : n - 1);
                t = __builtin_fmaf(ct[k], in ? v : 0.f, t);
            }
            acc = __builtin_fmaf(SYM == SYM_EVEN ? sg : -sg, t, acc);
        }
    } else if (bc == DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE) {
        if (c == 0) {
            for (int tt = 1; tt <= r; ++tt)
                for (int k = tt; k <= r; ++k) acc = __builtin_fmaf(ct[k], src(k - tt), acc);
        }
        if (c == n - 1) {
            float t = 0.f;
            for (int tt = 1; tt <= r; ++tt)
                for (int k = tt; k <= r; ++k) t = __builtin_fmaf(ct[k], src(n - 1 + tt - k), t);
            acc += SYM == SYM_EVEN ? t : -t;
        }
    }
    return acc;
}

template <int SYM>
__device__ __forceinline__ float pair(float hi, float lo) {
    return SYM == SYM_EVEN ? hi + lo : hi - lo;
}

// Tap arrays are read through a pointer the compiler cannot see through, renewed in
// each stage of each step: the taps are then loaded (scalar loads from the constant
// address space) where a stage needs them instead of all five arrays being hoisted
// out of the loop, which would exceed the SGPRs and spill them into VGPR lanes.
// The pointer is formed from the kernel-argument segment (the BwdArgs argument sits at
// offset 0), so the argument itself never has to be addressable.
using ctap = const __attribute__((address_space(4))) float;
__device__ __forceinline__ ctap *taps_here(size_t offset) {
    const __attribute__((address_space(4))) char *seg =
        (const __attribute__((address_space(4))) char *)__builtin_amdgcn_kernarg_segment_ptr();
    ctap *q = reinterpret_cast<ctap *>(seg + offset);
    asm volatile("" : "+s"(q));
    return q;
}

// Raw buffer resource over `bytes` bytes from `base` (no stride, no swizzle; the last
// word is the data format CDNA expects for 32-bit raw accesses).  Every access made
// through it lies inside the plane it describes.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float *base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), static_cast<short>(0), bytes, 0x00020000);
}
__device__ __forceinline__ float buffer_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void buffer_st(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}

}  // namespace

template <int RX, int RY>
struct BwdGeom {
    static constexpr int S = kBwdStep;        // rows per step
    static constexpr int P = 8;               // outputs per thread and window
    static constexpr int W = 128;             // staged columns
    static constexpr int TXC = W - 2 * RX;    // output columns per tile
    static constexpr int GH = S + 2 * RY;     // staged g rows
    static constexpr int Hp = GH | 1;         // column pitch (column-major g)
    static constexpr int NWX = P + 2 * RX;    // x-stage window length
    static constexpr int NWY = P + 2 * RY;    // y-stage window length
    static constexpr int LP = ((W + 2 * RX + P) / P) | 1;  // interleave line length
    static constexpr int GROWS = (GH + 3) / 4;
    static constexpr int FOLDW = 2 * (RX + 1);  // x fold terms per row: targets 0..RX, nx-1-RX..nx-1
    // Row pitch of the interleaved rows: the x stage's wave reads four rows at once,
    // sixteen consecutive words from each; a pitch of 16 (mod 32) words puts the four
    // runs on disjoint banks (32 or 64 banks).  Unpadded when the padding would exceed
    // 64 KB or leave fewer workgroups resident on a CU (160 KB of LDS) than without it:
    // occupancy matters more than the conflicts.
    static constexpr int ROWF0 = P * LP;
    static constexpr int ROWFP = ((ROWF0 - 16 + 31) / 32) * 32 + 16;  // smallest >= ROWF0, = 16 mod 32
    static constexpr size_t FIXED = static_cast<size_t>(W) * Hp + static_cast<size_t>(S) * FOLDW;
    static constexpr size_t LDS_P = (FIXED + 3 * static_cast<size_t>(S) * ROWFP) * sizeof(float);
    static constexpr size_t LDS_U = (FIXED + 3 * static_cast<size_t>(S) * ROWF0) * sizeof(float);
    static constexpr size_t kCuLds = 160 * 1024;
    static constexpr bool PAD = LDS_P <= 65536 && kCuLds / LDS_P >= kCuLds / LDS_U;
    static constexpr int ROWF = PAD ? ROWFP : ROWF0;  // floats per interleaved row
    static constexpr size_t lds_floats =
        static_cast<size_t>(W) * Hp + 3 * static_cast<size_t>(S) * ROWF + static_cast<size_t>(S) * FOLDW;
};

// Interleaved position of staged sample i in a row of P = 8 lines of LP words.
template <int LP>
__device__ __forceinline__ int ipos_lp(int i) {
    return (i & 7) * LP + (i >> 3);
}

template <int SX, int SY, int RX, int RY>
__global__ __launch_bounds__(kBlock) void k_bwd2d(const BwdArgs a) {
    using Gm = BwdGeom<RX, RY>;
    constexpr int S = Gm::S, P = Gm::P, W = Gm::W, TXC = Gm::TXC, GH = Gm::GH, Hp = Gm::Hp, NWX = Gm::NWX, NWY = Gm::NWY,
                  LP = Gm::LP, ROWF = Gm::ROWF, GROWS = Gm::GROWS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *gbuf = reinterpret_cast<float *>(smem_raw);  // W columns x Hp
    float *hbuf = gbuf + W * Hp;                        // S interleaved rows
    float *h2buf = hbuf + S * ROWF;
    float *xbuf = h2buf + S * ROWF;
    float *foldx = xbuf + S * ROWF;  // S rows x FOLDW: folded x halo terms of the targets
    const int rx = a.rx, ry = a.ry, nx = a.nx, ny = a.ny;

    const int64_t bid = xcd_block();
    const int tile = static_cast<int>(bid % a.ntiles);
    const int64_t rest = bid / a.ntiles;
    const int chunk = static_cast<int>(rest % a.nchunks);
    const int64_t plane = rest / a.nchunks;
    int64_t ox = 0, odx = 0, og = 0;
    {
        int64_t idx = plane;
        for (int d = 0; d < a.outer.n; ++d) {
            const int64_t c = idx % a.outer.size[d];
            idx /= a.outer.size[d];
            ox += c * a.outer.sin[d];
            odx += c * a.outer.sout[d];
            og += c * a.gplane[d];
        }
    }
    const float *X = a.x + ox;
    const float *G = a.g + og;
    float *DX = a.dx + odx;
    const int c0 = tile * TXC;        // image column of output column 0
    const int cs0 = c0 - RX;          // image column of staged column 0
    int z0, zend;
    if (a.ytop > 0) {
        if (chunk == 0) {
            z0 = 0;
            zend = a.ytop;
        } else if (chunk == a.nchunks - 1) {
            z0 = a.ybot;
            zend = ny;
        } else {
            z0 = a.ytop + (chunk - 1) * a.chunk;
            zend = min(z0 + a.chunk, a.ybot);
        }
    } else {
        z0 = chunk * a.chunk;
        zend = min(z0 + a.chunk, ny);
    }
    const int tid = threadIdx.x, lane = tid & 63, wave = wave_id();

    // Global memory is read and written through raw buffer resources over the plane:
    // a lane's column is a loop-invariant byte offset and a row's offset a scalar, so
    // no per-access 64-bit address arithmetic.
    const __amdgpu_buffer_rsrc_t gres = plane_rsrc(G, a.gbytes);
    const __amdgpu_buffer_rsrc_t xres = plane_rsrc(X, a.xbytes);
    const int syg4 = static_cast<int>(a.syg) * 4, syx4 = static_cast<int>(a.syx) * 4;

    // staged column slots of this lane (lane, lane + 64): source column (byte offset)
    // and factor of g (the adjoint's extension) and of x (the forward extension; a
    // constant column takes the constant's bits instead of the sample's)
    int gvo[2], xvo[2];
    float gmul[2], xmul[2];
    uint32_t xkeep[2], xcbits[2];
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
        const int j = cs0 + lane + 64 * sl;
        int gj = j, xj = j, m;
        float gm = 1.f, xm = 1.f;
        uint32_t keep = 0xffffffffu, cb = 0u;
        if (j < 0 || j >= nx) {
            if (a.modex == kBwdPeriodic) {
                gj = bc_fold(j, nx, a.bcx, &m);
                gm = static_cast<float>(m);
            } else {
                gj = 0;
                gm = 0.f;
            }
            xj = bc_fold(j, nx, a.bcx, &m);
            if (xj < 0) {
                xj = 0;
                keep = 0u;
                cb = __builtin_bit_cast(uint32_t, a.cvalx);
            } else {
                xm = static_cast<float>(m);
            }
        }
        gvo[sl] = gj * 4;
        xvo[sl] = xj * 4;
        gmul[sl] = gm;
        xmul[sl] = xm;
        xkeep[sl] = keep;
        xcbits[sl] = cb;
    }
    // no staged column outside the line: the tile's g and x go to LDS unchanged
    const bool tile_plain = cs0 >= 0 && cs0 + W <= nx;
    // x stage: this thread's row of the step, its first output column, how many of its
    // P outputs lie in the tile and the line, and dL/dx's byte offset of output 0
    const int xrow = tid >> 4, xl = tid & 15;
    const int cb = c0 + 8 * xl;
    const int nv = max(0, min(P, min(TXC, nx - c0) - 8 * xl));
    const __amdgpu_buffer_rsrc_t dres = plane_rsrc(DX, a.dxbytes);
    const int sydx4 = static_cast<int>(a.sydx) * 4;
    const int dvo = xrow * sydx4 + cb * 4;

    // register prefetch of one step (raw values; factors applied at the LDS store).
    // Rows are wave-uniform: their folds and offsets are scalar.
    float pg[GROWS][2], px[4][2];
    float pgm[GROWS];
    auto load_step = [&](int i0) {
#pragma unroll
        for (int rr = 0; rr < GROWS; ++rr) {
            const int hrow = wave + 4 * rr;
            if (rr < GROWS - 1 || hrow < GH) {
                int j = i0 - RY + hrow;
                float rm = 1.f;
                if (j < 0 || j >= ny) {
                    if (a.modey == kBwdPeriodic) {
                        int m;
                        j = __builtin_amdgcn_readfirstlane(bc_fold(j, ny, a.bcy, &m));
                        rm = static_cast<float>(__builtin_amdgcn_readfirstlane(m));
                    } else {
                        j = 0;
                        rm = 0.f;
                    }
                }
                pgm[rr] = rm;
                const int so = j * syg4;
#pragma unroll
                for (int sl = 0; sl < 2; ++sl) pg[rr][sl] = buffer_ld(gres, gvo[sl], so);
            }
        }
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            int j = i0 + wave + 4 * rr;
            j = j < ny ? j : ny - 1;
            const int so = j * syx4;
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) px[rr][sl] = buffer_ld(xres, xvo[sl], so);
        }
    };

    double dsx = 0.0, dsy = 0.0;
    const int nsteps = (zend - z0 + S - 1) / S;
    // Blocks with no folded halo terms (neither an edge tile nor a chunk holding rows
    // within ry of the top or bottom, or conditions without folds) run a copy of the
    // step loop without the fold code: its branches, registers and barriers stay out of
    // the common path.  Rows past zend are computed but never stored.
    const bool fold_x = a.modex == kBwdFold && (c0 <= rx || c0 + TXC - 1 >= nx - 1 - rx);
    const bool fold_y = a.modey == kBwdFold && (z0 <= ry || zend - 1 >= ny - 1 - ry);
    auto steps = [&](auto fold_tag) {
    constexpr bool FOLD = decltype(fold_tag)::value;
    load_step(z0);
    for (int st = 0; st < nsteps; ++st) {
        const int i0 = z0 + st * S;
        // --- stage g (column-major) and x (interleaved rows) ---
        const bool plain = tile_plain && i0 - RY >= 0 && i0 + S + RY <= ny;
        if (plain) {
#pragma unroll
            for (int rr = 0; rr < GROWS; ++rr) {
                const int hrow = wave + 4 * rr;
                if (rr < GROWS - 1 || hrow < GH) {
#pragma unroll
                    for (int sl = 0; sl < 2; ++sl) gbuf[(lane + 64 * sl) * Hp + hrow] = pg[rr][sl];
                }
            }
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                float *xr = xbuf + (wave + 4 * rr) * ROWF;
#pragma unroll
                for (int sl = 0; sl < 2; ++sl) xr[ipos_lp<LP>(lane + 64 * sl)] = px[rr][sl];
            }
        } else {
#pragma unroll
            for (int rr = 0; rr < GROWS; ++rr) {
                const int hrow = wave + 4 * rr;
                if (rr < GROWS - 1 || hrow < GH) {
#pragma unroll
                    for (int sl = 0; sl < 2; ++sl)
                        gbuf[(lane + 64 * sl) * Hp + hrow] = pg[rr][sl] * gmul[sl] * pgm[rr];
                }
            }
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                float *xr = xbuf + (wave + 4 * rr) * ROWF;
#pragma unroll
                for (int sl = 0; sl < 2; ++sl) {
                    const uint32_t b = __builtin_bit_cast(uint32_t, px[rr][sl] * xmul[sl]);
                    xr[ipos_lp<LP>(lane + 64 * sl)] = __builtin_bit_cast(float, (b & xkeep[sl]) | xcbits[sl]);
                }
            }
        }
        lds_barrier();  // staged; the previous step's x stage is done with hbuf / h2buf
        if (st + 1 < nsteps) load_step(i0 + S);

        // --- y stage: column tid % 128, rows 8 * (tid / 128) .. +7 of the step ---
        {
            const int col = tid & (W - 1), rg = tid >> 7;  // rg is wave-uniform
            const float *gc = gbuf + col * Hp + rg * P;
            float v[NWY];
#pragma unroll
            for (int q = 0; q < NWY; ++q) v[q] = gc[q];
            float h[P], h2[P];
            ctap *ty = taps_here(offsetof(BwdArgs, tya));
            ctap *tdy = taps_here(offsetof(BwdArgs, tdya));
#pragma unroll
            for (int p = 0; p < P; ++p) {
                float s1 = ty[0] * v[p + RY], s2 = tdy[0] * v[p + RY];
#pragma unroll
                for (int k = 1; k <= RY; ++k) {
                    const float t = pair<SY>(v[p + RY + k], v[p + RY - k]);
                    s1 = __builtin_fmaf(ty[k], t, s1);
                    s2 = __builtin_fmaf(tdy[k], t, s2);
                }
                h[p] = s1;
                h2[p] = s2;
            }
            const int r0 = i0 + rg * P;
            const int ip = ipos_lp<LP>(col);
#pragma unroll
            for (int p = 0; p < P; ++p) {
                hbuf[(rg * P + p) * ROWF + ip] = h[p];
                h2buf[(rg * P + p) * ROWF + ip] = h2[p];
            }
        }
        lds_barrier();  // h, h' ready
        // --- folded halo terms (edge steps and edge tiles only; uniform branches) ---
        // y: halo rows fold onto rows within ry of the top / bottom edge; the rows they
        // read are in this step's staging.  Thread: column tid % 128, rows rg, rg+2, ...
        const bool yfold = FOLD && a.modey == kBwdFold && (i0 <= ry || i0 + S - 1 >= ny - 1 - ry);
        if (yfold) {
            const int col = tid & (W - 1);
            const float *gcol0 = gbuf + col * Hp;
            auto gsrc = [&](int q) { return gcol0[q - (i0 - RY)]; };
            for (int row = tid >> 7; row < S; row += 2) {
                const int t = i0 + row;
                if (t >= ny || (t > ry && t < ny - 1 - ry)) continue;
                const int ip = row * ROWF + ipos_lp<LP>(col);
                hbuf[ip] += fold_sum<SY, RY>(a.tya, ry, a.bcy, ny, t, gsrc);
                h2buf[ip] += fold_sum<SY, RY>(a.tdya, ry, a.bcy, ny, t, gsrc);
            }
            lds_barrier();  // folded h rows ready
        }
        // x: per row, the folded terms of the target columns 0..rx and nx-1-rx..nx-1
        // that fall in this tile, from the (y-folded) h row
        const bool xfold = FOLD && fold_x;
        if (xfold) {
            // only the side(s) whose targets lie in this tile: entries 0..rx (left
            // targets) and rx+1..2rx+1 (right); the x stage reads no others
            const int klo = c0 <= rx ? 0 : rx + 1;
            const int khi = c0 + TXC - 1 >= nx - 1 - rx ? 2 * (rx + 1) : rx + 1;
            const int FW = khi - klo;
            for (int w = tid; w < S * FW; w += kBlock) {
                const int row = w / FW, k = klo + w % FW;
                const int t = k <= rx ? k : nx - 1 - rx + (k - rx - 1);  // target column
                float v = 0.f;
                if (t >= c0 && t < c0 + TXC && t < nx) {
                    const float *hrow0 = hbuf + row * ROWF;
                    auto hsrc = [&](int q) { return hrow0[ipos_lp<LP>(q - cs0)]; };
                    v = fold_sum<SX, RX>(a.txa, rx, a.bcx, nx, t, hsrc);
                }
                foldx[row * Gm::FOLDW + k] = v;
            }
            lds_barrier();  // fold terms ready
        }

        // --- x stage: row tid / 16, staged columns RX + 8 * (tid % 16) .. +7 ---
        {
            const int row = xrow, l = xl;
            const int i = i0 + row;
            const float *hr = hbuf + row * ROWF + l;
            const float *xr = xbuf + row * ROWF + l;
            float hv[NWX], xv[NWX];
#pragma unroll
            for (int q = 0; q < NWX; ++q) {
                hv[q] = hr[(q & 7) * LP + (q >> 3)];
                xv[q] = xr[(q & 7) * LP + (q >> 3)];
            }
            float d[P], u[P], u2[P];
            ctap *txa = taps_here(offsetof(BwdArgs, txa));
            ctap *tx = taps_here(offsetof(BwdArgs, tx));
            ctap *tdx = taps_here(offsetof(BwdArgs, tdx));
#pragma unroll
            for (int p = 0; p < P; ++p) {
                float sd = txa[0] * hv[p + RX], su = tx[0] * xv[p + RX], su2 = tdx[0] * xv[p + RX];
#pragma unroll
                for (int k = 1; k <= RX; ++k) {
                    sd = __builtin_fmaf(txa[k], pair<SX>(hv[p + RX + k], hv[p + RX - k]), sd);
                    const float t = pair<SX>(xv[p + RX + k], xv[p + RX - k]);
                    su = __builtin_fmaf(tx[k], t, su);
                    su2 = __builtin_fmaf(tdx[k], t, su2);
                }
                d[p] = sd;
                u[p] = su;
                u2[p] = su2;
            }
            if (xfold) {
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const int t = cb + p;
                    if (t <= rx) d[p] += foldx[row * Gm::FOLDW + t];
                    else if (t >= nx - 1 - rx && t < nx) d[p] += foldx[row * Gm::FOLDW + rx + 1 + (t - (nx - 1 - rx))];
                }
            }
            if (i < zend) {
                // outputs p < nv of this lane are in the tile and the line; the count is
                // made opaque so its comparisons stay here instead of being hoisted out
                // of the step loop as 64-bit lane masks
                int nvv = nv;
                asm volatile("" : "+v"(nvv));
                const float *h2r = h2buf + row * ROWF + l;
                float fx = 0.f, fy = 0.f;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const bool ok = p < nvv;
                    const float h2v = h2r[((p + RX) & 7) * LP + ((p + RX) >> 3)];
                    // a select, not a zero factor: unused window words may hold any bits
                    const float fxn = __builtin_fmaf(hv[p + RX], u2[p], fx);
                    const float fyn = __builtin_fmaf(h2v, u[p], fy);
                    fx = ok ? fxn : fx;
                    fy = ok ? fyn : fy;
                }
                dsx += static_cast<double>(fx);
                dsy += static_cast<double>(fy);
                if (a.want_dx) {
                    const int so = i0 * sydx4;
                    if (nvv == P) {
#pragma unroll
                        for (int p = 0; p < P; ++p) buffer_st(d[p], dres, dvo + 4 * p, so);
                    } else {
#pragma unroll
                        for (int p = 0; p < P; ++p)
                            if (p < nvv) buffer_st(d[p], dres, dvo + 4 * p, so);
                    }
                }
            }
        }
    }
    };
    if (fold_x || fold_y)
        steps(std::true_type{});
    else
        steps(std::false_type{});

    // block sums in a fixed order: waves by shuffles, then the four waves
    __shared__ double red[2][4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        dsx += __shfl_xor(dsx, off);
        dsy += __shfl_xor(dsy, off);
    }
    if (lane == 0) {
        red[0][wave] = dsx;
        red[1][wave] = dsy;
    }
    __syncthreads();
    if (tid == 0) {
        a.partials[2 * bid] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        a.partials[2 * bid + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    }
}

template <int RX, int RY>
size_t bwd2d_lds_of() {
    return BwdGeom<RX, RY>::lds_floats * sizeof(float);
}

// Launches k_bwd2d<., ., RX, RY> for the filters' symmetries; false for any other pair.
template <int RX, int RY>
bool bwd2d_launch_pair(int symx, int symy, const BwdArgs &a, int64_t nblocks, hipStream_t stream) {
    const dim3 grid(static_cast<unsigned>(nblocks)), block(kBlock);
    const size_t lds = bwd2d_lds_of<RX, RY>();
    if (symx == SYM_EVEN && symy == SYM_EVEN)
        hipLaunchKernelGGL((k_bwd2d<SYM_EVEN, SYM_EVEN, RX, RY>), grid, block, lds, stream, a);
    else if (symx == SYM_EVEN && symy == SYM_ODD)
        hipLaunchKernelGGL((k_bwd2d<SYM_EVEN, SYM_ODD, RX, RY>), grid, block, lds, stream, a);
    else if (symx == SYM_ODD && symy == SYM_EVEN)
        hipLaunchKernelGGL((k_bwd2d<SYM_ODD, SYM_EVEN, RX, RY>), grid, block, lds, stream, a);
    else if (symx == SYM_ODD && symy == SYM_ODD)
        hipLaunchKernelGGL((k_bwd2d<SYM_ODD, SYM_ODD, RX, RY>), grid, block, lds, stream, a);
    else
        return false;
    return true;
}

// Instances with adjacent, unequal x and y classes (k_bwd2d_xy.hip).
bool bwd2d_launch_xy(int cx, int cy, int symx, int symy, const BwdArgs &a, int64_t nblocks, hipStream_t stream);
size_t bwd2d_lds_xy(int cx, int cy);

}  // namespace dipamd
