// K3Z: the x, y and z passes of a float volume in one kernel (gfx950).
//
// Framework::Separable runs the x pass over the whole volume, then y, then z
// (framework_separable.cpp:146-206), each rounding to float; the two-pass path here
// fuses x and y (K3) and sweeps z over the intermediate, so the intermediate makes a
// full round trip through HBM.  This kernel keeps it on chip: a workgroup owns a tile
// of 64 columns x 48 rows and a chunk of planes, and streams the planes through it:
//   * plane q's tile with a halo of RXC columns and ry rows (boundary-folded, the
//     source offsets computed once per workgroup) is staged in LDS, the next plane's
//     samples already in flight (register prefetch);
//   * x stage: every staged row is x-filtered, four outputs per item, the window read
//     with 128-bit LDS loads at compile-time offsets; rows outside the plane take the
//     y condition's sign;
//   * y stage: a thread owns one column and six rows and y-filters them from the
//     x-filtered rows; a plane outside the volume takes the z condition's sign;
//   * z stage: the six values enter a register ring of W = 2 RZ + 1 planes whose
//     slot indices are compile-time constants (the plane loop is unrolled by W); once
//     it is full the centre plane's outputs are z-filtered and stored.
// The conditions are index-based (mirrors, periodic, zero order): a sample outside is
// a sample inside, maybe negated.  Every output sample is computed in the reference's
// order on each axis (convolution.cpp:214-241) from the same float intermediates as
// the pass-by-pass path, so results are bit-identical to it.
#pragma once

#include <utility>

#include "fused_rc.hpp"

namespace dipamd {

struct Fused3Args {
    const void *in;
    float *out;
    int nx, ny, nz;
    int64_t syin, szin;  // input strides (samples); x is contiguous
    int64_t sxout, syout, szout;
    int rx, ry;
    int bcx, bcy, bcz;  // index-based conditions only
    int ntx, nty, zchunk, nzchunks;
    int vec_ok;  // input rows 16-byte aligned: vector loads in interior tiles
    Outer outer;  // dims beyond the three: sin (input), sout (output)
    float tx[kFusedCap], ty[kFusedCap], tz[8];
};

constexpr int kF3TX = 64, kF3TY = 48, kF3Threads = 512, kF3MaxRy = 8, kF3MaxPlanes = 1024;

__host__ __device__ constexpr int f3_stage_pitch(int rxc) { return rc_pitch<float>(kF3TX + 2 * rxc); }
constexpr int kF3XPitch = kF3TX + 4;
// LDS of a workgroup: the staged rows, the x-filtered rows and the plane map of its chunk
__host__ __device__ constexpr size_t f3_lds_bytes(int rxc, int planes) {
    return (static_cast<size_t>(kF3TY + 2 * kF3MaxRy) * (f3_stage_pitch(rxc) + kF3XPitch) + 4) * sizeof(float) +
           static_cast<size_t>(planes) * sizeof(int);
}

namespace {

// A staged column of x-filtered rows (row pitch `pitch`).
struct F3ColSrc {
    const float *col;
    int pitch, ri, li;
    __device__ float at(int q) const { return col[q * pitch]; }
    __device__ void rbegin(int q) { ri = q * pitch; }
    __device__ float rnext() {
        const float v = col[ri];
        ri += pitch;
        return v;
    }
    __device__ void lbegin(int q) { li = q * pitch; }
    __device__ float lnext() {
        const float v = col[li];
        li -= pitch;
        return v;
    }
};

template <class F, int... U>
__device__ __forceinline__ void f3_for_each(std::integer_sequence<int, U...>, F &&f) {
    (f(std::integral_constant<int, U>{}), ...);
}

}  // namespace

template <typename TI, int SX, int SY, int SZ, int RXC, int RZ, bool FAST, bool INTERIOR>
__device__ __forceinline__ void fused3d_body(const Fused3Args &a, const TI *__restrict__ in, float *__restrict__ out,
                                             float *lds, int x0, int y0, int z0, int zend) {
    constexpr int TX = kF3TX, TY = kF3TY, NT = kF3Threads, RYM = kF3MaxRy;
    constexpr int W = 2 * RZ + 1;
    constexpr int SW = TX + 2 * RXC;       // staged columns
    constexpr int SPT = f3_stage_pitch(RXC);
    constexpr int XPT = kF3XPitch;
    constexpr int WN = 4 + 2 * RXC;        // x window of an item (four outputs)
    constexpr int SW4 = SW / 4;            // float4 per staged row
    constexpr int MAXR = TY + 2 * RYM;     // staged rows at most
    constexpr int PY = TY * TX / NT;       // rows per thread in the y stage
    constexpr int NXI = (MAXR * (TX / 4) + NT - 1) / NT;  // x-stage items per thread at most
    constexpr int NPF = INTERIOR ? (MAXR * SW4 + NT - 1) / NT : (MAXR * SW + NT - 1) / NT;
    static_assert(RXC % 4 == 0, "x class: a multiple of 4");
    static_assert(std::is_same<TI, float>::value, "float volumes");
    const int nx = a.nx, ny = a.ny, nz = a.nz, rx = a.rx, ry = a.ry;
    const int NR = TY + 2 * ry;            // staged rows
    float *stg = lds;                                   // MAXR rows x SPT
    float *xb = lds + MAXR * SPT + 4;                   // (a dummy float4), MAXR rows x XPT
    int *zmap = reinterpret_cast<int *>(lds + MAXR * (SPT + XPT) + 4);  // plane map of the chunk
    const int tid = threadIdx.x;
    const int E = (zend - z0) + 2 * RZ;               // planes streamed

    // --- set-up, once per workgroup: where each staged sample comes from ---
    // The conditions are index-based (mirrors, periodic, zero-order): a sample outside the
    // volume is a sample inside, negated for the asymmetric ones.  A thread's staging items
    // are the same (row, column) in every plane: source element offset within the plane
    // << 1 | negate, and the LDS slot it goes to (items past the tile: a dummy slot, their
    // loads read the plane's first sample, so neither needs a branch).
    int goff[NPF], sdst[NPF];
    {
        const int nitems = INTERIOR ? NR * SW4 : NR * SW;
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
            const int it = tid + NT * i;
            int off = 0, dst = MAXR * SPT;  // the dummy slot
            if (it < nitems) {
                const int r = INTERIOR ? it / SW4 : it / SW;
                const int c = INTERIOR ? 4 * (it - r * SW4) : it - r * SW;
                int y = y0 - ry + r;
                if (y < 0 || y >= ny) {
                    int m;
                    y = max(bc_fold(y, ny, a.bcy, &m), 0);  // the y sign applies after the x stage
                }
                int x = x0 - RXC + c, neg = 0;
                if (!INTERIOR && (x < 0 || x >= nx)) {
                    int m;
                    x = max(bc_fold(x, nx, a.bcx, &m), 0);
                    neg = m < 0;
                }
                off = ((y * static_cast<int>(a.syin) + x) << 1) | neg;
                dst = r * SPT + c;
            }
            goff[i] = off;
            sdst[i] = dst;
        }
    }
    // x-stage rows outside the plane: the y condition's sign, per item (bit 31)
    unsigned ysgn[NXI];
#pragma unroll
    for (int k = 0; k < NXI; ++k) {
        const int it = tid + NT * k;
        const int y = y0 - ry + (it >> 4);
        unsigned sg = 0;
        if (it < NR * (TX / 4) && (y < 0 || y >= ny)) {
            int m;
            bc_fold(y, ny, a.bcy, &m);
            sg = m < 0 ? 0x80000000u : 0u;
        }
        ysgn[k] = sg;
    }
    // plane map of the chunk: source plane << 1 | negate
    for (int e = tid; e < E; e += NT) {
        const int zs = z0 - RZ + e;
        int q = zs, neg = 0;
        if (zs < 0 || zs >= nz) {
            int m;
            q = max(bc_fold(zs, nz, a.bcz, &m), 0);
            neg = m < 0;
        }
        zmap[e] = (q << 1) | neg;
    }
    lds_barrier();

    // --- staging: a plane into registers, then into LDS ---
    typedef float f4v __attribute__((ext_vector_type(4)));  // (kept in registers, unlike float4 arrays)
    using PV = typename std::conditional<INTERIOR, f4v, float>::type;
    PV pre[NPF];
    auto load_plane = [&](int e) __attribute__((always_inline)) {
        const int zm = __builtin_amdgcn_readfirstlane(zmap[e]);
        const TI *pl = in + static_cast<int64_t>(zm >> 1) * a.szin;
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
            int g = goff[i];
            asm volatile("" : "+v"(g));  // keeps the sign test inside the loop (no hoisted masks)
            if constexpr (INTERIOR) {
                pre[i] = *reinterpret_cast<const f4v *>(pl + (g >> 1));
            } else {
                const unsigned v = __float_as_uint(static_cast<float>(pl[g >> 1]));
                pre[i] = __uint_as_float(v ^ (static_cast<unsigned>(g) << 31));
            }
        }
    };
    auto store_plane = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
            if constexpr (INTERIOR)
                *reinterpret_cast<f4v *>(stg + sdst[i]) = pre[i];
            else
                stg[sdst[i]] = pre[i];
        }
    };

    const int col = tid & (TX - 1), rg = tid >> 6;  // y stage: column, group of PY rows
    float zr[W][PY];
#pragma unroll
    for (int u = 0; u < W; ++u)
#pragma unroll
        for (int p = 0; p < PY; ++p) zr[u][p] = 0.f;

    load_plane(0);
    auto plane = [&](int e, auto slot) __attribute__((always_inline)) {
        constexpr int u = decltype(slot)::value;
        // (the previous x stage finished reading `stg` before its barrier below, and no
        // thread writes `xb` again before the next barrier, which follows its y stage)
        store_plane();
        lds_barrier();  // staged
        if (e + 1 < E) load_plane(e + 1);
        // --- x stage: items (row, group of 4 columns) ---
        {
            const float *txp = a.tx;
#pragma unroll
            for (int k2 = 0; k2 < NXI; ++k2) {
                const int it = tid + NT * k2;
                if (it < NR * (TX / 4)) {
                    const int r = it >> 4, g = it & 15;
                    float w[WN];
                    const float4 *wv = reinterpret_cast<const float4 *>(stg + r * SPT + 4 * g);
#pragma unroll
                    for (int i = 0; i < WN / 4; ++i) {
                        const float4 t = wv[i];
                        w[4 * i] = t.x;
                        w[4 * i + 1] = t.y;
                        w[4 * i + 2] = t.z;
                        w[4 * i + 3] = t.w;
                    }
                    float acc[4];
                    const float f0 = txp[0];
#pragma unroll
                    for (int p = 0; p < 4; ++p) acc[p] = f0 * w[RXC + p];
#pragma unroll
                    for (int k = 1; k <= RXC; ++k) {
                        if (k <= rx) {
                            const float f = txp[k];
#pragma unroll
                            for (int p = 0; p < 4; ++p) {
                                const float hi = w[RXC + p + k], lo = w[RXC + p - k];
                                acc[p] = mac<float, FAST>(acc[p], f, SX == SYM_EVEN ? hi + lo : hi - lo);
                            }
                        }
                    }
                    unsigned sg = ysgn[k2];
                    asm volatile("" : "+v"(sg));
#pragma unroll
                    for (int p = 0; p < 4; ++p) acc[p] = __uint_as_float(__float_as_uint(acc[p]) ^ sg);
                    *reinterpret_cast<float4 *>(xb + r * XPT + 4 * g) = make_float4(acc[0], acc[1], acc[2], acc[3]);
                }
            }
        }
        lds_barrier();  // x-filtered rows ready
        // --- y stage: column col, rows PY rg .. PY rg + PY - 1 ---
        float yv[PY];
        {
            F3ColSrc src{xb + col + (rg * PY + ry) * XPT, XPT, 0, 0};
            fir_window<float, PY, SY, FAST>(a.ty, ry + 1, 0, src, yv);
        }
        {
            const int zm = __builtin_amdgcn_readfirstlane(zmap[e]);
            if (zm & 1) {  // a plane outside the volume, negated by the z condition
#pragma unroll
                for (int p = 0; p < PY; ++p) yv[p] = -yv[p];
            }
        }
#pragma unroll
        for (int p = 0; p < PY; ++p) zr[u][p] = yv[p];
        // --- z stage: centre plane once the ring is full ---
        if (e >= 2 * RZ) {
            constexpr int cs = ((u - RZ) % W + W) % W;
            const int o = z0 + e - 2 * RZ;
            const int x = x0 + col;
            if (x < nx) {
                float *dst = out + static_cast<int64_t>(o) * a.szout + static_cast<int64_t>(x) * a.sxout;
                const float f0 = a.tz[0];
#pragma unroll
                for (int p = 0; p < PY; ++p) {
                    float acc = f0 * zr[cs][p];
#pragma unroll
                    for (int k = 1; k <= RZ; ++k) {
                        const float hi = zr[(cs + k) % W][p], lo = zr[(cs - k + W) % W][p];
                        acc = mac<float, FAST>(acc, a.tz[k], SZ == SYM_EVEN ? hi + lo : hi - lo);
                    }
                    const int y = y0 + rg * PY + p;
                    if (y < ny) dst[static_cast<int64_t>(y) * a.syout] = acc;
                }
            }
        }
    };
    for (int eb = 0; eb < E; eb += W) {
        f3_for_each(std::make_integer_sequence<int, W>{}, [&](auto slot) __attribute__((always_inline)) {
            if (eb + decltype(slot)::value < E) plane(eb + decltype(slot)::value, slot);
        });
    }
}

template <typename TI, int SX, int SY, int SZ, int RXC, int RZ, bool FAST>
__global__ __launch_bounds__(kF3Threads) void k_fused3d(const Fused3Args a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *lds = reinterpret_cast<float *>(smem_raw);
    const int64_t bid = xcd_block();
    const int tx = static_cast<int>(bid % a.ntx);
    int64_t rest = bid / a.ntx;
    const int ty = static_cast<int>(rest % a.nty);
    rest /= a.nty;
    const int zc = static_cast<int>(rest % a.nzchunks);
    const int64_t outer = rest / a.nzchunks;
    int64_t oin, oout;
    outer_offsets(a.outer, outer, &oin, &oout);
    const TI *in = static_cast<const TI *>(a.in) + oin;
    float *out = a.out + oout;
    const int x0 = tx * kF3TX, y0 = ty * kF3TY;
    const int z0 = zc * a.zchunk;
    const int zend = min(z0 + a.zchunk, a.nz);
    if (a.vec_ok && x0 - RXC >= 0 && x0 + kF3TX + RXC <= a.nx)
        fused3d_body<TI, SX, SY, SZ, RXC, RZ, FAST, true>(a, in, out, lds, x0, y0, z0, zend);
    else
        fused3d_body<TI, SX, SY, SZ, RXC, RZ, FAST, false>(a, in, out, lds, x0, y0, z0, zend);
}

}  // namespace dipamd
