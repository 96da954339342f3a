// Streaming backward pass of the separable Gaussian over stacks of 2-D planes (B3).
//
// The forward pass is y = A_y A_x x with A_d = C_d E_d (E_d: boundary extension along
// axis d, C_d: the symmetric FIR on the extended line).  The two axes act on different
// tensor dims, so A_x and A_y commute, and for g = dL/dy one pass over g gives
//     v = A_x^T g,   w = A'_x^T g               (x stage: two filters, shared pairs)
//     dL/dx       = A_y^T v                       (y stage)
//     dL/dsigma_x = < g, A'_x A_y x > = < E_y^T C_y^T w , x > = sum over the extended
//                   rows j of (C_y^T w)[j] (E_y x)[j]
//     dL/dsigma_y = < g, A'_y A_x x > = sum over j of (C'_y^T v)[j] (E_y x)[j]
// where C^T is the correlation with reversed taps (the filter itself for even taps,
// its negative for odd ones) over the zero-padded line (periodic conditions: over the
// periodic extension, no folds), and E^T adds every halo position into the pixel the
// extension copied it from, with the extension's sign.
//
// One wave owns a strip of 64 columns of one plane and a chunk of rows, and streams
// the rows of g down the chunk with no workgroup barrier:
//   * row q of g (with an x halo of zeros, or the periodic wrap) is staged in a
//     wave-private LDS row; each lane reads its window of 1 + 2 RXC samples and
//     computes v and w at its column; in the strips at the left and right edge of a
//     mirror condition the lanes of the fold targets also correlate at their halo
//     position and add that (E_x^T), with the condition's sign;
//   * v and w enter register rings of W = 2 RYC + 1 rows whose slot indices are
//     compile-time constants (the row loop is unrolled by W);
//   * once the ring is full, the centre row o gets dL/dx = C_y^T v, C'_y^T v and
//     C_y^T w (five operations per tap pair); the dot products take x at row o.
// With a folding condition along y the first chunk also produces the halo rows
// o = -ry..-1 and the last chunk o = ny..ny-1+ry: their dL/dx values are added into
// the rows they fold onto (stored by the same lane, so plain loads and stores order
// them), and their dot products take the extended x row (E_y x)[o].
//
// RXC and RYC are compile-time radius classes (taps beyond rx, ry are zero), so every
// window and ring index is an immediate.  The row loop makes no calls: rows outside the
// plane take their source row and sign from a table the host fills (scalar loads), and
// the zero-order x fold is a dot of the end samples with suffix sums of the adjoint taps
// (also from the host).  Arithmetic is fused multiply-add; the gradients are checked
// against FP64 references.
#pragma once

#include <utility>

#include "k_bwd2d.hpp"

namespace dipamd {

namespace {

// Orders this wave's LDS accesses at this point (the wave's own rows only: no other
// wave reads them, and LDS operations of one wave complete in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// A tap array in the kernel-argument segment, addressed through a pointer that is
// defined only once `dep` is: its scalar loads cannot be hoisted above the code that
// computes dep (which keeps one stage's taps in scalar registers at a time).
__device__ __forceinline__ ctap *taps_after(size_t offset, float dep) {
    const __attribute__((address_space(4))) char *seg =
        (const __attribute__((address_space(4))) char *)__builtin_amdgcn_kernarg_segment_ptr();
    ctap *q = reinterpret_cast<ctap *>(seg + offset);
    asm volatile("" : "+s"(q) : "v"(dep));
    return q;
}

// Calls f(std::integral_constant<int, U>) for U = 0, 1, ...: the ring slot of each row
// of a period is a compile-time constant however large the unrolled body grows.
template <class F, int... U>
__device__ __forceinline__ void for_each_slot(std::integer_sequence<int, U...>, F &&f) {
    (f(std::integral_constant<int, U>{}), ...);
}

template <int P>
struct B2sVec;
template <>
struct B2sVec<1> {
    __device__ static void ld(float (&o)[1], __amdgpu_buffer_rsrc_t r, int voff, int soff, bool) {
        o[0] = buffer_ld(r, voff, soff);
    }
};
template <>
struct B2sVec<2> {
    // whole: both columns lie in the line (else only the first, read alone)
    __device__ static void ld(float (&o)[2], __amdgpu_buffer_rsrc_t r, int voff, int soff, bool whole) {
        if (whole) {
            const float2 v = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
            o[0] = v.x;
            o[1] = v.y;
        } else {
            o[0] = buffer_ld(r, voff, soff);
            o[1] = 0.f;
        }
    }
};

}  // namespace

template <int SX, int SY, int RXC, int RYC, int P>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_bwd2s(const Bwd2sArgs a) {
    constexpr int W = 2 * RYC + 1;       // ring rows
    constexpr int TC = 64 * P;           // columns of a strip
    // staged columns: the strip with 2 RXC columns before it and 3 RXC after it (the
    // windows of the folded halo positions of a mirror condition reach that far)
    constexpr int LPAD = 2 * RXC;
    constexpr int SW = TC + 5 * RXC;
    constexpr int NWIN = 1 + 2 * RXC;    // x window of a lane
    constexpr int NSL = (SW + 63) / 64;  // staged slots per lane
    static_assert(P == 1, "one column per lane");
    __shared__ __align__(16) float stage[4][SW];

    const int lane = threadIdx.x & 63, wave = wave_id();
    // waves in dispatch order; the strips at the left and right edge (the slower rows of
    // boundary handling) come first, so that none of them starts last
    const int64_t gw = static_cast<int64_t>(blockIdx.x) * 4 + wave;
    const bool active = gw < a.nwaves;
    double dsx = 0.0, dsy = 0.0;
    if (active) {
        const int ne = a.nstrips > 1 ? 2 : 1;  // edge strips
        const int64_t nedge = static_cast<int64_t>(ne) * a.nchunks_e * a.nplanes;
        int strip;
        int64_t rest;
        const bool edge = gw < nedge;
        if (edge) {
            strip = (gw % ne) == 0 ? 0 : a.nstrips - 1;
            rest = gw / ne;
        } else {
            const int64_t j = gw - nedge;
            const int ni = a.nstrips - ne;
            strip = 1 + static_cast<int>(j % ni);
            rest = j / ni;
        }
        // edge strips may take shorter chunks (their rows cost more under an x fold)
        const int nch = edge ? a.nchunks_e : a.nchunks;
        const int chl = edge ? a.chunk_e : a.chunk;
        const int chunk = static_cast<int>(rest % nch);
        const int64_t plane = rest / nch;
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
        const int nx = a.nx, ny = a.ny, rx = a.rx, ry = a.ry;
        const __amdgpu_buffer_rsrc_t gres = plane_rsrc(a.g + og, a.gbytes);
        const __amdgpu_buffer_rsrc_t xres = plane_rsrc(a.x + ox, a.xbytes);
        const __amdgpu_buffer_rsrc_t dres = plane_rsrc(a.want_dx ? a.dx + odx : a.x + ox, a.want_dx ? a.dxbytes : 4);
        const int syg4 = static_cast<int>(a.syg) * 4, syx4 = static_cast<int>(a.syx) * 4;
        const int sydx4 = static_cast<int>(a.sydx) * 4;
        const int c0 = strip * TC;           // first column of the strip
        const int c = c0 + P * lane;         // this lane's first column
        const int cs0 = c0 - LPAD;           // image column of staged column 0
        const bool lane_in = c < nx;
        const int cvo = (lane_in ? c : 0) * 4;  // byte offset of the lane's columns (x and dL/dx)
        const int nvalid = lane_in ? min(P, nx - c) : 0;

        // staged slots of this lane: source column byte offset and factor (the adjoint's
        // extension of g: zero outside the line, or the periodic wrap with its sign)
        int gvo[NSL];
        float gmul[NSL];
        bool plain = cs0 >= 0 && cs0 + SW <= nx;
#pragma unroll
        for (int i = 0; i < NSL; ++i) {
            const int s = lane + 64 * i;
            const int j = cs0 + s;
            int src = j, m = 1;
            float f = 1.f;
            if (s >= SW) {
                src = 0;
            } else if (j < 0 || j >= nx) {
                if (a.modex == kBwdPeriodic) {
                    src = bc_fold(j, nx, a.bcx, &m);
                    f = static_cast<float>(m);
                } else {
                    src = 0;
                    f = 0.f;
                }
            }
            gvo[i] = src * 4;
            gmul[i] = f;
        }
        // x-folded terms: strips holding targets within rx of a folding edge; for the
        // mirror conditions each target's fold term is the correlation at its halo
        // position (-c, or 2(nx-1) - c), from the same staged row
        const bool xfold = a.modex == kBwdFold && (c0 <= rx || c0 + TC - 1 >= nx - 1 - rx);
        const bool xmirror = a.bcx == DIPAMD_BC_SYMMETRIC_MIRROR || a.bcx == DIPAMD_BC_ASYMMETRIC_MIRROR;
        const float xsg = a.bcx == DIPAMD_BC_ASYMMETRIC_MIRROR ? -1.f : 1.f;
        const bool tleft = c >= 1 && c <= rx;
        const bool tright = c >= nx - 1 - rx && c <= nx - 2;
        const int lbase = tleft ? (c0 - c) - cs0 - RXC : lane + RXC;  // window start, staged index
        const int rbase = tright ? (2 * (nx - 1) - c) - cs0 - RXC : lane + RXC;
        const bool yfold = a.modey == kBwdFold;

        // rows: outputs ostart .. oend-1 (halo rows of a folding condition included);
        // streamed rows ostart - RYC .. oend - 1 + RYC
        const int z0 = chunk * chl;
        const int zend = chunk == nch - 1 ? ny : z0 + chl;
        const int ostart = (yfold && chunk == 0) ? -ry : z0;
        const int oend = (yfold && chunk == nch - 1) ? ny + ry : zend;
        const int E = (oend - ostart) + 2 * RYC;
        const int q0 = ostart - RYC;  // streamed row of e = 0

        float *srow = &stage[wave][0];
        float vr[W][P], wr[W][P];
#pragma unroll
        for (int u = 0; u < W; ++u)
#pragma unroll
            for (int p = 0; p < P; ++p) vr[u][p] = wr[u][p] = 0.f;

        // a row q outside 0..ny-1 (wave-uniform): its source row under the y extension
        // (-1: a constant zero row) and sign, from the table the host filled for rows
        // -kB2sFold..-1 and ny..ny+kB2sFold-1 (no call, a scalar load)
        auto yext = [&](int q, int *m) -> int {
            const int i = __builtin_amdgcn_readfirstlane(q < 0 ? -q - 1 : kB2sFold + (q - ny));
            const int e = a.yfold[i];
            *m = (e & 1) ? -1 : 1;
            return e < 0 ? -1 : (e >> 1);
        };
        // row q of g: source row (wave-uniform) and factor; a zero row needs no load
        auto grow = [&](int q, int *src) -> float {
            if (q >= 0 && q < ny) {
                *src = q;
                return 1.f;
            }
            if (a.modey == kBwdPeriodic) {
                int m;
                *src = yext(q, &m);
                return static_cast<float>(m);
            }
            *src = 0;
            return 0.f;
        };
        // (E_y x) row o: source row and factor (the forward extension along y)
        auto xrow = [&](int o, int *src) -> float {
            if (o >= 0 && o < ny) {
                *src = o;
                return 1.f;
            }
            int m;
            const int s = yext(o, &m);
            if (s < 0) {
                *src = 0;
                return 0.f;
            }
            *src = s;
            return static_cast<float>(m);
        };
        float pg0[NSL], pg1[NSL], pg2[NSL];  // prefetched rows e, e + 1, e + 2
        float rm0, rm1, rm2;
        auto load_g = [&](int e, float (&dst)[NSL], float *rm) {
            int src;
            const float f = grow(q0 + (e < E ? e : E - 1), &src);
            *rm = f;
            const int so = src * syg4;
#pragma unroll
            for (int i = 0; i < NSL; ++i) dst[i] = buffer_ld(gres, gvo[i], so);
        };
        float px0[P], px1[P];  // x at the output rows of e, e + 1
        float xm0, xm1;
        auto load_x = [&](int e, float (&dst)[P], float *xm) {
            int src;
            const int o = ostart + e - 2 * RYC;
            *xm = xrow(o < ostart ? ostart : o, &src);
            B2sVec<P>::ld(dst, xres, cvo, src * syx4, nvalid == P);
        };
        load_g(0, pg0, &rm0);
        load_g(1, pg1, &rm1);
        load_g(2, pg2, &rm2);
        load_x(2 * RYC, px0, &xm0);
        load_x(2 * RYC + 1, px1, &xm1);
        float fx = 0.f, fy = 0.f;  // dot products of the current period

        // One streamed row e in ring slot u.  RP ("row plain"): the g row, the output row
        // and the rows prefetched for them lie inside the plane, the ring is full and the
        // output row receives no folded halo terms.  SP ("strip plain"): every staged
        // column lies inside the line and every lane's columns are outputs.
        auto row = [&](int e, auto slot, auto rp_tag, auto sp_tag, auto warm_tag) __attribute__((always_inline)) {
            constexpr int u = decltype(slot)::value;
            // an opaque row index: the row offsets are computed here from it, not hoisted
            // out of the loop as one scalar per slot
            asm volatile("" : "+s"(e));
            constexpr bool RP = decltype(rp_tag)::value, SP = decltype(sp_tag)::value;
            constexpr bool WARM = decltype(warm_tag)::value;  // RP rows of the first period
            // --- this row's g (prefetched), next prefetch ---
            float gcur[NSL];
            const float rm = RP ? 1.f : rm0;
#pragma unroll
            for (int i = 0; i < NSL; ++i) {
                gcur[i] = pg0[i];
                pg0[i] = pg1[i];
                pg1[i] = pg2[i];
            }
            if constexpr (RP) {
                const int so = (q0 + e + 3) * syg4;
#pragma unroll
                for (int i = 0; i < NSL; ++i) pg2[i] = buffer_ld(gres, gvo[i], so);
            } else {
                rm0 = rm1;
                rm1 = rm2;
                load_g(e + 3, pg2, &rm2);
            }
            float v[P], w[P];
            if (!RP && rm == 0.f) {  // a zero row (wave-uniform)
#pragma unroll
                for (int p = 0; p < P; ++p) v[p] = w[p] = 0.f;
            } else {
                // --- stage the row ---
#pragma unroll
                for (int i = 0; i < NSL; ++i) {
                    const int s = lane + 64 * i;
                    if (SW % 64 == 0 || s < SW) srow[s] = SP ? gcur[i] : gcur[i] * gmul[i];
                }
                wave_lds_sync();
                float win[NWIN];
#pragma unroll
                for (int i = 0; i < NWIN; ++i) win[i] = srow[lane + RXC + i];
                // taps beyond rx are zero; loaded per stage (scalar loads ordered after the
                // window reads, so the four tap arrays are never live together)
                ctap *txa = taps_after(offsetof(Bwd2sArgs, txa), win[0]);
                ctap *tdxa = taps_after(offsetof(Bwd2sArgs, tdxa), win[0]);
                // the two x filters at one centre from a window
                auto xcorr = [&](const float (&wn)[NWIN], float *sv_, float *sw_) __attribute__((always_inline)) {
                    float sv = txa[0] * wn[RXC], sw = tdxa[0] * wn[RXC];
#pragma unroll
                    for (int k = 1; k <= RXC; ++k) {
                        const float t = pair<SX>(wn[RXC + k], wn[RXC - k]);
                        sv = __builtin_fmaf(txa[k], t, sv);
                        sw = __builtin_fmaf(tdxa[k], t, sw);
                    }
                    *sv_ = sv;
                    *sw_ = sw;
                };
                xcorr(win, &v[0], &w[0]);
                if (!SP && xfold) {  // edge strips of a folding condition (wave-uniform)
                    if (xmirror) {
                        // each side: the correlation at the targets' halo positions (other
                        // lanes read their own window and add nothing)
                        auto side = [&](int base, bool tgt) __attribute__((always_inline)) {
                            float hw[NWIN];
#pragma unroll
                            for (int i = 0; i < NWIN; ++i) hw[i] = srow[base + i];
                            float hv_, hw_;
                            xcorr(hw, &hv_, &hw_);
                            v[0] = tgt ? __builtin_fmaf(xsg, hv_, v[0]) : v[0];
                            w[0] = tgt ? __builtin_fmaf(xsg, hw_, w[0]) : w[0];
                        };
                        if (c0 <= rx) side(lbase, tleft);
                        if (c0 + TC - 1 >= nx - 1 - rx) side(rbase, tright);
                    } else {
                        // zero order: every halo position folds onto the end sample; the
                        // sum of their correlations over the zero-padded line is
                        // sum_m g[m] H[m] at the left end (g[n-1-m] at the right), H[m] the
                        // sum of the adjoint taps m+1..rx (the host's hzv, hzw)
                        ctap *hv = taps_after(offsetof(Bwd2sArgs, hzv), win[0]);
                        ctap *hw = taps_after(offsetof(Bwd2sArgs, hzw), win[0]);
                        if (c == 0) {
                            float sv = 0.f, sw = 0.f;
#pragma unroll
                            for (int m = 0; m < RXC; ++m) {
                                const float gm = srow[m - cs0];
                                sv = __builtin_fmaf(hv[m], gm, sv);
                                sw = __builtin_fmaf(hw[m], gm, sw);
                            }
                            v[0] += sv;
                            w[0] += sw;
                        }
                        if (c == nx - 1) {
                            float sv = 0.f, sw = 0.f;
#pragma unroll
                            for (int m = 0; m < RXC; ++m) {
                                const float gm = srow[nx - 1 - m - cs0];
                                sv = __builtin_fmaf(hv[m], gm, sv);
                                sw = __builtin_fmaf(hw[m], gm, sw);
                            }
                            v[0] += SX == SYM_EVEN ? sv : -sv;
                            w[0] += SX == SYM_EVEN ? sw : -sw;
                        }
                    }
                }
                if (!RP && rm != 1.f) {  // periodic wrap with a sign
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        v[p] *= rm;
                        w[p] *= rm;
                    }
                }
                wave_lds_sync();  // window reads done before the next row is staged
            }
#pragma unroll
            for (int p = 0; p < P; ++p) {
                // pinned here: computed now, not sunk towards their far uses (which
                // would keep every pair sum alive instead of one value)
                asm volatile("" : "+v"(v[p]), "+v"(w[p]));
                vr[u][p] = v[p];
                wr[u][p] = w[p];
            }
            // --- output row o = centre of the ring ---
            float xcur[P];
            const float xm = RP ? 1.f : xm0;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                xcur[p] = px0[p];
                px0[p] = px1[p];
            }
            if constexpr (RP) {
                // (in the warm-up period the rows before the first output are not needed)
                const int xo = (WARM && u + 2 < 2 * RYC) ? ostart : ostart + e + 2 - 2 * RYC;
                B2sVec<P>::ld(px1, xres, cvo, xo * syx4, SP || nvalid == P);
            } else {
                xm0 = xm1;
                load_x(e + 2, px1, &xm1);
            }
            if (RP ? (!WARM || u >= 2 * RYC) : e >= 2 * RYC) {
                const int o = ostart + e - 2 * RYC;
                constexpr int cs = ((u - RYC) % W + W) % W;
                float d[P], dv[P], dw[P];
                ctap *tya = taps_after(offsetof(Bwd2sArgs, tya), vr[u][0]);
                ctap *tdya = taps_after(offsetof(Bwd2sArgs, tdya), vr[u][0]);
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    float s1 = tya[0] * vr[cs][p], s2 = tdya[0] * vr[cs][p], s3 = tya[0] * wr[cs][p];
#pragma unroll
                    for (int k = 1; k <= RYC; ++k) {  // taps beyond ry are zero
                        const float tv = pair<SY>(vr[(cs + k) % W][p], vr[(cs - k + W) % W][p]);
                        const float tw = pair<SY>(wr[(cs + k) % W][p], wr[(cs - k + W) % W][p]);
                        s1 = __builtin_fmaf(tya[k], tv, s1);
                        s2 = __builtin_fmaf(tdya[k], tv, s2);
                        s3 = __builtin_fmaf(tya[k], tw, s3);
                    }
                    d[p] = s1;
                    dv[p] = s2;
                    dw[p] = s3;
                }
#pragma unroll
                for (int p = 0; p < P; ++p) asm volatile("" : "+v"(d[p]), "+v"(dv[p]), "+v"(dw[p]));
                // dot products with (E_y x)[o]; lanes past the line add nothing
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const float xv = RP ? xcur[p] : xm * xcur[p];
                    if constexpr (SP) {
                        fx = __builtin_fmaf(dw[p], xv, fx);
                        fy = __builtin_fmaf(dv[p], xv, fy);
                    } else {
                        const bool ok = p < nvalid;
                        fx = ok ? __builtin_fmaf(dw[p], xv, fx) : fx;
                        fy = ok ? __builtin_fmaf(dv[p], xv, fy) : fy;
                    }
                }
                if (a.want_dx && (SP || nvalid > 0)) {
                    if (RP || (o >= 0 && o < ny)) {
                        const int so = o * sydx4;
                        // rows receiving top halo terms already hold them (this lane stored them)
                        const bool pend = !RP && yfold && chunk == 0 &&
                                          (a.bcy == DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE ? o == 0 : (o >= 1 && o <= ry));
                        if (pend) {
#pragma unroll
                            for (int p = 0; p < P; ++p)
                                if (p < nvalid) d[p] += buffer_ld(dres, cvo + 4 * p, so);
                        }
                        if (SP || nvalid == P) {
                            if constexpr (P == 2)
                                __builtin_amdgcn_raw_buffer_store_b64(
                                    __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(dres, 0, 0, 0)),
                                                       make_float2(d[0], d[1])),
                                    dres, cvo, so, 0);
                            else
                                buffer_st(d[0], dres, cvo, so);
                        } else {
                            buffer_st(d[0], dres, cvo, so);  // the one column of a lane at the line's end
                        }
                    } else {  // a halo row: fold onto its target
                        int m;
                        const int t = yext(o, &m);
                        const int so = t * sydx4;
                        const bool first = o < 0 && (a.bcy != DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE || o == -ry);
#pragma unroll
                        for (int p = 0; p < P; ++p) {
                            if (p < nvalid) {
                                const float add = m < 0 ? -d[p] : d[p];
                                const float prev = first ? 0.f : buffer_ld(dres, cvo + 4 * p, so);
                                buffer_st(prev + add, dres, cvo + 4 * p, so);
                            }
                        }
                    }
                }
            }
        };
        // one period of W rows from row eb; GUARD: the period may end the chunk early
        auto period = [&](int eb, auto rp_tag, auto sp_tag, auto warm_tag, auto guard_tag)
                          __attribute__((always_inline)) {
            for_each_slot(std::make_integer_sequence<int, W>{}, [&](auto slot) __attribute__((always_inline)) {
                if (!decltype(guard_tag)::value || eb + decltype(slot)::value < E)
                    row(eb + decltype(slot)::value, slot, rp_tag, sp_tag, warm_tag);
            });
            dsx += static_cast<double>(fx);
            dsy += static_cast<double>(fy);
            fx = fy = 0.f;
        };
        using F = std::false_type;
        using T = std::true_type;
        // Whole periods whose rows are all plain (the g rows and their prefetches inside
        // the plane, the outputs inside it, past the rows receiving folded terms, and the
        // x rows prefetched for them inside it) run the lean rows; the others (the ends of
        // the first and last chunks, a partial last period) check every row.  Period 0 is
        // a lean warm-up when its rows are plain.
        const int olim = (yfold && chunk == 0) ? ry + 1 : 0;  // first output row without folds
        const bool warm0 = q0 >= 0 && ostart >= olim;
        const int elo = max(-q0, 2 * RYC + olim - ostart);
        const int p1 = warm0 ? 0 : (elo + W - 1) / W;
        const int ehi = min(E, min(ny - 3 - q0, ny - 2 - ostart + 2 * RYC));
        const int p2 = max(p1, ehi / W);
        const int nper = (E + W - 1) / W;
        auto run = [&](auto sp_tag) __attribute__((always_inline)) {
            for (int pb = 0; pb < nper; ++pb) {
                if (pb >= p1 && pb < p2) {
                    if (pb == 0)
                        period(0, T{}, sp_tag, T{}, F{});
                    else
                        period(pb * W, T{}, sp_tag, F{}, F{});
                } else {
                    period(pb * W, F{}, sp_tag, F{}, T{});
                }
            }
        };
        if (plain && c0 + TC <= nx)
            run(T{});
        else
            run(F{});
    }

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
    if (threadIdx.x == 0) {
        const int64_t bid = blockIdx.x;
        a.partials[2 * bid] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        a.partials[2 * bid + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    }
}

// Launch of one instance (symmetries, classes, columns per lane).
template <int SX, int SY, int RXC, int RYC, int P>
void bwd2s_launch_one(const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream) {
    hipLaunchKernelGGL((k_bwd2s<SX, SY, RXC, RYC, P>), dim3(static_cast<unsigned>(nblocks)), dim3(kBlock), 0, stream,
                       a);
}

template <int RXC, int RYC, int P>
bool bwd2s_launch_sym(int symx, int symy, const Bwd2sArgs &a, int64_t nblocks, hipStream_t stream) {
    if (symx == SYM_EVEN && symy == SYM_EVEN)
        bwd2s_launch_one<SYM_EVEN, SYM_EVEN, RXC, RYC, P>(a, nblocks, stream);
    else if (symx == SYM_EVEN && symy == SYM_ODD)
        bwd2s_launch_one<SYM_EVEN, SYM_ODD, RXC, RYC, P>(a, nblocks, stream);
    else if (symx == SYM_ODD && symy == SYM_EVEN)
        bwd2s_launch_one<SYM_ODD, SYM_EVEN, RXC, RYC, P>(a, nblocks, stream);
    else if (symx == SYM_ODD && symy == SYM_ODD)
        bwd2s_launch_one<SYM_ODD, SYM_ODD, RXC, RYC, P>(a, nblocks, stream);
    else
        return false;
    return true;
}

}  // namespace dipamd
