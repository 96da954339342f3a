// Host-side launchers of the separable FIR kernels (defined in sep_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sep_common.hpp"

namespace dipamd {

// Input sample type of a kernel's first pass; other input types are converted to
// the buffer type by launch_cast_in first.
enum InType : int { IN_U8 = 0, IN_BUF = 1 };

// Fused 2-D tile configurations: columns per workgroup and outputs per lane.
struct FusedConfig {
    int txc;
    int p;
};
constexpr int kNumFusedConfigs = 3;
constexpr FusedConfig kFusedConfigs[kNumFusedConfigs] = {{64, 4}, {128, 8}, {256, 16}};

// LDS bytes of the line kernel for a given border.
size_t line_lds_bytes(int border, bool is_double);
// LDS bytes of the fused kernel for a given config and radii.
size_t fused_lds_bytes(int cfg, int rx, int ry, bool is_double);
// Rows per workgroup step of the fused kernel.
constexpr int kFusedStep = 16;

hipError_t launch_line(InType ti, bool is_double, int sym, bool fast, const void *args, int64_t n, int64_t nlines,
                       int border, hipStream_t stream);
hipError_t launch_sweep(InType ti, bool is_double, int sym, bool fast, const void *args, int64_t e, int64_t nchunks,
                        int64_t nouter, hipStream_t stream);
// Register-window sweep for symmetric radii 1..kZSweepMaxR (k_zsweep.hip).
hipError_t launch_zsweep(InType ti, bool is_double, int r, int sym, bool fast, const void *args, int64_t e,
                         int64_t nchunks, int64_t nouter, hipStream_t stream);
// The same sweep of a buffer-type image `args.in` (no output written) into
// *out = <u, sweep(in)> in FP64, deterministic: one partial per workgroup in
// `partials` (at least as many doubles as workgroups), then their fixed-order sum.
hipError_t launch_zsweep_dot(bool is_double, int r, int sym, bool fast, const void *args, int64_t e, int64_t nchunks,
                             int64_t nouter, const void *u, double *partials, double *out, hipStream_t stream);
// Register-streaming fused kernel for uint8 planes (k_stream2d.hip).
hipError_t launch_stream2d(int symx, int symy, bool fast, const void *args, hipStream_t stream);
// The derivative family's streaming form (k_stream2d_deriv.hip); args: DerivStreamArgs.
hipError_t launch_stream2d_deriv(int mode, bool fast, const void *args, hipStream_t stream);
// Column-tile sweep: rows per block for a border (0 when the block does not fit).
int colsweep_rows(int border, bool is_double);
hipError_t launch_colsweep(InType ti, bool is_double, int sym, bool fast, const void *args, int64_t e,
                           int64_t nchunks, int64_t nouter, hipStream_t stream);
hipError_t launch_fused(InType ti, bool is_double, int cfg, int symx, int symy, bool fast, const void *args,
                        int64_t nx, int64_t nchunks, int64_t nouter, int rx, int ry, hipStream_t stream);

// The fused kernel with a compile-time x radius class rxc (a multiple of 4 up to
// kFusedRcMaxClass, or 0 with symx == SYM_NONE; fused_rc.hpp).  Needs a unit x stride;
// hipErrorNotSupported for combinations it does not cover.
constexpr int kFusedRcMaxClass = 24;
// Radius class of an x radius for the radius-class kernels: the next multiple of 4, at
// least 4 (a radius of 0 runs the centre tap only).
inline int rc_class_of(int r) { return r <= 4 ? 4 : ((r + 3) / 4) * 4; }
hipError_t launch_fused_rc(InType ti, bool is_double, int cfg, int symx, int symy, bool fast, const void *args,
                           int64_t nx, int64_t nchunks, int64_t nouter, int rxc, int ry, hipStream_t stream);
// LDS bytes of the radius-class fused kernel.
size_t fused_rc_lds_bytes(int cfg, int rxc, int ry, bool is_double);

// Extended lines for the boundary conditions computed from edge values (k_expand.hip):
// `border` samples computed on each side, then `zpad` samples of 0 beyond them.
hipError_t launch_expand(InType ti, bool is_double, const void *in, void *out, const void *geo, int d, int n,
                         int border, int zpad, int bc, hipStream_t stream);

hipError_t launch_cast_in(int dt_in, bool to_double, const void *in, void *out, const void *geo, int64_t total,
                          hipStream_t stream);
hipError_t launch_cast_out(bool from_double, int dt_out, const void *in, void *out, const void *geo, int64_t total,
                           hipStream_t stream);

}  // namespace dipamd
