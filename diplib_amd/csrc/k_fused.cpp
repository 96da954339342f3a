// Dispatch of the fused 2-D kernel over its per-type translation units.
#include "sep_kernels.hpp"

namespace dipamd {

#define DIPAMD_FUSED_ENTRY(name)                                                                              \
    hipError_t name(int cfg, int sx, int sy, const void *args, int64_t nx, int64_t nc, int64_t no, int rx, int ry, \
                    hipStream_t st);
DIPAMD_FUSED_ENTRY(fused_entry_u8_0)
DIPAMD_FUSED_ENTRY(fused_entry_u8_1)
DIPAMD_FUSED_ENTRY(fused_entry_f32_0)
DIPAMD_FUSED_ENTRY(fused_entry_f32_1)
DIPAMD_FUSED_ENTRY(fused_entry_f64_0)
DIPAMD_FUSED_ENTRY(fused_entry_f64_1)

#define DIPAMD_FUSED_RC_ENTRY(name)                                                                      \
    hipError_t name(int cfg, int sx, int sy, int rxc, const void *args, int64_t nx, int64_t nc, int64_t no, \
                    int ry, hipStream_t st);
DIPAMD_FUSED_RC_ENTRY(fused_rc_entry_u8_0)
DIPAMD_FUSED_RC_ENTRY(fused_rc_entry_u8_1)
DIPAMD_FUSED_RC_ENTRY(fused_rc_entry_f32_0)
DIPAMD_FUSED_RC_ENTRY(fused_rc_entry_f32_1)
DIPAMD_FUSED_RC_ENTRY(fused_rc_entry_f64_0)
DIPAMD_FUSED_RC_ENTRY(fused_rc_entry_f64_1)

hipError_t launch_fused_rc(InType ti, bool is_double, int cfg, int symx, int symy, bool fast, const void *args,
                           int64_t nx, int64_t nchunks, int64_t nouter, int rxc, int ry, hipStream_t stream) {
    if (is_double)
        return (fast ? fused_rc_entry_f64_1 : fused_rc_entry_f64_0)(cfg, symx, symy, rxc, args, nx, nchunks, nouter,
                                                                    ry, stream);
    if (ti == IN_U8)
        return (fast ? fused_rc_entry_u8_1 : fused_rc_entry_u8_0)(cfg, symx, symy, rxc, args, nx, nchunks, nouter, ry,
                                                                  stream);
    return (fast ? fused_rc_entry_f32_1 : fused_rc_entry_f32_0)(cfg, symx, symy, rxc, args, nx, nchunks, nouter, ry,
                                                                stream);
}

hipError_t launch_fused(InType ti, bool is_double, int cfg, int symx, int symy, bool fast, const void *args,
                        int64_t nx, int64_t nchunks, int64_t nouter, int rx, int ry, hipStream_t stream) {
    if (is_double)
        return (fast ? fused_entry_f64_1 : fused_entry_f64_0)(cfg, symx, symy, args, nx, nchunks, nouter, rx, ry, stream);
    if (ti == IN_U8)
        return (fast ? fused_entry_u8_1 : fused_entry_u8_0)(cfg, symx, symy, args, nx, nchunks, nouter, rx, ry, stream);
    return (fast ? fused_entry_f32_1 : fused_entry_f32_0)(cfg, symx, symy, args, nx, nchunks, nouter, rx, ry, stream);
}

}  // namespace dipamd
