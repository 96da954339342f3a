// Radius-class fused 2-D kernel instantiations: input double, buffer double, fast=false.
#include "k_fusedrc_common.hpp"

namespace dipamd {

hipError_t fused_rc_entry_f64_0(int cfg, int sx, int sy, int rxc, const void *args, int64_t nx, int64_t nc,
                                    int64_t no, int ry, hipStream_t st) {
    return fused_rc_cfg<double, double, false>(cfg, sx, sy, rxc, args, nx, nc, no, ry, st);
}

}  // namespace dipamd
