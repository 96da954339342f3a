// Radius-class batched kernel instantiations: input double, buffer double.
#include "k_batchrc_common.hpp"

namespace dipamd {

// Launches the group's persistent kernel; with occ non-null, only reports the
// workgroups per CU for lds bytes of dynamic LDS.
hipError_t batched_rc_entry_f64(int cfg, bool fast, int rxc, const void *args, int grid, size_t lds, hipStream_t st,
                                 int *occ) {
    return fast ? batched_rc_cfg<double, double, true>(cfg, rxc, args, grid, lds, st, occ)
                : batched_rc_cfg<double, double, false>(cfg, rxc, args, grid, lds, st, occ);
}

}  // namespace dipamd
