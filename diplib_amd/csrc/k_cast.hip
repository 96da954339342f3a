// Launch tables (see sep_kernels_impl.hpp).
#include "sep_kernels_impl.hpp"

namespace dipamd {
namespace {
template <typename TS, typename TB>
hipError_t cast_in_t(const void *in, void *out, const CastGeo &g, int64_t total, hipStream_t st) {
    int64_t nblk = (total + kBlock - 1) / kBlock;
    if (nblk > 65536) nblk = 65536;
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL((k_cast_in<TS, TB>), dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, st,
                       static_cast<const TS *>(in), static_cast<TB *>(out), g);
    return hipGetLastError();
}
template <typename TS, typename TD>
hipError_t cast_out_t(const void *in, void *out, const CastGeo &g, int64_t total, hipStream_t st) {
    int64_t nblk = (total + kBlock - 1) / kBlock;
    if (nblk > 65536) nblk = 65536;
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL((k_cast_out<TS, TD>), dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, st,
                       static_cast<const TS *>(in), static_cast<TD *>(out), g);
    return hipGetLastError();
}
template <typename TB>
hipError_t cast_in_dt(int dt, const void *in, void *out, const CastGeo &g, int64_t total, hipStream_t st) {
    switch (dt) {
        case DIPAMD_DT_UINT8: return cast_in_t<uint8_t, TB>(in, out, g, total, st);
        case DIPAMD_DT_UINT16: return cast_in_t<uint16_t, TB>(in, out, g, total, st);
        case DIPAMD_DT_UINT32: return cast_in_t<uint32_t, TB>(in, out, g, total, st);
        case DIPAMD_DT_SINT8: return cast_in_t<int8_t, TB>(in, out, g, total, st);
        case DIPAMD_DT_SINT16: return cast_in_t<int16_t, TB>(in, out, g, total, st);
        case DIPAMD_DT_SINT32: return cast_in_t<int32_t, TB>(in, out, g, total, st);
        case DIPAMD_DT_SFLOAT: return cast_in_t<float, TB>(in, out, g, total, st);
        case DIPAMD_DT_DFLOAT: return cast_in_t<double, TB>(in, out, g, total, st);
        case DIPAMD_DT_BIN: return cast_in_t<bin_t, TB>(in, out, g, total, st);
    }
    return hipErrorInvalidValue;
}
template <typename TS>
hipError_t cast_out_dt(int dt, const void *in, void *out, const CastGeo &g, int64_t total, hipStream_t st) {
    switch (dt) {
        case DIPAMD_DT_UINT8: return cast_out_t<TS, uint8_t>(in, out, g, total, st);
        case DIPAMD_DT_UINT16: return cast_out_t<TS, uint16_t>(in, out, g, total, st);
        case DIPAMD_DT_UINT32: return cast_out_t<TS, uint32_t>(in, out, g, total, st);
        case DIPAMD_DT_SINT8: return cast_out_t<TS, int8_t>(in, out, g, total, st);
        case DIPAMD_DT_SINT16: return cast_out_t<TS, int16_t>(in, out, g, total, st);
        case DIPAMD_DT_SINT32: return cast_out_t<TS, int32_t>(in, out, g, total, st);
        case DIPAMD_DT_SFLOAT: return cast_out_t<TS, float>(in, out, g, total, st);
        case DIPAMD_DT_DFLOAT: return cast_out_t<TS, double>(in, out, g, total, st);
        case DIPAMD_DT_BIN: return cast_out_t<TS, bin_t>(in, out, g, total, st);
    }
    return hipErrorInvalidValue;
}
}  // namespace

hipError_t launch_cast_in(int dt_in, bool to_double, const void *in, void *out, const void *geo, int64_t total,
                          hipStream_t stream) {
    const CastGeo &g = *static_cast<const CastGeo *>(geo);
    return to_double ? cast_in_dt<double>(dt_in, in, out, g, total, stream)
                     : cast_in_dt<float>(dt_in, in, out, g, total, stream);
}

hipError_t launch_cast_out(bool from_double, int dt_out, const void *in, void *out, const void *geo, int64_t total,
                           hipStream_t stream) {
    const CastGeo &g = *static_cast<const CastGeo *>(geo);
    return from_double ? cast_out_dt<double>(dt_out, in, out, g, total, stream)
                       : cast_out_dt<float>(dt_out, in, out, g, total, stream);
}

}  // namespace dipamd
