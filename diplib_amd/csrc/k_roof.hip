// Streaming copy kernels that measure what the device's memory system sustains for
// the FIR path's traffic shapes: uint8 -> sfloat (1 byte read, 4 written per sample, the
// shape of the 2-D uint8 cases) and sfloat -> sfloat (4 + 4).  Each thread moves 16
// samples with 16-byte loads and stores; one thread per 16 samples, no loop.  A
// measurement helper (bench.py reports the kernels' fractions against it); it is no
// part of the drop-in interface, so include/ does not declare it.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dipamd.h"

namespace {

template <bool NT>
__device__ __forceinline__ void put4(float4 *p, float4 v) {
    if constexpr (NT) {
        __builtin_nontemporal_store(v.x, &p->x);
        __builtin_nontemporal_store(v.y, &p->y);
        __builtin_nontemporal_store(v.z, &p->z);
        __builtin_nontemporal_store(v.w, &p->w);
    } else {
        *p = v;
    }
}

template <bool NT>
__global__ __launch_bounds__(256) void k_copy_u8_f32_il(const uint32_t *__restrict__ in, float4 *__restrict__ out,
                                                        int64_t n16) {
    const int64_t base = blockIdx.x * static_cast<int64_t>(blockDim.x) * 4;
    const int64_t nw = n16 * 4;
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t w = base + k * blockDim.x + threadIdx.x;
        v[k] = w < nw ? in[w] : 0u;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t w = base + k * blockDim.x + threadIdx.x;
        if (w < nw)
            put4<NT>(out + w, make_float4(static_cast<float>(v[k] & 0xffu), static_cast<float>((v[k] >> 8) & 0xffu),
                                          static_cast<float>((v[k] >> 16) & 0xffu), static_cast<float>(v[k] >> 24)));
    }
}

template <bool NT>
__global__ __launch_bounds__(256) void k_copy_f32_f32_il(const float4 *__restrict__ in, float4 *__restrict__ out,
                                                         int64_t n16) {
    const int64_t base = blockIdx.x * static_cast<int64_t>(blockDim.x) * 4;
    const int64_t nv = n16 * 4;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t w = base + k * blockDim.x + threadIdx.x;
        if (w < nv) v[k] = in[w];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t w = base + k * blockDim.x + threadIdx.x;
        if (w < nv) put4<NT>(out + w, v[k]);
    }
}

template <bool NT>
__global__ __launch_bounds__(256) void k_copy_u8_f32(const uint4 *__restrict__ in, float4 *__restrict__ out, int64_t n16) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n16) return;
    const uint4 w = in[i];
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t v = ws[k];
        put4<NT>(out + 4 * i + k, make_float4(static_cast<float>(v & 0xffu), static_cast<float>((v >> 8) & 0xffu),
                                              static_cast<float>((v >> 16) & 0xffu), static_cast<float>(v >> 24)));
    }
}

template <bool NT>
__global__ __launch_bounds__(256) void k_copy_f32_f32(const float4 *__restrict__ in, float4 *__restrict__ out, int64_t n16) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n16) return;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = in[4 * i + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) put4<NT>(out + 4 * i + k, v[k]);
}

}  // namespace

// n samples (a multiple of 16, both buffers 16-byte aligned); flags bit 0: non-temporal
// stores, bit 1: the interleaved layout.  Returns 0 or a negative DIPAMD_E_* code.
extern "C" int dipamd_measure_copy(const void *src, int32_t src_dt, void *dst, int64_t n, uint32_t flags,
                                   void *stream) {
    if (!src || !dst || n <= 0 || (n & 15) || (reinterpret_cast<uintptr_t>(src) & 15) ||
        (reinterpret_cast<uintptr_t>(dst) & 15))
        return DIPAMD_E_PARAMETER;
    const int64_t n16 = n / 16;
    const dim3 grid(static_cast<unsigned>((n16 + 255) / 256));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const bool nt = flags & 1u;
    if (flags & 2u) {
        if (src_dt == DIPAMD_DT_UINT8 && !nt)
            hipLaunchKernelGGL(k_copy_u8_f32_il<false>, grid, dim3(256), 0, st, static_cast<const uint32_t *>(src),
                               static_cast<float4 *>(dst), n16);
        else if (src_dt == DIPAMD_DT_UINT8)
            hipLaunchKernelGGL(k_copy_u8_f32_il<true>, grid, dim3(256), 0, st, static_cast<const uint32_t *>(src),
                               static_cast<float4 *>(dst), n16);
        else if (src_dt == DIPAMD_DT_SFLOAT && !nt)
            hipLaunchKernelGGL(k_copy_f32_f32_il<false>, grid, dim3(256), 0, st, static_cast<const float4 *>(src),
                               static_cast<float4 *>(dst), n16);
        else if (src_dt == DIPAMD_DT_SFLOAT)
            hipLaunchKernelGGL(k_copy_f32_f32_il<true>, grid, dim3(256), 0, st, static_cast<const float4 *>(src),
                               static_cast<float4 *>(dst), n16);
        else
            return DIPAMD_E_PARAMETER;
        return hipGetLastError() == hipSuccess ? DIPAMD_OK : DIPAMD_E_HIP;
    }
    if (src_dt == DIPAMD_DT_UINT8) {
        if (nt)
            hipLaunchKernelGGL(k_copy_u8_f32<true>, grid, dim3(256), 0, st, static_cast<const uint4 *>(src),
                               static_cast<float4 *>(dst), n16);
        else
            hipLaunchKernelGGL(k_copy_u8_f32<false>, grid, dim3(256), 0, st, static_cast<const uint4 *>(src),
                               static_cast<float4 *>(dst), n16);
    } else if (src_dt == DIPAMD_DT_SFLOAT) {
        if (nt)
            hipLaunchKernelGGL(k_copy_f32_f32<true>, grid, dim3(256), 0, st, static_cast<const float4 *>(src),
                               static_cast<float4 *>(dst), n16);
        else
            hipLaunchKernelGGL(k_copy_f32_f32<false>, grid, dim3(256), 0, st, static_cast<const float4 *>(src),
                               static_cast<float4 *>(dst), n16);
    } else {
        return DIPAMD_E_PARAMETER;
    }
    return hipGetLastError() == hipSuccess ? DIPAMD_OK : DIPAMD_E_HIP;
}
