// Element-wise combinations of the derivative family (derivative.cpp:190-209,
// 404-416), over contiguous buffer-type images, in the reference's operation order
// and type: each product and sum rounded in the buffer type (no contraction), the
// square root correctly rounded (computed in double, which rounds once more exactly
// for a square root of a float), dip::Abs for a single dim.
#include "sep_kernels_impl.hpp"

namespace dipamd {

namespace {

struct EwArgs {
    const void *d[DIPAMD_MAX_DIMS];  // the derivative images
    void *out;
    int64_t n;
    int32_t nd;
    int32_t op;  // 0: Laplace (sum), 1: gradient magnitude
};

template <typename T>
__device__ __forceinline__ T ew_sqrt(T v);
template <>
__device__ __forceinline__ float ew_sqrt<float>(float v) {
    return static_cast<float>(__builtin_sqrt(static_cast<double>(v)));
}
template <>
__device__ __forceinline__ double ew_sqrt<double>(double v) {
    return __builtin_sqrt(v);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_combine(const EwArgs a) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < a.n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const T d0 = static_cast<const T *>(a.d[0])[i];
        T s;
        if (a.op == 0) {
            s = d0;
            for (int k = 1; k < a.nd; ++k) s = s + static_cast<const T *>(a.d[k])[i];
        } else if (a.nd == 1) {
            s = static_cast<T>(__builtin_fabs(static_cast<double>(d0)));  // sign bit cleared, as std::abs
        } else {
            s = d0 * d0;
            for (int k = 1; k < a.nd; ++k) {
                const T dk = static_cast<const T *>(a.d[k])[i];
                const T q = dk * dk;
                s = s + q;
            }
            s = ew_sqrt<T>(s);
        }
        static_cast<T *>(a.out)[i] = s;
    }
}

}  // namespace

hipError_t launch_combine(bool is_double, int op, const void *const *d, int nd, void *out, int64_t n,
                          hipStream_t stream) {
    EwArgs a;
    for (int k = 0; k < DIPAMD_MAX_DIMS; ++k) a.d[k] = k < nd ? d[k] : nullptr;
    a.out = out;
    a.n = n;
    a.nd = nd;
    a.op = op;
    int64_t nblk = (n + kBlock - 1) / kBlock;
    if (nblk > 65536) nblk = 65536;
    if (nblk < 1) nblk = 1;
    if (is_double)
        hipLaunchKernelGGL(k_combine<double>, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, stream, a);
    else
        hipLaunchKernelGGL(k_combine<float>, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

}  // namespace dipamd
