// Element-wise combinations of the derivative family (derivative.cpp:190-209,
// 404-416, 424-484) and of the structure tensor (structure_tensor.cpp:55;
// arithmetic.cpp:116-170), over contiguous buffer-type images, in the reference's order
// and type: each product and sum rounded in the buffer type (no contraction), the
// square root correctly rounded (hipcc's default for sqrtf; on gfx950 it equals the
// double square root rounded to float for every one of the 2^32 float inputs, NaNs
// included), dip::Abs for a single dim.
#include "sep_kernels_impl.hpp"

namespace dipamd {

namespace {

constexpr int kEwMaxIn = 2 * DIPAMD_MAX_DIMS + DIPAMD_MAX_DIMS * (DIPAMD_MAX_DIMS + 1) / 2;

struct EwArgs {
    const void *d[kEwMaxIn];  // the derivative images (gradient, then Hessian for Dgg)
    void *out[DIPAMD_MAX_DIMS * (DIPAMD_MAX_DIMS + 1) / 2];
    int64_t n;
    int32_t nd;
    int32_t op;  // 0: Laplace (sum), 1: gradient magnitude, 2/3/4: Dgg, Laplace + Dgg,
                 // Laplace - Dgg, 5: the symmetric products g g^T of the structure tensor
};

template <typename T>
__device__ __forceinline__ T ew_sqrt(T v);
template <>
__device__ __forceinline__ float ew_sqrt<float>(float v) {
    return __builtin_sqrtf(v);
}
template <>
__device__ __forceinline__ double ew_sqrt<double>(double v) {
    return __builtin_sqrt(v);
}

template <typename T>
__device__ __forceinline__ T ew_in(const EwArgs &a, int k, int64_t i) {
    return static_cast<const T *>(a.d[k])[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_combine(const EwArgs a) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < a.n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int n = a.nd;
        if (a.op == 5) {
            // StructureTensor's g * g^T (arithmetic.cpp:116-170): a sum started at 0 for each
            // element, the diagonal first, then above the diagonal column by column
            int e = 0;
            for (int c = 0; c < n; ++c) {
                const T g = ew_in<T>(a, c, i);
                static_cast<T *>(a.out[e++])[i] = T(0) + g * g;
            }
            for (int c = 1; c < n; ++c)
                for (int r = 0; r < c; ++r)
                    static_cast<T *>(a.out[e++])[i] = T(0) + ew_in<T>(a, r, i) * ew_in<T>(a, c, i);
            continue;
        }
        if (a.op >= 2) {
            // the Dgg family (derivative.cpp:424-484): g (n images), then H (diagonal, then
            // above the diagonal column by column); each image operation rounded in T
            const T g0 = ew_in<T>(a, 0, i);
            T gs = g0 * g0;
            T o = gs * ew_in<T>(a, n, i);
            for (int k = 1; k < n; ++k) {
                const T gk = ew_in<T>(a, k, i);
                T t = gk * gk;
                gs = gs + t;
                t = t * ew_in<T>(a, n + k, i);
                o = o + t;
            }
            for (int r = 1; r < n - 1; ++r) {  // the reference's loop starts at 1
                for (int c = r + 1; c < n; ++c) {
                    T t = ew_in<T>(a, r, i) * ew_in<T>(a, c, i);
                    t = t * ew_in<T>(a, n + n + c * (c - 1) / 2 + r, i);
                    o = o * T(1) + t * T(2);  // LinearCombination(out, tmp, 1, 2)
                }
            }
            o = gs == T(0) ? T(0) : o / gs;  // SafeDivide
            if (a.op != 2) {
                T tr = T(0);  // Trace: std::accumulate from 0 over the diagonal
                for (int k = 0; k < n; ++k) tr = tr + ew_in<T>(a, n + k, i);
                o = a.op == 3 ? tr + o : tr - o;
            }
            static_cast<T *>(a.out[0])[i] = o;
            continue;
        }
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
        static_cast<T *>(a.out[0])[i] = s;
    }
}

}  // namespace

// d: the input images (op 0/1: nd derivatives; op 2-4: nd gradient images then the
// nd(nd+1)/2 Hessian elements; op 5: nd gradient images); out: one image (op 5:
// nd(nd+1)/2 images).
hipError_t launch_combine(bool is_double, int op, const void *const *d, int nd, void *const *out, int64_t n,
                          hipStream_t stream) {
    EwArgs a;
    const int nin = op >= 2 && op <= 4 ? nd + nd * (nd + 1) / 2 : nd;
    const int nout = op == 5 ? nd * (nd + 1) / 2 : 1;
    for (int k = 0; k < kEwMaxIn; ++k) a.d[k] = k < nin ? d[k] : nullptr;
    for (int k = 0; k < DIPAMD_MAX_DIMS * (DIPAMD_MAX_DIMS + 1) / 2; ++k) a.out[k] = k < nout ? out[k] : nullptr;
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

namespace {

// The masked forms (derivative.cpp:522-600), over contiguous buffer-type arrays, each
// image operation rounded in the buffer type.  op 0: w = m != 0 ? x : 0 (a binary mask,
// Image::Mask); op 1: w = x * m (MultiplySampleWise); op 2: a = SafeDivide(a, den);
// op 3: out = SafeDivide(dw - a * dm, den) with a = NC (the NDC's last lines).
template <typename T>
__device__ __forceinline__ T safe_div(T a, T b) {
    return b == T(0) ? T(0) : a / b;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_masked(int op, const T *x, const T *m, T *a, const T *den, const T *dw,
                                                   const T *dm, T *out, int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (op == 0) {
            a[i] = m[i] != T(0) ? x[i] : T(0);
        } else if (op == 1) {
            a[i] = x[i] * m[i];
        } else if (op == 2) {
            a[i] = safe_div(a[i], den[i]);
        } else {
            const T p = a[i] * dm[i];
            const T d = dw[i] - p;
            out[i] = safe_div(d, den[i]);
        }
    }
}

}  // namespace

hipError_t launch_masked(bool is_double, int op, const void *x, const void *m, void *a, const void *den, const void *dw,
                         const void *dm, void *out, int64_t n, hipStream_t stream) {
    int64_t nblk = (n + kBlock - 1) / kBlock;
    if (nblk > 65536) nblk = 65536;
    if (nblk < 1) nblk = 1;
    if (is_double)
        hipLaunchKernelGGL(k_masked<double>, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, stream, op,
                           static_cast<const double *>(x), static_cast<const double *>(m), static_cast<double *>(a),
                           static_cast<const double *>(den), static_cast<const double *>(dw),
                           static_cast<const double *>(dm), static_cast<double *>(out), n);
    else
        hipLaunchKernelGGL(k_masked<float>, dim3(static_cast<unsigned>(nblk)), dim3(kBlock), 0, stream, op,
                           static_cast<const float *>(x), static_cast<const float *>(m), static_cast<float *>(a),
                           static_cast<const float *>(den), static_cast<const float *>(dw),
                           static_cast<const float *>(dm), static_cast<float *>(out), n);
    return hipGetLastError();
}

}  // namespace dipamd
