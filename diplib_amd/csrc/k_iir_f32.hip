// Instances of the recursive Gaussian line kernel: float input samples, strided or unaligned filter axis.
#include "k_iir.hpp"

namespace dipamd {

namespace {

template <typename TO, int NN>
hipError_t iir_nn(int order, const IirArgs &a, hipStream_t st) {
    const unsigned nblk = static_cast<unsigned>((a.nlines + kIirBlock - 1) / kIirBlock);
    switch (order) {
        case 0: hipLaunchKernelGGL((k_iir<float, TO, NN, 0, false>), dim3(nblk), dim3(kIirBlock), 0, st, a); break;
        case 1: hipLaunchKernelGGL((k_iir<float, TO, NN, 1, false>), dim3(nblk), dim3(kIirBlock), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_iir<float, TO, NN, 2, false>), dim3(nblk), dim3(kIirBlock), 0, st, a); break;
        case 3: hipLaunchKernelGGL((k_iir<float, TO, NN, 3, false>), dim3(nblk), dim3(kIirBlock), 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_iir<float, TO, NN, 4, false>), dim3(nblk), dim3(kIirBlock), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename TO>
hipError_t iir_to(int nn, int order, const IirArgs &a, hipStream_t st) {
    switch (nn) {
        case 1: return iir_nn<TO, 1>(order, a, st);
        case 2: return iir_nn<TO, 2>(order, a, st);
        case 3: return iir_nn<TO, 3>(order, a, st);
        case 4: return iir_nn<TO, 4>(order, a, st);
        case 5: return iir_nn<TO, 5>(order, a, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_iir_f32(int nn, int order, const IirArgs &a, hipStream_t st) {
    return a.out_double ? iir_to<double>(nn, order, a, st) : iir_to<float>(nn, order, a, st);
}

}  // namespace dipamd
