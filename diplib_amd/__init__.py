"""diplib_amd -- MI355X-native Gaussian FIR filtering with DIPlib's semantics.

The compute path is libdipamd.so (HIP kernels for gfx950 behind the C ABI in
include/dipamd.h); this package is its PyDIP-style front end on PyTorch-ROCm
device tensors.  Importing the package does not touch the GPU.
"""
from ._lib import DipError, lib  # noqa: F401
from .autograd import gauss_fir_autograd, gauss_fir_backward  # noqa: F401
from .batch import GaussBatch, gauss_batch_noalloc  # noqa: F401
from .derivative import (Derivative, Dgg, Gradient, GradientMagnitude, Hessian, Laplace,  # noqa: F401
                         LaplaceMinusDgg, LaplacePlusDgg, NormalizedConvolution, NormalizedDifferentialConvolution,
                         StructureTensor)
from .gauss import (Gauss, GaussFIR, GaussFT, GaussIIR, MakeHalfGaussian, OneDimensionalFilter,  # noqa: F401
                    SeparableConvolution, autotune, best_gauss_method, gauss_radius, image_struct)

__all__ = ["Gauss", "GaussFIR", "GaussFT", "GaussIIR", "SeparableConvolution", "OneDimensionalFilter", "MakeHalfGaussian", "DipError",
           "gauss_radius", "best_gauss_method", "lib", "Derivative", "Gradient", "GradientMagnitude", "Hessian",
           "Laplace", "Dgg", "LaplacePlusDgg", "LaplaceMinusDgg", "StructureTensor", "NormalizedConvolution",
           "NormalizedDifferentialConvolution"]
