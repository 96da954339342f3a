"""ctypes binding of libdipamd.so, the C ABI declared in include/dipamd.h.

The library is the only compute path: when it is missing or fails to load, every
call raises.  There is no CPU fallback.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdipamd.so")

MAX_DIMS = 8

DT_UINT8, DT_UINT16, DT_UINT32, DT_SINT8, DT_SINT16, DT_SINT32, DT_SFLOAT, DT_DFLOAT, DT_BIN = range(9)
DT_SCOMPLEX, DT_DCOMPLEX = 9, 10

BC_CODES = {
    "": 0, "default": 0, "mirror": 0, "asym mirror": 1, "antisym reflect": 2, "periodic": 3,
    "asym periodic": 4, "add zeros": 5, "add max": 6, "add min": 7, "zero order": 8,
    "first order": 9, "second order": 10, "third order": 11, "already expanded": 12,
}
SYM_CODES = {"": 0, "general": 0, "even": 1, "odd": 2, "conj": 3, "d-even": 4, "d-odd": 5, "d-conj": 6}

EXACT, FAST, NO_FUSION, FORCE_FUSION = 0, 1, 2, 4

E_PARAMETER, E_NOT_FORGED, E_ARRAY_LENGTH, E_SIZES, E_UNSUPPORTED, E_HIP, E_ORDER = -1, -2, -3, -4, -5, -6, -7


class DipamdImage(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_void_p), ("dtype", ctypes.c_int32), ("ndim", ctypes.c_int32),
                ("sizes", ctypes.c_int64 * MAX_DIMS), ("strides", ctypes.c_int64 * MAX_DIMS)]


class DipamdFilter(ctypes.Structure):
    _fields_ = [("taps", ctypes.POINTER(ctypes.c_double)), ("ntaps", ctypes.c_int32), ("origin", ctypes.c_int32),
                ("symmetry", ctypes.c_int32), ("is_complex", ctypes.c_int32)]


class DipamdBatchImage(ctypes.Structure):
    _fields_ = [("inp", ctypes.c_void_p), ("out", ctypes.c_void_p), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("in_stride_y", ctypes.c_int64), ("out_stride_y", ctypes.c_int64),
                ("sigma_x", ctypes.c_double), ("sigma_y", ctypes.c_double), ("order_x", ctypes.c_int32),
                ("order_y", ctypes.c_int32)]


class DipamdTraceRecord(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("ms", ctypes.c_double), ("bytes", ctypes.c_int64),
                ("ops", ctypes.c_int64)]


class DipError(RuntimeError):
    """Error raised by the library (dip::ParameterError / dip::RunTimeError counterpart)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    """Loads libdipamd.so (once).  Raises when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m diplib_amd.build` "
                           "(or __graft_entry__.build()); there is no fallback path")
    L = ctypes.CDLL(LIB_PATH)
    P, I32, U32, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_double
    IMG = ctypes.POINTER(DipamdImage)
    L.dipamd_gauss_fir.argtypes = [IMG, IMG, P, I32, P, I32, P, I32, D, U32, P]
    L.dipamd_separable_convolution.argtypes = [IMG, IMG, ctypes.POINTER(DipamdFilter), I32, P, I32, P, U32, P]
    L.dipamd_make_half_gaussian.argtypes = [D, I32, D, I32, P, I32]
    L.dipamd_gauss_radius.argtypes = [D, I32, D, I32]
    L.dipamd_last_error.restype = ctypes.c_char_p
    L.dipamd_trace_enable.argtypes = [I32]
    L.dipamd_trace_read.argtypes = [ctypes.POINTER(DipamdTraceRecord), I32]
    L.dipamd_version.restype = ctypes.c_char_p
    if hasattr(L, "dipamd_batch_create"):
        L.dipamd_batch_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(DipamdBatchImage), I32, I32, I32, P, I32,
                                          D, U32, I32, P]
        L.dipamd_batch_run.argtypes = [P, P]
        L.dipamd_batch_run_graph.argtypes = [P, P]
        L.dipamd_batch_run_per_image.argtypes = [P, P]
        L.dipamd_batch_destroy.argtypes = [P]
        L.dipamd_batch_info.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(I32)]
    if hasattr(L, "dipamd_batch_gauss_noalloc"):
        L.dipamd_batch_gauss_noalloc.argtypes = [ctypes.POINTER(DipamdBatchImage), I32, I32, I32, P, I32, D, U32, I32,
                                                 P]
    L.dipamd_derivative.argtypes = [I32, IMG, IMG, I32, P, I32, P, I32, P, I32, D, U32, P]
    L.dipamd_gauss_iir.argtypes = [IMG, IMG, P, I32, P, I32, P, I32, P, I32, I32, D, U32, P]
    L.dipamd_gauss_ft.argtypes = [IMG, IMG, P, I32, P, I32, P, I32, D, U32, P]
    L.dipamd_derivative_outputs.argtypes = [I32, I32]
    L.dipamd_structure_tensor.argtypes = [IMG, IMG, I32, P, I32, P, I32, P, I32, D, U32, P]
    L.dipamd_normalized_convolution.argtypes = [IMG, IMG, IMG, P, I32, P, I32, D, U32, P]
    L.dipamd_normalized_differential_convolution.argtypes = [IMG, IMG, IMG, I32, P, I32, P, I32, D, U32, P]
    L.dipamd_structure_tensor_masked.argtypes = [IMG, IMG, IMG, I32, P, I32, P, I32, P, I32, D, U32, P]
    if hasattr(L, "dipamd_autotune"):
        L.dipamd_autotune.argtypes = [U32, P]
    if hasattr(L, "dipamd_measure_copy"):  # measurement helper, not part of the drop-in interface
        L.dipamd_measure_copy.argtypes = [P, I32, P, ctypes.c_int64, U32, P]
    if hasattr(L, "dipamd_gauss_fir_backward"):
        L.dipamd_gauss_fir_backward.argtypes = [IMG, IMG, IMG, P, P, I32, P, I32, P, I32, D, U32, P]
    L.dipamd_gauss_fir_keep.argtypes = [IMG, IMG, P, I32, P, I32, P, I32, D, U32, IMG, ctypes.POINTER(I32), P]
    L.dipamd_gauss_fir_backward_kept.argtypes = [IMG, IMG, IMG, P, P, I32, P, I32, P, I32, D, U32, IMG, P]
    _lib = L
    return L


def check(status):
    if status < 0:
        msg = lib().dipamd_last_error().decode(errors="replace")
        raise DipError(status, msg)
    return status


EXPORTED_SYMBOLS = [
    "dipamd_gauss_fir", "dipamd_separable_convolution", "dipamd_make_half_gaussian", "dipamd_batch_create",
    "dipamd_batch_run", "dipamd_batch_run_graph", "dipamd_batch_run_per_image", "dipamd_batch_destroy", "dipamd_batch_num_configs",
    "dipamd_batch_info", "dipamd_gauss_fir_backward", "dipamd_gauss_radius", "dipamd_last_error", "dipamd_version",
    "dipamd_trace_enable", "dipamd_trace_read", "dipamd_derivative", "dipamd_derivative_outputs",
    "dipamd_structure_tensor", "dipamd_gauss_iir", "dipamd_gauss_ft", "dipamd_batch_gauss_noalloc",
    "dipamd_autotune", "dipamd_normalized_convolution", "dipamd_normalized_differential_convolution",
    "dipamd_structure_tensor_masked", "dipamd_separable_lines_count", "dipamd_separable_lines_gather",
    "dipamd_separable_lines_scatter", "dipamd_gauss_fir_keep", "dipamd_gauss_fir_backward_kept",
]
TUNE_QUICK = 256
FT_IN_FREQUENCY, FT_OUT_FREQUENCY = 256, 512

FN_GRADIENT, FN_GRADIENT_MAGNITUDE, FN_HESSIAN, FN_LAPLACE = 0, 1, 2, 3
IIR_FORWARD_BACKWARD, IIR_DISCRETE_TIME_FIT = 1, 2
GAUSS_IIR, TENSOR_IIR, GAUSS_FT, TENSOR_FT = 8, 16, 32, 64
GRAD_SIGMA_TOTAL = 128  # dipamd_gauss_fir_backward: one dL/dsigma, summed over the dims
FN_DGG, FN_LAPLACE_PLUS_DGG, FN_LAPLACE_MINUS_DGG, FN_STRUCTURE_TENSOR = 4, 5, 6, 7
