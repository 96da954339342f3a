"""PyDIP-style front end of the Gaussian FIR path on PyTorch-ROCm device tensors.

Mirrors PyDIP's `dip.Gauss` / `dip.GaussFIR` / `dip.SeparableConvolution`
(pydip/src/filtering.cpp:121-136) with DIPlib's conventions:

* axis order: a tensor of shape (..., H, W) is a dip image of sizes (W, H, ...);
  sigma / order / boundary lists are in dip order (x first), as in PyDIP
  (pydip/src/image.cpp:142-146 reverses the axes);
* output type: `SuggestFlex(in)` -- sfloat, or dfloat for 32-bit integer and dfloat
  input (datatype.cpp:109-124); a caller-supplied `out` keeps its own type and is
  written with clamp_cast, like a protected dip::Image;
* method "best" resolves as BestGaussMethod (derivative.cpp:35-58).

Device tensors are processed in place on the current HIP stream; numpy arrays are
copied to the GPU and back (the copies are not part of the kernel path).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import DipamdFilter, DipamdImage, check, lib

_TORCH_DT = {
    torch.uint8: _lib.DT_UINT8, torch.int8: _lib.DT_SINT8, torch.int16: _lib.DT_SINT16, torch.int32: _lib.DT_SINT32,
    torch.float32: _lib.DT_SFLOAT, torch.float64: _lib.DT_DFLOAT, torch.bool: _lib.DT_BIN,
    torch.complex64: _lib.DT_SCOMPLEX, torch.complex128: _lib.DT_DCOMPLEX,
}
for _name, _code in (("uint16", _lib.DT_UINT16), ("uint32", _lib.DT_UINT32)):
    if hasattr(torch, _name):
        _TORCH_DT[getattr(torch, _name)] = _code


def _flex_torch(dt):
    """SuggestFlex (datatype.cpp:109-124)."""
    if dt in (torch.complex64, torch.complex128):
        return dt
    if dt in (torch.float64, torch.int32) or (hasattr(torch, "uint32") and dt == torch.uint32):
        return torch.float64
    return torch.float32


def _complex_torch(dt):
    """SuggestComplex (datatype.cpp:95-106)."""
    if dt in (torch.float64, torch.int32, torch.complex128) or (hasattr(torch, "uint32") and dt == torch.uint32):
        return torch.complex128
    return torch.complex64


def image_struct(t):
    """DipamdImage describing torch tensor `t` (dip axis order: reversed)."""
    if t.dtype not in _TORCH_DT:
        raise TypeError(f"data type {t.dtype} not supported")
    if t.dim() < 1 or t.dim() > _lib.MAX_DIMS:
        raise ValueError("dimensionality not supported")
    im = DipamdImage()
    im.origin = ctypes.c_void_p(t.data_ptr())
    im.dtype = _TORCH_DT[t.dtype]
    im.ndim = t.dim()
    for d in range(t.dim()):
        ax = t.dim() - 1 - d
        im.sizes[d] = t.shape[ax]
        im.strides[d] = t.stride(ax)
    return im


def _stream_handle(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _as_list(v):
    if v is None:
        return []
    if isinstance(v, (str, bytes)):
        return [v]
    if hasattr(v, "__len__"):
        return list(v)
    return [v]


def _bc_codes(bc):
    out = []
    for b in _as_list(bc):
        b = b.decode() if isinstance(b, bytes) else b
        if b not in _lib.BC_CODES:
            raise _lib.DipError(_lib.E_PARAMETER, "Boundary condition not recognized: " + str(b))
        out.append(_lib.BC_CODES[b])
    return out


def best_gauss_method(sigmas, orders):
    """BestGaussMethod (derivative.cpp:35-58)."""
    if any(o > 3 for o in orders):
        return "ft"
    if any(0.0 < s < 0.8 for s in sigmas):
        return "ft"
    if any(s > 10 for s in sigmas):
        return "iir"
    return "fir"


def _to_device(x):
    """(device tensor, was_numpy)."""
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            return x.cuda(), True
        return x, False
    a = np.asarray(x)
    if a.dtype == np.uint16:
        t = torch.from_numpy(a.astype(np.int32)).cuda()  # widened: same values, same output type (sfloat)
        return t.to(torch.uint16) if hasattr(torch, "uint16") else t, True
    return torch.from_numpy(np.ascontiguousarray(a)).cuda(), True


def _prepare_out(x, out, default_dtype):
    if out is None:
        return torch.empty(x.shape, dtype=default_dtype, device=x.device)
    if tuple(out.shape) != tuple(x.shape):
        raise _lib.DipError(_lib.E_SIZES, "Image is protected")  # sizes of a protected out cannot change
    return out


def GaussFIR(in_, sigmas=(1.0,), derivativeOrder=(0,), boundaryCondition=(), truncation=3.0, out=None, fast=False,
             fusion=None):
    """dip::GaussFIR (linear.h:300-318).  Returns the output tensor (or numpy array
    for numpy input).  `fast=True` allows FMA (not bit-identical to the reference)."""
    x, was_np = _to_device(in_)
    sig = [float(s) for s in _as_list(sigmas)]
    ords = [int(o) for o in _as_list(derivativeOrder)]
    nd = x.dim()
    if len(sig) not in (1, nd) or len(ords) not in (0, 1, nd):
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    bcs = _bc_codes(boundaryCondition)
    y = _prepare_out(x, out, _flex_torch(x.dtype))
    L = lib()
    a_in, a_out = image_struct(x), image_struct(y)
    s_arr = (ctypes.c_double * max(1, len(sig)))(*sig)
    o_arr = (ctypes.c_int32 * max(1, len(ords)))(*ords)
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    flags = _lib.FAST if fast else _lib.EXACT
    if fusion is False:
        flags |= _lib.NO_FUSION
    elif fusion is True:
        flags |= _lib.FORCE_FUSION
    with torch.cuda.device(x.device):
        check(L.dipamd_gauss_fir(ctypes.byref(a_in), ctypes.byref(a_out), s_arr, len(sig), o_arr, len(ords), b_arr,
                                 len(bcs), float(truncation), flags, _stream_handle(x.device)))
    if was_np and out is None:
        return y.cpu().numpy()
    return y


IIR_DESIGN_METHODS = {"forward backward": _lib.IIR_FORWARD_BACKWARD, "discrete time fit": _lib.IIR_DISCRETE_TIME_FIT}


def GaussIIR(in_, sigmas=(1.0,), derivativeOrder=(0,), boundaryCondition=(), filterOrder=(),
             designMethod="discrete time fit", truncation=3.0, out=None):
    """dip::GaussIIR (linear.h:407-429; gaussiir.cpp:741-806): the recursive Gaussian
    and its derivatives up to order 4, computed in double, bit-identical to the
    reference.  An empty `filterOrder` means 3 + order, capped at 5."""
    x, was_np = _to_device(in_)
    sig = [float(s) for s in _as_list(sigmas)]
    ords = [int(o) for o in _as_list(derivativeOrder)]
    fos = [int(f) for f in _as_list(filterOrder)]
    nd = x.dim()
    if len(sig) not in (0, 1, nd) or len(ords) not in (0, 1, nd) or len(fos) not in (0, 1, nd):
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    if designMethod not in IIR_DESIGN_METHODS:
        raise _lib.DipError(_lib.E_PARAMETER, "Invalid flag: " + str(designMethod))
    bcs = _bc_codes(boundaryCondition)
    y = _prepare_out(x, out, _flex_torch(x.dtype))
    L = lib()
    a_in, a_out = image_struct(x), image_struct(y)
    s_arr = (ctypes.c_double * max(1, len(sig)))(*sig)
    o_arr = (ctypes.c_int32 * max(1, len(ords)))(*ords)
    f_arr = (ctypes.c_int32 * max(1, len(fos)))(*fos)
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    with torch.cuda.device(x.device):
        check(L.dipamd_gauss_iir(ctypes.byref(a_in), ctypes.byref(a_out), s_arr, len(sig), o_arr, len(ords), b_arr,
                                 len(bcs), f_arr, len(fos), IIR_DESIGN_METHODS[designMethod], float(truncation), 0,
                                 _stream_handle(x.device)))
    if was_np and out is None:
        return y.cpu().numpy()
    return y


def GaussFT(in_, sigmas=(1.0,), derivativeOrder=(0,), truncation=3.0, inRepresentation="spatial",
            outRepresentation="spatial", boundaryCondition=(), out=None):
    """dip::GaussFT (linear.h:355-364; gauss.cpp:397-487): the Gaussian and its
    derivatives of any order through the frequency domain (rocFFT), equal to the
    reference to FFT rounding.  An empty `boundaryCondition` transforms the image as it
    is (periodic).  inRepresentation / outRepresentation "frequency": the input is / the
    result is the centred spectrum (dip::FourierTransform's), complex; those forms take
    no boundary condition."""
    for r in (inRepresentation, outRepresentation):
        if r not in ("spatial", "frequency"):
            raise _lib.DipError(_lib.E_PARAMETER, "Invalid flag: " + r)
    in_freq, out_freq = inRepresentation == "frequency", outRepresentation == "frequency"
    x, was_np = _to_device(in_)
    sig = [float(s) for s in _as_list(sigmas)]
    ords = [int(o) for o in _as_list(derivativeOrder)]
    nd = x.dim()
    if len(sig) not in (0, 1, nd) or len(ords) not in (0, 1, nd):
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    bcs = _bc_codes(boundaryCondition)
    freq = in_freq or out_freq
    y = _prepare_out(x, out, _complex_torch(x.dtype) if freq else _flex_torch(x.dtype))
    L = lib()
    a_in, a_out = image_struct(x), image_struct(y)
    s_arr = (ctypes.c_double * max(1, len(sig)))(*sig)
    o_arr = (ctypes.c_int32 * max(1, len(ords)))(*ords)
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    flags = (_lib.FT_IN_FREQUENCY if in_freq else 0) | (_lib.FT_OUT_FREQUENCY if out_freq else 0)
    with torch.cuda.device(x.device):
        check(L.dipamd_gauss_ft(ctypes.byref(a_in), ctypes.byref(a_out), s_arr, len(sig), o_arr, len(ords), b_arr,
                                len(bcs), float(truncation), flags, _stream_handle(x.device)))
    if was_np and out is None:
        return y.cpu().numpy()
    return y


def Gauss(in_, sigmas=(1.0,), derivativeOrder=(0,), method="best", boundaryCondition=(), truncation=3.0, out=None,
          fast=False):
    """dip::Gauss (linear.h:452-472; derivative.cpp:62-86): "FIR", "IIR" and "FT", and
    "best" choosing among them as the reference does."""
    m = method[5:] if method[:5] == "gauss" else method
    sig = [float(s) for s in _as_list(sigmas)]
    ords = [int(o) for o in _as_list(derivativeOrder)]
    if m == "best":
        m = best_gauss_method(sig, ords)
    if m in ("FIR", "fir"):
        return GaussFIR(in_, sig, ords, boundaryCondition, truncation, out=out, fast=fast)
    if m in ("IIR", "iir"):
        return GaussIIR(in_, sig, ords, boundaryCondition, (), "discrete time fit", truncation, out=out)
    if m in ("FT", "ft"):
        # the empty condition is not the default for GaussFT (derivative.cpp:74-77)
        bc = boundaryCondition if _as_list(boundaryCondition) else ["default"]
        return GaussFT(in_, sig, ords, truncation, "spatial", "spatial", bc, out=out)
    raise _lib.DipError(_lib.E_PARAMETER, "Invalid flag: " + method)


class OneDimensionalFilter:
    """dip::OneDimensionalFilter (linear.h:84-90).  A complex filter is given as
    Python complex values, or with isComplex=True as (re, im) pairs of floats as the
    reference stores it."""

    def __init__(self, filter, origin=-1, symmetry="general", isComplex=False):
        vals = list(filter)
        if any(isinstance(v, complex) for v in vals) and not isComplex:
            isComplex = True
            flat = []
            for v in vals:
                c = complex(v)
                flat += [c.real, c.imag]
            vals = flat
        self.filter = [float(v) for v in vals]
        if isComplex and len(self.filter) % 2:
            raise _lib.DipError(_lib.E_PARAMETER, "Complex filter must have an even number of values")
        self.origin = int(origin)
        self.symmetry = symmetry
        self.isComplex = bool(isComplex)


def SeparableConvolution(in_, filter, boundaryCondition=(), process=(), out=None, fast=False, fusion=None):
    """dip::SeparableConvolution (linear.h:114-130) with real 1-D filters.  `fusion`
    False runs one pass per axis, True the fused 2-D kernel wherever it applies, None
    lets the planner choose."""
    x, was_np = _to_device(in_)
    filters = filter if isinstance(filter, (list, tuple)) else [filter]
    nd = x.dim()
    if len(filters) not in (1, nd):
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    keep = []
    farr = (DipamdFilter * len(filters))()
    for i, f in enumerate(filters):
        if f.symmetry not in _lib.SYM_CODES:
            raise _lib.DipError(_lib.E_PARAMETER, "Symmetry string not recognized: " + f.symmetry)
        taps = (ctypes.c_double * max(1, len(f.filter)))(*f.filter)
        keep.append(taps)
        cplx = getattr(f, "isComplex", False)
        farr[i].taps = ctypes.cast(taps, ctypes.POINTER(ctypes.c_double))
        farr[i].ntaps = len(f.filter) // 2 if cplx else len(f.filter)
        farr[i].origin = f.origin
        farr[i].symmetry = _lib.SYM_CODES[f.symmetry]
        farr[i].is_complex = 1 if cplx else 0
    proc = list(process) if process is not None else []
    if len(proc) not in (0, nd):
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    p_arr = (ctypes.c_uint8 * nd)(*[1 if p else 0 for p in proc]) if proc else None
    bcs = _bc_codes(boundaryCondition)
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    any_complex = any(getattr(f, "isComplex", False) for f in filters)
    y = _prepare_out(x, out, _complex_torch(x.dtype) if any_complex else _flex_torch(x.dtype))
    a_in, a_out = image_struct(x), image_struct(y)
    flags = _lib.FAST if fast else _lib.EXACT
    if fusion is False:
        flags |= _lib.NO_FUSION
    elif fusion is True:
        flags |= _lib.FORCE_FUSION
    with torch.cuda.device(x.device):
        check(lib().dipamd_separable_convolution(ctypes.byref(a_in), ctypes.byref(a_out), farr, len(filters), b_arr,
                                                 len(bcs), p_arr, flags, _stream_handle(x.device)))
    if was_np and out is None:
        return y.cpu().numpy()
    return y


def MakeHalfGaussian(sigma, derivativeOrder=0, truncation=3.0, dataType="SFLOAT"):
    """dip::MakeHalfGaussian (generation.h:638-653): r+1 taps, x = 0 last."""
    buf = (ctypes.c_double * 4096)()
    n = check(lib().dipamd_make_half_gaussian(float(sigma), int(derivativeOrder), float(truncation),
                                              1 if dataType.upper() == "DFLOAT" else 0, buf, 4096))
    return [buf[i] for i in range(n)]


def gauss_radius(sigma, order=0, truncation=3.0, is_double=False):
    return lib().dipamd_gauss_radius(float(sigma), int(order), float(truncation), int(is_double))


def autotune(fast=False, quick=False):
    """Measures the plain path's schedule choices on this GPU (dipamd_autotune) and
    stores them in the persistent tuning cache the planner reads on every call.
    Results never depend on the choices; only the speed does."""
    dev = torch.device("cuda", torch.cuda.current_device())
    flags = (_lib.FAST if fast else _lib.EXACT) | (_lib.TUNE_QUICK if quick else 0)
    with torch.cuda.device(dev):
        check(lib().dipamd_autotune(flags, _stream_handle(dev)))
