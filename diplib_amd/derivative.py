"""The derivative family on the Gaussian FIR path, with PyDIP's names and arguments:
dip.Derivative, dip.Gradient, dip.GradientMagnitude, dip.Hessian, dip.Laplace, dip.Dgg,
dip.LaplacePlusDgg, dip.LaplaceMinusDgg (derivative.cpp:88-521; linear.h:666, 697,
862, 904) and dip.StructureTensor (structure_tensor.cpp:32-57; analysis.h:454).

Tensor results (Gradient: the partial derivatives in dim order; Hessian: the
symmetric matrix as dip stores it, diagonal first, then the upper triangle column by
column) come back as a tensor whose LAST axis holds the elements, as numpy shows a
PyDIP tensor image; the elements are stored as planes, so each is a contiguous image.
"""
import ctypes

import torch

from . import _lib
from ._lib import check, lib
from .gauss import (Gauss, _as_list, _bc_codes, _flex_torch, _stream_handle, _to_device, best_gauss_method,
                    image_struct)


def _gauss_method(method, sigmas, orders, tensor=False):
    """The Gaussian method the derivatives run (dip::Derivative passes `method` on to
    dip::Gauss, derivative.cpp:111-114), as C ABI flags: 0 for FIR, the IIR or FT flag
    otherwise (the structure tensor's smoothing has flags of its own)."""
    m = "best" if method == "gauss" else (method[5:] if method[:5] == "gauss" else method)
    if m == "best":
        m = best_gauss_method(sigmas, orders)
    if m in ("FIR", "fir"):
        return 0
    if m in ("IIR", "iir"):
        return _lib.TENSOR_IIR if tensor else _lib.GAUSS_IIR
    if m in ("FT", "ft"):
        return _lib.TENSOR_FT if tensor else _lib.GAUSS_FT
    if m == "finitediff":
        raise NotImplementedError("method 'finitediff' is outside this library (Gaussian methods only)")
    raise _lib.DipError(_lib.E_PARAMETER, "Invalid flag: " + method)


def _diff_dims(x, sigmas, process):
    nd = x.dim()
    sig = [float(s) for s in _as_list(sigmas)] or [1.0]
    sig = sig * nd if len(sig) == 1 else sig
    proc = [bool(p) for p in _as_list(process)] or [True]
    proc = proc * nd if len(proc) == 1 else proc
    if len(sig) != nd or len(proc) != nd:
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    sizes = list(reversed(x.shape))  # dip order
    return sig, proc, [d for d in range(nd) if proc[d] and sizes[d] > 1 and sig[d] > 0.0]


def _run(function, x, outs, sig, proc, bc, truncation, fast, method_flags=0):
    nd = x.dim()
    bcs = _bc_codes(bc)
    imgs = (_lib.DipamdImage * len(outs))(*[image_struct(o) for o in outs])
    s_arr = (ctypes.c_double * nd)(*sig)
    p_arr = (ctypes.c_uint8 * nd)(*[1 if p else 0 for p in proc])
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    a_in = image_struct(x)
    with torch.cuda.device(x.device):
        check(lib().dipamd_derivative(function, ctypes.byref(a_in), imgs, len(outs), s_arr, nd, b_arr, len(bcs),
                                      p_arr, nd, float(truncation),
                                      (_lib.FAST if fast else _lib.EXACT) | method_flags,
                                      _stream_handle(x.device)))


def _tensor_result(function, in_, sigmas, method, boundaryCondition, process, truncation, fast):
    x, was_np = _to_device(in_)
    sig, proc, dims = _diff_dims(x, sigmas, process)
    if not dims:
        raise _lib.DipError(_lib.E_PARAMETER, "Dimensionality not supported")
    mflags = _gauss_method(method, sig, [2] * x.dim() if function == _lib.FN_HESSIAN else [1] * x.dim())
    n = check(lib().dipamd_derivative_outputs(function, len(dims)))
    planes = torch.empty((n,) + tuple(x.shape), dtype=_flex_torch(x.dtype), device=x.device)
    _run(function, x, [planes[i] for i in range(n)], sig, proc, boundaryCondition, truncation, fast, mflags)
    y = planes.movedim(0, -1)
    return y.cpu().numpy() if was_np else y


def _scalar_result(function, in_, sigmas, method, boundaryCondition, process, truncation, out, fast):
    x, was_np = _to_device(in_)
    sig, proc, dims = _diff_dims(x, sigmas, process)
    if not dims:
        raise _lib.DipError(_lib.E_PARAMETER, "Dimensionality not supported")
    mflags = _gauss_method(method, sig, [2] * x.dim())
    y = torch.empty(x.shape, dtype=_flex_torch(x.dtype), device=x.device) if out is None else out
    _run(function, x, [y], sig, proc, boundaryCondition, truncation, fast, mflags)
    return y.cpu().numpy() if (was_np and out is None) else y


def Derivative(in_, derivativeOrder=(0,), sigmas=(1.0,), method="best", boundaryCondition=(), truncation=3.0,
               out=None, fast=False):
    """dip::Derivative (derivative.cpp:88-115), Gaussian methods: dip::Gauss."""
    return Gauss(in_, sigmas, derivativeOrder, method, boundaryCondition, truncation, out=out, fast=fast)


def Gradient(in_, sigmas=(1.0,), method="best", boundaryCondition=(), process=(), truncation=3.0, fast=False):
    """dip::Gradient (derivative.cpp:142-172)."""
    return _tensor_result(_lib.FN_GRADIENT, in_, sigmas, method, boundaryCondition, process, truncation, fast)


def Hessian(in_, sigmas=(1.0,), method="best", boundaryCondition=(), process=(), truncation=3.0, fast=False):
    """dip::Hessian (derivative.cpp:330-372)."""
    return _tensor_result(_lib.FN_HESSIAN, in_, sigmas, method, boundaryCondition, process, truncation, fast)


def GradientMagnitude(in_, sigmas=(1.0,), method="best", boundaryCondition=(), process=(), truncation=3.0, out=None,
                      fast=False):
    """dip::GradientMagnitude (derivative.cpp:174-210)."""
    return _scalar_result(_lib.FN_GRADIENT_MAGNITUDE, in_, sigmas, method, boundaryCondition, process, truncation, out,
                          fast)


def Laplace(in_, sigmas=(1.0,), method="best", boundaryCondition=(), process=(), truncation=3.0, out=None,
            fast=False):
    """dip::Laplace (derivative.cpp:374-418), Gaussian methods."""
    return _scalar_result(_lib.FN_LAPLACE, in_, sigmas, method, boundaryCondition, process, truncation, out, fast)


def Dgg(in_, sigmas=(1.0,), method="best", boundaryCondition=(), process=(), truncation=3.0, out=None, fast=False):
    """dip::Dgg (derivative.cpp:486-496): the second derivative in the gradient
    direction, as the reference computes it (for two dims its sum over the off-diagonal
    Hessian terms starts at index 1 and so leaves the mixed term out)."""
    return _scalar_result(_lib.FN_DGG, in_, sigmas, method, boundaryCondition, process, truncation, out, fast)


def LaplacePlusDgg(in_, sigmas=(1.0,), method="best", boundaryCondition=(), process=(), truncation=3.0, out=None,
                   fast=False):
    """dip::LaplacePlusDgg (derivative.cpp:498-508)."""
    return _scalar_result(_lib.FN_LAPLACE_PLUS_DGG, in_, sigmas, method, boundaryCondition, process, truncation, out,
                          fast)


def LaplaceMinusDgg(in_, sigmas=(1.0,), method="best", boundaryCondition=(), process=(), truncation=3.0, out=None,
                    fast=False):
    """dip::LaplaceMinusDgg (derivative.cpp:510-520)."""
    return _scalar_result(_lib.FN_LAPLACE_MINUS_DGG, in_, sigmas, method, boundaryCondition, process, truncation,
                          out, fast)


def _masked_args(x, mask, sigmas, boundaryCondition):
    m, _ = _to_device(mask)
    if m.device != x.device:
        m = m.to(x.device)
    nd = x.dim()
    sig = [float(s) for s in _as_list(sigmas)] or [1.0]
    sig = sig * nd if len(sig) == 1 else sig
    if len(sig) != nd:
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    bcs = _bc_codes(boundaryCondition)
    return m, sig, (ctypes.c_double * nd)(*sig), (ctypes.c_int32 * max(1, len(bcs)))(*bcs), len(bcs)


def NormalizedConvolution(in_, mask, sigmas=(1.0,), method="best", boundaryCondition=(), truncation=3.0, out=None):
    """dip::NormalizedConvolution (derivative.cpp:522-542): Gauss(in x mask) / Gauss(mask),
    0 where the denominator is 0; a boolean mask zeroes the pixels outside it."""
    x, was_np = _to_device(in_)
    m, sig, s_arr, b_arr, nbc = _masked_args(x, mask, sigmas, boundaryCondition)
    y = torch.empty(x.shape, dtype=_flex_torch(x.dtype), device=x.device) if out is None else out
    flags = _gauss_method(method, sig, [0] * x.dim())
    a_in, a_m, a_out = image_struct(x), image_struct(m), image_struct(y)
    with torch.cuda.device(x.device):
        check(lib().dipamd_normalized_convolution(ctypes.byref(a_in), ctypes.byref(a_m), ctypes.byref(a_out), s_arr,
                                                  len(sig), b_arr, nbc, float(truncation), flags,
                                                  _stream_handle(x.device)))
    return y.cpu().numpy() if (was_np and out is None) else y


def NormalizedDifferentialConvolution(in_, mask, dimension=0, sigmas=(1.0,), method="best", boundaryCondition=(),
                                      truncation=3.0, out=None):
    """dip::NormalizedDifferentialConvolution (derivative.cpp:544-600): the first
    derivative along `dimension` (dip order) of the masked input, normalized by the
    mask's Gaussian."""
    x, was_np = _to_device(in_)
    m, sig, s_arr, b_arr, nbc = _masked_args(x, mask, sigmas, boundaryCondition)
    y = torch.empty(x.shape, dtype=_flex_torch(x.dtype), device=x.device) if out is None else out
    flags = _gauss_method(method, sig, [1] * x.dim())
    a_in, a_m, a_out = image_struct(x), image_struct(m), image_struct(y)
    with torch.cuda.device(x.device):
        check(lib().dipamd_normalized_differential_convolution(
            ctypes.byref(a_in), ctypes.byref(a_m), ctypes.byref(a_out), int(dimension), s_arr, len(sig), b_arr, nbc,
            float(truncation), flags, _stream_handle(x.device)))
    return y.cpu().numpy() if (was_np and out is None) else y


def StructureTensor(in_, mask=None, gradientSigmas=(1.0,), tensorSigmas=(5.0,), method="best",
                    boundaryCondition=(), truncation=3.0, fast=False):
    """dip::StructureTensor (structure_tensor.cpp:32-57): the outer product of the
    gradient, each element smoothed with `tensorSigmas`; with a mask the gradient is the
    normalized differential convolution of every dim.  Elements on the last axis, as dip
    stores a symmetric matrix (diagonal first, then the upper triangle column by
    column)."""
    x, was_np = _to_device(in_)
    if mask is not None:
        return _structure_tensor_masked(x, was_np, mask, gradientSigmas, tensorSigmas, method, boundaryCondition,
                                        truncation, fast)
    nd = x.dim()
    gs = [float(s) for s in _as_list(gradientSigmas)] or [1.0]
    ts = [float(s) for s in _as_list(tensorSigmas)] or [5.0]
    gs = gs * nd if len(gs) == 1 else gs
    ts = ts * nd if len(ts) == 1 else ts
    if len(gs) != nd or len(ts) != nd:
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    sizes = list(reversed(x.shape))
    dims = [d for d in range(nd) if sizes[d] > 1 and gs[d] > 0.0]
    if not dims:
        raise _lib.DipError(_lib.E_PARAMETER, "Dimensionality not supported")
    flags = (_lib.FAST if fast else _lib.EXACT) | _gauss_method(method, gs, [1] * nd) | \
        _gauss_method(method, ts, [0] * nd, tensor=True)
    n = check(lib().dipamd_derivative_outputs(_lib.FN_STRUCTURE_TENSOR, len(dims)))
    planes = torch.empty((n,) + tuple(x.shape), dtype=_flex_torch(x.dtype), device=x.device)
    bcs = _bc_codes(boundaryCondition)
    imgs = (_lib.DipamdImage * n)(*[image_struct(planes[i]) for i in range(n)])
    g_arr = (ctypes.c_double * nd)(*gs)
    t_arr = (ctypes.c_double * nd)(*ts)
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    a_in = image_struct(x)
    with torch.cuda.device(x.device):
        check(lib().dipamd_structure_tensor(ctypes.byref(a_in), imgs, n, g_arr, nd, t_arr, nd, b_arr, len(bcs),
                                            float(truncation), flags, _stream_handle(x.device)))
    y = planes.movedim(0, -1)
    return y.cpu().numpy() if was_np else y


def _structure_tensor_masked(x, was_np, mask, gradientSigmas, tensorSigmas, method, boundaryCondition, truncation,
                             fast):
    nd = x.dim()
    m, gs, g_arr, b_arr, nbc = _masked_args(x, mask, gradientSigmas, boundaryCondition)
    ts = [float(s) for s in _as_list(tensorSigmas)] or [5.0]
    ts = ts * nd if len(ts) == 1 else ts
    if len(ts) != nd:
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "Array parameter has the wrong number of elements")
    flags = (_lib.FAST if fast else _lib.EXACT) | _gauss_method(method, gs, [1] * nd) | \
        _gauss_method(method, ts, [0] * nd, tensor=True)
    n = nd * (nd + 1) // 2
    planes = torch.empty((n,) + tuple(x.shape), dtype=_flex_torch(x.dtype), device=x.device)
    imgs = (_lib.DipamdImage * n)(*[image_struct(planes[i]) for i in range(n)])
    t_arr = (ctypes.c_double * nd)(*ts)
    a_in, a_m = image_struct(x), image_struct(m)
    with torch.cuda.device(x.device):
        check(lib().dipamd_structure_tensor_masked(ctypes.byref(a_in), ctypes.byref(a_m), imgs, n, g_arr, nd, t_arr,
                                                   nd, b_arr, nbc, float(truncation), flags,
                                                   _stream_handle(x.device)))
    y = planes.movedim(0, -1)
    return y.cpu().numpy() if was_np else y
