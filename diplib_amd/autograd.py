"""Differentiable dip::Gauss (FIR) on PyTorch-ROCm: a torch.autograd.Function whose
forward is the library's exact drop-in and whose backward is the library's adjoint
kernel path (C ABI: dipamd_gauss_fir_keep, dipamd_gauss_fir_backward_kept: for images of
three or more dims the forward keeps its intermediate after dims 0 and 1, which the
sigma derivative of the later dims needs).

    y = gauss_fir_autograd(x, sigma, orders=(0,), bc=("mirror",), truncation=3.0, sigma_dims=None)

`sigma` is a tensor: a scalar (one sigma for the dims in `sigma_dims`, all dims by
default; the other dims are not filtered) or one value per dip dim.  Dims are in
dip order (dim 0 = the last tensor axis).  dL/dx is the exact adjoint of the
boundary-extended separable convolution; dL/dsigma differentiates the normalised
discrete taps with the radius held at its forward value (the radius is a step
function of sigma, so the derivative is taken on the piece the forward used).
"""
import ctypes

import torch

from . import _lib
from ._lib import check, lib
from .gauss import GaussFIR, _bc_codes, _flex_torch, _stream_handle, image_struct


def _per_dim_sigmas(sigma, nd, sigma_dims, sigma_host=None):
    # the taps are computed on the host: a device-resident sigma is read back (which
    # waits for the stream) unless the caller passes its value as sigma_host
    if sigma_host is not None:
        s = [float(v) for v in (sigma_host if hasattr(sigma_host, "__len__") else [sigma_host])]
        if len(s) != sigma.numel():
            raise _lib.DipError(_lib.E_ARRAY_LENGTH, "sigma_host must have as many values as sigma")
    else:
        s = sigma.detach().double().cpu().reshape(-1).tolist()
    if len(s) == 1:
        dims = range(nd) if sigma_dims is None else sigma_dims
        out = [0.0] * nd
        for d in dims:
            out[d] = s[0]
        return out
    if len(s) != nd:
        raise _lib.DipError(_lib.E_ARRAY_LENGTH, "sigma must have one value or one per dimension")
    return s


def _gauss_fir_keep(x, sig, orders, bc, truncation, fast):
    """GaussFIR that also returns the intermediate after dims 0 and 1 (or None when the
    first pass is not the fused one over them)."""
    y = torch.empty(x.shape, dtype=_flex_torch(x.dtype), device=x.device)
    keep = torch.empty(x.shape, dtype=_flex_torch(x.dtype), device=x.device)
    a_in, a_out, a_keep = image_struct(x), image_struct(y), image_struct(keep)
    s_arr = (ctypes.c_double * len(sig))(*sig)
    o_arr = (ctypes.c_int32 * max(1, len(orders)))(*orders)
    bcs = _bc_codes(bc)
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    kept = ctypes.c_int32(0)
    with torch.cuda.device(x.device):
        check(lib().dipamd_gauss_fir_keep(ctypes.byref(a_in), ctypes.byref(a_out), s_arr, len(sig), o_arr,
                                          len(orders), b_arr, len(bcs), float(truncation),
                                          _lib.FAST if fast else _lib.EXACT, ctypes.byref(a_keep),
                                          ctypes.byref(kept), _stream_handle(x.device)))
    return y, (keep if kept.value else None)


class _GaussFIRFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, sigma, orders, bc, truncation, sigma_dims, fast, sigma_host):
        nd = x.dim()
        sig = _per_dim_sigmas(sigma, nd, sigma_dims, sigma_host)
        kept = None
        if nd >= 3 and ctx.needs_input_grad[1]:
            # the forward's intermediate after dims 0 and 1 serves the sigma derivative of
            # the later dims in the backward, which would otherwise compute it again
            y, kept = _gauss_fir_keep(x, sig, list(orders), list(bc), truncation, fast)
        else:
            y = GaussFIR(x, sig, list(orders), list(bc), truncation, fast=fast)
        if kept is not None:
            ctx.save_for_backward(x, sigma, kept)
        else:
            ctx.save_for_backward(x, sigma)
        ctx.cfg = (sig, list(orders), list(bc), float(truncation), sigma_dims, fast)
        return y

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        x, sigma = saved[0], saved[1]
        kept = saved[2] if len(saved) > 2 else None
        sig, orders, bc, truncation, sigma_dims, fast = ctx.cfg
        nd = x.dim()
        want_dx = ctx.needs_input_grad[0]
        want_ds = ctx.needs_input_grad[1]
        g = g.to(_flex_torch(x.dtype))
        dx = torch.empty(x.shape, dtype=_flex_torch(x.dtype), device=x.device) if want_dx else None
        # a scalar sigma: the library sums dL/dsigma over the dims itself (the dims it
        # does not cover have sigma 0, are not filtered and contribute nothing)
        total = want_ds and sigma.numel() == 1
        ds = torch.empty(1 if total else nd, dtype=torch.float64, device=x.device) if want_ds else None
        flags = (_lib.FAST if fast else _lib.EXACT) | (_lib.GRAD_SIGMA_TOTAL if total else 0)
        a_x, a_g = image_struct(x), image_struct(g)
        a_dx = image_struct(dx) if want_dx else _lib.DipamdImage()
        s_arr = (ctypes.c_double * nd)(*sig)
        o_arr = (ctypes.c_int32 * max(1, len(orders)))(*orders)
        bcs = _bc_codes(bc)
        b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
        a_k = image_struct(kept) if kept is not None else None
        with torch.cuda.device(x.device):
            check(lib().dipamd_gauss_fir_backward_kept(
                ctypes.byref(a_x), ctypes.byref(a_g), ctypes.byref(a_dx),
                ctypes.c_void_p(ds.data_ptr()) if want_ds else None, s_arr, nd, o_arr, len(orders), b_arr, len(bcs),
                truncation, flags, ctypes.byref(a_k) if a_k is not None else None, _stream_handle(x.device)))
        # on sigma's own device: a CPU sigma (e.g. a CPU nn.Parameter) with a GPU image
        dsigma = ds.to(device=sigma.device, dtype=sigma.dtype).reshape(sigma.shape) if want_ds else None
        if want_dx:
            dx = dx.to(x.dtype)
        return dx, dsigma, None, None, None, None, None, None


def gauss_fir_autograd(x, sigma, orders=(0,), bc=("mirror",), truncation=3.0, sigma_dims=None, fast=False,
                       sigma_host=None):
    """Differentiable GaussFIR: gradients with respect to `x` (float tensors) and `sigma`.
    `sigma_host` (optional): the value(s) of `sigma` on the host, which the caller
    guarantees to match, so that a device-resident sigma is not read back each call."""
    if not torch.is_tensor(sigma):
        sigma = torch.tensor(float(sigma), dtype=torch.float64, device=x.device)
    orders = [int(o) for o in (orders if hasattr(orders, "__len__") else [orders])]
    bc = [bc] if isinstance(bc, str) else list(bc)
    dims = None if sigma_dims is None else tuple(int(d) for d in sigma_dims)
    host = None if sigma_host is None else tuple(float(v) for v in (sigma_host if hasattr(sigma_host, "__len__")
                                                                     else [sigma_host]))
    return _GaussFIRFunction.apply(x, sigma, tuple(orders), tuple(bc), float(truncation), dims, bool(fast), host)


def gauss_fir_backward(x, grad_out, sigmas, orders=(0,), bc=("mirror",), truncation=3.0, fast=False,
                       want_dx=True, want_dsigma=True, sigma_total=False):
    """Direct call of the backward kernels: returns (dL/dx, dL/dsigma per dip dim), or
    with sigma_total dL/dsigma as one value summed over the dims."""
    nd = x.dim()
    sig = [float(s) for s in (sigmas if hasattr(sigmas, "__len__") else [sigmas])]
    if len(sig) == 1:
        sig = sig * nd
    orders = [int(o) for o in (orders if hasattr(orders, "__len__") else [orders])]
    bcs = _bc_codes(bc)
    dx = torch.empty(x.shape, dtype=_flex_torch(x.dtype), device=x.device) if want_dx else None
    ds = torch.empty(1 if sigma_total else nd, dtype=torch.float64, device=x.device) if want_dsigma else None
    flags = (_lib.FAST if fast else _lib.EXACT) | (_lib.GRAD_SIGMA_TOTAL if sigma_total else 0)
    a_x, a_g = image_struct(x), image_struct(grad_out)
    a_dx = image_struct(dx) if want_dx else _lib.DipamdImage()
    s_arr = (ctypes.c_double * nd)(*sig)
    o_arr = (ctypes.c_int32 * max(1, len(orders)))(*orders)
    b_arr = (ctypes.c_int32 * max(1, len(bcs)))(*bcs)
    with torch.cuda.device(x.device):
        check(lib().dipamd_gauss_fir_backward(
            ctypes.byref(a_x), ctypes.byref(a_g), ctypes.byref(a_dx),
            ctypes.c_void_p(ds.data_ptr()) if want_dsigma else None, s_arr, nd, o_arr, len(orders), b_arr, len(bcs),
            float(truncation), flags, _stream_handle(x.device)))
    return dx, ds
