"""ctypes bindings for the CPU oracle (oracle/liborc.so) and, where it was built,
the reference library wrapper (oracle/_ref/libdipref.so).  Test infrastructure:
imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg only.
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libdipref.so")

# data-type codes shared by the oracle, the reference wrapper and include/dipamd.h
DT = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.uint32): 2,
      np.dtype(np.int8): 3, np.dtype(np.int16): 4, np.dtype(np.int32): 5,
      np.dtype(np.float32): 6, np.dtype(np.float64): 7, np.dtype(np.bool_): 8,
      np.dtype(np.complex64): 9, np.dtype(np.complex128): 10}
BC = {"": 0, "default": 0, "mirror": 0, "asym mirror": 1, "antisym reflect": 2, "periodic": 3,
      "asym periodic": 4, "add zeros": 5, "add max": 6, "add min": 7, "zero order": 8,
      "first order": 9, "second order": 10, "third order": 11}
SYM = {"general": 0, "": 0, "even": 1, "odd": 2, "conj": 3, "d-even": 4, "d-odd": 5, "d-conj": 6}

_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        lib = ctypes.CDLL(ORC_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        lib.orc_gauss_radius.argtypes = [D, I, D, I]
        lib.orc_half_gaussian.argtypes = [D, I, D, I, P, I]
        lib.orc_half_gaussian_n.argtypes = [D, I, I, P]
        lib.orc_gauss_fir.argtypes = [P, I, P, I, I, P, P, I, P, I, P, I, D, I]
        lib.orc_separable.argtypes = [P, I, P, I, I, P, P, P, P, P, I, P, I, P, I]
        lib.orc_bc_map.argtypes = [ctypes.c_long, ctypes.c_long, I, ctypes.POINTER(I)]
        lib.orc_bc_map.restype = ctypes.c_long
        lib.orc_gauss_iir.argtypes = [P, I, P, I, I, P, P, I, P, I, P, I, P, I, I, D, I]
        lib.orc_expand_line.argtypes = [P, I, ctypes.c_long, ctypes.c_long, I]
        _orc = lib
    return _orc


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        lib.dipref_gauss.argtypes = [P, I, P, I, I, P, P, I, P, I, P, I, D, ctypes.c_char_p, I]
        lib.dipref_gauss_inplace.argtypes = [P, I, I, P, P, I, P, I, ctypes.c_char_p, D]
        lib.dipref_half_gaussian.argtypes = [D, I, D, I, P, I]
        lib.dipref_separable.argtypes = [P, I, P, I, I, P, P, P, P, I, P, I, P, P, P]
        lib.dipref_last_error.restype = ctypes.c_char_p
        if hasattr(lib, "dipref_gauss_ft"):
            lib.dipref_gauss_ft.argtypes = [P, I, P, I, I, P, P, I, P, I, P, I, D]
            lib.dipref_gauss_ft_repr.argtypes = [P, I, P, I, I, P, P, I, P, I, D, ctypes.c_char_p, ctypes.c_char_p]
        if hasattr(lib, "dipref_gauss_iir"):
            lib.dipref_gauss_iir.argtypes = [P, I, P, I, I, P, P, I, P, I, P, I, P, I, ctypes.c_char_p, D, I]
        if hasattr(lib, "dipref_derivative_masked"):
            lib.dipref_derivative_masked.argtypes = [ctypes.c_char_p, P, I, P, I, P, I, I, ctypes.POINTER(I), I, P, P,
                                                     I, P, I, ctypes.c_char_p, D]
        if hasattr(lib, "dipref_derivative"):
            lib.dipref_derivative.argtypes = [ctypes.c_char_p, P, I, P, I, I, ctypes.POINTER(I), I, P, P, I, P, I,
                                              ctypes.c_char_p, D]
        _ref = lib
    return _ref


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sizes(arr):
    # numpy (..., H, W) -> dip sizes (W, H, ...): dim 0 is the contiguous axis
    return np.ascontiguousarray(arr.shape[::-1], dtype=np.int64)


def _flex(dt):
    dt = np.dtype(dt)
    return np.dtype(np.float64) if dt in (np.dtype(np.uint32), np.dtype(np.int32), np.dtype(np.float64),
                                          np.dtype(np.complex128)) else np.dtype(np.float32)


def _as_list(v, n):
    v = list(v) if hasattr(v, "__len__") else [v]
    return v


def half_gaussian(sigma, order, truncation=3.0, is_double=False, use_ref=False):
    buf = np.zeros(4096, np.float64)
    if use_ref:
        n = ref().dipref_half_gaussian(sigma, order, truncation, int(is_double), _ptr(buf), 4096)
    else:
        n = orc().orc_half_gaussian(sigma, order, truncation, int(is_double), _ptr(buf), 4096)
    if n < 0:
        raise ValueError("half_gaussian failed")
    return buf[:n].copy()


def gauss_radius(sigma, order, truncation=3.0, is_double=False):
    return orc().orc_gauss_radius(sigma, order, truncation, int(is_double))


def gauss_fir(x, sigmas, orders=(0,), bc=("mirror",), truncation=3.0, out_dtype=None, nthreads=1,
              use_ref=False):
    """dip::Gauss(x, sigmas, orders, "FIR", bc, truncation) on a C-contiguous numpy array.
    sigmas / orders / bc are in dip axis order (dim 0 = last numpy axis)."""
    x = np.ascontiguousarray(x)
    out_dtype = np.dtype(out_dtype) if out_dtype is not None else _flex(x.dtype)
    out = np.zeros(x.shape, out_dtype)
    sizes = _sizes(x)
    nd = x.ndim
    sig = np.ascontiguousarray(list(sigmas) if hasattr(sigmas, "__len__") else [sigmas], np.float64)
    ords = np.ascontiguousarray(list(orders) if hasattr(orders, "__len__") else [orders], np.int32)
    bcl = list(bc) if not isinstance(bc, str) else [bc]
    if use_ref:
        arr = (ctypes.c_char_p * max(1, len(bcl)))(*[b.encode() for b in bcl])
        st = ref().dipref_gauss(_ptr(x), DT[x.dtype], _ptr(out), DT[out_dtype], nd, _ptr(sizes),
                                _ptr(sig), len(sig), _ptr(ords), len(ords), arr, len(bcl), truncation,
                                b"FIR", nthreads)
        if st != 0:
            raise RuntimeError(ref().dipref_last_error().decode())
    else:
        bci = np.ascontiguousarray([BC[b] for b in bcl], np.int32)
        st = orc().orc_gauss_fir(_ptr(x), DT[x.dtype], _ptr(out), DT[out_dtype], nd, _ptr(sizes),
                                 _ptr(sig), len(sig), _ptr(ords), len(ords), _ptr(bci), len(bcl),
                                 truncation, nthreads)
        if st != 0:
            raise RuntimeError(f"oracle gauss_fir failed ({st})")
    return out


IIR_METHODS = {"forward backward": 1, "discrete time fit": 2}


def gauss_iir(x, sigmas, orders=(0,), bc=("mirror",), filter_orders=(), method="discrete time fit",
              truncation=3.0, out_dtype=None, nthreads=1, use_ref=False):
    """dip::GaussIIR(x, sigmas, orders, bc, filter_orders, method, truncation) on a
    C-contiguous numpy array (arguments in dip axis order)."""
    x = np.ascontiguousarray(x)
    out_dtype = np.dtype(out_dtype) if out_dtype is not None else _flex(x.dtype)
    out = np.zeros(x.shape, out_dtype)
    sizes = _sizes(x)
    sig = np.ascontiguousarray(list(sigmas) if hasattr(sigmas, "__len__") else [sigmas], np.float64)
    ords = np.ascontiguousarray(list(orders) if hasattr(orders, "__len__") else [orders], np.int32)
    fo = np.ascontiguousarray(list(filter_orders) if hasattr(filter_orders, "__len__") else [filter_orders],
                              np.int32)
    bcl = list(bc) if not isinstance(bc, str) else [bc]
    if use_ref:
        arr = (ctypes.c_char_p * max(1, len(bcl)))(*[b.encode() for b in bcl])
        fop = _ptr(fo) if len(fo) else None
        st = ref().dipref_gauss_iir(_ptr(x), DT[x.dtype], _ptr(out), DT[out_dtype], x.ndim, _ptr(sizes),
                                    _ptr(sig), len(sig), _ptr(ords), len(ords), arr, len(bcl), fop, len(fo),
                                    method.encode(), truncation, nthreads)
        if st != 0:
            raise RuntimeError(ref().dipref_last_error().decode())
    else:
        bci = np.ascontiguousarray([BC[b] for b in bcl] or [0], np.int32)
        fop = _ptr(fo) if len(fo) else None
        st = orc().orc_gauss_iir(_ptr(x), DT[x.dtype], _ptr(out), DT[out_dtype], x.ndim, _ptr(sizes),
                                 _ptr(sig), len(sig), _ptr(ords), len(ords), _ptr(bci), len(bcl), fop, len(fo),
                                 IIR_METHODS[method], truncation, nthreads)
        if st != 0:
            raise RuntimeError(f"oracle gauss_iir failed ({st})")
    return out


def ref_gauss_ft(x, sigmas, orders=(0,), bc=(), truncation=3.0, out_dtype=None):
    """The reference's dip::GaussFT(x, sigmas, orders, truncation, "spatial", "spatial",
    bc) on a C-contiguous numpy array (arguments in dip axis order)."""
    x = np.ascontiguousarray(x)
    out_dtype = np.dtype(out_dtype) if out_dtype is not None else _flex(x.dtype)
    out = np.zeros(x.shape, out_dtype)
    sizes = _sizes(x)
    sig = np.ascontiguousarray(list(sigmas), np.float64)
    ords = np.ascontiguousarray(list(orders), np.int32)
    bcl = list(bc) if not isinstance(bc, str) else [bc]
    arr = (ctypes.c_char_p * max(1, len(bcl)))(*[b.encode() for b in bcl])
    st = ref().dipref_gauss_ft(_ptr(x), DT[x.dtype], _ptr(out), DT[out_dtype], x.ndim, _ptr(sizes), _ptr(sig),
                               len(sig), _ptr(ords), len(ords), arr, len(bcl), truncation)
    if st != 0:
        raise RuntimeError(ref().dipref_last_error().decode())
    return out


def separable(x, filters, symmetries, bc=("mirror",), process=None, out_dtype=None, origins=None,
              nthreads=1, use_ref=False, is_complex=None):
    """dip::SeparableConvolution with 1-D filters given as tap lists (user order).  A
    complex filter (`is_complex[i]`, reference only) is a complex tap list."""
    x = np.ascontiguousarray(x)
    cplx = [bool(c) for c in is_complex] if is_complex is not None else [False] * len(filters)
    if out_dtype is None:
        out_dtype = _flex(x.dtype)
        if any(cplx) or x.dtype.kind == "c":
            out_dtype = np.dtype(np.complex128) if out_dtype == np.float64 else np.dtype(np.complex64)
    out_dtype = np.dtype(out_dtype)
    out = np.zeros(x.shape, out_dtype)
    sizes = _sizes(x)
    nd = x.ndim
    taps = [np.ascontiguousarray(np.asarray(f, np.complex128).view(np.float64) if c else np.asarray(f, np.float64))
            for f, c in zip(filters, cplx)]
    tptrs = (ctypes.c_void_p * len(taps))(*[t.ctypes.data for t in taps])
    ntaps = np.ascontiguousarray([len(t) for t in taps], np.int32)
    bcl = list(bc) if not isinstance(bc, str) else [bc]
    proc = None if process is None else np.ascontiguousarray(process, np.uint8)
    org = None if origins is None else np.ascontiguousarray(origins, np.int32)
    if use_ref:
        syms = (ctypes.c_char_p * len(taps))(*[s.encode() for s in symmetries])
        arr = (ctypes.c_char_p * max(1, len(bcl)))(*[b.encode() for b in bcl])
        cx = np.ascontiguousarray([int(c) for c in cplx], np.int32)
        st = ref().dipref_separable(_ptr(x), DT[x.dtype], _ptr(out), DT[out_dtype], nd, _ptr(sizes),
                                    tptrs, _ptr(ntaps), syms, len(taps), arr, len(bcl),
                                    None if proc is None else _ptr(proc), None if org is None else _ptr(org),
                                    _ptr(cx))
        if st != 0:
            raise RuntimeError(ref().dipref_last_error().decode())
    else:
        if any(cplx) or x.dtype.kind == "c" or out_dtype.kind == "c":
            raise ValueError("the C oracle takes real filters and images only")
        syms = np.ascontiguousarray([SYM[s] for s in symmetries], np.int32)
        bci = np.ascontiguousarray([BC[b] for b in bcl], np.int32)
        st = orc().orc_separable(_ptr(x), DT[x.dtype], _ptr(out), DT[out_dtype], nd, _ptr(sizes),
                                 tptrs, _ptr(ntaps), _ptr(syms), None if org is None else _ptr(org),
                                 len(taps), _ptr(bci), len(bcl), None if proc is None else _ptr(proc),
                                 nthreads)
        if st != 0:
            raise RuntimeError(f"oracle separable failed ({st})")
    return out


def bc_map(j, n, bc):
    s = ctypes.c_int(0)
    idx = orc().orc_bc_map(j, n, BC[bc] if isinstance(bc, str) else bc, ctypes.byref(s))
    return idx, s.value


def half_gaussian_n(sigma, order, n):
    """Taps at a FIXED half size n (x = 0 last): the oracle's tap formula with the
    radius frozen, for finite differences in sigma."""
    buf = np.zeros(n, np.float64)
    if orc().orc_half_gaussian_n(sigma, order, n, _ptr(buf)) != n:
        raise ValueError("half_gaussian_n failed")
    return buf


def gauss_fixed_radius(x, sigmas, orders, bc, radii):
    """dip::Gauss FIR in FP64 with the per-dim radii held fixed (dip dim order)."""
    nd = x.ndim
    taps, syms, proc = [], [], []
    for d in range(nd):
        s = sigmas[d]
        o = orders[d]
        if s > 0 and x.shape[nd - 1 - d] > 1:
            taps.append(half_gaussian_n(s, o, radii[d] + 1))
            syms.append("odd" if o % 2 else "even")
            proc.append(1)
        else:
            taps.append(np.array([1.0]))
            syms.append("general")
            proc.append(0)
    return separable(np.asarray(x, np.float64), taps, syms, bc=bc, process=proc)


def jacobian_t(shape, sigmas, orders, bc, w, truncation=3.0):
    """J^T w for dip::Gauss FIR (FP64), J assembled from impulse responses."""
    n = int(np.prod(shape))
    out = np.zeros(n)
    e = np.zeros(shape)
    ef = e.reshape(-1)
    for j in range(n):
        ef[j] = 1.0
        out[j] = float((gauss_fir(e, sigmas, orders, bc, truncation) * w).sum())
        ef[j] = 0.0
    return out.reshape(shape)


DERIVATIVE_FUNCTIONS = ("gradient", "gradmag", "hessian", "laplace")


def ref_derivative(which, x, sigmas, bc=("mirror",), method="gaussfir", truncation=3.0):
    """The reference's dip::Gradient / GradientMagnitude / Hessian / Laplace; returns the
    tensor elements stacked on a new first axis (element order as dip stores them)."""
    x = np.ascontiguousarray(x)
    odt = _flex(x.dtype)
    cap = 64
    out = np.zeros((cap,) + x.shape, odt)
    sizes = _sizes(x)
    sig = np.ascontiguousarray(list(sigmas), np.float64)
    bcl = list(bc) if not isinstance(bc, str) else [bc]
    arr = (ctypes.c_char_p * max(1, len(bcl)))(*[b.encode() for b in bcl])
    n = ctypes.c_int(0)
    st = ref().dipref_derivative(which.encode(), _ptr(x), DT[x.dtype], _ptr(out), DT[odt], cap, ctypes.byref(n),
                                 x.ndim, _ptr(sizes), _ptr(sig), len(sig), arr, len(bcl), method.encode(), truncation)
    if st != 0:
        raise RuntimeError(ref().dipref_last_error().decode())
    return out[:n.value].copy()


def ref_derivative_masked(which, x, mask, sigmas, bc=("mirror",), method="gaussfir", truncation=3.0):
    """The reference's dip::NormalizedConvolution ("nc"), NormalizedDifferentialConvolution
    ("ndc:<dim>") and masked StructureTensor ("structure:<tensor sigma>"), elements on a
    new first axis."""
    x = np.ascontiguousarray(x)
    mask = np.ascontiguousarray(mask)
    odt = _flex(x.dtype)
    cap = 64
    out = np.zeros((cap,) + x.shape, odt)
    sizes = _sizes(x)
    sig = np.ascontiguousarray(list(sigmas), np.float64)
    bcl = list(bc) if not isinstance(bc, str) else [bc]
    arr = (ctypes.c_char_p * max(1, len(bcl)))(*[b.encode() for b in bcl])
    n = ctypes.c_int(0)
    st = ref().dipref_derivative_masked(which.encode(), _ptr(x), DT[x.dtype], _ptr(mask), DT[mask.dtype], _ptr(out),
                                        DT[odt], cap, ctypes.byref(n), x.ndim, _ptr(sizes), _ptr(sig), len(sig), arr,
                                        len(bcl), method.encode(), truncation)
    if st != 0:
        raise RuntimeError(ref().dipref_last_error().decode())
    return out[:n.value].copy()


def _safe_divide(a, b):
    """dip::SafeDivide: a / b, 0 where b is 0 (arithmetic.h)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(b == 0, a.dtype.type(0), a / np.where(b == 0, b.dtype.type(1), b))


def derivative_masked(which, x, mask, sigmas, bc=("mirror",), truncation=3.0):
    """The oracle's restatement of the masked forms (derivative.cpp:522-600,
    structure_tensor.cpp:46-56) on the FIR Gaussian, each image operation rounded in the
    buffer type: the mask's Gaussian as the denominator; the input weighted by the mask
    (a binary mask zeroes the pixels outside it, a real one multiplies); for "nc" the
    weighted Gaussian over the denominator; for "ndc:<dim>" the derivative of the
    weighted input minus (weighted Gaussian / denominator) x (the mask's derivative), over
    the denominator; "structure:<t>" the tensor of the per-dim NDCs (every dim), each
    product then Gaussian-smoothed with sigma t."""
    x = np.asarray(x)
    nd = x.ndim
    dt = _flex(x.dtype)
    sig = list(sigmas) if len(sigmas) == nd else list(sigmas) * nd
    den = gauss_fir(mask, sig, [0], bc, truncation).astype(dt)
    if mask.dtype == np.bool_:
        w = np.where(mask, x.astype(dt), dt.type(0))
    else:
        w = (x.astype(dt) * mask.astype(dt)).astype(dt)
    if which == "nc":
        return _safe_divide(gauss_fir(w, sig, [0], bc, truncation).astype(dt), den)[None]

    def ndc(dim):
        nc = _safe_divide(gauss_fir(w, sig, [0], bc, truncation).astype(dt), den)
        o = [0] * nd
        o[dim] = 1
        dm = gauss_fir(mask, sig, o, bc, truncation).astype(dt)
        dw = gauss_fir(w, sig, o, bc, truncation).astype(dt)
        return _safe_divide((dw - (nc * dm).astype(dt)).astype(dt), den)

    if which.startswith("ndc:"):
        return ndc(int(which.split(":")[1]))[None]
    if which.startswith("structure:"):
        tsig = float(which.split(":")[1])
        g = [ndc(d) for d in range(nd)]
        zero = dt.type(0)
        el = [zero + g[c] * g[c] for c in range(nd)]
        for c in range(1, nd):
            for r in range(c):
                el.append(zero + g[r] * g[c])
        return np.stack([gauss_fir(e, [tsig], [0], bc, truncation) for e in el])
    raise ValueError(which)


def _ft_oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gauss_ft_oracle
    return gauss_ft_oracle


def ref_gauss_ft_repr(x, sigmas, orders=(0,), truncation=3.0, in_repr="spatial", out_repr="frequency",
                      out_dtype=None):
    """The reference's dip::GaussFT(x, sigmas, orders, truncation, in_repr, out_repr, {})
    (no boundary extension; the result has the input's sizes)."""
    x = np.ascontiguousarray(x)
    if out_dtype is None:
        out_dtype = np.complex128 if x.dtype in (np.float64, np.complex128, np.int32, np.uint32) else np.complex64
        if out_repr == "spatial" and in_repr == "spatial":
            out_dtype = _flex(x.dtype)
    out = np.zeros(x.shape, np.dtype(out_dtype))
    sizes = _sizes(x)
    sig = np.ascontiguousarray(list(sigmas), np.float64)
    ords = np.ascontiguousarray(list(orders), np.int32)
    st = ref().dipref_gauss_ft_repr(_ptr(x), DT[x.dtype], _ptr(out), DT[out.dtype], x.ndim, _ptr(sizes), _ptr(sig),
                                    len(sig), _ptr(ords), len(ords), truncation, in_repr.encode(), out_repr.encode())
    if st != 0:
        raise RuntimeError(ref().dipref_last_error().decode())
    return out


def separable_complex(x, filters, symmetries, bc=("mirror",), process=None, out_dtype=None, origins=None,
                      is_complex=None):
    """dip::SeparableConvolution with a complex buffer type: the numpy restatement
    (oracle/separable_complex_oracle.py), boundary extension by the C oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import separable_complex_oracle
    return separable_complex_oracle.separable(x, filters, symmetries, bc, process, origins, is_complex, orc(),
                                              out_dtype)


def derivative(which, x, sigmas, bc=("mirror",), process=None, truncation=3.0, iir=False, tensor_iir=False,
               ft=False, tensor_ft=False):
    """The oracle's restatement of the derivative family (derivative.cpp:120-418): Gauss
    derivatives from the oracle (FIR, or the IIR Gaussian with `iir`; the structure
    tensor's smoothing with `tensor_iir`), combined in the reference's order and type
    (numpy float32 / float64 arithmetic rounds each operation as the reference does)."""
    nd = x.ndim
    sig = list(sigmas) if len(sigmas) == nd else list(sigmas) * nd
    proc = [True] * nd if process is None else list(process)
    sizes = list(reversed(x.shape))
    dims = [d for d in range(nd) if proc[d] and sizes[d] > 1 and sig[d] > 0]
    def d_of(order):
        if ft:  # dip::Gauss's FT branch: an empty condition means "default"
            return _ft_oracle().gauss_ft(x, sig, order, list(bc) or ["default"], truncation)
        return gauss_iir(x, sig, order, bc, (), "discrete time fit", truncation) if iir else \
            gauss_fir(x, sig, order, bc, truncation)
    def unit(pairs):
        o = [0] * nd
        for d, v in pairs:
            o[d] += v
        return o
    if which == "gradient":
        return np.stack([d_of(unit([(d, 1)])) for d in dims])
    if which == "hessian":
        el = [d_of(unit([(d, 2)])) for d in dims]
        for j in range(1, len(dims)):
            for i in range(j):
                el.append(d_of(unit([(dims[i], 1), (dims[j], 1)])))
        return np.stack(el)
    if which == "laplace":
        out = d_of(unit([(dims[0], 2)]))
        for d in dims[1:]:
            out = out + d_of(unit([(d, 2)]))
        return out[None]
    if which == "gradmag":
        out = d_of(unit([(dims[0], 1)]))
        if len(dims) == 1:
            return np.abs(out)[None]
        out = out * out
        for d in dims[1:]:
            t = d_of(unit([(d, 1)]))
            out = out + t * t
        return np.sqrt(out)[None]
    if which in ("dgg", "laplace+dgg", "laplace-dgg"):
        # derivative.cpp:424-484, image operation by image operation
        g = derivative("gradient", x, sigmas, bc, process, truncation, iir, ft=ft)
        H = derivative("hessian", x, sigmas, bc, process, truncation, iir, ft=ft)
        n = g.shape[0]
        one = g.dtype.type(1)
        two = g.dtype.type(2)
        zero = g.dtype.type(0)
        gs = g[0] * g[0]
        out = gs * H[0]
        for k in range(1, n):
            t = g[k] * g[k]
            gs = gs + t
            t = t * H[k]
            out = out + t
        for r in range(1, n - 1):  # the reference's loop starts at 1
            for c in range(r + 1, n):
                t = g[r] * g[c]
                t = t * H[n + c * (c - 1) // 2 + r]
                out = out * one + t * two
        with np.errstate(divide="ignore", invalid="ignore"):
            out = np.where(gs == 0, zero, out / np.where(gs == 0, one, gs))
        if which == "dgg":
            return out[None]
        tr = np.zeros_like(out)
        for k in range(n):
            tr = tr + H[k]
        return (tr + out if which == "laplace+dgg" else tr - out)[None]
    if which.startswith("structure:"):
        # structure_tensor.cpp:53-56: g g^T (a sum started at 0), then Gauss per element
        tsig = float(which.split(":")[1])
        g = derivative("gradient", x, sigmas, bc, None, truncation, iir, ft=ft)
        n = g.shape[0]
        zero = g.dtype.type(0)
        el = [zero + g[c] * g[c] for c in range(n)]
        for c in range(1, n):
            for r in range(c):
                el.append(zero + g[r] * g[c])
        if tensor_ft:
            return np.stack([_ft_oracle().gauss_ft(e, [tsig], [0], list(bc) or ["default"], truncation) for e in el])
        if tensor_iir:
            return np.stack([gauss_iir(e, [tsig], [0], bc, (), "discrete time fit", truncation) for e in el])
        return np.stack([gauss_fir(e, [tsig], [0], bc, truncation) for e in el])
    raise ValueError(which)
