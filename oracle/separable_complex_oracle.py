"""CPU restatement of dip::SeparableConvolution over a complex buffer type -- test
infrastructure only (imported by tests/, never by the product).

Follows, for a complex filter array (buffer SuggestComplex(in)) or a complex image
(buffer SuggestFlex(in)):
  * convolution.cpp:319-421 -- buffer type, filter internalisation (reversed taps,
    sizes 2n-1 / 2n, origin -> size - origin - 1, border max(origin, size - origin - 1);
    "conj"/"d-conj" are "even"/"d-even" unless a filter of the array is complex, and a
    real filter of a complex array is stored as (f, 0));
  * convolution.cpp:184-297 -- the line filter with TPI complex: a real tap multiplies
    both parts; a complex tap multiplies as std::complex does for finite values,
    re = a.re*b.re - a.im*b.im, im = a.re*b.im + a.im*b.re, each product rounded in the
    buffer's precision; CONJ adds f*x[c+k] + conj(f)*x[c-k] before the sum; the d-*
    forms read one sample past the border for an origin left of the centre (read as 0,
    as in the real restatement);
  * copy_buffer.cpp:802-808 -- the boundary extension of a complex line is the real
    extension of its real and imaginary parts, done here by the C restatement
    (liborc.so: orc_expand_line);
  * framework_separable.cpp:146-163 -- passes in dim order (normal strides).
Arithmetic is numpy float32 / float64, one rounding per operation.
"""
import ctypes

import numpy as np

SYMS = {"general": 0, "": 0, "even": 1, "odd": 2, "conj": 3, "d-even": 4, "d-odd": 5, "d-conj": 6}
BCS = {"": 0, "default": 0, "mirror": 0, "asym mirror": 1, "antisym reflect": 2, "periodic": 3,
       "asym periodic": 4, "add zeros": 5, "add max": 6, "add min": 7, "zero order": 8,
       "first order": 9, "second order": 10, "third order": 11}


def suggest_complex(dt):
    dt = np.dtype(dt)
    return np.dtype(np.complex128) if dt in (np.uint32, np.int32, np.float64, np.complex128) else \
        np.dtype(np.complex64)


def _intern(taps, sym, origin, any_complex, ts):
    """(re, im, size, internal origin, border, symmetry code, zpad)."""
    t = np.asarray(taps, dtype=np.complex128)
    n = t.size
    code = SYMS[sym]
    if code == 3 and not any_complex:
        code = 1
    if code == 6 and not any_complex:
        code = 4
    size = n
    if code in (1, 2, 3):
        size = 2 * n - 1
    elif code in (4, 5, 6):
        size = 2 * n
    o = size // 2 if origin < 0 else origin
    if n and o >= size:
        raise ValueError("Origin outside of filter")
    rev = t[::-1]
    oi = size - o - 1 if n else 0
    border = max(oi, size - oi - 1) if n else 0
    zpad = 0
    if code in (4, 5, 6) and oi >= n:
        border, zpad = oi + 1, 1
    return (rev.real.astype(ts), rev.imag.astype(ts), size, oi, border, code, zpad)


def _meaningless(F, ts):
    re, im, size = F[0], F[1], F[2]
    return size == 0 or (size == 1 and re[0] == ts(1) and im[0] == ts(0))


def _extend(lines, border, zpad, bc, ts, orc):
    """lines (L, n) of ts -> (L, n + 2*border) extended, the zpad outermost left samples 0."""
    L, n = lines.shape
    ref = border - zpad
    out = np.zeros((L, n + 2 * border), dtype=ts)
    buf = np.zeros(n + 2 * ref, dtype=ts)
    for i in range(L):
        buf[:] = 0
        buf[ref:ref + n] = lines[i]
        if ref > 0:
            st = orc.orc_expand_line(buf[ref:].ctypes.data_as(ctypes.c_void_p), int(ts == np.float64), n, ref, bc)
            if st != 0:
                raise RuntimeError("extension failed")
        out[i, zpad:zpad + n + 2 * ref] = buf
    return out


def _recover(a, b, c, d, ts):
    """The infinity recovery of C11 Annex G.5.1 (the product of a + ib and c + id when
    the plain formula gave NaN + iNaN), which the compiled std::complex product applies
    (GCC: __mulsc3 / __muldc3)."""
    one, zero, inf = ts(1), ts(0), ts(np.inf)
    recalc = False
    if np.isinf(a) or np.isinf(b):
        a = np.copysign(one if np.isinf(a) else zero, a)
        b = np.copysign(one if np.isinf(b) else zero, b)
        c = np.copysign(zero, c) if np.isnan(c) else c
        d = np.copysign(zero, d) if np.isnan(d) else d
        recalc = True
    if np.isinf(c) or np.isinf(d):
        c = np.copysign(one if np.isinf(c) else zero, c)
        d = np.copysign(one if np.isinf(d) else zero, d)
        a = np.copysign(zero, a) if np.isnan(a) else a
        b = np.copysign(zero, b) if np.isnan(b) else b
        recalc = True
    if not recalc and (np.isinf(a * c) or np.isinf(b * d) or np.isinf(a * d) or np.isinf(b * c)):
        a, b, c, d = [np.copysign(zero, v) if np.isnan(v) else v for v in (a, b, c, d)]
        recalc = True
    if recalc:
        return inf * (a * c - b * d), inf * (a * d + b * c)
    return None


def _cmul(fr, fi, xr, xi, ts):
    """(fr + i fi) * (xr + i xi) as std::complex multiplies: the plain formula, each
    product rounded, and the Annex G recovery where both parts came out NaN."""
    fr, fi = ts(fr), ts(fi)
    with np.errstate(all="ignore"):
        re = fr * xr - fi * xi
        im = fr * xi + fi * xr
        bad = np.isnan(re) & np.isnan(im)
        if bad.any():
            re, im = re.copy(), im.copy()
            for idx in zip(*np.nonzero(bad)):
                r = _recover(fr, fi, xr[idx], xi[idx], ts)
                if r is not None:
                    re[idx], im[idx] = r
    return re, im


def _filter_lines(xr, xi, F, fc, ts):
    """The line filter on extended lines (L, n + 2*border); returns (L, n) re, im."""
    re, im, size, oi, border, code, zpad = F
    L = xr.shape[0]
    n = xr.shape[1] - 2 * border
    ds = re.size
    idx0 = border + np.arange(n)  # sample i of the line in the extended array
    zero = np.zeros((L, n), dtype=ts)

    def x(j):
        return xr[:, idx0 + j], xi[:, idx0 + j]

    def tap(k, tr, ti):
        if fc:
            return _cmul(re[k], im[k], tr, ti, ts)
        return (ts(re[k]) * tr, ts(re[k]) * ti)

    if code == 0:
        sr, si = zero.copy(), zero.copy()
        for k in range(ds):
            pr, pi = tap(k, *x(-oi + k))
            sr, si = sr + pr, si + pi
        return sr, si
    c = -oi + ds - 1
    if code in (1, 2, 3):
        sr, si = tap(0, *x(c))
        for k in range(1, ds):
            rr, ri = x(c + k)
            lr, li = x(c - k)
            if code == 3:
                ar, ai = _cmul(re[k], im[k], rr, ri, ts)
                br, bi = _cmul(re[k], -im[k], lr, li, ts)
                pr, pi = ar + br, ai + bi
            else:
                pr, pi = tap(k, rr + lr, ri + li) if code == 1 else tap(k, rr - lr, ri - li)
            sr, si = sr + pr, si + pi
        return sr, si
    sr, si = zero.copy(), zero.copy()
    for k in range(ds):
        rr, ri = x(c + k)
        lr, li = x(c - 1 - k)
        if code == 6:
            ar, ai = _cmul(re[k], im[k], rr, ri, ts)
            br, bi = _cmul(re[k], -im[k], lr, li, ts)
            pr, pi = ar + br, ai + bi
        else:
            pr, pi = tap(k, rr + lr, ri + li) if code == 4 else tap(k, rr - lr, ri - li)
        sr, si = sr + pr, si + pi
    return sr, si


def separable(x, filters, symmetries, bc=("mirror",), process=None, origins=None, is_complex=None, orc=None,
              out_dtype=None):
    """dip::SeparableConvolution of numpy array x (dip dims = reversed numpy axes) with
    a complex buffer type; filters are tap lists (complex values for complex filters)."""
    with np.errstate(all="ignore"):
        return _separable(x, filters, symmetries, bc, process, origins, is_complex, orc, out_dtype)


def _separable(x, filters, symmetries, bc, process, origins, is_complex, orc, out_dtype):
    x = np.ascontiguousarray(x)
    nd = x.ndim
    cplx = [bool(c) for c in is_complex] if is_complex is not None else [False] * len(filters)
    any_complex = any(cplx)
    buf = suggest_complex(x.dtype) if any_complex else np.dtype(x.dtype)
    if buf.kind != "c":
        raise ValueError("real buffer: use the real oracle")
    ts = np.float64 if buf == np.complex128 else np.float32
    org = list(origins) if origins is not None else [-1] * len(filters)
    F = [_intern(f, s, o, any_complex, ts) for f, s, o in zip(filters, symmetries, org)]
    bcl = list(bc) if not isinstance(bc, str) else [bc]
    proc = [True] * nd if process is None else [bool(p) for p in process]
    sizes = list(x.shape[::-1])
    for d in range(nd):
        f = F[0] if len(F) == 1 else F[d]
        if len(F) == 1:
            if _meaningless(F[0], ts):
                proc[d] = False
        elif sizes[d] <= 1 or _meaningless(f, ts):
            proc[d] = False
        if sizes[d] == 1:
            proc[d] = False
    if x.dtype.kind == "c":
        xr, xi = x.real.astype(ts), x.imag.astype(ts)
    else:
        xr, xi = x.astype(ts), np.zeros(x.shape, ts)
    for d in range(nd):
        if not proc[d]:
            continue
        f = F[0] if len(F) == 1 else F[d]
        b = BCS[bcl[0] if len(bcl) == 1 else bcl[d]]
        ax = nd - 1 - d
        r = np.moveaxis(xr, ax, -1)
        i = np.moveaxis(xi, ax, -1)
        shp = r.shape
        r2 = r.reshape(-1, shp[-1])
        i2 = i.reshape(-1, shp[-1])
        er = _extend(r2, f[4], f[6], b, ts, orc)
        ei = _extend(i2, f[4], f[6], b, ts, orc)
        yr, yi = _filter_lines(er, ei, f, any_complex, ts)
        xr = np.moveaxis(yr.reshape(shp), -1, ax)
        xi = np.moveaxis(yi.reshape(shp), -1, ax)
    out = (xr.astype(np.float64) + 1j * xi.astype(np.float64)).astype(buf)
    out.real = xr
    out.imag = xi
    if out_dtype is not None:
        out = out.astype(out_dtype)
    return np.ascontiguousarray(out)
