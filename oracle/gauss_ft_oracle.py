"""CPU oracle for dip::GaussFT (src/linear/gauss.cpp:290-487) -- TEST INFRASTRUCTURE ONLY.

A numpy restatement: the image is extended dim after dim with the boundary condition
(ExtendImageToSize with the image in the centred window: boundary.cpp:245-319,
image_manip.cpp:556-564) to the next 2-3-5-smooth size holding the Gaussian border
(ExpandInput, gauss.cpp:380-393), transformed, multiplied by the separable frequency
response (GaussFTLineFilter, gauss.cpp:290-378, computed in the transform's precision
as the reference computes it), transformed back and cropped.  numpy's FFT runs in
double precision; the reference's runs in single precision for 8-bit and sfloat images,
so the two agree to FFT rounding (checked with a tolerance against the reference's own
outputs in tests/test_oracle.py) -- except where the real-to-complex dim has an odd
length: there the reference's result is not shift invariant (an impulse at the edge of
an odd line comes out distorted, at the centre it does not), and this restatement
keeps the exact convolution.  Only the tests use this module, to compare against.
"""
import math

import numpy as np

PI = 3.14159265358979323846264338327950288


def _max_trunc(is_double):
    eps = np.finfo(np.float64 if is_double else np.float32).eps
    return math.sqrt(-2.0 * math.log(float(eps)))


def half_gaussian_size(sigma, order, truncation, is_double):
    """HalfGaussianSize (gauss.cpp:38-51)."""
    if truncation <= 0:
        truncation = 3.0
    t = min(truncation, _max_trunc(is_double)) + 0.5 * order
    v = math.ceil(t * sigma)
    return max(0, int(v))


def next_smooth5(n):
    """GetOptimalDFTSize(n, true, 5) (dft_optimal_size.cpp:63-108)."""
    if n <= 6:
        return n
    m = n
    while True:
        r = m
        for f in (2, 3, 5):
            while r % f == 0:
                r //= f
        if r == 1:
            return m
        m += 1


def _bc_index(j, n, bc):
    """Source index and sign of coordinate j (any integer) of a line of n samples, or
    (None, constant) for the constant conditions."""
    if 0 <= j < n:
        return j, 1
    if bc in ("mirror", "", "default", "asym mirror"):
        if n == 1:
            return 0, (-1 if bc == "asym mirror" else 1)
        P = 2 * (n - 1)
        p = j % P
        idx = p if p < n else P - p
        sign = 1
        if bc == "asym mirror":
            m = -j if j < 0 else j - (n - 1)
            c = (m - 1) // (n - 1) + 1
            sign = -1 if c & 1 else 1
        return idx, sign
    if bc in ("periodic", "asym periodic"):
        q = j // n
        sign = -1 if (bc == "asym periodic" and (q & 1)) else 1
        return j - q * n, sign
    if bc == "zero order":
        return (0 if j < 0 else n - 1), 1
    if bc == "add zeros":
        return None, 0.0
    raise ValueError("boundary condition not supported by the FT oracle: " + bc)


def _extend_axis(a, axis, left, size, bc):
    n = a.shape[axis]
    parts = []
    for j in range(-left, size - left):
        idx, s = _bc_index(j, n, bc)
        if idx is None:
            sl = np.full(np.take(a, [0], axis=axis).shape, s, a.dtype)
        else:
            sl = np.take(a, [idx], axis=axis) * a.dtype.type(s)
        parts.append(sl)
    return np.concatenate(parts, axis=axis)


def _response(size, sigma, order, truncation, single):
    """GaussFTLineFilter's response of one dim, centred order (gauss.cpp:294-345)."""
    f = np.float32 if single else np.float64
    c = np.complex64 if single else np.complex128
    lut = np.zeros(size, c)
    origin = size // 2
    b = f(f(2.0 * PI * sigma) / f(size))
    b = f(-f(0.5) * b * b)
    if b == 0:
        N = size
    else:
        N = half_gaussian_size(size / (2.0 * PI * sigma), order, truncation, not single)
    begin, end = max(0, origin - N), min(size, origin + N + 1)
    x = (np.arange(begin, end) - origin).astype(f)
    if order > 0:
        a = c(complex(0.0, float(f(2.0 * PI) / f(size)))) ** order
        lut[begin:end] = a * (x ** f(order)) * (np.exp(b * x * x) if b != 0 else f(1))
    elif b != 0:
        lut[begin:end] = np.exp(b * x * x)
    else:
        lut[:] = 1
    return lut


def optimal_dim(sizes, strides):
    """The dimension DIPlib's FourierTransform runs its real-to-complex pass along
    (fourier.cpp:548-562): the first of the smallest |stride| among sizes > 2."""
    opt = 0
    for d in range(1, len(sizes)):
        if sizes[d] > 2 and (sizes[opt] <= 2 or abs(strides[d]) < abs(strides[opt])):
            opt = d
    return opt


def gauss_ft(x, sigmas, orders=(0,), bc=(), truncation=3.0):
    """dip::GaussFT(x, sigmas, orders, truncation, "spatial", "spatial", bc) on a real
    numpy image (numpy axis order; sigmas / orders / bc in dip order, dim 0 = last axis).
    Returns sfloat for 8-bit / sfloat inputs and dfloat for dfloat.

    The reference transforms every dim, the real-to-complex pass along
    optimal_dim(); its inverse uses the non-positive half of that dim's centred
    spectrum, which equals the standard complex-to-real transform of the non-negative
    half multiplied by conj(H(-k)) instead of H(k) -- the two differ only at the
    Nyquist bin of even-sized dims with odd derivative orders, where the response has
    no partner.  Dims that are neither filtered nor the real-to-complex dim are left
    out (their response is 1: transforming them is the identity)."""
    x = np.asarray(x)
    nd = x.ndim
    single = x.dtype != np.float64
    sig = list(sigmas) if len(sigmas) == nd else list(sigmas) * nd
    ords = list(orders) if len(orders) == nd else (list(orders) * nd if orders else [0] * nd)
    bcs = list(bc) if len(bc) in (0, nd) else list(bc) * nd
    sizes = list(reversed(x.shape))  # dip order
    for d in range(nd):
        if sizes[d] == 1:
            sig[d], ords[d] = 0.0, 0
        elif sig[d] < 0:
            sig[d] = 0.0
    a = x.astype(np.float64 if not single else np.float32)
    filt = [d for d in range(nd) if sig[d] > 0 or ords[d] > 0]
    if not filt:
        return a
    P = []
    for d in range(nd):
        n = sizes[d]
        P.append(next_smooth5(n + 2 * half_gaussian_size(sig[d], ords[d], truncation, not single)) if bcs else n)
    dstrides = list(reversed([s // a.itemsize for s in a.strides]))
    opt = optimal_dim(P, dstrides if not bcs else list(np.cumprod([1] + P[:-1])))
    tdims = sorted(set(filt) | {opt})
    offs = {d: (P[d] - sizes[d]) // 2 + (1 if (P[d] % 2 == 0 and sizes[d] % 2 == 1) else 0) for d in tdims}
    if bcs:
        for d in tdims:  # one dim after the other, each over the extended data
            a = _extend_axis(a, nd - 1 - d, offs[d], P[d], bcs[d])
    else:
        offs = {d: 0 for d in tdims}
    axes = [nd - 1 - d for d in tdims if d != opt] + [nd - 1 - opt]  # the half axis last
    spec = np.fft.rfftn(a.astype(np.float64), axes=axes)
    for d in tdims:
        lut = _response(P[d], sig[d], ords[d], truncation, single).astype(np.complex128)
        lut = np.roll(lut, -(P[d] // 2))  # centred order -> FFT order
        lut = np.conj(lut[(-np.arange(P[d])) % P[d]])
        if d == opt:
            lut = lut[:P[d] // 2 + 1]
        shape = [1] * nd
        shape[nd - 1 - d] = lut.size
        spec = spec * lut.reshape(shape)
    y = np.fft.irfftn(spec, s=[a.shape[ax] for ax in axes], axes=axes)
    sl = [slice(None)] * nd
    for d in tdims:
        sl[nd - 1 - d] = slice(offs[d], offs[d] + sizes[d])
    return y[tuple(sl)].astype(np.float32 if single else np.float64)


def centred_fft(a, inverse=False):
    """DIPlib's FourierTransform of every dim with the origin in the middle
    (fourier.cpp:116-150 CopyForDFT with shift, ShiftCornerToCenter): fftshift of the
    DFT of the ifftshift of the input; the forward transform unscaled, the inverse scaled
    by 1/N (fourier.cpp:555-565)."""
    a = np.fft.ifftshift(np.asarray(a, np.complex128))
    b = np.fft.ifftn(a) if inverse else np.fft.fftn(a)
    return np.fft.fftshift(b)


def gauss_ft_frequency(x, sigmas, orders=(0,), truncation=3.0, in_frequency=False, out_frequency=True):
    """dip::GaussFT(x, sigmas, orders, truncation, inRepr, outRepr) without a boundary
    condition (gauss.cpp:397-487) where a domain is "frequency": the centred spectrum of
    the input (or the input itself when it is a spectrum) times the separable response
    of each dim (GaussFTLineFilter in its centred order), returned as a spectrum or
    transformed back (complex: the inverse of a spectrum input is not taken as real).
    Complex result of the input's complex type (SuggestComplex)."""
    x = np.asarray(x)
    nd = x.ndim
    single = x.dtype not in (np.float64, np.complex128, np.int32, np.uint32)
    sig = list(sigmas) if len(sigmas) == nd else list(sigmas) * nd
    ords = list(orders) if len(orders) == nd else (list(orders) * nd if orders else [0] * nd)
    sizes = list(reversed(x.shape))
    for d in range(nd):
        if sizes[d] == 1:
            sig[d], ords[d] = 0.0, 0
        elif sig[d] < 0:
            sig[d] = 0.0
    spec = np.asarray(x, np.complex128) if in_frequency else centred_fft(x)
    for d in range(nd):
        if sig[d] > 0 or ords[d] > 0:
            lut = _response(sizes[d], sig[d], ords[d], truncation, single).astype(np.complex128)
            shape = [1] * nd
            shape[nd - 1 - d] = lut.size
            spec = spec * lut.reshape(shape)
    y = spec if out_frequency else centred_fft(spec, inverse=True)
    return y.astype(np.complex64 if single else np.complex128)
