"""Pins the CPU oracle (oracle/gauss_oracle.c) before the GPU results are compared with it.

* taps: bit-identical to the reference's own MakeHalfGaussian (golden fixtures),
  plus the moment invariants of gauss.cpp:496-531 that fix the formulas;
* boundary extension: the oracle's closed-form map against np.pad's definitions
  (reflect = dip "mirror", wrap = "periodic") and the pointer walks via a delta filter;
* pipeline: bit-identical to the reference's dip::Gauss outputs (golden fixtures);
  a dense FP64 operator built from the taps and the boundary map reproduces it;
  the result does not depend on the number of threads (gauss.cpp:547-569).
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import cases
import oracle_ctypes as oc

HERE = os.path.dirname(os.path.abspath(__file__))
TAPS = np.load(os.path.join(HERE, "golden", "taps.npz"))
GOLD = np.load(os.path.join(HERE, "golden", "small_cases.npz"))
EXTRAP = np.load(os.path.join(HERE, "golden", "extrap_cases.npz"))
DERIV = np.load(os.path.join(HERE, "golden", "deriv_cases.npz"))
DERIV2 = np.load(os.path.join(HERE, "golden", "deriv2_cases.npz"))
SEP = np.load(os.path.join(HERE, "golden", "sep_cases.npz"))
CSEP = np.load(os.path.join(HERE, "golden", "csep_cases.npz"))


@pytest.mark.parametrize("sigma", cases.TAP_SIGMAS)
@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("is_double", [0, 1])
def test_taps_match_reference(sigma, order, is_double):
    got = oc.half_gaussian(sigma, order, 3.0, bool(is_double))
    exp = TAPS[f"s{sigma}_o{order}_d{is_double}"]
    assert np.array_equal(got, exp)


def test_radius_table():
    # radii of the reference's MakeHalfGaussian at truncation 3 (orders 0/1/2)
    table = {1: (3, 4, 4), 2: (6, 7, 8), 3: (9, 11, 12), 4: (12, 14, 16), 8: (24, 28, 32), 16: (48, 56, 64)}
    for s, rs in table.items():
        assert tuple(oc.gauss_radius(s, o) for o in range(3)) == rs
    assert oc.gauss_radius(0.5, 0) == 2
    # truncation is capped at sqrt(-2 ln eps): 5.646 (sfloat), 8.49 (dfloat)
    assert oc.gauss_radius(1.0, 0, 100.0, False) == 6
    assert oc.gauss_radius(1.0, 0, 100.0, True) == 9


@pytest.mark.parametrize("sigma", [1.0, 2.0, 3.0, 5.0])
def test_tap_invariants(sigma):
    # order 0: sums to 1
    f0 = oc.half_gaussian(sigma, 0)
    full0 = np.concatenate([f0, f0[-2::-1]])
    assert abs(full0.sum() - 1.0) < 1e-12
    # order 1 (correlation with x[i+k] - x[i-k]): a ramp has derivative exactly 1
    f1 = oc.half_gaussian(sigma, 1)[::-1]  # f1[k] = tap at offset k
    k = np.arange(len(f1))
    assert abs((f1 * 2 * k).sum() - 1.0) < 1e-12
    # order 2: sums to 0 and a parabola x^2 has second derivative 2
    f2 = oc.half_gaussian(sigma, 2)[::-1]
    k = np.arange(len(f2))
    assert abs(f2[0] + 2 * f2[1:].sum()) < 1e-12
    assert abs((f2 * 2 * k * k).sum() - 2.0) < 1e-12
    # order 3: a cubic x^3 has third derivative 6
    f3 = oc.half_gaussian(sigma, 3)[::-1]
    k = np.arange(len(f3))
    assert abs((f3 * 2 * k ** 3).sum() - 6.0) < 1e-9


def _np_ext(n, bc, j):
    """Reference definition of the extension via np.pad on an index ramp."""
    pad = abs(j) + n + 2
    idx = np.arange(n)
    if bc == "mirror":
        e = np.pad(idx, pad, mode="reflect")
    elif bc == "periodic":
        e = np.pad(idx, pad, mode="wrap")
    else:
        raise ValueError
    return int(e[pad + j])


@pytest.mark.parametrize("bc", ["mirror", "periodic"])
@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_bc_map_closed_form(bc, n):
    for j in range(-40, 40 + n):
        idx, s = oc.bc_map(j, n, bc)
        assert idx == _np_ext(n, bc, j), (j, n)
        assert s == 1


@pytest.mark.parametrize("bc", ["mirror", "asym mirror", "periodic", "asym periodic", "zero order", "add zeros"])
@pytest.mark.parametrize("n", [2, 3, 4, 7])
def test_bc_walk_matches_closed_form(bc, n):
    """The oracle's restated pointer walks (copy_buffer.cpp) against its closed form,
    probed with a GENERAL filter that picks the single sample x[i+k]."""
    x = (np.arange(n, dtype=np.float64) + 1.0) * 10.0
    for k in list(range(-3 * n, -n)) + list(range(n, 3 * n)):
        # a general filter of length 2|k|+1 with a single 1 at the far end reads x[i + k]
        L = 2 * abs(k) + 1
        taps = np.zeros(L)
        taps[0 if k > 0 else L - 1] = 1.0  # user (convolution) order: reversed internally
        y = oc.separable(x[None, :], [taps, [1.0]], ["general", "general"], bc=[bc], process=[1, 0])
        for i in range(n):
            j = i + k
            idx, s = oc.bc_map(j, n, bc)
            exp = 0.0 if idx < 0 else s * x[idx]
            assert y[0, i] == exp, (bc, n, k, i, y[0, i], exp)


def _dense_1d(n, taps_half, odd, bc):
    """Dense FP64 operator of one filter pass along a line of n samples."""
    r = len(taps_half) - 1
    f = taps_half[::-1]  # f[k], k = 0..r
    M = np.zeros((n, n))
    for i in range(n):
        for k in range(-r, r + 1):
            c = f[abs(k)] * (np.sign(k) if odd else 1.0)
            idx, s = oc.bc_map(i + k, n, bc)
            if idx >= 0:
                M[i, idx] += s * c
    return M


@pytest.mark.parametrize("bc", ["mirror", "periodic", "asym mirror"])
@pytest.mark.parametrize("shape,sig,ords", [((6, 9), 2.0, (1, 0)), ((13, 5), 3.0, (0, 2)), ((4, 4), 1.0, (0, 0))])
def test_dense_operator(bc, shape, sig, ords):
    x = np.random.default_rng(3).random(shape)
    y = oc.gauss_fir(x, [sig], list(ords), [bc])
    H, W = shape
    Mx = _dense_1d(W, oc.half_gaussian(sig, ords[0], is_double=True), ords[0] % 2 == 1, bc)
    My = _dense_1d(H, oc.half_gaussian(sig, ords[1], is_double=True), ords[1] % 2 == 1, bc)
    exp = My @ x @ Mx.T
    assert np.allclose(y, exp, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("case", cases.SMALL_CASES, ids=[c[0] for c in cases.SMALL_CASES])
def test_oracle_matches_reference_outputs(case):
    name, shape, dt, sig, ords, bc, odt = case
    x = GOLD[name + "__in"]
    exp = GOLD[name + "__out"]
    got = oc.gauss_fir(x, sig, ords, bc, 3.0, out_dtype=odt)
    assert got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


@pytest.mark.parametrize("case", cases.SEP_CASES, ids=[c[0] for c in cases.SEP_CASES])
def test_oracle_separable_matches_reference(case):
    """dip::SeparableConvolution with user filters of every symmetry, origins, process
    masks and boundary conditions: the oracle against the reference's outputs, bit for bit
    (signed zeros included)."""
    name, shape, dt, filters, bc, proc, odt = case
    x = SEP[name + "__in"]
    exp = SEP[name + "__out"]
    got = oc.separable(x, [f[0] for f in filters], [f[1] for f in filters], bc, proc, odt,
                       origins=[f[2] for f in filters])
    assert got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def _full_filter(taps, sym):
    """The full convolution kernel a symmetric filter description stands for
    (linear.h:70-80): even [a,b,c,b,a], odd [a,b,c,-b,-a], d-even [a,b,c,c,b,a],
    d-odd [a,b,c,-c,-b,-a]; conj and d-conj are even and d-even for real taps."""
    t = list(taps)
    sym = {"conj": "even", "d-conj": "d-even"}.get(sym, sym)
    if sym == "general":
        return t
    if sym == "even":
        return t + t[-2::-1]
    if sym == "odd":
        return t + [-v for v in t[-2::-1]]
    if sym == "d-even":
        return t + t[::-1]
    return t + [-v for v in t[::-1]]


@pytest.mark.parametrize("sym", ["general", "even", "odd", "conj", "d-even", "d-odd", "d-conj"])
@pytest.mark.parametrize("n,origin", [(3, -1), (4, -1), (4, 0), (4, 2), (5, 1)])
@pytest.mark.parametrize("bc", ["mirror", "periodic", "add zeros"])
def test_separable_dense_convolution(sym, n, origin, bc):
    """The oracle's line filter against the textbook definition: out[i] = sum_j h[j]
    x[i + o - j] with the full kernel h of the symmetry table and the user origin o
    (centre = size // 2), samples outside the line taken from the boundary map.  A
    transposed kernel, a wrong origin or a wrong d-* pairing fails this.

    Two behaviours of the reference are part of the definition: the d-even / d-odd /
    d-conj forms act one sample left of the documented origin (convolution.cpp:256-283
    pair x[i] with x[i-1] for the centre taps), and for an origin left of the centre
    they reach one sample past the border the framework extends (:355-368), which
    reads as 0 (the reference's fixtures, test_oracle_separable_matches_reference)."""
    taps = list(np.random.default_rng(n * 7 + max(origin, 0)).standard_normal(n))
    h = _full_filter(taps, sym)
    size = len(h)
    if origin >= size:
        pytest.skip("origin outside the filter")
    o = size // 2 if origin < 0 else origin
    oi = size - o - 1
    border = max(oi, size - oi - 1)
    if sym.startswith("d-"):
        o -= 1
    W = 11
    x = np.random.default_rng(5).standard_normal((1, W))
    exp = np.zeros(W)
    for i in range(W):
        for j in range(size):
            q = i + o - j
            if q < -border:
                assert sym.startswith("d-") and q == -border - 1 and i == 0
                continue
            idx, sgn = oc.bc_map(q, W, bc)
            if idx >= 0:
                exp[i] += h[j] * sgn * x[0, idx]
    got = oc.separable(x, [taps, [1.0]], [sym, "general"], [bc], process=[1, 0], origins=[origin, -1])
    assert np.allclose(got[0], exp, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("case", cases.CSEP_CASES, ids=[c[0] for c in cases.CSEP_CASES])
def test_oracle_complex_separable_matches_reference(case):
    """dip::SeparableConvolution with a complex buffer type (complex filters, complex
    images, real filters in a complex array, value-based boundary conditions): the numpy
    restatement against the reference's outputs, bit for bit."""
    name, shape, dt, filters, bc, proc, odt = case
    x = CSEP[name + "__in"]
    exp = CSEP[name + "__out"]
    got = oc.separable_complex(x, [f[0] for f in filters], [f[1] for f in filters], bc, proc, odt,
                               origins=[f[2] for f in filters], is_complex=[f[3] for f in filters])
    assert got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def _full_complex_filter(taps, sym):
    """The full kernel of a complex filter description (linear.h:70-80): conj
    [a,b,c,b*,a*], d-conj [a,b,c,c*,b*,a*]; the others as for real taps."""
    t = list(taps)
    if sym == "conj":
        return t + [np.conj(v) for v in t[-2::-1]]
    if sym == "d-conj":
        return t + [np.conj(v) for v in t[::-1]]
    return _full_filter(t, sym)


@pytest.mark.parametrize("sym", ["general", "even", "odd", "conj", "d-even", "d-odd", "d-conj"])
@pytest.mark.parametrize("n,origin", [(3, -1), (4, -1), (4, 0), (4, 5)])
@pytest.mark.parametrize("bc", ["mirror", "periodic", "first order"])
def test_complex_separable_dense_convolution(sym, n, origin, bc):
    """The complex restatement against the textbook complex convolution out[i] =
    sum_j h[j] x[i + o - j] with the full complex kernel, on a complex line; boundary
    samples from the real extension of each part (the C restatement's, checked against
    np.pad above).  A conjugate on the wrong side, a swapped real/imaginary product or a
    wrong d-* pairing fails this."""
    rng = np.random.default_rng(n * 11 + max(origin, 0))
    taps = list(rng.standard_normal(n) + 1j * rng.standard_normal(n))
    h = _full_complex_filter(taps, sym)
    size = len(h)
    if origin >= size:
        pytest.skip("origin outside the filter")
    o = size // 2 if origin < 0 else origin
    oi = size - o - 1
    border = max(oi, size - oi - 1)
    if sym.startswith("d-"):
        o -= 1
    W = 13
    x = rng.standard_normal((1, W)) + 1j * rng.standard_normal((1, W))
    ext = np.zeros(W + 2 * border, np.complex128)
    for part, f in ((0, np.real), (1, np.imag)):
        buf = np.zeros(W + 2 * border)
        buf[border:border + W] = f(x[0])
        assert oc.orc().orc_expand_line(buf[border:].ctypes.data_as(oc.ctypes.c_void_p), 1, W, border,
                                        oc.BC[bc]) == 0
        ext += buf if part == 0 else 1j * buf
    exp = np.zeros(W, np.complex128)
    for i in range(W):
        for j in range(size):
            q = i + o - j
            if q < -border:
                assert sym.startswith("d-") and q == -border - 1 and i == 0
                continue
            exp[i] += h[j] * ext[q + border]
    got = oc.separable_complex(x, [taps, [1.0]], [sym, "general"], [bc], process=[1, 0], origins=[origin, -1],
                               is_complex=[True, False])
    assert np.allclose(got[0], exp, rtol=1e-12, atol=1e-12)


def test_thread_independence():
    x = cases.make_input((59, 192, 256), "float64", seed=5)
    a = oc.gauss_fir(x, [2.0], [0], ["mirror"], nthreads=1)
    b = oc.gauss_fir(x, [2.0], [0], ["mirror"], nthreads=4)
    assert np.array_equal(a, b)


@pytest.mark.slow
def test_oracle_large_hash():
    with open(os.path.join(HERE, "golden", "large_cases.json")) as f:
        large = json.load(f)
    name = "u8_8192sq_s1_o10"
    spec = [c for c in cases.LARGE_CASES if c[0] == name][0]
    _, shape, dt, seed, imp, sig, ords, bc, odt = spec
    x = cases.make_input(shape, dt, seed, imp)
    y = oc.gauss_fir(x, sig, ords, bc, nthreads=8)
    assert hashlib.sha256(y.tobytes()).hexdigest() == large[name]["output_sha256"]


def test_jacobian_fixture_consistent():
    """J^T w from the reference's impulse responses equals the dense operator's
    transpose applied to w (FP64), which pins the adjoint the backward must match."""
    jac = np.load(os.path.join(HERE, "golden", "jacobian.npz"))
    meta = json.loads(str(jac["o1_mirror_64__meta"]))
    w = jac["o1_mirror_64__w"]
    jtw = jac["o1_mirror_64__jtw"]
    H, W = meta["shape"]
    sig = meta["sigmas"][0]
    ox, oy = meta["orders"]
    Mx = _dense_1d(W, oc.half_gaussian(sig, ox, is_double=True), ox % 2 == 1, meta["bc"])
    My = _dense_1d(H, oc.half_gaussian(sig, oy, is_double=True), oy % 2 == 1, meta["bc"])
    exp = My.T @ w @ Mx
    assert np.allclose(jtw, exp, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("case", cases.EXTRAP_CASES, ids=[c[0] for c in cases.EXTRAP_CASES])
def test_oracle_matches_reference_extrapolation(case):
    """Boundary conditions computed from the edge values (antisym reflect, first to
    third order): the oracle against the reference's own outputs, bit for bit."""
    name, shape, dt, sig, ords, bc, odt = case
    x = EXTRAP[name + "__in"]
    exp = EXTRAP[name + "__out"]
    got = oc.gauss_fir(x, sig, ords, bc, 3.0, out_dtype=odt)
    assert got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def extension(line, border, bc):
    """The extended line [-border, n + border) as the separable framework builds it,
    read back through shift filters: a GENERAL filter of 2*border+1 taps (centre
    origin, so the border is `border`) with a single 1 at tap j outputs
    x[i + border - j]; output 0 with j = border + k is x[-k], output n-1 with
    j = border - k is x[n-1+k]."""
    n = line.size
    left, right = [], []
    for k in range(1, border + 1):
        e = np.zeros(2 * border + 1)
        e[border + k] = 1.0
        left.append(oc.separable(line, [e], ["general"], [bc])[0])
        e = np.zeros(2 * border + 1)
        e[border - k] = 1.0
        right.append(oc.separable(line, [e], ["general"], [bc])[n - 1])
    return np.concatenate([left[::-1], line, right]).astype(np.float64)


def test_extension_readback_is_the_mirror():
    x = np.arange(1, 8, dtype=np.float32)
    ext = extension(x, 3, "mirror")
    assert np.array_equal(ext, [4, 3, 2, 1, 2, 3, 4, 5, 6, 7, 6, 5, 4])


@pytest.mark.parametrize("border", [1, 3, 6])
@pytest.mark.parametrize("n", [2, 4, 9])
def test_extrapolation_closed_forms(border, n):
    """The polynomials the reference documents (boundary.h:58-60, copy_buffer.cpp
    comments), solved here independently as linear systems: first order falls
    linearly from the edge value to 0 at distance border+1; second order passes
    through the two edge values (t = 0, -1) and 0 at t = border+1; third order also
    has zero slope there.  Antisymmetric reflection is 2 f(0) - f(k) before it bounces."""
    rng = np.random.default_rng(n * 10 + border)
    x = rng.random(n).astype(np.float32)
    b = border + 1
    t = np.arange(1, border + 1, dtype=np.float64)
    for side in (0, 1):
        xs = x if side == 0 else x[::-1]
        f0, f1 = float(xs[0]), float(xs[1])
        got_first = extension(x, border, "first order")
        got_second = extension(x, border, "second order")
        got_third = extension(x, border, "third order")
        got_anti = extension(x, border, "antisym reflect")
        pick = (lambda e: e[:border][::-1]) if side == 0 else (lambda e: e[n + border:])
        exp_first = f0 * (1.0 - t / b)
        A2 = np.array([[1, 0, 0], [1, -1, 1], [1, b, b * b]], np.float64)
        c2 = np.linalg.solve(A2, [f0, f1, 0.0])
        exp_second = c2[0] + c2[1] * t + c2[2] * t * t
        A3 = np.array([[1, 0, 0, 0], [1, -1, 1, -1], [1, b, b * b, b ** 3], [0, 1, 2 * b, 3 * b * b]], np.float64)
        c3 = np.linalg.solve(A3, [f0, f1, 0.0, 0.0])
        exp_third = c3[0] + c3[1] * t + c3[2] * t * t + c3[3] * t ** 3
        for got, exp in ((got_first, exp_first), (got_second, exp_second), (got_third, exp_third)):
            assert np.allclose(pick(got), exp, rtol=1e-6, atol=1e-6), (side, pick(got), exp)
        k = np.arange(1, min(border, n - 1) + 1)
        assert np.allclose(pick(got_anti)[:k.size], 2 * f0 - xs[k].astype(np.float64), rtol=1e-6, atol=1e-6)


def test_antisym_reflect_bounces():
    """Past the far end the reflection turns around the value written there: on a
    line of 2 samples a, b the left extension is 2a-b, then a-b+... (the walk of
    copy_buffer.cpp:551-589), which for this input is an arithmetic progression."""
    x = np.array([1.0, 3.0], np.float32)  # slope 2: the reflection continues it
    ext = extension(x, 4, "antisym reflect")
    assert np.array_equal(ext, [-7, -5, -3, -1, 1, 3, 5, 7, 9, 11])


@pytest.mark.parametrize("case", cases.DERIV_CASES, ids=[c[0] for c in cases.DERIV_CASES])
def test_oracle_derivative_family_matches_reference(case):
    """dip::Gradient / GradientMagnitude / Hessian / Laplace: the oracle's composition
    of its Gauss derivatives against the reference's own results, bit for bit."""
    name, fn, shape, dt, sig, bc = case
    x = DERIV[name + "__in"]
    exp = DERIV[name + "__out"]
    got = oc.derivative(fn, x, sig, bc)
    assert got.shape == exp.shape and got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def test_derivative_family_identities():
    """What the mathematics fixes: on a quadratic ramp the Gaussian gradient is the
    exact derivative away from the edges, the Laplace is the constant trace, and the
    gradient magnitude is the norm of the gradient."""
    yy, xx = np.mgrid[0:40, 0:48].astype(np.float64)
    img = 0.5 * xx ** 2 + 2.0 * yy ** 2 + 3.0 * xx * yy
    g = oc.derivative("gradient", img, [2.0], ["mirror"])
    h = oc.derivative("hessian", img, [2.0], ["mirror"])
    lap = oc.derivative("laplace", img, [2.0], ["mirror"])[0]
    mag = oc.derivative("gradmag", img, [2.0], ["mirror"])[0]
    inner = (slice(12, -12), slice(12, -12))
    assert np.allclose(g[0][inner], (xx + 3 * yy)[inner], rtol=1e-9, atol=1e-6)
    assert np.allclose(g[1][inner], (4 * yy + 3 * xx)[inner], rtol=1e-9, atol=1e-6)
    assert np.allclose(h[0][inner], 1.0, atol=1e-6) and np.allclose(h[1][inner], 4.0, atol=1e-6)
    assert np.allclose(h[2][inner], 3.0, atol=1e-6)
    assert np.allclose(lap[inner], 5.0, atol=1e-6)
    assert np.allclose(mag, np.sqrt(g[0] ** 2 + g[1] ** 2), rtol=1e-12)


@pytest.mark.parametrize("case", cases.DERIV2_CASES, ids=[c[0] for c in cases.DERIV2_CASES])
def test_oracle_dgg_and_structure_tensor_match_reference(case):
    """dip::Dgg / LaplacePlusDgg / LaplaceMinusDgg / StructureTensor: the oracle's
    image-by-image restatement against the reference's own results, bit for bit."""
    name, fn, shape, dt, sig, bc = case
    x = DERIV2[name + "__in"]
    exp = DERIV2[name + "__out"]
    got = oc.derivative(fn, x, sig, bc)
    assert got.shape == exp.shape and got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def test_dgg_and_structure_tensor_identities():
    """On f = (a x + b y)^2 / 2 (g = s (a, b), s = a x + b y; H = [[a^2, ab], [ab, b^2]]):
    the reference's Dgg sums g_i g_j H_ij over the diagonal and over i < j starting at
    i = 1 (derivative.cpp:463-469), which in 2-D leaves out the mixed term: Dgg is
    (a^4 + b^4) / (a^2 + b^2) there, not the curvature a^2 + b^2.  Laplace -+ Dgg follow.
    The structure tensor of a ramp is the constant g g^T."""
    yy, xx = np.mgrid[0:40, 0:48].astype(np.float64)
    a, b = 0.6, -0.8
    inner = (slice(14, -14), slice(14, -14))
    q = 0.5 * (a * xx + b * yy) ** 2
    dgg = oc.derivative("dgg", q, [2.0], ["mirror"])[0]
    lmd = oc.derivative("laplace-dgg", q, [2.0], ["mirror"])[0]
    lpd = oc.derivative("laplace+dgg", q, [2.0], ["mirror"])[0]
    ref2d = (a ** 4 + b ** 4) / (a * a + b * b)
    away = np.abs(a * xx + b * yy)[inner] > 1.0  # the gradient vanishes on the line s = 0
    assert np.allclose(dgg[inner][away], ref2d, atol=1e-6)
    assert np.allclose(lmd[inner][away], (a * a + b * b) - ref2d, atol=1e-6)
    assert np.allclose(lpd[inner][away], (a * a + b * b) + ref2d, atol=1e-6)
    st = oc.derivative("structure:2", a * xx + b * yy, [1.0], ["mirror"])
    inner2 = (slice(16, -16), slice(16, -16))
    assert np.allclose(st[0][inner2], a * a, atol=1e-6)
    assert np.allclose(st[1][inner2], b * b, atol=1e-6)
    assert np.allclose(st[2][inner2], a * b, atol=1e-6)


# ---------------------------------------------------------------------------
# dip::GaussIIR
# ---------------------------------------------------------------------------
IIR = np.load(os.path.join(HERE, "golden", "iir_cases.npz"))


@pytest.mark.parametrize("case", cases.IIR_CASES, ids=[c[0] for c in cases.IIR_CASES])
def test_oracle_iir_matches_reference(case):
    """The oracle's GaussIIR against the reference's own outputs, bit for bit
    (tests/golden/iir_cases.npz: every pole count and design method, derivative orders
    0-4, every boundary condition, lines down to 2 samples)."""
    name, shape, dt, sig, ords, bc, fo, method, trunc, odt = case
    x = IIR[name + "__in"]
    y = oc.gauss_iir(x, sig, ords, bc, fo, method, trunc, out_dtype=odt)
    assert np.array_equal(y, IIR[name + "__out"])


def _iir1(x, sigma, order, fo, method):
    return oc.gauss_iir(x, [sigma], [order], ["mirror"], [fo], method)


def test_iir_gaussian_properties():
    """What the mathematics fixes, as the reference's own test checks it
    (gaussiir.cpp:819-876): the impulse response of the smoothing filter peaks at
    1/(sqrt(2 pi) sigma) and sums to 1, the filter orders agree within their
    accuracy, and the derivative filters give 1, 2 and 6 on x, x^2 and x^3."""
    sigma = 5.0
    imp = np.zeros(256)
    imp[128] = 1.0
    r1 = _iir1(imp, sigma, 0, 5, "discrete time fit")
    assert abs(r1[128] - 1.0 / (np.sqrt(2 * np.pi) * sigma)) < 0.00015
    assert r1.sum() == pytest.approx(1.0, rel=1e-5)
    for fo, method, tol in ((4, "discrete time fit", 0.0003), (3, "discrete time fit", 0.002),
                            (2, "discrete time fit", 0.004), (5, "forward backward", 0.003),
                            (4, "forward backward", 0.002), (3, "forward backward", 0.009)):
        r2 = _iir1(imp, sigma, 0, fo, method)
        assert np.max(np.abs(r1 - r2)) <= tol, (fo, method)
    x = np.arange(-128.0, 128.0)
    for fo, method in ((5, "discrete time fit"), (4, "discrete time fit"), (3, "discrete time fit"),
                       (5, "forward backward"), (4, "forward backward"), (3, "forward backward")):
        assert _iir1(x, sigma, 1, fo, method)[128] == pytest.approx(1.0, rel=1e-5)
        assert _iir1(x * x, sigma, 2, fo, method)[128] == pytest.approx(2.0, rel=1e-5)
    for fo, method in ((5, "discrete time fit"), (4, "discrete time fit"), (5, "forward backward"),
                       (4, "forward backward")):
        assert _iir1(x * x * x, sigma, 3, fo, method)[128] == pytest.approx(6.0, rel=1e-5)


def test_iir_separable_and_linear():
    """Size-independent properties: a 2-D result equals the two 1-D passes applied in
    turn (x first, in double), and the filter is linear up to rounding."""
    rng = np.random.default_rng(11)
    x = rng.random((40, 60))
    y = oc.gauss_iir(x, [6.0, 4.0], [1, 0], ["mirror"])
    tx = oc.gauss_iir(x, [6.0, 0.0], [1, 0], ["mirror"])
    ty = oc.gauss_iir(tx, [0.0, 4.0], [0, 0], ["mirror"])
    assert np.array_equal(y, ty)
    a = oc.gauss_iir(x, [7.0], [0], ["periodic"])
    b = oc.gauss_iir(2.0 * x, [7.0], [0], ["periodic"])
    assert np.array_equal(2.0 * a, b)  # scaling by 2 is exact in every step


@pytest.mark.skipif(not oc.ref_available(), reason="reference library not built")
def test_oracle_iir_random_cases_against_reference():
    """Seeded random sweep over shapes, types, sigmas, orders, filter orders, design
    methods and boundary conditions, oracle against the reference library."""
    rng = np.random.default_rng(2024)
    bcs = ["mirror", "asym mirror", "periodic", "asym periodic", "add zeros", "zero order", "antisym reflect",
           "first order", "second order", "third order"]
    for it in range(60):
        nd = int(rng.integers(1, 4))
        shape = tuple(int(v) for v in rng.integers(2, 40 if nd < 3 else 12, size=nd))
        dt = [np.uint8, np.float32, np.float64][int(rng.integers(0, 3))]
        x = cases.make_input(shape, dt, seed=7000 + it)
        sig = [float(v) for v in rng.uniform(0.5, 15.0, size=nd)]
        ords = [int(v) for v in rng.integers(0, 5, size=nd)]
        fo = [] if rng.random() < 0.5 else [int(rng.integers(1, 6))]
        method = "discrete time fit" if rng.random() < 0.7 else "forward backward"
        bc = [bcs[int(rng.integers(0, len(bcs)))]]
        r = oc.gauss_iir(x, sig, ords, bc, fo, method, use_ref=True)
        q = oc.gauss_iir(x, sig, ords, bc, fo, method)
        assert np.array_equal(r, q), (shape, dt, sig, ords, fo, method, bc)


DERIV_IIR = np.load(os.path.join(HERE, "golden", "deriv_iir_cases.npz"))


@pytest.mark.parametrize("case", cases.DERIV_IIR_CASES, ids=[c[0] for c in cases.DERIV_IIR_CASES])
def test_oracle_derivative_iir_matches_reference(case):
    """The derivative family and the structure tensor with the IIR method ("best" with
    sigma > 10, "gaussiir"): the oracle against the reference's own results, bit for bit
    (the structure tensor's smoothing resolves "best" with its own sigmas)."""
    name, fn, shape, dt, sig, bc, method = case
    x = DERIV_IIR[name + "__in"]
    exp = DERIV_IIR[name + "__out"]
    iir, tensor_iir = cases.iir_flags(fn, sig, method)
    got = oc.derivative(fn, x, sig, bc, None, 3.0, iir, tensor_iir)
    assert got.shape == exp.shape and got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


# ---------------------------------------------------------------------------
# dip::GaussFT (numpy restatement in oracle/gauss_ft_oracle.py)
# ---------------------------------------------------------------------------
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import gauss_ft_oracle as fto  # noqa: E402

FT = np.load(os.path.join(HERE, "golden", "ft_cases.npz"))


def ft_tolerance(dtype):
    """Relative to the largest output magnitude: FFT rounding in single precision for
    8-bit / sfloat images, double for dfloat."""
    return 1e-5 if np.dtype(dtype) != np.float64 else 1e-12


@pytest.mark.parametrize("case", cases.FT_CASES, ids=[c[0] for c in cases.FT_CASES])
def test_ft_oracle_matches_reference(case):
    """The FT oracle against the reference's own dip::GaussFT outputs (tolerance: the
    FFTs round differently), including odd derivative orders on even sizes, where the
    reference's inverse takes the non-positive half of the spectrum."""
    name, shape, dt, sig, ords, bc, trunc = case
    x = FT[name + "__in"]
    exp = FT[name + "__out"]
    got = fto.gauss_ft(x, sig, ords, bc, trunc)
    assert got.dtype == exp.dtype and got.shape == exp.shape
    assert np.max(np.abs(got.astype(np.float64) - exp)) <= ft_tolerance(exp.dtype) * np.max(np.abs(exp))


def test_ft_oracle_properties():
    """What the mathematics fixes: the response to a constant image is that constant
    (order 0, unit DC gain), filtering is shift invariant (periodic, any length),
    sigma 0 with order 0 is the identity, and for sigma 2 the FT Gaussian equals the
    sampled FIR Gaussian to within the FIR's own truncation and sampling error."""
    c = np.full((20, 30), 7.0)
    assert np.allclose(fto.gauss_ft(c, [1.3], [0], ["mirror"]), 7.0, rtol=1e-12)
    for n in (9, 10, 41):
        base = np.zeros(n)
        base[0] = 1.0
        r0 = fto.gauss_ft(base, [0.6], [1], [])
        for s in (1, n // 2, n - 1):
            r = fto.gauss_ft(np.roll(base, s), [0.6], [1], [])
            assert np.allclose(r, np.roll(r0, s), atol=1e-12)
    x = np.random.default_rng(3).random((12, 17))
    assert np.array_equal(fto.gauss_ft(x, [0.0], [0], ["mirror"]), x)
    y_ft = fto.gauss_ft(x.astype(np.float64), [2.0], [0], ["mirror"])
    y_fir = oc.gauss_fir(x.astype(np.float64), [2.0], [0], ["mirror"])
    assert np.max(np.abs(y_ft - y_fir)) < 2e-3


@pytest.mark.skipif(not oc.ref_available(), reason="reference library not built")
def test_ft_reference_odd_length_defect():
    """Recorded behaviour of the reference: along an odd-length real-to-complex dim its
    output is not shift invariant (an impulse at the edge of a periodic line comes out
    distorted, one at the centre does not), so odd lengths are compared with the
    oracle's exact convolution instead (FT_ODD_CASES)."""
    n = 9
    a = np.zeros(n, np.float32)
    a[0] = 1.0
    b = np.zeros(n, np.float32)
    b[n // 2] = 1.0
    ra = oc.ref_gauss_ft(a, [0.5], [0], [])
    rb = oc.ref_gauss_ft(b, [0.5], [0], [])
    assert np.max(np.abs(np.roll(rb, -(n // 2)) - ra)) > 0.05  # not a shifted copy
    assert np.allclose(rb, fto.gauss_ft(b, [0.5], [0], []), atol=1e-6)  # the centre one is right


FT_REPR = np.load(os.path.join(HERE, "golden", "ft_repr_cases.npz"))


@pytest.mark.parametrize("case", cases.FT_REPR_CASES, ids=[c[0] for c in cases.FT_REPR_CASES])
def test_ft_frequency_oracle_matches_reference(case):
    """The frequency-domain forms of dip::GaussFT (centred spectrum out, spectrum in,
    both): the oracle against the reference's own outputs, to FFT rounding."""
    name, shape, dt, sig, ords, ir, orr = case
    x = FT_REPR[name + "__in"]
    exp = FT_REPR[name + "__out"]
    got = fto.gauss_ft_frequency(x, sig, ords, 3.0, ir == "frequency", orr == "frequency")
    assert got.dtype == exp.dtype and got.shape == exp.shape
    tol = 1e-5 if exp.dtype == np.complex64 else 1e-12
    assert np.max(np.abs(got.astype(np.complex128) - exp)) <= tol * np.max(np.abs(exp))


def test_ft_frequency_oracle_properties():
    """What the mathematics fixes for the centred spectrum: an impulse at the origin
    (the centre pixel) has the flat spectrum 1; the spectrum of a real image is
    conjugate-symmetric about the centre; sigma 0, order 0 returns the spectrum itself;
    spectrum out followed by spectrum in and back gives the spatial-domain result (an
    even order: an odd order's response at the unpaired Nyquist bin of an even size
    makes the inverse complex);
    the Gaussian's spectrum at the centre (DC) is 1."""
    x = np.zeros((8, 10))
    x[4, 5] = 1.0
    assert np.allclose(fto.gauss_ft_frequency(x, [0.0], [0]), 1.0, atol=1e-12)
    r = np.random.default_rng(4).random((8, 10))
    X = fto.gauss_ft_frequency(r, [0.0], [0])
    assert np.allclose(X[1:, 1:], np.conj(X[1:, 1:][::-1, ::-1]), atol=1e-9)
    assert np.allclose(X, np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(r))), atol=1e-12)
    Y = fto.gauss_ft_frequency(r, [1.2], [2, 0])
    back = fto.gauss_ft_frequency(Y.astype(np.complex128), [0.0], [0], in_frequency=True, out_frequency=False)
    direct = fto.gauss_ft(r, [1.2], [2, 0], [])
    assert np.allclose(back.real, direct, atol=1e-10) and np.max(np.abs(back.imag)) < 1e-10
    g = fto.gauss_ft_frequency(x, [2.0], [0])
    assert abs(g[4, 5] - 1.0) < 1e-12


@pytest.mark.skipif(not oc.ref_available(), reason="reference library not built")
def test_ft_frequency_reference_odd_length_defect():
    """Recorded behaviour of the reference: the spectrum of a REAL image is not the DFT
    where its real-to-complex dim (x) has an odd length, nor at the Nyquist plane of a
    third dim (the complex input of the same values gives the DFT), so the real-input
    fixtures are 1-D or 2-D with x even."""
    for shape in ((44, 31), (10, 16, 20)):  # odd x; a third dim (its Nyquist plane)
        x = cases.make_input(shape, "float32", seed=3)
        r_real = oc.ref_gauss_ft_repr(x, [1.0], [0], 3.0, "spatial", "frequency")
        r_cplx = oc.ref_gauss_ft_repr(x.astype(np.complex64), [1.0], [0], 3.0, "spatial", "frequency")
        exact = fto.gauss_ft_frequency(x, [1.0], [0])
        assert np.max(np.abs(r_cplx - exact)) <= 1e-5 * np.max(np.abs(exact))
        assert np.max(np.abs(r_real - exact)) > 5e-5 * np.max(np.abs(exact))


DERIV_FT = np.load(os.path.join(HERE, "golden", "deriv_ft_cases.npz"))


@pytest.mark.parametrize("case", cases.DERIV_FT_CASES, ids=[c[0] for c in cases.DERIV_FT_CASES])
def test_oracle_derivative_ft_matches_reference(case):
    """The derivative family and the structure tensor through the FT method ("best"
    with sigma < 0.8, "gaussft") against the reference, to FFT rounding."""
    name, fn, shape, dt, sig, bc, method = case
    x = DERIV_FT[name + "__in"]
    exp = DERIV_FT[name + "__out"]
    m, tm = cases.method_flags(fn, sig, method)
    got = oc.derivative(fn, x, sig, bc, None, 3.0, m == "iir", tm == "iir", m == "ft", tm == "ft")
    assert got.shape == exp.shape and got.dtype == exp.dtype
    assert np.max(np.abs(got.astype(np.float64) - exp)) <= ft_tolerance(exp.dtype) * np.max(np.abs(exp))


MASKED = np.load(os.path.join(HERE, "golden", "masked_cases.npz"))


@pytest.mark.parametrize("case", cases.MASKED_CASES, ids=[c[0] for c in cases.MASKED_CASES])
def test_oracle_masked_matches_reference(case):
    """dip::NormalizedConvolution, NormalizedDifferentialConvolution and the masked
    StructureTensor: the oracle's restatement (FIR Gaussians of the C oracle, image
    operations in numpy in the buffer type) against the reference's outputs, bit for bit."""
    name, fn, shape, dt, mk, sig, bc = case
    x, m = MASKED[name + "__in"], MASKED[name + "__mask"]
    exp = MASKED[name + "__out"]
    got = oc.derivative_masked(fn, x, m, sig, bc)
    assert got.dtype == exp.dtype and got.shape == exp.shape
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8))


def test_oracle_masked_properties():
    """What the definitions fix: with a mask of ones the normalized convolution is the
    Gaussian divided by the Gaussian of ones (1 away from the border effects of a
    mirror condition: exactly the Gaussian there), the NDC is the derivative minus
    0 x (the mask's derivative, 0) over 1; a mask that is zero everywhere gives 0."""
    x = cases.make_input((30, 40), "float64", seed=7)
    ones = np.ones(x.shape, np.float64)
    g = oc.gauss_fir(x, [2.0], [0], ["mirror"])
    assert np.allclose(oc.derivative_masked("nc", x, ones, [2.0], ["mirror"])[0], g, rtol=1e-12, atol=1e-12)
    d = oc.gauss_fir(x, [2.0], [1, 0], ["mirror"])
    assert np.allclose(oc.derivative_masked("ndc:0", x, ones, [2.0], ["mirror"])[0], d, rtol=1e-9, atol=1e-12)
    zero = np.zeros(x.shape, bool)
    assert not oc.derivative_masked("ndc:1", x, zero, [2.0], ["mirror"]).any()
