"""Parity of the HIP path (through the C ABI) against the reference's own outputs.

Small cases: full outputs written by the reference's dip::Gauss (tests/golden/
small_cases.npz), compared BIT FOR BIT in the default exact mode.  Full-size
workloads: SHA-256 of the exact output bytes written by the reference
(tests/golden/large_cases.json) for inputs regenerated from their seeds.
The fast (FMA) mode is held to max relative error 1e-5 (sfloat) / 1e-12 (dfloat)
against the same outputs.
"""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

import cases
import diplib_amd as dip

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "small_cases.npz"))
EXTRAP = np.load(os.path.join(HERE, "golden", "extrap_cases.npz"))
with open(os.path.join(HERE, "golden", "large_cases.json")) as f:
    LARGE = json.load(f)


def _t(a, dev):
    if a.dtype == np.uint16:
        return torch.from_numpy(a.astype(np.int32)).to(dev).to(torch.uint16)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.astype(np.int64)).to(dev).to(torch.uint32)
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _np(t):
    if hasattr(torch, "uint16") and t.dtype == torch.uint16:
        return t.to(torch.int32).cpu().numpy().astype(np.uint16)
    if hasattr(torch, "uint32") and t.dtype == torch.uint32:
        return t.to(torch.int64).cpu().numpy().astype(np.uint32)
    return t.cpu().numpy()


def dip_mre(a, b):
    """dip::MaximumRelativeError (src/statistics/error.cpp:77-109)."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    den = 0.5 * (np.abs(a) + np.abs(b))
    num = np.abs(a - b)
    r = np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)
    return float(r.max()) if r.size else 0.0


def _out_dtype(odt, x):
    if odt is not None:
        return {"uint8": torch.uint8, "uint16": torch.uint16, "float64": torch.float64, "float32": torch.float32,
                "int16": torch.int16}[odt]
    return None


@pytest.mark.parametrize("case", cases.SMALL_CASES, ids=[c[0] for c in cases.SMALL_CASES])
@pytest.mark.parametrize("fusion", [None, False])
def test_small_bitexact(cuda, case, fusion):
    name, shape, dt, sig, ords, bc, odt = case
    x = GOLD[name + "__in"]
    ref = GOLD[name + "__out"]
    xt = _t(x, cuda)
    out = None
    if odt is not None:
        out = torch.empty(x.shape, dtype=_out_dtype(odt, x), device=cuda)
    y = dip.Gauss(xt, sig, ords, "FIR", bc, 3.0, out=out) if fusion is None else \
        dip.GaussFIR(xt, sig, ords, bc, 3.0, out=out, fusion=fusion)
    torch.cuda.synchronize()
    yn = _np(y)
    assert yn.dtype == ref.dtype
    assert np.array_equal(yn.view(np.uint8), ref.view(np.uint8)), \
        f"{name}: max abs diff {np.abs(yn.astype(np.float64) - ref.astype(np.float64)).max()}"


@pytest.mark.parametrize("case", [c for c in cases.SMALL_CASES if c[2] in ("float32", "uint8", "float64")],
                         ids=[c[0] for c in cases.SMALL_CASES if c[2] in ("float32", "uint8", "float64")])
def test_small_fast_mode(cuda, case):
    name, shape, dt, sig, ords, bc, odt = case
    if odt is not None:
        pytest.skip("clamp-cast outputs are compared in exact mode")
    if bc[0] in ("add max", "add min"):
        pytest.skip("FLT_MAX boundary values overflow; their inf/nan pattern is pinned in exact mode only")
    x = GOLD[name + "__in"]
    ref = GOLD[name + "__out"]
    y = dip.GaussFIR(_t(x, cuda), sig, ords, bc, 3.0, fast=True)
    yn = _np(y).astype(np.float64)
    r = ref.astype(np.float64)
    # The bound is relative to the larger of max|ref| and the conditioning scale
    # max|x| * prod_d sum|taps_d| (outputs that nearly cancel, e.g. derivatives of
    # short periodic lines, cannot be held to 1e-5 of their own magnitude).
    nd = x.ndim
    l1 = 1.0
    for d in range(nd):
        s = sig[0] if len(sig) == 1 else sig[d]
        o = ords[0] if len(ords) == 1 else (ords[d] if ords else 0)
        if s > 0 and x.shape[nd - 1 - d] > 1:
            h = np.asarray(dip.MakeHalfGaussian(s, o, 3.0))
            l1 *= np.abs(h).sum() * 2 - abs(h[-1])
    scale = max(np.abs(r).max(), np.abs(x.astype(np.float64)).max() * l1, 1e-30)
    tol = 1e-12 if ref.dtype == np.float64 else 1e-5
    assert np.abs(yn - r).max() / scale < tol
    if max(ords) == 0 and bc[0] in ("mirror", "periodic", "zero order", "add zeros"):
        assert dip_mre(yn, r) < tol * 10


@pytest.mark.parametrize("name", list(LARGE.keys()))
def test_large_hash(cuda, name):
    spec = [c for c in cases.LARGE_CASES if c[0] == name][0]
    _, shape, dt, seed, imp, sig, ords, bc, odt = spec
    x = cases.make_input(shape, dt, seed, imp)
    exp = LARGE[name]
    assert hashlib.sha256(x.tobytes()).hexdigest() == exp["input_sha256"], "input generator changed"
    y = dip.GaussFIR(_t(x, cuda), sig, ords, bc, 3.0)
    torch.cuda.synchronize()
    yn = y.cpu().numpy()
    flat = yn.reshape(-1)
    idx = np.asarray(exp["sample_index"])
    np.testing.assert_array_equal(flat[idx], np.asarray(exp["sample_value"], dtype=yn.dtype))
    assert hashlib.sha256(yn.tobytes()).hexdigest() == exp["output_sha256"]


def test_inplace_and_views(cuda):
    """out aliasing in (in place) and non-contiguous / negative-stride views."""
    name = "f32_64x48_s2_o0_mirror"
    x = GOLD[name + "__in"]
    ref = GOLD[name + "__out"]
    xt = _t(x, cuda)
    buf = xt.clone()
    dip.GaussFIR(buf, [2.0], [0], ["mirror"], 3.0, out=buf)
    assert np.array_equal(buf.cpu().numpy(), ref)
    # transposed view: dip dims swap, so sigma/orders follow
    xtt = xt.t()
    yt = dip.GaussFIR(xtt, [2.0], [0], ["mirror"], 3.0)
    assert np.allclose(yt.cpu().numpy(), ref.T, rtol=1e-6, atol=1e-7)
    # ROI view inside a larger buffer
    big = torch.zeros((60, 80), dtype=torch.float32, device=cuda)
    big[5:53, 7:71] = xt
    roi = big[5:53, 7:71]
    yr = dip.GaussFIR(roi, [2.0], [0], ["mirror"], 3.0)
    assert np.array_equal(yr.cpu().numpy(), ref)
    # negative stride (flip along x): the result is the flipped filter of the flipped image
    yf = dip.GaussFIR(torch.flip(xt, [1]), [2.0], [0], ["mirror"], 3.0)
    assert np.allclose(np.flip(yf.cpu().numpy(), 1), ref, rtol=1e-6, atol=1e-7)


def test_errors(cuda):
    x = torch.rand((16, 16), device=cuda)
    with pytest.raises(dip.DipError):
        dip.GaussFIR(x, [1.0], [4])
    with pytest.raises(dip.DipError):
        dip.GaussFIR(x, [1.0, 2.0, 3.0])
    with pytest.raises(dip.DipError):
        dip.GaussFIR(x, [1.0], [0], ["bogus"])
    with pytest.raises(dip.DipError, match="not \\(yet\\) been implemented"):
        dip.GaussFIR(x, [1.0], [0], ["already expanded"])
    # the batched and backward forms take the conditions that are index maps only
    with pytest.raises(dip.DipError):
        dip.GaussBatch([x], [torch.empty_like(x)], [(1.0, 1.0)], boundary="first order")
    with pytest.raises(dip.DipError):
        dip.gauss_fir_backward(x, torch.ones_like(x), [1.0], [0], ["third order"])
    # "best" with sigma < 0.8 runs the FT method with the default boundary condition
    assert torch.equal(dip.Gauss(x, [0.5]), dip.GaussFT(x, [0.5], [0], 3.0, "spatial", "spatial", ["default"]))


@pytest.mark.parametrize("case", cases.EXTRAP_CASES, ids=[c[0] for c in cases.EXTRAP_CASES])
@pytest.mark.parametrize("fusion", [None, False])
def test_extrapolation_bitexact(cuda, case, fusion):
    """Antisymmetric reflection and first to third order extrapolation against the
    reference's own outputs (tests/golden/extrap_cases.npz), bit for bit."""
    name, shape, dt, sig, ords, bc, odt = case
    x = EXTRAP[name + "__in"]
    ref = EXTRAP[name + "__out"]
    out = None
    if odt is not None:
        out = torch.empty(x.shape, dtype=_out_dtype(odt, x), device=cuda)
    y = dip.Gauss(_t(x, cuda), sig, ords, "FIR", bc, 3.0, out=out) if fusion is None else \
        dip.GaussFIR(_t(x, cuda), sig, ords, bc, 3.0, out=out, fusion=fusion)
    torch.cuda.synchronize()
    yn = _np(y)
    assert yn.dtype == ref.dtype
    assert np.array_equal(yn.view(np.uint8), ref.view(np.uint8)), \
        f"{name}: max abs diff {np.abs(yn.astype(np.float64) - ref.astype(np.float64)).max()}"


@pytest.mark.parametrize("bc", cases.VALUE_BCS)
def test_extrapolation_larger_vs_oracle(cuda, bc):
    """Sizes the fixtures do not reach (a column-tile sweep along y at sigma 6, a 3-D
    volume, a uint8 input, a general filter with an off-centre origin) against the
    oracle, which the fixtures pin to the reference, bit for bit."""
    import oracle_ctypes as oc
    x = cases.make_input((300, 257), "float32", seed=41)
    y = dip.GaussFIR(_t(x, cuda), [1.5, 6.0], [1, 0], [bc], 3.0)
    assert np.array_equal(_np(y), oc.gauss_fir(x, [1.5, 6.0], [1, 0], [bc], 3.0))
    v = cases.make_input((17, 40, 33), "float32", seed=42)
    y = dip.GaussFIR(_t(v, cuda), [2.0, 1.0, 3.0], [0, 2, 1], [bc, "periodic", bc], 3.0)
    assert np.array_equal(_np(y), oc.gauss_fir(v, [2.0, 1.0, 3.0], [0, 2, 1], [bc, "periodic", bc], 3.0))
    u = cases.make_input((70, 90), "uint8", seed=43)
    y = dip.GaussFIR(_t(u, cuda), [3.0], [0], [bc], 3.0)
    assert np.array_equal(_np(y), oc.gauss_fir(u, [3.0], [0], [bc], 3.0))
    f = [0.1, -0.4, 0.25, 0.9, 0.3]
    y = dip.SeparableConvolution(_t(x, cuda), [dip.OneDimensionalFilter(f, origin=1)], [bc])
    assert np.array_equal(_np(y), oc.separable(x, [f, f], ["general", "general"], [bc], origins=[1, 1]))
