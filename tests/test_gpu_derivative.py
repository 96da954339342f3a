"""The derivative family (dip::Gradient, GradientMagnitude, Hessian, Laplace with the
FIR method) through the C ABI (dipamd_derivative) against the reference's own results
(tests/golden/deriv_cases.npz), bit for bit, and against the oracle at sizes the
fixtures do not reach."""
import os

import numpy as np
import pytest
import torch

import cases
import diplib_amd as dip
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DERIV = np.load(os.path.join(HERE, "golden", "deriv_cases.npz"))
DERIV2 = np.load(os.path.join(HERE, "golden", "deriv2_cases.npz"))
DERIV_IIR = np.load(os.path.join(HERE, "golden", "deriv_iir_cases.npz"))
DERIV_FT = np.load(os.path.join(HERE, "golden", "deriv_ft_cases.npz"))
FN = {"gradient": dip.Gradient, "gradmag": dip.GradientMagnitude, "hessian": dip.Hessian, "laplace": dip.Laplace,
      "dgg": dip.Dgg, "laplace+dgg": dip.LaplacePlusDgg, "laplace-dgg": dip.LaplaceMinusDgg}


def _elements(fn, y):
    """Tensor elements on the first axis, as the fixtures hold them."""
    y = y.cpu().numpy()
    if fn in ("gradient", "hessian") or fn.startswith("structure"):
        return np.moveaxis(y, -1, 0)
    return y[None]


def _call(fn, x, sig, method, bc):
    if fn.startswith("structure:"):
        return dip.StructureTensor(x, None, sig, [float(fn.split(":")[1])], method, bc)
    return FN[fn](x, sig, method, bc)


@pytest.mark.parametrize("case", cases.DERIV_CASES, ids=[c[0] for c in cases.DERIV_CASES])
def test_derivative_family_bitexact(cuda, case):
    name, fn, shape, dt, sig, bc = case
    x = DERIV[name + "__in"]
    exp = DERIV[name + "__out"]
    y = FN[fn](torch.from_numpy(x).to(cuda), sig, "FIR", bc)
    torch.cuda.synchronize()
    got = _elements(fn, y)
    assert got.shape == exp.shape and got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), \
        f"{name}: max abs diff {np.abs(got.astype(np.float64) - exp.astype(np.float64)).max()}"


@pytest.mark.parametrize("fn", sorted(FN))
def test_derivative_family_larger_vs_oracle(cuda, fn):
    x = cases.make_input((300, 257), "uint8", seed=51)
    y = FN[fn](torch.from_numpy(x).to(cuda), [4.0], "best", ["mirror"])
    assert np.array_equal(_elements(fn, y), oc.derivative(fn, x, [4.0], ["mirror"]))
    v = cases.make_input((24, 40, 33), "float32", seed=52)
    y = FN[fn](torch.from_numpy(v).to(cuda), [1.0, 2.0, 1.5], "gaussFIR", ["periodic"])
    assert np.array_equal(_elements(fn, y), oc.derivative(fn, v, [1.0, 2.0, 1.5], ["periodic"]))


@pytest.mark.parametrize("fn", sorted(FN))
def test_derivative_one_launch_vs_oracle(cuda, fn):
    """2-D uint8 images with every Gaussian derivative of radius <= 4 take the one-launch
    form: bit-identical to the oracle's separate derivatives and combination, over three
    strips (the last one partial), isotropic and per-dim sigmas and boundary conditions."""
    x = cases.make_input((520, 700), "uint8", seed=55)
    for sig, bc in (([1.0], ["mirror"]), ([0.9, 1.0], ["periodic", "zero order"])):
        y = FN[fn](torch.from_numpy(x).to(cuda), sig, "FIR", bc)
        exp = oc.derivative(fn, x, sig, bc)
        assert np.array_equal(_elements(fn, y).view(np.uint8), exp.view(np.uint8)), (fn, sig, bc)


def test_derivative_family_process_and_errors(cuda):
    x = torch.rand((30, 40), device=cuda)
    g = dip.Gradient(x, [2.0], "FIR", ["mirror"], process=[True, False])
    assert g.shape == (30, 40, 1)
    exp = oc.derivative("gradient", x.cpu().numpy(), [2.0], ["mirror"], process=[True, False])
    assert np.array_equal(np.moveaxis(g.cpu().numpy(), -1, 0), exp)
    with pytest.raises(dip.DipError):
        dip.Gradient(x, [0.0])  # no dim to differentiate
    # "best" with sigma < 0.8 resolves to FT: the same as the FT method named
    assert torch.equal(dip.Laplace(x, [0.5]), dip.Laplace(x, [0.5], "FT"))
    with pytest.raises(NotImplementedError):
        dip.Laplace(x, [2.0], "finitediff")
    # a protected output of another type is refused for the combined forms
    with pytest.raises(dip.DipError):
        dip.Laplace(x, [2.0], "FIR", out=torch.empty((30, 40), dtype=torch.float64, device=cuda))
    # Derivative is dip::Gauss
    d = dip.Derivative(x, [1, 0], [2.0], "FIR")
    assert torch.equal(d, dip.GaussFIR(x, [2.0], [1, 0]))


@pytest.mark.parametrize("case", cases.DERIV2_CASES, ids=[c[0] for c in cases.DERIV2_CASES])
def test_dgg_and_structure_tensor_bitexact(cuda, case):
    name, fn, shape, dt, sig, bc = case
    x = DERIV2[name + "__in"]
    exp = DERIV2[name + "__out"]
    y = _call(fn, torch.from_numpy(x).to(cuda), sig, "FIR", bc)
    torch.cuda.synchronize()
    got = _elements(fn, y)
    assert got.shape == exp.shape and got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), \
        f"{name}: max abs diff {np.abs(got.astype(np.float64) - exp.astype(np.float64)).max()}"


@pytest.mark.parametrize("fn", ["dgg", "laplace+dgg", "laplace-dgg", "structure:4"])
def test_dgg_and_structure_tensor_larger_vs_oracle(cuda, fn):
    x = cases.make_input((300, 257), "uint8", seed=53)
    y = _call(fn, torch.from_numpy(x).to(cuda), [2.0], "best", ["mirror"])
    assert np.array_equal(_elements(fn, y), oc.derivative(fn, x, [2.0], ["mirror"]))
    v = cases.make_input((20, 33, 28), "float32", seed=54)
    y = _call(fn, torch.from_numpy(v).to(cuda), [1.0, 2.0, 1.5], "gaussFIR", ["periodic"])
    assert np.array_equal(_elements(fn, y), oc.derivative(fn, v, [1.0, 2.0, 1.5], ["periodic"]))


@pytest.mark.parametrize("case", cases.DERIV_IIR_CASES, ids=[c[0] for c in cases.DERIV_IIR_CASES])
def test_derivative_iir_bitexact(cuda, case):
    """The derivative family and the structure tensor through the IIR Gaussian ("best"
    with sigma > 10, "gaussiir") against the reference's own results, bit for bit."""
    name, fn, shape, dt, sig, bc, method = case
    x = DERIV_IIR[name + "__in"]
    exp = DERIV_IIR[name + "__out"]
    y = _call(fn, torch.from_numpy(x).to(cuda), sig, method, bc)
    torch.cuda.synchronize()
    got = _elements(fn, y)
    assert got.shape == exp.shape and got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), \
        f"{name}: max abs diff {np.abs(got.astype(np.float64) - exp.astype(np.float64)).max()}"


@pytest.mark.parametrize("case", cases.DERIV_FT_CASES, ids=[c[0] for c in cases.DERIV_FT_CASES])
def test_derivative_ft_vs_reference(cuda, case):
    """The derivative family and the structure tensor through the FT Gaussian ("best"
    with sigma < 0.8, "gaussft") against the reference: max |d| <= 1e-5 max |ref|
    (single-precision transforms), 1e-12 for dfloat."""
    name, fn, shape, dt, sig, bc, method = case
    x = DERIV_FT[name + "__in"]
    exp = DERIV_FT[name + "__out"]
    y = _call(fn, torch.from_numpy(x).to(cuda), sig, method, bc)
    torch.cuda.synchronize()
    got = _elements(fn, y)
    assert got.shape == exp.shape and got.dtype == exp.dtype
    tol = 1e-5 if exp.dtype != np.float64 else 1e-12
    err = np.max(np.abs(got.astype(np.float64) - exp)) / np.max(np.abs(exp))
    assert err <= tol, f"{name}: {err:.3e}"


MASKED = np.load(os.path.join(HERE, "golden", "masked_cases.npz"))


def _masked_call(fn, x, m, sig, bc):
    if fn == "nc":
        return dip.NormalizedConvolution(x, m, sig, "FIR", bc)[None]
    if fn.startswith("ndc:"):
        return dip.NormalizedDifferentialConvolution(x, m, int(fn[4:]), sig, "FIR", bc)[None]
    y = dip.StructureTensor(x, m, sig, [float(fn.split(":")[1])], "FIR", bc)
    return y.movedim(-1, 0)


@pytest.mark.parametrize("case", cases.MASKED_CASES, ids=[c[0] for c in cases.MASKED_CASES])
def test_masked_bitexact(cuda, case):
    """dip::NormalizedConvolution, NormalizedDifferentialConvolution and the masked
    StructureTensor against the reference's own outputs, bit for bit."""
    name, fn, shape, dt, mk, sig, bc = case
    x = torch.from_numpy(MASKED[name + "__in"]).to(cuda)
    m = torch.from_numpy(MASKED[name + "__mask"]).to(cuda)
    exp = MASKED[name + "__out"]
    got = _masked_call(fn, x, m, sig, bc).cpu().numpy()
    assert got.dtype == exp.dtype and got.shape == exp.shape
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), \
        f"{name}: max abs diff {np.nanmax(np.abs(got.astype(np.float64) - exp.astype(np.float64)))}"


def test_masked_larger_vs_oracle_and_errors(cuda):
    x = cases.make_input((300, 257), "uint8", seed=61)
    m = cases.make_mask((300, 257), "bin", seed=62)
    exp = oc.derivative_masked("structure:4", x, m, [2.0], ["mirror"])
    got = dip.StructureTensor(torch.from_numpy(x).to(cuda), torch.from_numpy(m).to(cuda), [2.0], [4.0], "FIR",
                              ["mirror"]).movedim(-1, 0).cpu().numpy()
    assert np.array_equal(got, exp)
    xt = torch.from_numpy(x).to(cuda)
    with pytest.raises(dip.DipError):  # mask of other sizes
        dip.NormalizedConvolution(xt, torch.ones((10, 10), device=cuda), [2.0])
    with pytest.raises(dip.DipError):  # dimension out of range
        dip.NormalizedDifferentialConvolution(xt, torch.from_numpy(m).to(cuda), 2, [2.0])
