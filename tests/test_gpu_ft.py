"""dip::GaussFT through the C ABI (dipamd_gauss_ft, rocFFT) against the reference's own
results (tests/golden/ft_cases.npz) and the oracle (oracle/gauss_ft_oracle.py), to
FFT rounding: max |got - expected| <= 1e-5 max |expected| for single-precision
transforms (8-bit and sfloat images), 1e-12 for dfloat."""
import os
import sys

import numpy as np
import pytest
import torch

import cases
import diplib_amd as dip

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import gauss_ft_oracle as fto  # noqa: E402

pytestmark = pytest.mark.gpu
FT = np.load(os.path.join(HERE, "golden", "ft_cases.npz"))


def _close(got, exp):
    tol = 1e-5 if exp.dtype != np.float64 else 1e-12
    err = np.max(np.abs(got.astype(np.float64) - exp)) / max(np.max(np.abs(exp)), 1e-300)
    return err <= tol, err


@pytest.mark.parametrize("case", cases.FT_CASES, ids=[c[0] for c in cases.FT_CASES])
def test_ft_vs_reference(cuda, case):
    name, shape, dt, sig, ords, bc, trunc = case
    x = FT[name + "__in"]
    exp = FT[name + "__out"]
    y = dip.GaussFT(torch.from_numpy(x).to(cuda), sig, ords, trunc, "spatial", "spatial", bc)
    got = y.cpu().numpy()
    assert got.dtype == exp.dtype and got.shape == exp.shape
    ok, err = _close(got, exp)
    assert ok, f"{name}: {err:.3e}"


@pytest.mark.parametrize("case", cases.FT_ODD_CASES, ids=[c[0] for c in cases.FT_ODD_CASES])
def test_ft_odd_lengths_vs_oracle(cuda, case):
    """Odd lengths along the real-to-complex dim: the exact convolution (the oracle)."""
    name, shape, dt, sig, ords, bc, trunc = case
    x = cases.make_input(shape, dt, seed=9700)
    exp = fto.gauss_ft(x, sig, ords, bc, trunc)
    got = dip.GaussFT(torch.from_numpy(x).to(cuda), sig, ords, trunc, "spatial", "spatial", bc).cpu().numpy()
    ok, err = _close(got, exp)
    assert ok, f"{name}: {err:.3e}"


@pytest.mark.parametrize("shape,dt,sig,ords,bc", [
    ((512, 768), np.float32, [0.6], [0, 0], ["mirror"]),
    ((300, 400), np.uint8, [0.5], [1, 2], ["periodic"]),
    ((64, 96, 80), np.float32, [0.7], [0, 0, 1], ["mirror"]),
    ((128, 160), np.float64, [1.5], [4, 0], ["mirror"]),
    ((256, 256), np.float32, [0.0], [2, 1], ["mirror"]),
    ((30, 32), np.float32, [20.0], [0, 0], ["mirror"]),  # padding far larger than the image
    ((30, 32), np.float32, [8.0], [2, 0], ["periodic"]),
])
def test_ft_larger_vs_oracle(cuda, shape, dt, sig, ords, bc):
    x = cases.make_input(shape, dt, seed=9800 + len(shape))
    exp = fto.gauss_ft(x, sig, ords, bc)
    got = dip.GaussFT(torch.from_numpy(x).to(cuda), sig, ords, 3.0, "spatial", "spatial", bc).cpu().numpy()
    ok, err = _close(got, exp)
    assert ok, f"{err:.3e}"


def test_ft_shift_invariance_and_routing(cuda):
    """A periodic line of odd length: every impulse position gives a shifted copy of the
    same response.  dip::Gauss "FT" and "best" (sigma < 0.8 or order > 3) run GaussFT
    with the default boundary condition."""
    n = 41
    base = torch.zeros(n, device=cuda)
    base[0] = 1.0
    r0 = dip.GaussFT(base, [0.6], [1], 3.0, "spatial", "spatial", []).cpu().numpy()
    for s in (1, 20, 40):
        r = dip.GaussFT(torch.roll(base, s), [0.6], [1], 3.0, "spatial", "spatial", []).cpu().numpy()
        assert np.allclose(r, np.roll(r0, s), atol=1e-6)
    x = torch.rand((60, 80), device=cuda)
    ref = dip.GaussFT(x, [0.5], [0, 0], 3.0, "spatial", "spatial", ["mirror"])
    assert torch.equal(dip.Gauss(x, [0.5], [0, 0], "FT"), ref)
    assert torch.equal(dip.Gauss(x, [0.5], [0, 0], "best"), ref)
    y4 = dip.GaussFT(x, [2.0], [4, 0], 3.0, "spatial", "spatial", ["mirror"])
    assert torch.equal(dip.Gauss(x, [2.0], [4, 0], "best"), y4)  # order > 3
    with pytest.raises(dip.DipError):  # a real image is no spectrum (gauss.cpp:412-414)
        dip.GaussFT(x, [1.0], [0], 3.0, "frequency", "spatial", [])
    with pytest.raises(dip.DipError):
        dip.GaussFT(x, [1.0], [0], 3.0, "spatial", "spatial", ["third order"])


FT_REPR = np.load(os.path.join(HERE, "golden", "ft_repr_cases.npz"))


def _close_c(got, exp):
    tol = 1e-5 if exp.dtype == np.complex64 else 1e-12
    err = np.max(np.abs(got.astype(np.complex128) - exp)) / max(np.max(np.abs(exp)), 1e-300)
    return err <= tol, err


@pytest.mark.parametrize("case", cases.FT_REPR_CASES, ids=[c[0] for c in cases.FT_REPR_CASES])
def test_ft_frequency_vs_reference(cuda, case):
    """dip::GaussFT with a frequency-domain input and / or output (the centred
    spectrum) against the reference's own results, to FFT rounding."""
    name, shape, dt, sig, ords, ir, orr = case
    x = FT_REPR[name + "__in"]
    exp = FT_REPR[name + "__out"]
    y = dip.GaussFT(torch.from_numpy(x).to(cuda), sig, ords, 3.0, ir, orr)
    got = y.cpu().numpy()
    assert got.dtype == exp.dtype and got.shape == exp.shape
    ok, err = _close_c(got, exp)
    assert ok, f"{name}: {err:.3e}"


@pytest.mark.parametrize("shape,dt", [((44, 31), "float32"), ((10, 16, 20), "float32"), ((9,), "float64")])
def test_ft_frequency_exact_dft_vs_oracle(cuda, shape, dt):
    """Where the reference's real-input spectrum is not the DFT (odd real-to-complex
    length, a third dim), the result is the DFT: against the oracle."""
    x = cases.make_input(shape, dt, seed=31)
    exp = fto.gauss_ft_frequency(x, [1.0], [1], 3.0)
    got = dip.GaussFT(torch.from_numpy(x).to(cuda), [1.0], [1], 3.0, "spatial", "frequency").cpu().numpy()
    ok, err = _close_c(got, exp)
    assert ok, f"{shape}: {err:.3e}"


def test_ft_frequency_errors(cuda):
    x = torch.rand((20, 30), device=cuda)
    with pytest.raises(dip.DipError):
        dip.GaussFT(x, [1.0], [0], 3.0, "spatial", "frequency", ["mirror"])  # no boundary condition here
    with pytest.raises(dip.DipError):
        dip.GaussFT(x, [1.0], [0], 3.0, "frequency", "spatial")  # a spectrum is complex
    with pytest.raises(dip.DipError):
        dip.GaussFT(x, [1.0], [0], 3.0, "spatial", "fourier")
