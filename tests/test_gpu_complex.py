"""dip::SeparableConvolution with a complex buffer type on the GPU: complex filters of
every symmetry (conj and d-conj acting as such), complex images with real or complex
filters, real filters in a complex filter array, every boundary condition (value-based
ones on each part), origins, process masks, 1-D to 3-D, single and double precision,
a narrower complex output.  Bit for bit against the reference's own outputs
(tests/golden/csep_cases.npz, written by the reference library); larger images and
strided views against the oracle, which those fixtures pin.
"""
import os

import numpy as np
import pytest
import torch

import cases
import diplib_amd as dip
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CSEP = np.load(os.path.join(HERE, "golden", "csep_cases.npz"))

_ODT = {"complex64": torch.complex64, "complex128": torch.complex128}


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _filters(spec):
    return [dip.OneDimensionalFilter(f[0], origin=f[2], symmetry=f[1]) for f in spec]


def _same_bits(got, exp, name):
    assert got.dtype == exp.dtype, name
    if not np.array_equal(got.view(np.uint8), exp.view(np.uint8)):
        d = np.abs(got.astype(np.complex128) - exp.astype(np.complex128))
        raise AssertionError(f"{name}: {np.count_nonzero(d)} samples differ, max abs diff {np.nanmax(d)}")


@pytest.mark.parametrize("case", cases.CSEP_CASES, ids=[c[0] for c in cases.CSEP_CASES])
def test_complex_separable_bitexact_vs_reference(cuda, case):
    name, shape, dt, filters, bc, proc, odt = case
    x = CSEP[name + "__in"]
    ref = CSEP[name + "__out"]
    out = None if odt is None else torch.empty(x.shape, dtype=_ODT[odt], device=cuda)
    y = dip.SeparableConvolution(_t(x, cuda), _filters(filters), bc, proc if proc is not None else (), out=out)
    torch.cuda.synchronize()
    _same_bits(y.cpu().numpy(), ref, name)


LARGE_CSEP = [
    ("conj_n9_f32", (300, 520), "float32", [(cases.csep_taps(9, 1), "conj", -1, True)], ["mirror"]),
    ("dconj_n6_o2_c64", (257, 333), "complex64", [(cases.csep_taps(6, 2), "d-conj", 2, True)], ["periodic"]),
    ("even_real_c128", (130, 270), "complex128", [(cases.sep_taps(7, 3), "even", -1, False)], ["first order"]),
    ("mixed_3d_c64", (20, 41, 70), "complex64",
     [(cases.csep_taps(4, 4), "general", 1, True), (cases.sep_taps(5, 5), "odd", -1, False),
      (cases.csep_taps(3, 6), "conj", -1, True)], ["mirror", "asym periodic", "third order"]),
]


@pytest.mark.parametrize("case", LARGE_CSEP, ids=[c[0] for c in LARGE_CSEP])
def test_complex_separable_larger_vs_oracle(cuda, case):
    name, shape, dt, filters, bc = case
    x = cases.make_input(shape, dt, seed=78)
    ref = oc.separable_complex(x, [f[0] for f in filters], [f[1] for f in filters], bc,
                               origins=[f[2] for f in filters], is_complex=[f[3] for f in filters])
    y = dip.SeparableConvolution(_t(x, cuda), _filters(filters), bc)
    torch.cuda.synchronize()
    _same_bits(y.cpu().numpy(), ref, name)


def test_complex_separable_strided_view(cuda):
    """A transposed complex view: the passes follow the input strides, so the result is
    the base image filtered with the per-dim filters swapped, transposed back."""
    xb = cases.make_input((90, 130), "complex64", seed=79)
    f0, f1 = (cases.csep_taps(5, 7), "general", 1, True), (cases.csep_taps(3, 8), "conj", -1, True)
    xt = _t(xb, cuda).t()
    y = dip.SeparableConvolution(xt, _filters([f0, f1]), ["mirror"])
    torch.cuda.synchronize()
    ref = oc.separable_complex(xb, [f1[0], f0[0]], [f1[1], f0[1]], ["mirror"], origins=[f1[2], f0[2]],
                               is_complex=[True, True]).T
    _same_bits(y.cpu().numpy(), np.ascontiguousarray(ref), "strided")


def test_complex_filter_needs_complex_output(cuda):
    x = _t(cases.make_input((20, 30), "float32", seed=80), cuda)
    with pytest.raises(dip.DipError):
        dip.SeparableConvolution(x, [dip.OneDimensionalFilter([1 + 1j, 0.5j], symmetry="conj")], ["mirror"],
                                 out=torch.empty((20, 30), dtype=torch.float32, device=cuda))


@pytest.mark.parametrize("dt", ["complex64", "complex128"])
def test_gauss_fir_complex_image(cuda, dt):
    """dip::GaussFIR on a complex image (gauss.cpp:254: taps in double for dcomplex) is
    the complex separable convolution with the real half-Gaussian filters."""
    x = cases.make_input((60, 90), dt, seed=82)
    dbl = dt == "complex128"
    t0 = list(oc.half_gaussian(2.0, 0, 3.0, dbl))
    t1 = list(oc.half_gaussian(1.5, 1, 3.0, dbl))
    ref = oc.separable_complex(x, [t0, t1], ["even", "odd"], ["mirror"], is_complex=[False, False])
    y = dip.GaussFIR(_t(x, cuda), [2.0, 1.5], [0, 1], ["mirror"])
    torch.cuda.synchronize()
    _same_bits(y.cpu().numpy(), ref, "gauss_fir_" + dt)


def test_iir_rejects_complex(cuda):
    x = _t(cases.make_input((20, 30), "complex64", seed=81), cuda)
    with pytest.raises(dip.DipError):
        dip.GaussIIR(x, [2.0])
