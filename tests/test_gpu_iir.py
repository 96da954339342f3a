"""dip::GaussIIR through the C ABI (dipamd_gauss_iir) against the reference's own
results (tests/golden/iir_cases.npz), bit for bit, and against the oracle at sizes the
fixtures do not reach (the oracle is itself bit-identical to the reference: the CPU
suite pins it)."""
import os

import numpy as np
import pytest
import torch

import cases
import diplib_amd as dip
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
IIR = np.load(os.path.join(HERE, "golden", "iir_cases.npz"))
TORCH_DT = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64, np.dtype(np.uint8): torch.uint8,
            np.dtype(np.int16): torch.int16}


def _dev(x, cuda):
    return torch.from_numpy(np.ascontiguousarray(x)).to(cuda)


@pytest.mark.parametrize("case", cases.IIR_CASES, ids=[c[0] for c in cases.IIR_CASES])
def test_iir_bitexact(cuda, case):
    name, shape, dt, sig, ords, bc, fo, method, trunc, odt = case
    x = IIR[name + "__in"]
    exp = IIR[name + "__out"]
    out = torch.empty(exp.shape, dtype=TORCH_DT[exp.dtype], device=cuda)
    y = dip.GaussIIR(_dev(x, cuda), sig, ords, bc, list(fo), method, trunc, out=out)
    assert np.array_equal(y.cpu().numpy(), exp), name


@pytest.mark.parametrize("shape,dt,sig,ords,bc", [
    ((384, 512), np.float32, [16.0], [0, 0], "mirror"),
    ((384, 512), np.float32, [16.0], [1, 0], "mirror"),
    ((384, 512), np.float32, [16.0], [0, 2], "periodic"),
    ((700, 1000), np.uint8, [20.0], [0, 0], "mirror"),
    ((32, 64, 64), np.float32, [12.0, 12.0, 11.0], [0], "mirror"),
    ((64, 48, 40), np.float64, [11.0], [1, 0, 1], "add zeros"),
    ((300, 257), np.float32, [35.0], [3, 0], "third order"),
    # borders far longer than the image: boundary walks that bounce many times
    ((20, 30), np.float32, [100.0], [0, 0], "mirror"),
    ((20, 30), np.float32, [60.0], [1, 0], "third order"),
    ((20, 30), np.float64, [100.0], [4, 0], "add zeros"),
])
def test_iir_larger_vs_oracle(cuda, shape, dt, sig, ords, bc):
    x = cases.make_input(shape, dt, seed=9100 + len(shape))
    exp = oc.gauss_iir(x, sig, ords, [bc], nthreads=8)
    y = dip.GaussIIR(_dev(x, cuda), sig, ords, [bc])
    assert np.array_equal(y.cpu().numpy(), exp)


def test_iir_views_and_inplace(cuda):
    """Strided and transposed views (the pass order follows the input strides, as the
    framework's), and the output being the input."""
    base = cases.make_input((200, 300), np.float32, seed=9200)
    xb = _dev(base, cuda)
    # every other column: x stride 2
    v = xb[:, ::2]
    exp = oc.gauss_iir(np.ascontiguousarray(base[:, ::2]), [13.0], [1, 0], ["mirror"])
    assert np.array_equal(dip.GaussIIR(v, [13.0], [1, 0]).cpu().numpy(), exp)
    # a transposed view: dim 0 of the view has the large stride, so its dim 1 (the
    # contiguous one) runs first -- the same passes as on the untransposed image
    t = xb.t()
    exp_t = oc.gauss_iir(base, [5.0, 12.0], [1, 0], ["mirror"]).T
    yt = dip.GaussIIR(t, [12.0, 5.0], [0, 1])
    assert np.array_equal(yt.cpu().numpy(), exp_t)
    # in place, float and double
    for dt in (np.float32, np.float64):
        a = cases.make_input((90, 120), dt, seed=9300)
        exp = oc.gauss_iir(a, [11.5], [0, 0], ["mirror"])
        ta = _dev(a, cuda)
        dip.GaussIIR(ta, [11.5], [0, 0], out=ta)
        assert np.array_equal(ta.cpu().numpy(), exp)


def test_iir_through_gauss(cuda):
    """dip::Gauss with method "IIR", and "best" resolving to IIR for sigma > 10, call
    GaussIIR with the discrete-time-fit design and default filter orders."""
    x = _dev(cases.make_input((120, 160), np.float32, seed=9400), cuda)
    ref = dip.GaussIIR(x, [12.0], [1, 0])
    assert torch.equal(dip.Gauss(x, [12.0], [1, 0], "IIR"), ref)
    assert torch.equal(dip.Gauss(x, [12.0], [1, 0], "best"), ref)
    assert torch.equal(dip.Gauss(x, [12.0], [1, 0], "gaussiir"), ref)
    exp = oc.gauss_iir(x.cpu().numpy(), [12.0], [1, 0], ["mirror"], use_ref=False)
    assert np.array_equal(ref.cpu().numpy(), exp)


def test_iir_errors(cuda):
    x = torch.rand((30, 40), device=cuda)
    with pytest.raises(dip.DipError, match="not \\(yet\\) been implemented"):
        dip.GaussIIR(x, [3.0], [5, 0])  # derivative order above 4
    with pytest.raises(dip.DipError):
        dip.GaussIIR(x, [3.0], [0], designMethod="nonsense")
    with pytest.raises(dip.DipError):
        dip.GaussIIR(x, [3.0, 3.0, 3.0])  # wrong array length
    with pytest.raises(dip.DipError, match="not \\(yet\\) been implemented"):
        dip.GaussIIR(x, [3.0], [0], ["already expanded"])
    # no dim processed: a copy with conversion
    y = dip.GaussIIR(x, [0.0])
    assert torch.equal(y, x)
