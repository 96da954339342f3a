"""dip::SeparableConvolution with user filters on the GPU, against the reference's own
outputs (tests/golden/sep_cases.npz, written by the reference library): every filter
symmetry (general, even, odd, conj, d-even, d-odd, d-conj), odd and even stored lengths,
origins 0 / last / off-centre / centre, process masks, meaningless and empty filters,
every boundary condition (lines shorter than the filter included), uint8 zeros times
negative taps (signed zeros), narrow output types, 1-D to 3-D.  Bit for bit, with the
planner's choice of kernels and with one pass per axis; larger images and strided views
against the oracle.
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
SEP = np.load(os.path.join(HERE, "golden", "sep_cases.npz"))

_ODT = {"uint8": torch.uint8, "int16": torch.int16, "float32": torch.float32, "float64": torch.float64}


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _filters(spec):
    return [dip.OneDimensionalFilter(f[0], origin=f[2], symmetry=f[1]) for f in spec]


@pytest.mark.parametrize("case", cases.SEP_CASES, ids=[c[0] for c in cases.SEP_CASES])
@pytest.mark.parametrize("fusion", [None, False])
def test_separable_bitexact_vs_reference(cuda, case, fusion):
    name, shape, dt, filters, bc, proc, odt = case
    x = SEP[name + "__in"]
    ref = SEP[name + "__out"]
    out = None if odt is None else torch.empty(x.shape, dtype=_ODT[odt], device=cuda)
    y = dip.SeparableConvolution(_t(x, cuda), _filters(filters), bc, proc if proc is not None else (), out=out,
                                 fusion=fusion)
    torch.cuda.synchronize()
    yn = y.cpu().numpy()
    assert yn.dtype == ref.dtype
    assert np.array_equal(yn.view(np.uint8), ref.view(np.uint8)), \
        f"{name}: max abs diff {np.abs(yn.astype(np.float64) - ref.astype(np.float64)).max()}"


# Larger images (the tiled kernels' interior and edge tiles, several workgroups per
# line) against the oracle, which the fixtures above pin to the reference.
LARGE_SEP = [
    ("general_n7_o2", (300, 700), "float32", [(cases.sep_taps(7, 1), "general", 2)], ["mirror"]),
    ("deven_n5_o1", (257, 513), "float32", [(cases.sep_taps(5, 2), "d-even", 1)], ["periodic"]),
    ("dodd_n4_o0_u8", (190, 1030), "uint8", [(cases.sep_taps(4, 3), "d-odd", 0)], ["asym mirror"]),
    ("odd_n6_o3_f64", (120, 333), "float64", [(cases.sep_taps(6, 4), "odd", 3)], ["mirror"]),
    ("mixed_3d", (40, 61, 130), "float32",
     [(cases.sep_taps(5, 5), "even", -1), (cases.sep_taps(6, 6), "d-even", -1), (cases.sep_taps(3, 7), "general", 0)],
     ["mirror", "periodic", "zero order"]),
    ("even_r30_y", (400, 300), "float32", [([1.0], "general", -1), (cases.sep_taps(31, 8), "even", -1)],
     ["mirror"]),
    ("general_third_order", (150, 260), "float32", [(cases.sep_taps(8, 9), "general", 5)], ["third order"]),
    ("dodd_antisym", (90, 170), "float32", [(cases.sep_taps(4, 10), "d-odd", 1)], ["antisym reflect"]),
]


@pytest.mark.parametrize("case", LARGE_SEP, ids=[c[0] for c in LARGE_SEP])
@pytest.mark.parametrize("fusion", [None, False])
def test_separable_larger_vs_oracle(cuda, case, fusion):
    name, shape, dt, filters, bc = case
    x = cases.make_input(shape, dt, seed=77)
    ref = oc.separable(x, [f[0] for f in filters], [f[1] for f in filters], bc, origins=[f[2] for f in filters])
    y = dip.SeparableConvolution(_t(x, cuda), _filters(filters), bc, fusion=fusion)
    torch.cuda.synchronize()
    yn = y.cpu().numpy()
    assert np.array_equal(yn.view(np.uint8), ref.view(np.uint8)), \
        f"{name}: max abs diff {np.abs(yn.astype(np.float64) - ref.astype(np.float64)).max()}"


def test_separable_strided_views(cuda):
    """Sub-sampled, cropped, reversed and transposed views of the input, written into a
    transposed output view (strides in samples, negative included), against the oracle
    on the same pixels.  The passes run in order of increasing input stride
    (framework_separable.cpp:146-163), so a transposed input filters its dim 1 first:
    its oracle is the contiguous image with the filters swapped, transposed back."""
    base = cases.make_input((96, 140), "float32", seed=78)
    filters = [(cases.sep_taps(5, 11), "d-even", 0), (cases.sep_taps(4, 12), "general", 1)]
    xt = _t(base, cuda)
    taps, syms, orgs = [f[0] for f in filters], [f[1] for f in filters], [f[2] for f in filters]
    views = [(lambda t: t[::2, ::3], False), (lambda t: t.flip(1), False), (lambda t: t[5:80, 7:129], False),
             (lambda t: t.t(), True)]
    for view, transposed in views:
        v = view(xt)
        if transposed:
            ref = oc.separable(base, taps[::-1], syms[::-1], ["mirror"], origins=orgs[::-1]).T
        else:
            ref = oc.separable(np.ascontiguousarray(v.cpu().numpy()), taps, syms, ["mirror"], origins=orgs)
        out = torch.full((v.shape[1], v.shape[0]), -1.0, device=cuda).t()  # a transposed output view
        y = dip.SeparableConvolution(v, _filters(filters), ["mirror"], out=out)
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy().view(np.uint32), np.ascontiguousarray(ref).view(np.uint32))


def test_separable_in_place(cuda):
    x = cases.make_input((64, 90), "float32", seed=79)
    filters = [(cases.sep_taps(6, 13), "general", 4)]
    ref = oc.separable(x, [filters[0][0]], ["general"], ["periodic"], origins=[4])
    t = _t(x, cuda)
    dip.SeparableConvolution(t, _filters(filters), ["periodic"], out=t)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy().view(np.uint32), ref.view(np.uint32))
