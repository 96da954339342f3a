"""The streaming 2-D backward kernel (float32 stacks of planes at least 128 columns
wide, k_bwd2s.hpp) against the FP64 backward (the per-axis adjoint passes, pinned by
the reference Jacobian and by central differences in test_gpu_backward.py): every
boundary condition it takes, both symmetries along each axis, radius classes 3..13,
strips that end inside a plane, several row chunks (the first and last holding the
folded halo rows), stacks of planes.  Tolerances are relative to the largest gradient."""
import numpy as np
import pytest
import torch

import oracle_ctypes as oc
import diplib_amd as dip

pytestmark = pytest.mark.gpu

BCS = ["mirror", "asym mirror", "periodic", "asym periodic", "zero order", "add zeros"]
# (numpy shape: planes, y, x), sigmas (x, y), orders (x, y)
# (radius classes equal along x and y, or adjacent: a first derivative along one axis)
CASES = [((1, 300, 256), [3.0, 3.0], [0, 0]),
         ((2, 180, 280), [3.0, 3.5], [1, 0]),
         ((1, 150, 257), [1.0, 1.0], [1, 1]),
         ((1, 600, 320), [3.5, 3.0], [0, 1]),
         ((3, 90, 270), [4.0, 4.0], [0, 0]),
         ((1, 1030, 260), [2.0, 2.0], [0, 0]),
         # adjacent classes: 11 / 9, 9 / 11, 5 / 3, 11 / 13
         ((2, 200, 300), [3.0, 3.0], [1, 0]),
         ((1, 260, 290), [3.0, 3.0], [0, 1]),
         ((1, 120, 256), [1.0, 1.0], [1, 0]),
         ((1, 300, 270), [3.5, 3.5], [0, 1]),
         # two strips (128..255 columns), short chunks of a cached stack
         ((4, 96, 128), [3.0, 3.0], [0, 0]),
         ((3, 80, 200), [2.0, 3.0], [1, 0])]


def _kernels_of(fn):
    """Names of the traced kernel launches of fn()."""
    from diplib_amd import _lib
    L = _lib.lib()
    L.dipamd_trace_enable(1)
    try:
        fn()
        torch.cuda.synchronize()
        recs = (_lib.DipamdTraceRecord * 256)()
        n = L.dipamd_trace_read(recs, 256)
        return [recs[i].name.decode() for i in range(n)]
    finally:
        L.dipamd_trace_enable(0)


def _run(x, w, sig, ords, bc):
    s3 = list(sig) + [0.0]
    o3 = list(ords) + [0]
    return dip.gauss_fir_backward(x, w, s3, o3, [bc])


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}_s{c[1]}_o{c[2]}" for c in CASES])
def test_streaming_backward_vs_fp64(cuda, bc, case):
    shape, sig, ords = case
    print(f"[stream bwd] {shape} sigma {sig} orders {ords} {bc}", flush=True)
    g = torch.Generator(device=cuda)
    g.manual_seed(11)
    x = torch.rand(shape, generator=g, device=cuda, dtype=torch.float32)
    w = torch.randn(shape, generator=g, device=cuda, dtype=torch.float32)
    out = {}
    names = _kernels_of(lambda: out.setdefault("r", _run(x, w, sig, ords, bc)))
    assert "bwd2s" in names, names  # the streaming kernel took the case
    dx32, ds32 = out["r"]
    dx64, ds64 = _run(x.double(), w.double(), sig, ords, bc)
    scale = dx64.abs().max().item()
    err = (dx32.double() - dx64).abs().max().item() / scale
    assert err < 2e-6, err
    ds32, ds64 = ds32.cpu().numpy(), ds64.cpu().numpy()
    # |<w, dy/dsigma>| <= |w| |dy/dsigma|: the error of a sum that may cancel, bounded by
    # its terms (dy/dsigma by FP64 central differences of the oracle at fixed radii)
    xn, wn = x.double().cpu().numpy(), w.double().cpu().numpy()
    s3, o3 = list(sig) + [0.0], list(ords) + [0]
    radii = [oc.gauss_radius(s, o, 3.0, True) if s > 0 else 0 for s, o in zip(s3, o3)]
    for d in range(2):
        sp, sm = list(s3), list(s3)
        sp[d] += 1e-4
        sm[d] -= 1e-4
        dyds = (oc.gauss_fixed_radius(xn, sp, o3, [bc], radii) - oc.gauss_fixed_radius(xn, sm, o3, [bc], radii)) / 2e-4
        bound = np.linalg.norm(wn) * np.linalg.norm(dyds)
        assert abs(ds32[d] - ds64[d]) <= 1e-6 * bound, (d, ds32[d], ds64[d], bound)
    assert ds32[2] == 0.0


def test_streaming_backward_strided_views(cuda):
    """Row-strided views of x, dL/dy and dL/dx (every third row of a taller buffer)."""
    g = torch.Generator(device=cuda)
    g.manual_seed(12)
    xb = torch.rand((2, 3 * 160, 300), generator=g, device=cuda, dtype=torch.float32)
    wb = torch.randn((2, 3 * 160, 300), generator=g, device=cuda, dtype=torch.float32)
    x, w = xb[:, ::3, :], wb[:, ::3, :]
    for bc in ("mirror", "periodic"):
        dx, ds = _run(x, w, [3.0, 3.0], [0, 0], bc)
        dx64, ds64 = _run(x.contiguous().double(), w.contiguous().double(), [3.0, 3.0], [0, 0], bc)
        err = (dx.double() - dx64).abs().max().item() / dx64.abs().max().item()
        assert err < 2e-6, (bc, err)
        assert np.allclose(ds.cpu().numpy(), ds64.cpu().numpy(), rtol=1e-4, atol=1e-3)
