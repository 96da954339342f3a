"""The fused 2-D backward kernel (float32 stacks of planes, k_bwd2d.hip) against the
FP64 backward (the per-axis adjoint passes, pinned by the reference Jacobian and by
central differences in test_gpu_backward.py), over every boundary condition it takes,
both symmetries, radii that put fold targets across chunk and tile borders, partial
tiles and stacks of planes.  Tolerances are relative to the largest gradient."""
import numpy as np
import pytest
import torch

import oracle_ctypes as oc
import diplib_amd as dip

pytestmark = pytest.mark.gpu

BCS = ["mirror", "asym mirror", "periodic", "asym periodic", "zero order", "add zeros"]
# (numpy shape: planes, y, x), sigmas (x, y), orders (x, y)
CASES = [((1, 64, 64), [2.0, 2.0], [0, 0]),
         ((2, 70, 130), [2.0, 3.0], [1, 0]),
         ((1, 300, 70), [8.0, 8.0], [0, 0]),
         ((1, 300, 70), [3.0, 5.0], [0, 2]),
         ((3, 53, 97), [1.0, 1.5], [1, 1]),
         ((1, 129, 200), [3.0, 3.0], [0, 1]),
         # fold rows in chunks of their own: y radius above the 16-row step, a middle
         # part that is not a whole number of steps; and a plane too short for them
         ((1, 100, 90), [2.0, 6.0], [0, 0]),
         ((2, 30, 90), [3.0, 3.0], [0, 1]),
         # unequal x and y radius classes: 11 and 9, 16 and 13
         ((2, 70, 130), [3.0, 3.0], [1, 0]),
         ((1, 80, 160), [4.0, 4.0], [1, 0])]


def _run(x, w, sig, ords, bc):
    s3 = list(sig) + [0.0]
    o3 = list(ords) + [0]
    return dip.gauss_fir_backward(x, w, s3, o3, [bc])


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}_s{c[1]}_o{c[2]}" for c in CASES])
def test_fused_backward_vs_fp64(cuda, bc, case):
    shape, sig, ords = case
    g = torch.Generator(device=cuda)
    g.manual_seed(7)
    x = torch.rand(shape, generator=g, device=cuda, dtype=torch.float32)
    w = torch.randn(shape, generator=g, device=cuda, dtype=torch.float32)
    dx32, ds32 = _run(x, w, sig, ords, bc)
    dx64, ds64 = _run(x.double(), w.double(), sig, ords, bc)
    scale = dx64.abs().max().item()
    err = (dx32.double() - dx64).abs().max().item() / scale
    assert err < 2e-6, err
    ds32, ds64 = ds32.cpu().numpy(), ds64.cpu().numpy()
    # dL/dsigma is a sum over every pixel and can cancel to far below its terms: bound
    # the error by |<w, dy/dsigma>| <= |w| |dy/dsigma| (dy/dsigma by FP64 central
    # differences of the oracle at fixed radii)
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
        assert abs(ds64[d] - (wn * dyds).sum()) <= 1e-6 * bound, (d, ds64[d], (wn * dyds).sum())
    assert ds32[2] == 0.0


def test_fused_backward_against_reference_jacobian_2d(cuda):
    """A 2-D float32 case straight against the oracle's Jacobian (FP64)."""
    shape = (40, 52)
    sig, ords = [2.2, 3.1], [1, 0]
    rng = np.random.default_rng(5)
    w = rng.standard_normal(shape)
    for bc in ("mirror", "periodic", "zero order", "asym mirror"):
        exp = oc.jacobian_t(shape, sig, ords, [bc], w)
        x = torch.rand(shape, dtype=torch.float32, device=cuda)
        dx, _ = dip.gauss_fir_backward(x, torch.from_numpy(w).float().to(cuda), sig, ords, [bc], want_dsigma=False)
        err = np.abs(dx.double().cpu().numpy() - exp).max() / np.abs(exp).max()
        assert err < 2e-6, (bc, err)


def test_fused_backward_dx_only_and_dsigma_only(cuda):
    x = torch.rand((2, 80, 90), device=cuda)
    w = torch.randn((2, 80, 90), device=cuda)
    dx_a, ds_a = _run(x, w, [2.5, 2.5], [0, 0], "mirror")
    dx_b, _ = dip.gauss_fir_backward(x, w, [2.5, 2.5, 0.0], [0, 0, 0], ["mirror"], want_dsigma=False)
    _, ds_c = dip.gauss_fir_backward(x, w, [2.5, 2.5, 0.0], [0, 0, 0], ["mirror"], want_dx=False)
    assert torch.equal(dx_a, dx_b)
    assert torch.equal(ds_a, ds_c)


@pytest.mark.parametrize("bc", BCS)
def test_fused_backward_3d_vs_fp64(cuda, bc):
    """3-D: the z adjoint pass and sigma_z by passes, then dims x and y through the fused
    kernel on the z adjoint of dL/dy."""
    shape = (20, 30, 40)  # z, y, x
    sig, ords = [2.0, 2.5, 1.5], [0, 1, 0]
    g = torch.Generator(device=cuda)
    g.manual_seed(9)
    x = torch.rand(shape, generator=g, device=cuda, dtype=torch.float32)
    w = torch.randn(shape, generator=g, device=cuda, dtype=torch.float32)
    dx32, ds32 = dip.gauss_fir_backward(x, w, sig, ords, [bc])
    dx64, ds64 = dip.gauss_fir_backward(x.double(), w.double(), sig, ords, [bc])
    err = (dx32.double() - dx64).abs().max().item() / dx64.abs().max().item()
    assert err < 2e-6, err
    ds32, ds64 = ds32.cpu().numpy(), ds64.cpu().numpy()
    xn, wn = x.double().cpu().numpy(), w.double().cpu().numpy()
    radii = [oc.gauss_radius(s, o, 3.0, True) for s, o in zip(sig, ords)]
    for d in range(3):
        sp, sm = list(sig), list(sig)
        sp[d] += 1e-4
        sm[d] -= 1e-4
        dyds = (oc.gauss_fixed_radius(xn, sp, ords, [bc], radii) - oc.gauss_fixed_radius(xn, sm, ords, [bc], radii)) / 2e-4
        bound = np.linalg.norm(wn) * np.linalg.norm(dyds)
        assert abs(ds32[d] - ds64[d]) <= 1e-6 * bound, (d, ds32[d], ds64[d], bound)


@pytest.mark.parametrize("shape,sig,ords,dt", [
    ((1, 90, 120), [3.0, 3.0, 0.0], [0, 1, 0], torch.float32),  # fused 2-D kernel
    ((12, 40, 50), [2.0, 2.0, 1.5], [1, 0, 0], torch.float32),  # adjoint pass along z, then fused
    ((30, 40), [2.0, 2.5], [0, 0], torch.float64),             # per-axis passes
])
def test_backward_sigma_total(cuda, shape, sig, ords, dt):
    """DIPAMD_GRAD_SIGMA_TOTAL: one dL/dsigma, the per-dim values summed in dim order;
    dL/dx unchanged."""
    g = torch.Generator(device=cuda)
    g.manual_seed(11)
    x = torch.rand(shape, generator=g, device=cuda, dtype=dt)
    w = torch.randn(shape, generator=g, device=cuda, dtype=dt)
    dx1, ds = dip.gauss_fir_backward(x, w, sig, ords, ["mirror"])
    dx2, tot = dip.gauss_fir_backward(x, w, sig, ords, ["mirror"], sigma_total=True)
    assert tot.shape == (1,)
    vals = ds.cpu().tolist()
    acc = 0.0
    for v in vals:
        acc += v
    assert tot.item() == acc
    assert torch.equal(dx1, dx2)
    _, tot_only = dip.gauss_fir_backward(x, w, sig, ords, ["mirror"], want_dx=False, sigma_total=True)
    assert tot_only.item() == acc
