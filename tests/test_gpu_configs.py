"""The benchmark's batched and differentiation workloads at their full sizes.

Batched uneven form: the 512 sfloat images of 64..2048 px per axis (seeded exactly as
bench.py seeds them: numpy rng 1234 for sizes and sigmas, a device generator seeded 99
for the pixels), sigma log-uniform in [0.5, 8], mirror: the persistent kernel replayed
from its hipGraph against the per-image loop of the plain drop-in (every image, bit for
bit) and against the oracle (the largest image, the smallest and largest sigma, and
more).

Differentiation form: the stack of 32 sfloat 1024x1024 images (sigma 3 on x and y) and
the 128^3 volume (sigma 3 on every axis), derivative orders 0 and 1 along x:
  * the FP64 adjoint identity <K x, w> = <x, K^T w> at full size, to 1e-12 of
    sum |K x| |w|;
  * the FP32 dL/dx (the fused backward kernel) against the FP64 one, within 2e-6 of its
    largest value, and the FP32 dL/dsigma against the FP64 one, within 2e-6 of
    sum |w| |dy/dsigma| (dy/dsigma by FP64 central differences at the forward's radii);
  * dL/dsigma against FP64 central differences of the oracle's forward with the radius
    held at the forward's value, on a 64x64 crop (16^3 for the volume) at sigma = 3,
    where the order-0 radius steps from 9 to 10.
"""
import numpy as np
import pytest
import torch

import oracle_ctypes as oc
import diplib_amd as dip

pytestmark = pytest.mark.gpu


def test_batched_uneven_512_full_size(cuda):
    rng = np.random.default_rng(1234)
    n = 512
    ws = rng.integers(64, 2049, n)
    hs = rng.integers(64, 2049, n)
    sig = np.exp(rng.uniform(np.log(0.5), np.log(8.0), n))
    g = torch.Generator(device=cuda)
    g.manual_seed(99)
    ins = [torch.rand((int(h), int(w)), generator=g, device=cuda, dtype=torch.float32) for w, h in zip(ws, hs)]
    outs_b = [torch.full_like(t, float("nan")) for t in ins]
    outs_l = [torch.empty_like(t) for t in ins]
    sigmas = [float(s) for s in sig]
    print(f"[cfg4] {n} images, {int(sum(int(w) * int(h) for w, h in zip(ws, hs)))} px", flush=True)
    plan = dip.GaussBatch(ins, outs_b, sigmas, boundary="mirror")
    loop = dip.GaussBatch(ins, outs_l, sigmas, boundary="mirror", tile_config=1)
    plan.run_graph()
    plan.run_graph()  # a replay of the captured graph
    loop.run_per_image()
    torch.cuda.synchronize()
    bad = [i for i, (a, b) in enumerate(zip(outs_b, outs_l)) if not torch.equal(a, b)]
    assert not bad, f"images differing from the per-image loop: {bad[:10]}"
    areas = ws.astype(np.int64) * hs
    pick = {int(np.argmax(areas)), int(np.argmin(sig)), int(np.argmax(sig)), int(np.argmin(areas)), 0, 1, 2, 3, 4}
    for i in sorted(pick):
        x = ins[i].cpu().numpy()
        ref = oc.gauss_fir(x, [sigmas[i]], [0], ["mirror"], 3.0, nthreads=8)
        assert np.array_equal(outs_b[i].cpu().numpy().view(np.uint32), ref.view(np.uint32)), \
            f"image {i} {x.shape} sigma {sigmas[i]} vs oracle"
    plan.close()
    loop.close()


def _frozen_forward(x, sig, orders, radii, bc):
    """dip::Gauss FIR with the radius held fixed, on the device in the input's type,
    through SeparableConvolution with taps from the oracle's tap formula."""
    filters = []
    for s, o, r in zip(sig, orders, radii):
        if s > 0:
            filters.append(dip.OneDimensionalFilter(list(oc.half_gaussian_n(s, o, r + 1)), -1,
                                                    "odd" if o % 2 else "even"))
        else:
            filters.append(dip.OneDimensionalFilter([1.0], -1, "general"))
    return dip.SeparableConvolution(x, filters, [bc])


CFG5 = [("stack_32x1024sq", (32, 1024, 1024), [3.0, 3.0, 0.0]), ("volume_128cube", (128, 128, 128), [3.0, 3.0, 3.0])]


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("case", CFG5, ids=[c[0] for c in CFG5])
def test_differentiation_full_size(cuda, case, order):
    name, shape, sig = case
    bc = "mirror"
    orders = [order, 0, 0]
    sdims = tuple(d for d in range(3) if sig[d] > 0)
    g = torch.Generator(device=cuda)
    g.manual_seed(5)
    x = torch.rand(shape, generator=g, device=cuda, dtype=torch.float32)
    w = torch.randn(shape, generator=g, device=cuda, dtype=torch.float32)
    print(f"[cfg5] {name} order {order}", flush=True)

    # FP64 adjoint identity at full size
    x64, w64 = x.double(), w.double()
    kx = dip.GaussFIR(x64, sig, orders, [bc], 3.0)
    ktw, _ = dip.gauss_fir_backward(x64, w64, sig, orders, [bc], want_dsigma=False)
    lhs = (kx * w64).sum().item()
    rhs = (x64 * ktw).sum().item()
    scale = (kx.abs() * w64.abs()).sum().item()
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)
    del kx

    # FP32 (autograd: the fused backward kernel) against FP64
    xr = x.clone().requires_grad_(True)
    s = torch.tensor(3.0, dtype=torch.float64, device=cuda, requires_grad=True)
    y = dip.gauss_fir_autograd(xr, s, orders, [bc], 3.0, sigma_dims=sdims)
    y.backward(w)
    dx32, ds32 = xr.grad.double(), s.grad.item()
    dx64, ds64 = dip.gauss_fir_backward(x64, w64, sig, orders, [bc], sigma_total=True)
    ds64 = ds64.item()
    err = (dx32 - dx64).abs().max().item() / dx64.abs().max().item()
    assert err < 2e-6, err
    # dL/dsigma: FP64 central differences of y at the forward's radii bound the error
    radii = [oc.gauss_radius(v, o, 3.0, True) if v > 0 else 0 for v, o in zip(sig, orders)]
    h = 1e-4
    sp = [v + h if v > 0 else 0.0 for v in sig]
    sm = [v - h if v > 0 else 0.0 for v in sig]
    dyds = (_frozen_forward(x64, sp, orders, radii, bc) - _frozen_forward(x64, sm, orders, radii, bc)) / (2 * h)
    bound = (w64.abs() * dyds.abs()).sum().item()
    fd_full = (w64 * dyds).sum().item()
    del dyds, dx32, dx64
    assert abs(ds32 - ds64) <= 2e-6 * bound, (ds32, ds64, bound)
    assert abs(ds64 - fd_full) <= 1e-6 * bound, (ds64, fd_full, bound)

    # dL/dsigma on a crop, against the oracle's frozen-radius forward in FP64
    crop = (slice(0, 1), slice(0, 64), slice(0, 64)) if shape[0] == 32 else (slice(0, 16),) * 3
    xc = x[crop].double().cpu().numpy()
    wc = w[crop].double().cpu().numpy()
    if shape[0] == 32:
        xc, wc = xc[0], wc[0]
        csig, cord, crad = sig[:2], orders[:2], radii[:2]
    else:
        csig, cord, crad = sig, orders, radii
    _, dsc = dip.gauss_fir_backward(torch.from_numpy(xc).to(cuda), torch.from_numpy(wc).to(cuda), csig, cord, [bc],
                                    want_dx=False, sigma_total=True)
    _, dsc32 = dip.gauss_fir_backward(torch.from_numpy(xc).float().to(cuda), torch.from_numpy(wc).float().to(cuda),
                                      csig, cord, [bc], want_dx=False, sigma_total=True)
    hc = 1e-5
    lp = (oc.gauss_fixed_radius(xc, [v + hc for v in csig], cord, [bc], crad) * wc).sum()
    lm = (oc.gauss_fixed_radius(xc, [v - hc for v in csig], cord, [bc], crad) * wc).sum()
    fd = (lp - lm) / (2 * hc)
    assert abs(dsc.item() - fd) <= 1e-7 * max(1.0, abs(fd)), (dsc.item(), fd)
    assert abs(dsc32.item() - fd) <= 1e-4 * max(1.0, abs(fd)), (dsc32.item(), fd)
