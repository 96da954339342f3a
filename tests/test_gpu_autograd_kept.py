"""The autograd path keeps the forward's intermediate after dims 0 and 1 for the sigma
derivative of the later dims (dipamd_gauss_fir_keep / dipamd_gauss_fir_backward_kept):
gradients bit-identical to the backward that computes the intermediate again, one
fused pass fewer in the backward, and the forward output unchanged."""
import numpy as np
import pytest
import torch

import diplib_amd as dip

pytestmark = pytest.mark.gpu


def _kernels_of(fn):
    from diplib_amd import _lib
    L = _lib.lib()
    L.dipamd_trace_enable(1)
    try:
        fn()
        torch.cuda.synchronize()
        recs = (_lib.DipamdTraceRecord * 512)()
        n = L.dipamd_trace_read(recs, 512)
        return [recs[i].name.decode() for i in range(n)]
    finally:
        L.dipamd_trace_enable(0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("bc", ["mirror", "periodic", "zero order"])
def test_kept_intermediate_gradients(cuda, dtype, bc):
    print(f"[kept] {dtype} {bc}", flush=True)
    g = torch.Generator(device=cuda)
    g.manual_seed(21)
    shape = (24, 40, 56)
    x0 = torch.rand(shape, generator=g, device=cuda, dtype=dtype)
    w = torch.randn(shape, generator=g, device=cuda, dtype=dtype)
    sig = [2.0, 1.5, 3.0]  # dip order: x, y, z (z takes the separate pass)
    x = x0.clone().requires_grad_(True)
    s = torch.tensor(sig, device=cuda, dtype=torch.float64, requires_grad=True)
    y = dip.gauss_fir_autograd(x, s, [0], [bc], 3.0, sigma_host=sig)
    ref_y = dip.GaussFIR(x0, sig, [0], [bc], 3.0)
    assert torch.equal(y.detach(), ref_y)
    names = _kernels_of(lambda: y.backward(w))
    dx_ref, ds_ref = dip.gauss_fir_backward(x0, w, sig, [0, 0, 0], [bc])
    names_ref = _kernels_of(lambda: dip.gauss_fir_backward(x0, w, sig, [0, 0, 0], [bc]))
    assert torch.equal(x.grad, dx_ref)
    assert np.array_equal(s.grad.cpu().numpy(), ds_ref.cpu().numpy())
    assert names.count("fused2d") + 1 == names_ref.count("fused2d"), (names, names_ref)
