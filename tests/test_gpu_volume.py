"""Volumes (three filtered dims: the fused x+y+z kernel, or the fused x+y pass and the
z sweep) against the oracle, bit for bit: aligned and unaligned rows, tiles cut by the
volume's edges,
volumes thinner than the z filter, every index-map boundary condition along each
axis, even and odd (derivative) filters, sfloat and dfloat, and against the
pass-by-pass path (fusion off) on the same inputs."""
import numpy as np
import pytest
import torch

import cases
import diplib_amd as dip
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu

VOL = [
    # (name, numpy shape (z, y, x), dtype, sigmas (dip order), orders, boundary list)
    ("f32_130x70x40_s215", (40, 70, 130), "float32", [2.0, 1.5, 1.0], [0], ["mirror"]),
    ("f32_128x64x33_periodic", (33, 64, 128), "float32", [1.0, 2.0, 0.8], [0], ["periodic"]),
    ("f32_96x50x20_o102", (20, 50, 96), "float32", [2.0, 1.0, 1.0], [1, 0, 2], ["mirror"]),
    ("f32_70x45x17_mixedbc", (17, 45, 70), "float32", [1.5, 2.5, 1.2], [0, 1, 1],
     ["asym mirror", "add zeros", "zero order"]),
    ("f32_64x40x12_asym_periodic", (12, 40, 64), "float32", [3.0, 3.0, 1.0], [0], ["asym periodic"]),
    ("f32_11x9x5_thin", (5, 9, 11), "float32", [1.0, 1.0, 1.3], [0], ["mirror"]),
    ("f32_33x20x7_addmax", (7, 20, 33), "float32", [1.0, 1.0, 1.0], [0], ["add max", "mirror", "add min"]),
    ("f32_200x70x24_s4", (24, 70, 200), "float32", [4.0, 4.0, 1.0], [0], ["mirror"]),
    ("f64_130x66x30", (30, 66, 130), "float64", [2.0, 2.0, 1.0], [0], ["mirror"]),
    # the fused x+y+z kernel (sfloat, z radius <= 4, y radius <= 8): interior and edge
    # tiles, odd x / y filters, several plane chunks, mixed conditions, thin z
    ("f32_300x100x40_s221", (40, 100, 300), "float32", [2.0, 2.0, 1.0], [0], ["mirror"]),
    ("f32_256x96x50_o010_periodic", (50, 96, 256), "float32", [2.0, 1.7, 1.0], [0, 1, 0], ["periodic"]),
    ("f32_192x64x36_o100_bcs", (36, 64, 192), "float32", [1.0, 2.0, 1.3], [1, 0, 0],
     ["asym mirror", "mirror", "zero order"]),
    # a constant condition along y: the separate passes (the fused kernel folds indices only)
    ("f32_192x64x36_o100_addzeros", (36, 64, 192), "float32", [1.0, 2.0, 1.3], [1, 0, 0],
     ["asym mirror", "add zeros", "zero order"]),
    # too many operations per sample for the fused kernel (y radius 8, z radius 4)
    ("f32_200x96x40_s25", (40, 96, 200), "float32", [2.5, 2.5, 1.3], [0], ["mirror"]),
    ("f32_260x70x90_zchunks", (90, 70, 260), "float32", [2.0, 1.0, 0.5], [0], ["mirror"]),
    # a long z axis (more planes than one chunk's plane map holds): many plane chunks
    ("f32_64x40x1100_longz", (1100, 40, 64), "float32", [1.0, 1.0, 1.0], [0], ["mirror"]),
    ("f32_130x64x9_asym_periodic", (9, 64, 130), "float32", [1.5, 1.5, 1.3], [0], ["asym periodic"]),
    ("f32_4d_128x64x20x2", (2, 20, 64, 128), "float32", [2.0, 2.0, 1.0, 0.0], [0], ["mirror"]),
    ("f64_64x32x16_o011", (16, 32, 64), "float64", [1.0, 2.0, 1.0], [0, 1, 1], ["periodic"]),
]


# the cases the fused x+y+z kernel takes
FUSED3 = {"f32_300x100x40_s221", "f32_256x96x50_o010_periodic", "f32_192x64x36_o100_bcs",
          "f32_260x70x90_zchunks", "f32_130x64x9_asym_periodic", "f32_4d_128x64x20x2",
          "f32_130x70x40_s215", "f32_128x64x33_periodic", "f32_96x50x20_o102", "f32_64x40x1100_longz"}


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


@pytest.mark.parametrize("case", VOL, ids=[c[0] for c in VOL])
def test_volume_vs_oracle(cuda, case):
    name, shape, dt, sig, ords, bc = case
    x = cases.make_input(shape, dt, seed=91, impulses=0.01)
    exp = oc.gauss_fir(x, sig, ords, bc)
    xt = torch.from_numpy(x).to(cuda)
    out = {}
    names = _kernels_of(lambda: out.setdefault("y", dip.GaussFIR(xt, sig, ords, bc)))
    y = out["y"]
    assert ("fused3d" in names) == (name in FUSED3), names
    got = y.cpu().numpy()
    assert got.dtype == exp.dtype
    assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), \
        f"{name}: max abs diff {np.nanmax(np.abs(got.astype(np.float64) - exp.astype(np.float64)))}"
    # the pass-by-pass path on the same input
    z = dip.GaussFIR(xt, sig, ords, bc, fusion=False)
    torch.cuda.synchronize()
    assert np.array_equal(z.cpu().numpy().view(np.uint8), got.view(np.uint8))


def test_volume_strided_output(cuda):
    """A view output (a sub-block of a larger volume) keeps the result bit-identical."""
    x = cases.make_input((20, 36, 72), "float32", seed=92)
    exp = oc.gauss_fir(x, [2.0, 2.0, 1.0], [0], ["mirror"])
    big = torch.zeros((22, 40, 80), dtype=torch.float32, device=cuda)
    view = big[1:21, 2:38, 4:76]
    dip.GaussFIR(torch.from_numpy(x).to(cuda), [2.0, 2.0, 1.0], [0], ["mirror"], out=view)
    torch.cuda.synchronize()
    assert np.array_equal(view.cpu().numpy().view(np.uint8), exp.view(np.uint8))
    assert float(big[0].abs().sum()) == 0.0 and float(big[:, 0].abs().sum()) == 0.0
