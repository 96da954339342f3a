"""The library's device scratch (a bump arena per stream): calls that need
intermediates do not wait for the device, results stay bit-identical whatever the
stream, the order of sizes (the arena grows and merges its blocks) and the number of
streams (beyond the arena limit idle arenas are released)."""
import time

import numpy as np
import pytest
import torch

import cases
import diplib_amd as dip
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu


def _vol(shape, seed):
    return cases.make_input(shape, "float32", seed=seed)


def test_scratch_sizes_and_streams(cuda):
    """Volumes of growing and shrinking sizes on two streams (every call needs an
    intermediate: sigma 3 along z takes the separate passes), against the oracle."""
    print("[scratch] sizes and streams", flush=True)
    shapes = [(20, 40, 72), (40, 90, 130), (12, 30, 50), (60, 100, 140), (20, 40, 72)]
    xs = [_vol(s, 40 + i) for i, s in enumerate(shapes)]
    exp = [oc.gauss_fir(x, [2.0, 2.0, 3.0], [0], ["mirror"]) for x in xs]
    side = torch.cuda.Stream()
    outs = []
    for rep in range(2):
        for i, x in enumerate(xs):
            st = side if (i + rep) % 2 else torch.cuda.current_stream()
            with torch.cuda.stream(st):
                xt = torch.from_numpy(x).to(cuda)
                outs.append((i, dip.GaussFIR(xt, [2.0, 2.0, 3.0], [0], ["mirror"])))
    torch.cuda.synchronize()
    for i, y in outs:
        assert np.array_equal(y.cpu().numpy().view(np.uint8), exp[i].view(np.uint8)), shapes[i]


def test_scratch_many_streams(cuda):
    """More streams than arenas: idle arenas are released, results unchanged."""
    print("[scratch] many streams", flush=True)
    x = _vol((16, 40, 80), 50)
    exp = oc.gauss_fir(x, [1.5, 1.5, 3.0], [0], ["mirror"])
    xt = torch.from_numpy(x).to(cuda)
    res = []
    streams = [torch.cuda.Stream() for _ in range(20)]
    for rep in range(2):
        for st in streams:
            with torch.cuda.stream(st):
                res.append(dip.GaussFIR(xt, [1.5, 1.5, 3.0], [0], ["mirror"]))
        torch.cuda.synchronize()
    for y in res:
        assert np.array_equal(y.cpu().numpy().view(np.uint8), exp.view(np.uint8))


def test_scratch_calls_do_not_wait_for_the_device(cuda):
    """Behind about 0.1 s of queued work, calls that need scratch return at once."""
    print("[scratch] host does not wait", flush=True)
    x = torch.rand((64, 128, 128), device=cuda)
    w = torch.randn((64, 128, 128), device=cuda)
    y = torch.empty_like(x)
    for _ in range(3):  # warm: the arena holds what these calls need
        dip.GaussFIR(x, [2.0, 2.0, 3.0], [0], ["mirror"], 3.0, out=y)
        dip.gauss_fir_backward(x, w, [2.0, 2.0, 2.0], [0, 0, 0], ["mirror"])
    torch.cuda.synchronize()
    a = torch.rand((8192, 8192), device=cuda)
    t0 = time.perf_counter()
    for _ in range(10):
        a = a @ a
        a = a / a.abs().max()
    torch.cuda.synchronize()
    busy = time.perf_counter() - t0
    for _ in range(10):
        a = a @ a
        a = a / a.abs().max()
    t0 = time.perf_counter()
    for _ in range(3):
        dip.GaussFIR(x, [2.0, 2.0, 3.0], [0], ["mirror"], 3.0, out=y)
        dip.gauss_fir_backward(x, w, [2.0, 2.0, 2.0], [0, 0, 0], ["mirror"])
    host = time.perf_counter() - t0
    torch.cuda.synchronize()
    assert host < 0.5 * busy, (host, busy)
