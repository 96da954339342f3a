"""Batched uneven form: one persistent launch (and its hipGraph replay) against the
per-image loop of the plain drop-in and against the oracle, bit for bit."""
import numpy as np
import pytest
import torch

import oracle_ctypes as oc
import diplib_amd as dip

pytestmark = pytest.mark.gpu


def _batch(dev, n, seed, dtype=torch.float32, lo=1, hi=300):
    rng = np.random.default_rng(seed)
    ws = rng.integers(lo, hi, n)
    hs = rng.integers(lo, hi, n)
    ws[:3] = [1, 2, 7]
    hs[:3] = [5, 1, 3]
    sig = np.exp(rng.uniform(np.log(0.5), np.log(8.0), n))
    ins_np = []
    for w, h in zip(ws, hs):
        if dtype == torch.uint8:
            ins_np.append(rng.integers(0, 256, (int(h), int(w))).astype(np.uint8))
        elif dtype == torch.float64:
            ins_np.append(rng.random((int(h), int(w))))
        else:
            ins_np.append(rng.random((int(h), int(w)), dtype=np.float32))
    ins = [torch.from_numpy(a).to(dev) for a in ins_np]
    return ins_np, ins, sig


@pytest.mark.parametrize("dtype", [torch.float32, torch.uint8, torch.float64])
@pytest.mark.parametrize("cfg", [-1, 0, 1])
@pytest.mark.parametrize("bc", ["mirror", "periodic"])
def test_batched_matches_loop_and_oracle(cuda, dtype, cfg, bc):
    ins_np, ins, sig = _batch(cuda, 24, 7)
    ins = [t.to(dtype) if dtype != torch.uint8 else torch.from_numpy((a * 255).astype(np.uint8)).to(cuda)
           for a, t in zip(ins_np, ins)]
    odt = torch.float64 if dtype == torch.float64 else torch.float32
    ob = [torch.empty(t.shape, dtype=odt, device=cuda) for t in ins]
    ol = [torch.empty(t.shape, dtype=odt, device=cuda) for t in ins]
    plan = dip.GaussBatch(ins, ob, [float(s) for s in sig], boundary=bc, tile_config=cfg)
    loop = dip.GaussBatch(ins, ol, [float(s) for s in sig], boundary=bc, tile_config=1)
    plan.run()
    loop.run_per_image()
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(ob, ol)):
        assert torch.equal(a, b), f"image {i} {tuple(a.shape)} sigma {sig[i]}"
    for i in (0, 1, 2, 5, 11):
        ref = oc.gauss_fir(ins[i].cpu().numpy(), [float(sig[i])], [0], [bc], 3.0)
        assert np.array_equal(ob[i].cpu().numpy(), ref), f"image {i} vs oracle"
    plan.close()
    loop.close()


def test_batched_graph_replays(cuda):
    ins_np, ins, sig = _batch(cuda, 40, 9, lo=16, hi=700)
    ob = [torch.empty_like(t) for t in ins]
    ref = [torch.empty_like(t) for t in ins]
    plan = dip.GaussBatch(ins, ob, [float(s) for s in sig])
    loop = dip.GaussBatch(ins, ref, [float(s) for s in sig], tile_config=0)
    loop.run_per_image()
    for rep in range(4):
        for t in ob:
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run_graph()
        torch.cuda.synchronize()
        junk = torch.rand(1000, 1000, device=cuda).sum()  # eager work between replays
        assert torch.isfinite(junk)
        for i, (a, b) in enumerate(zip(ob, ref)):
            assert torch.equal(a, b), f"replay {rep} image {i}"
    plan.close()
    loop.close()


def test_batched_errors(cuda):
    x = torch.rand((10, 10), device=cuda)
    with pytest.raises(Exception):
        dip.GaussBatch([x], [torch.empty((10, 11), device=cuda)], [1.0])
    with pytest.raises(dip.DipError):
        dip.GaussBatch([x], [torch.empty_like(x)], [30.0])  # radius above the fused-tile limit
