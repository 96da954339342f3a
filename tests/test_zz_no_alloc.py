"""The batched form's entry point that allocates nothing (dipamd_batch_gauss_noalloc):
its tables and work-queue counters live in the library's fixed device array.  Checked
through the C ABI against the plan path (dipamd_batch_create + dipamd_batch_run), bit
for bit, on two calls in a row -- the second with a smaller batch, so stale tables or
an uncleared counter from the first would show -- and against the oracle."""
import numpy as np
import pytest
import torch

import oracle_ctypes as oc
import diplib_amd as dip

pytestmark = pytest.mark.gpu


def _batch(dev, n, seed, dtype, lo, hi):
    rng = np.random.default_rng(seed)
    ws = rng.integers(lo, hi, n)
    hs = rng.integers(lo, hi, n)
    sig = np.exp(rng.uniform(np.log(0.5), np.log(8.0), n))
    ins = []
    for w, h in zip(ws, hs):
        a = rng.integers(0, 256, (int(h), int(w))).astype(np.uint8) if dtype == torch.uint8 else \
            rng.random((int(h), int(w)), dtype=np.float32)
        ins.append(torch.from_numpy(a).to(dev))
    return ins, [float(s) for s in sig]


@pytest.mark.parametrize("dtype", [torch.float32, torch.uint8])
def test_noalloc_matches_plan_twice(cuda, dtype):
    for call, (n, lo, hi, seed) in enumerate(((48, 8, 700, 31), (9, 1, 90, 32))):  # the second is smaller
        ins, sig = _batch(cuda, n, seed, dtype, lo, hi)
        got = [torch.full(t.shape, float("nan"), dtype=torch.float32, device=cuda) for t in ins]
        ref = [torch.empty(t.shape, dtype=torch.float32, device=cuda) for t in ins]
        print(f"[no-alloc] call {call}: {n} images {lo}..{hi} px, {dtype}", flush=True)
        dip.gauss_batch_noalloc(ins, got, sig, boundary="mirror")
        plan = dip.GaussBatch(ins, ref, sig, boundary="mirror")
        plan.run()
        torch.cuda.synchronize()
        plan.close()
        for i, (a, b) in enumerate(zip(got, ref)):
            assert torch.equal(a, b), f"call {call} image {i} {tuple(a.shape)} sigma {sig[i]}"
        for i in (0, n - 1):
            exp = oc.gauss_fir(ins[i].cpu().numpy(), [sig[i]], [0], ["mirror"], 3.0)
            assert np.array_equal(got[i].cpu().numpy(), exp), f"call {call} image {i} vs oracle"


def test_noalloc_rejects_tables_above_the_scratch(cuda):
    """A batch whose tables exceed the fixed array is refused before any launch."""
    x = torch.rand((4, 4), device=cuda)
    y = torch.empty_like(x)
    n = 70000
    with pytest.raises(dip.DipError) as e:
        dip.gauss_batch_noalloc([x] * n, [y] * n, [2.0] * n)
    assert e.value.code == dip._lib.E_UNSUPPORTED
    # the library still works after the refusal
    z = torch.empty_like(x)
    dip.gauss_batch_noalloc([x], [z], [2.0])
    torch.cuda.synchronize()
    assert np.array_equal(z.cpu().numpy(), oc.gauss_fir(x.cpu().numpy(), [2.0], [0], ["mirror"], 3.0))
