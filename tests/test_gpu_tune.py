"""The schedule tuner (dipamd_autotune): it measures every candidate schedule and keeps
the winners (in memory here: the tests do not write the persistent cache); whatever it
chooses, results stay bit for bit the reference's (the schedules differ in speed only)."""
import os

import numpy as np
import pytest
import torch

import cases
import diplib_amd as dip
import oracle_ctypes as oc

pytestmark = pytest.mark.gpu


def test_autotune_then_results_unchanged(cuda):
    print("[tune] quick autotune", flush=True)
    assert os.environ.get("DIPAMD_TUNE_WRITE") == "0"  # the choices stay in this process
    dip.autotune(quick=True)
    for shape, dt, sig, ords in (((300, 1024), "uint8", 1.0, (0, 1)), ((256, 520), "float32", 4.0, (1, 0)),
                                 ((128, 96), "float64", 2.0, (0, 2)), ((64, 130, 70), "float32", 3.0, (0, 0, 1))):
        x = cases.make_input(shape, dt, seed=101)
        ref = oc.gauss_fir(x, [sig], list(ords), ["mirror"], 3.0)
        y = dip.Gauss(torch.from_numpy(x).to(cuda), [sig], list(ords), "FIR", ["mirror"], 3.0)
        torch.cuda.synchronize()
        assert np.array_equal(y.cpu().numpy().view(np.uint8), ref.view(np.uint8)), (shape, dt, sig, ords)
