"""The C++ drop-in (include/diplib_amd.hpp) driven by a compiled C++ caller
(tests/cpp/test_dip_shim, built by __graft_entry__.build()), checked against the
reference's own dip::Gauss output for the 512x512 sfloat sigma=2 mirror case."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_dip_shim")


def test_cpp_drop_in(cuda):
    assert os.path.exists(BIN), "tests/cpp/test_dip_shim is built by __graft_entry__.build()"
    gold = np.load(os.path.join(HERE, "golden", "small_cases.npz"))
    name = "f32_512x512_s2_o0_mirror"
    x = np.ascontiguousarray(gold[name + "__in"], np.float32)
    y = np.ascontiguousarray(gold[name + "__out"], np.float32)
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in.f32"), os.path.join(d, "out.f32")
        x.tofile(fi)
        y.tofile(fo)
        p = subprocess.run([BIN, fi, str(x.shape[1]), str(x.shape[0]), fo, "2.0"], capture_output=True, text=True,
                           timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED" in p.stdout
