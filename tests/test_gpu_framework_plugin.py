"""The device separable framework with user line filters (include/diplib_amd_framework.hpp,
the plug-in operator API of dip::Framework::Separable): a compiled caller
(tests/cpp/test_framework_plugin, built by __graft_entry__.build()) runs the
reference's GENERAL convolution loop as a plug-in and checks it bit for bit against
dip::SeparableConvolution (pinned by the reference's fixtures) over boundary
conditions, input types, dims, process flags, a strided view and in-place use, and a
non-linear plug-in (running maximum) against a host loop."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_framework_plugin")


def test_framework_plugin(cuda):
    assert os.path.exists(BIN), "tests/cpp/test_framework_plugin is built by __graft_entry__.build()"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED" in p.stdout
