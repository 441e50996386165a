"""The C-ABI-only caller (t2-relion-kernel_amd/examples/coarse_search): no
Python between the caller and the library.  It builds a projector, filters
unfiltered images on the device, runs one batched search per form over
particles with different translation counts, and checks every particle's
best sample against the single-image drop-in on the same image; it exits 0
and prints "coarse_search ok" when every check holds."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "t2-relion-kernel_amd", "examples", "coarse_search")


def test_c_caller(torch_cuda):
    assert os.access(EXE, os.X_OK), f"{EXE} is not built (make -C t2-relion-kernel_amd)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "coarse_search ok" in r.stdout
    assert "fp32" in r.stdout and "bf16x3" in r.stdout
