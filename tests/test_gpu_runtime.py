"""The library and PyTorch share one HIP runtime whichever is used first: the
library loads torch's runtime library without importing torch, works, and a
torch imported afterwards still sees the GPU (a second HIP runtime would
leave torch with none).  Run in a fresh process, where nothing has imported
torch yet."""
import os
import subprocess
import sys

import pytest

SCRIPT = r"""
import sys
import numpy as np
from regridding_amd import _util, workloads
from regridding_amd._conservative import build_2d
from oracle import oracle
gi = workloads.warped_grid((41, 37), 0.3, 0.05, 0.02)
go = workloads.warped_grid((29, 31), 0.1, 0.03, 0.05)
ci, co, *_ = _util.normalize_coordinates(gi, go, perturb=True)
h = build_2d(ci, co)
got = h.export()
ref = oracle.weights_2d(ci, co)
assert all(np.array_equal(a, b) for a, b in zip(got, ref))
print("torch imported before the first build:", "torch" in sys.modules)
import torch
assert torch.cuda.is_available()
t = torch.arange(10, device="cuda:0", dtype=torch.float64)
assert float(t.sum()) == 45.0
h2 = build_2d(ci, co)
assert all(np.array_equal(a, b) for a, b in zip(h2.export(), ref))
print("ok")
"""


@pytest.mark.gpu
def test_library_first_then_torch():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True, timeout=240,
                       cwd=root)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torch imported before the first build: False" in r.stdout
    assert r.stdout.strip().endswith("ok")
