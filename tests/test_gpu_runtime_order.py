"""One HIP runtime per process: loading libskm.so before PyTorch touches the
GPU must leave PyTorch working, and the library must accept torch's tensors
afterwards (PyTorch's ROCm wheel bundles its own libamdhip64.so.7; the loader
makes sure that copy is the one both use).  Run in a fresh interpreter, since
this test process has imported torch already."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import numpy as np
from skmimageregistration_amd import GaussianSmoothing, Image, _capi
_capi.load()
img = Image(np.random.default_rng(0).random((16, 18, 20)).astype(np.float32))
ref = GaussianSmoothing(img, 1.5).data          # the library first
import torch
t = torch.from_numpy(img.data).cuda()           # then torch's own GPU use
assert float(t.sum().cpu()) > 0
from skmimageregistration_amd import warmstart  # and a torch tensor through the library
net = warmstart.WarmStartNet((16, 16, 16), (3, 3, 3), (1.0, 1.0, 1.0), (4.0, 4.0, 4.0))
out = net.forward(torch.zeros((1, 16, 16, 16, 2), device="cuda"))
assert tuple(out.shape) == (1, 3, 3, 3, 3)
print("ok")
"""


def test_library_first_then_torch():
    print("fresh interpreter: libskm, then torch", flush=True)
    r = subprocess.run([sys.executable, "-c", SCRIPT], cwd=REPO, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-2000:] + r.stderr[-2000:]
