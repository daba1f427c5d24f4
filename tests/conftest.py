"""Test configuration: import paths, the `gpu` marker and shared fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("hm-cupano_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import hm_oracle

    hm_oracle.lib()
    return hm_oracle


@pytest.fixture(scope="session")
def torch_gpu():
    """torch with a visible GPU; GPU tests fail (not skip) without one."""
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible MI355X"
    return torch


@pytest.fixture(scope="session")
def hm(torch_gpu):
    import hm_stitch

    hm_stitch.lib()  # raises if libhm_stitch.so is missing: no fallback path
    return hm_stitch


def frames_for(rig, batch, seed, channels=4):
    """Deterministic frames [B, H, W, C] for every camera of a rig."""
    import hm_synth

    out = []
    for i in range(rig.n):
        fr = np.stack([hm_synth.make_frame(rig.src_h, rig.src_w, seed + 101 * i + 7 * b, channels)
                       for b in range(batch)])
        out.append(fr)
    return out
