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


# ---- small rigs and images of the reference's pano tests ----
def identity_maps(w, h):
    return (np.tile(np.arange(w, dtype=np.uint16), (h, 1)), np.tile(np.arange(h, dtype=np.uint16)[:, None], (1, w)))


class IdRig:
    """Identity-map rig of the reference's pano tests."""

    def __init__(self, sizes, positions, seam):
        self.map_x = [identity_maps(w, h)[0] for w, h in sizes]
        self.map_y = [identity_maps(w, h)[1] for w, h in sizes]
        self.pos_x = [p[0] for p in positions]
        self.pos_y = [p[1] for p in positions]
        self.seam = seam
        self.n = len(sizes)


def const_u8(w, h, rgba):
    return np.broadcast_to(np.uint8(rgba), (1, h, w, 4)).copy()


def patterned_u8(w, h, idx):
    ys, xs = np.mgrid[0:h, 0:w]
    base = idx * 50
    img = np.zeros((1, h, w, 4), np.uint8)
    img[0, ..., 0] = base + xs % 17
    img[0, ..., 1] = base + ys % 19
    img[0, ..., 2] = base + (xs + ys) % 23
    img[0, ..., 3] = 255
    return img


def scatter_far_entries(rig, seed, rows=0.15, frac=0.3):
    """A copy of `rig` whose maps hold, on some rows, entries far outside the
    source (x or y up to 65534, or the 65535 sentinel) among the regular ones:
    out-of-range pixels of the reference's remap, and 64-entry segments whose
    spread is too wide for the compact map layout (its 4-byte fallback)."""
    import copy

    rng = np.random.default_rng(seed)
    r = copy.copy(rig)
    r.map_x = [m.copy() for m in rig.map_x]
    r.map_y = [m.copy() for m in rig.map_y]
    for mx, my in zip(r.map_x, r.map_y):
        h, w = mx.shape
        for y in rng.choice(h, size=max(1, int(rows * h)), replace=False):
            cols = rng.choice(w, size=max(1, int(frac * w)), replace=False)
            kind = rng.integers(0, 3, size=cols.size)
            mx[y, cols[kind == 0]] = rng.integers(1000, 65535, size=int((kind == 0).sum()))
            my[y, cols[kind == 1]] = rng.integers(1000, 65536, size=int((kind == 1).sum()))
            mx[y, cols[kind == 2]] = 65535
    return r
