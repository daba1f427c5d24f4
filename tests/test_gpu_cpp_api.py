"""Runs the C++ API test program (tests/cpp/pano_api_test.cpp): the reference's
class interfaces (ControlMasks*, CudaStitchPano, CudaStitchPanoN,
CudaStitchPano3, CudaMat, CudaStatusOr) driving the gfx950 library, with
uint8 and float pipelines.  The float canvases the program writes are compared
with the oracle run on the same float inputs."""
import os
import subprocess

import numpy as np
import pytest

from conftest import IdRig, patterned_u8

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hm-cupano_amd", "bin", "pano_api_test")


def pattern_f(w, h, idx, frac):
    img = patterned_u8(w, h, idx).astype(np.float32)
    img[..., :3] += np.float32(frac)
    return img


def load(d, name, w, h):
    return np.fromfile(os.path.join(d, name + ".f32"), np.float32).reshape(1, h, w, 4)


def test_cpp_pano_api(torch_gpu, oracle, tmp_path):
    assert os.path.exists(BIN), "build the C++ API test first (__graft_entry__.build())"
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 18

    # src/pano/cudaPano_test.cpp:141-204 in float: the minimized canvases against the oracle
    w, h, x2 = 384, 64, 192
    cw = w + x2
    seam = np.zeros((h, cw), np.uint8)
    seam[:, : cw // 2] = 1
    rig = IdRig([(w, h), (w, h)], [(0, 0), (x2, 0)], seam)
    for name, frac in (("pano2_float_mini", 0.0), ("pano2_float_mini_frac", 0.375)):
        a, b = pattern_f(w, h, 0, frac), pattern_f(w, h, 1, frac)
        np.testing.assert_array_equal(load(tmp_path, name, cw, h), oracle.pano2(rig, a, b, 4, minimize=True))
    a3, b3 = pattern_f(w, h, 0, 0.375)[..., :3].copy(), pattern_f(w, h, 1, 0.375)[..., :3].copy()
    want3 = oracle.pano2(rig, a3, b3, 4, minimize=True)
    np.testing.assert_array_equal(load(tmp_path, "pano2_float3_mini_frac", cw, h)[..., :3], want3)

    # src/pano/cudaPanoN_test.cpp:192-299 in float
    cases = {
        "two_way_overlap": (384, 384, [0, 0], (192, 384)),
        "no_overlap": (192, 384, [0, 192], (192, 384)),
        "gap": (128, 384, [0, 256], (192, 384)),
        "multi_seam_intersection": (384, 384, [0, 0, 0, 0], None),
    }
    for case, (iw, ih, xs, band) in cases.items():
        seam = np.zeros((ih, 384), np.uint8)
        if band:
            seam[:, band[0]: band[1]] = 1
        else:
            cx, cy = iw // 2, ih // 2
            for y in range(cy - 32, cy + 32):
                for x in range(cx - 32, cx + 32):
                    seam[y, x] = (2 if y >= cy else 0) + (1 if x >= cx else 0)
        rig = IdRig([(iw, ih)] * len(xs), [(x, 0) for x in xs], seam)
        imgs = [pattern_f(iw, ih, i, 0.0) for i in range(len(xs))]
        got = load(tmp_path, "panoN_float_" + case, 384, ih)
        np.testing.assert_array_equal(got, oracle.panoN(rig, imgs, 4, minimize=True), err_msg=case)
