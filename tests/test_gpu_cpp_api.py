"""Runs the C++ API test program (tests/cpp/pano_api_test.cpp): the reference's
class interfaces (ControlMasks*, CudaStitchPano, CudaStitchPanoN,
CudaStitchPano3, CudaMat, CudaStatusOr) driving the gfx950 library."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hm-cupano_amd", "bin", "pano_api_test")


def test_cpp_pano_api(torch_gpu):
    assert os.path.exists(BIN), "build the C++ API test first (__graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 10
