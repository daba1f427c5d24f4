"""The down pass's 2x2 mean divides a sum by the tap count with two fused
multiply-adds (div_count, hm-cupano_amd/csrc/hm_pyramid_kernels.hpp) instead of
the IEEE division sequence; the oracle divides (cudaBlend.cu:85-134 computes
sum / count).  This checks on the CPU, for every finite float (-0 aside, which
the sums never are), that the two give the same bits for each count 1..4, and
that the check catches the correction step left out (x times the rounded
reciprocal alone)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "numerics", "div_count_check.c")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    exe = str(tmp_path_factory.mktemp("divc") / "div_count_check")
    subprocess.run(["gcc", "-O3", "-mavx2", "-mfma", "-ffp-contract=off", "-pthread", SRC, "-o", exe, "-lm"],
                   check=True)
    return exe


def test_div_count_matches_ieee_division_for_every_float(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=300)
    lines = r.stdout.split()
    assert r.returncode == 0, r.stdout
    assert lines == ["d=1", "mismatches=0", "d=2", "mismatches=0", "d=3", "mismatches=0", "d=4", "mismatches=0"]


def test_div_count_check_catches_a_missing_correction(checker):
    r = subprocess.run([checker, "no-correction"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0
    d, n = r.stdout.split()
    assert d == "d=3" and int(n.split("=")[1]) > 1000
