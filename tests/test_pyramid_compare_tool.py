"""hm-cupano_amd/compare_pyramid_dumps.py: the semantics of the reference's
scripts/compare_pyramid_dumps.py:24-66 on dump directories written with the
library's TIFF writer (CPU only)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def tool():
    sys.path.insert(0, os.path.join(ROOT, "hm-cupano_amd"))
    import compare_pyramid_dumps

    return compare_pyramid_dumps


def make_dump(hs, d, arrays):
    for rel, a in arrays.items():
        p = os.path.join(d, rel)
        os.makedirs(os.path.dirname(p), exist_ok=True)
        hs.tiff_write(p, a, opencv_order=True)


def arrays(seed):
    rng = np.random.default_rng(seed)
    return {
        "blend/level_0_batch_0.tiff": rng.standard_normal((8, 12, 4)).astype(np.float32),
        "blend/level_1_batch_0.tiff": rng.standard_normal((4, 6, 4)).astype(np.float32),
        "mask/level_0.tiff": rng.random((8, 12)).astype(np.float32),
        "mask/level_1.tiff": rng.random((4, 6, 1)).astype(np.float32),
    }


def test_identical_dumps(tool, tmp_path):
    import hm_stitch as hs

    a = arrays(1)
    make_dump(hs, tmp_path / "ref", a)
    make_dump(hs, tmp_path / "cand", a)
    lines = []
    r = tool.compare(tmp_path / "ref", tmp_path / "cand", out=lines.append)
    assert [f for f, _, _ in r["files"]] == sorted(a)
    assert all(m == 0.0 and mm == 0.0 for _, m, mm in r["files"])
    # all ties at zero: the first file stays the worst, nothing is nonzero
    assert r["worst"] == sorted(a)[0] and r["first_nonzero"] is None
    assert lines[-1].startswith("worst=")


def test_differences_are_located(tool, tmp_path):
    import hm_stitch as hs

    a = arrays(2)
    b = {k: v.copy() for k, v in a.items()}
    b["mask/level_0.tiff"][3, 4] += 0.25
    b["blend/level_1_batch_0.tiff"][1, 2, 3] -= 0.5
    make_dump(hs, tmp_path / "ref", a)
    make_dump(hs, tmp_path / "cand", b)
    r = tool.compare(tmp_path / "ref", tmp_path / "cand", out=lambda s: None)
    got = {f: (m, mm) for f, m, mm in r["files"]}
    assert got["mask/level_0.tiff"][0] == np.float32(a["mask/level_0.tiff"][3, 4] + np.float32(0.25)) - a[
        "mask/level_0.tiff"][3, 4]
    assert r["worst"] == "blend/level_1_batch_0.tiff" and r["worst_max"] == 0.5
    assert r["first_nonzero"] == "blend/level_1_batch_0.tiff"  # sorted order: blend/ before mask/
    assert got["blend/level_1_batch_0.tiff"][1] == pytest.approx(0.5 / (4 * 6 * 4))


def test_errors(tool, tmp_path):
    import hm_stitch as hs

    a = arrays(3)
    make_dump(hs, tmp_path / "ref", a)
    os.makedirs(tmp_path / "empty")
    with pytest.raises(SystemExit, match="No TIFF files"):
        tool.compare(tmp_path / "empty", tmp_path / "ref", out=lambda s: None)
    b = dict(a)
    del b["mask/level_1.tiff"]
    make_dump(hs, tmp_path / "cand", b)
    with pytest.raises(SystemExit, match="Missing candidate"):
        tool.compare(tmp_path / "ref", tmp_path / "cand", out=lambda s: None)
    b = dict(a)
    b["mask/level_1.tiff"] = np.zeros((5, 6), np.float32)
    make_dump(hs, tmp_path / "cand2", b)
    with pytest.raises(SystemExit, match="Shape mismatch"):
        tool.compare(tmp_path / "ref", tmp_path / "cand2", out=lambda s: None)


def test_command_line(tmp_path):
    import subprocess

    import hm_stitch as hs

    a = arrays(4)
    make_dump(hs, tmp_path / "ref", a)
    make_dump(hs, tmp_path / "cand", a)
    out = subprocess.run([sys.executable, os.path.join(ROOT, "hm-cupano_amd", "compare_pyramid_dumps.py"),
                          "--reference", str(tmp_path / "ref"), "--candidate", str(tmp_path / "cand")],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "mask/level_0.tiff: max_abs=0.000000000 mean_abs=0.000000000" in out.stdout
