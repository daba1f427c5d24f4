#!/usr/bin/env python
"""Compare two pyramid dump directories (hmStitchPanoDumpPyramid /
dump_soft_blend_pyramid output), file by file.

The behaviour of scripts/compare_pyramid_dumps.py:24-66: every *.tiff under
the reference directory (sorted by relative path) is read from both sides as
float32, a trailing single channel is dropped, and per file the maximum and
mean absolute difference are printed; then the worst file (largest maximum,
ties broken by the mean) and the first file with a nonzero difference.  A
missing candidate file, a shape mismatch, an empty reference directory or
nothing compared ends the run with an error message.

The files are decoded by the library's TIFF reader (hm_stitch.tiff_read, in
OpenCV's channel order), so dumps written by either implementation compare.

  python hm-cupano_amd/compare_pyramid_dumps.py --reference DIR --candidate DIR
"""
from __future__ import annotations

import argparse
import os
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import hm_stitch  # noqa: E402


def load(path: Path) -> np.ndarray:
    if not path.exists():
        raise FileNotFoundError(path)
    a = hm_stitch.tiff_read(str(path), opencv_order=True).astype(np.float32, copy=False)
    if a.ndim == 3 and a.shape[-1] == 1:
        a = a[..., 0]
    return a


def compare(reference: Path, candidate: Path, out=print) -> dict:
    """Returns {"files": [(rel, max_abs, mean_abs)], "worst": rel, "worst_max", "worst_mean",
    "first_nonzero": rel or None}; raises SystemExit like the reference's script."""
    files = sorted(p.relative_to(reference) for p in reference.rglob("*.tiff"))
    if not files:
        raise SystemExit(f"No TIFF files found under {reference}")
    rows = []
    worst, worst_max, worst_mean, first_nonzero = None, -1.0, -1.0, None
    for rel in files:
        cand_path = candidate / rel
        if not cand_path.exists():
            raise SystemExit(f"Missing candidate file: {cand_path}")
        ref = load(reference / rel)
        cand = load(cand_path)
        if ref.shape != cand.shape:
            raise SystemExit(f"Shape mismatch for {rel}: {ref.shape} vs {cand.shape}")
        diff = np.abs(ref - cand)
        max_abs = float(diff.max(initial=0.0))
        mean_abs = float(diff.mean()) if diff.size else 0.0
        out(f"{rel}: max_abs={max_abs:.9f} mean_abs={mean_abs:.9f}")
        rows.append((str(rel), max_abs, mean_abs))
        if max_abs > 0.0 and first_nonzero is None:
            first_nonzero = rel
        if max_abs > worst_max or (max_abs == worst_max and mean_abs > worst_mean):
            worst, worst_max, worst_mean = rel, max_abs, mean_abs
    if worst is None:
        raise SystemExit("No files were compared")
    out(f"worst={worst} max_abs={worst_max:.9f} mean_abs={worst_mean:.9f}")
    if first_nonzero is not None:
        out(f"first_nonzero={first_nonzero}")
    return {"files": rows, "worst": str(worst), "worst_max": worst_max, "worst_mean": worst_mean,
            "first_nonzero": None if first_nonzero is None else str(first_nonzero)}


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description="Compare two pano pyramid dump directories.")
    parser.add_argument("--reference", required=True, help="Reference dump directory.")
    parser.add_argument("--candidate", required=True, help="Candidate dump directory.")
    args = parser.parse_args(argv)
    compare(Path(args.reference), Path(args.candidate))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
