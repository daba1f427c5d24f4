#!/usr/bin/env python
"""Stitch raw BGR video streams of a rig into one raw BGR panorama stream.

The frame loop of the reference's tests/stitch_two_videos.cpp (decode ->
stitch -> encode) with the codec left to a separate process: inputs and
output are raw bgr24 frame streams (files, FIFOs or '-' for stdin/stdout),
as ffmpeg reads and writes them:

  ffmpeg -i left.mp4 -f rawvideo -pix_fmt bgr24 left.bgr
  ffmpeg -i right.mp4 -f rawvideo -pix_fmt bgr24 right.bgr
  python hm-cupano_amd/stitch_raw_video.py --control rig/ --input left.bgr --input right.bgr \\
      --input-size 4096x2160 --levels 6 --output - | \\
      ffmpeg -f rawvideo -pix_fmt bgr24 -s <canvas WxH> -r 30 -i - out.mp4

The rig comes from the control directory (mapping_*.tif, seam_file.png);
two inputs use the 2-image class, more the N-image class.  Frames are
streamed through hmFrameLoopRun (pinned double buffering, BGR over PCIe).
The canvas size is printed on stderr before the first frame.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import hm_stitch  # noqa: E402


def parse_size(s: str):
    w, h = s.lower().split("x")
    return int(w), int(h)


def open_stream(path: str, mode: str):
    if path == "-":
        return (sys.stdin if "r" in mode else sys.stdout).buffer
    return open(path, mode + "b")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--control", required=True, help="control directory (mapping_*.tif, seam_file.png)")
    ap.add_argument("--input", action="append", required=True, help="raw bgr24 input stream (repeat per camera)")
    ap.add_argument("--input-size", action="append", required=True, type=parse_size,
                    help="WxH of each input's frames (one value for all, or one per --input)")
    ap.add_argument("--output", required=True, help="raw bgr24 output stream ('-' = stdout)")
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--channels", type=int, default=4, choices=(3, 4),
                    help="pipeline: 4 = BGRA (the reference's uchar4), 3 = BGR (uchar3)")
    ap.add_argument("--max-frames", type=int, default=-1)
    ap.add_argument("--slots", type=int, default=3, help="pinned host slots in flight")
    args = ap.parse_args(argv)

    n = len(args.input)
    if n < 2:
        ap.error("need at least two --input streams")
    sizes = args.input_size if len(args.input_size) == n else args.input_size[:1] * n
    if len(sizes) != n:
        ap.error("--input-size: give one value or one per --input")
    if n == 2:
        rig = hm_stitch.load_control_dir(args.control, hm_stitch.HM_CONTROL_TWO)
        pano = hm_stitch.StitchPano(1, args.levels, rig, channels=args.channels)
    else:
        rig = hm_stitch.load_control_dir(args.control, hm_stitch.HM_CONTROL_N, n)
        pano = hm_stitch.StitchPanoN(1, args.levels, rig, channels=args.channels)
    if pano.status != 0:
        raise SystemExit(f"pano construction failed: hipError {pano.status}")
    cw, ch = pano.canvas_size()
    print(f"Canvas: {cw}x{ch}", file=sys.stderr)

    ins = [open_stream(p, "r") for p in args.input]
    out = open_stream(args.output, "w")
    frame_bytes = [w * h * 3 for w, h in sizes]

    def source(k):
        frames = []
        for f, nb, (w, h) in zip(ins, frame_bytes, sizes):
            buf = f.read(nb)
            if len(buf) < nb:
                return None  # a stream ended
            frames.append(np.frombuffer(buf, np.uint8).reshape(h, w, 3))
        return frames

    def sink(k, bgr):
        out.write(bgr.tobytes())

    st = hm_stitch.frame_loop(pano, sizes, source, sink, slots=args.slots, max_frames=args.max_frames)
    out.flush()
    fps = st["frames"] / st["seconds"] if st["seconds"] > 0 else 0.0
    print(f"Processed frames: {st['frames']}, {st['seconds']:.2f} s ({fps:.1f} fps); per frame: upload "
          f"{st['upload_ms']:.3f} ms, stitch {st['stitch_ms']:.3f} ms, download {st['download_ms']:.3f} ms",
          file=sys.stderr)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
