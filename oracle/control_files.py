"""The oracle of the rig control-directory loaders -- TEST INFRASTRUCTURE.

Only tests/ import this module, as the checker of the library's
hmControlFilesLoad / hmTiffRead / hmPngReadSeam.  It decodes the files with
Pillow (its libtiff and zlib), independently of the library's own decoders,
and restates the reference's rules after decoding:

  maps       cv::imread(IMREAD_ANYDEPTH) of mapping_%04d_{x,y}.tif, uint16
             (src/pano/controlMasks.cpp:290-316, controlMasksN.cpp:137-152,
             controlMasks3.cpp:244-278)
  placement  XPOSITION x XRESOLUTION, YPOSITION x YRESOLUTION in float32; the
             2- and 3-image loaders use the resolutions only when both are
             present (controlMasks.cpp:67-118, controlMasks3.cpp:29-58), the
             N-image loader each on its own (controlMasksN.cpp:22-41); an
             unreadable file gives (0, 0); then the minimum is subtracted
             (controlMasks.cpp:42-57)
  seam, 2    8-bit paletted -> indices, else grayscale; then max -> 0 and
             min -> 1 (controlMasks.cpp:225-251)
  seam, 3    must be 8-bit paletted (controlMasks3.cpp:92-198)
  seam, N    paletted or grayscale; exactly N distinct values, mapped to
             their rank unless already 0..N-1 (controlMasksN.cpp:156-190)

Grayscale seams: 16-bit keeps the high byte, 1/2/4-bit values scale to
0..255, gray + alpha drops the alpha.  Colour seams are refused.
"""
from __future__ import annotations

import os
import struct

import numpy as np

# status codes of include/hm_files.h
OK, NOT_FOUND, BAD_FORMAT, UNSUPPORTED, WRONG_TYPE, SEAM_CLASSES, MISMATCH = range(7)
TWO, THREE, N = 2, 3, 0


class OracleFileError(Exception):
    def __init__(self, code: int, what: str):
        super().__init__(f"[{code}] {what}")
        self.code = code


def _png_header(path):
    with open(path, "rb") as f:
        head = f.read(33)
    if head[:8] != b"\x89PNG\r\n\x1a\n" or head[12:16] != b"IHDR":
        raise OracleFileError(BAD_FORMAT, f"{path}: not a PNG")
    w, h, depth, ctype = struct.unpack(">IIBB", head[16:26])
    return w, h, depth, ctype


def read_seam(path, paletted_only: bool = False):
    """(uint8 [h, w], paletted) -- the seam as the reference's readers see it."""
    from PIL import Image

    if not os.path.isfile(path):
        raise OracleFileError(NOT_FOUND, path)
    _, _, depth, ctype = _png_header(path)
    if ctype == 3 and depth == 8:
        with Image.open(path) as im:
            assert im.mode == "P"
            return np.array(im, dtype=np.uint8), True
    if paletted_only:
        raise OracleFileError(UNSUPPORTED, f"{path}: not 8-bit paletted")
    if ctype not in (0, 4):
        raise OracleFileError(UNSUPPORTED, f"{path}: colour PNG")
    with Image.open(path) as im:
        if im.mode == "1":
            return np.array(im, dtype=np.uint8) * np.uint8(255), False
        if im.mode == "LA":
            return np.array(im, dtype=np.uint8)[..., 0].copy(), False
        a = np.array(im)
        if depth == 16:
            return (a.astype(np.uint32) >> 8).astype(np.uint8), False
        return a.astype(np.uint8), False


def read_map(path):
    from PIL import Image

    if not os.path.isfile(path):
        raise OracleFileError(NOT_FOUND, path)
    with Image.open(path) as im:
        a = np.array(im)
    if a.dtype != np.uint16 or a.ndim != 2:
        raise OracleFileError(WRONG_TYPE, f"{path}: {a.dtype} {a.shape}")
    return a


def read_position(path, both_resolutions: bool):
    """(xpos, ypos, found) as float32, before normalisation."""
    from PIL import Image

    zero = np.float32(0)
    try:
        im = Image.open(path)
    except (OSError, ValueError):
        return zero, zero, False
    with im:
        tags = im.tag_v2

        def tag(k):
            if k not in tags:
                return None
            v = tags[k]
            num, den = (v.numerator, v.denominator) if hasattr(v, "numerator") else (float(v), 1)
            return np.float32(0.0 if (num == 0 or den == 0) else float(num) / float(den))

        xres, yres = tag(282), tag(283)
        if both_resolutions and (xres is None or yres is None):
            xres = yres = None
        xres = zero if xres is None else xres
        yres = zero if yres is None else yres
        xp, yp = tag(286), tag(287)
        xp = zero if xp is None else xp
        yp = zero if yp is None else yp
    return np.float32(xp * xres), np.float32(yp * yres), True


def invert_two_image_seam(seam: np.ndarray) -> np.ndarray:
    lo, hi = int(seam.min()), int(seam.max())
    out = seam.copy()
    out[seam == hi] = 0
    out[seam == lo] = 1
    return out


def relabel_n(seam: np.ndarray, n: int) -> np.ndarray:
    uniq = np.unique(seam)
    if uniq.size == 0 or uniq.size != n:
        raise OracleFileError(SEAM_CLASSES, f"{uniq.size} classes, expected {n}")
    if uniq[0] == 0 and uniq[-1] == n - 1:
        return seam
    lut = np.zeros(256, np.uint8)
    lut[uniq] = np.arange(n, dtype=np.uint8)
    return lut[seam]


def load(directory, kind: int, n_images: int = 0) -> dict:
    n = 2 if kind == TWO else 3 if kind == THREE else n_images
    d = str(directory)
    maps_x, maps_y = [], []
    for i in range(n):
        mx = read_map(os.path.join(d, f"mapping_{i:04d}_x.tif"))
        my = read_map(os.path.join(d, f"mapping_{i:04d}_y.tif"))
        if mx.shape != my.shape:
            raise OracleFileError(MISMATCH, f"image {i}")
        maps_x.append(mx)
        maps_y.append(my)
    seam, paletted = read_seam(os.path.join(d, "seam_file.png"), paletted_only=kind == THREE)
    if kind == TWO:
        seam = invert_two_image_seam(seam)
    elif kind == N:
        seam = relabel_n(seam, n)
    xs, ys, found = [], [], 0
    for i in range(n):
        x, y, ok = read_position(os.path.join(d, f"mapping_{i:04d}.tif"), both_resolutions=kind != N)
        xs.append(x)
        ys.append(y)
        found |= int(ok) << i
    mx_, my_ = min(xs), min(ys)
    xs = [np.float32(x - mx_) for x in xs]
    ys = [np.float32(y - my_) for y in ys]
    return {"map_x": maps_x, "map_y": maps_y, "xpos": xs, "ypos": ys, "found": found, "seam": seam,
            "paletted": paletted}
