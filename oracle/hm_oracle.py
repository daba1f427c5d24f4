"""ctypes wrapper of the CPU oracle (oracle/libhm_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module; the product (hm-cupano_amd/) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhm_oracle.so")
MAX_IMAGES = 8
MAX_LEVELS = 16

_P = ctypes.c_void_p
_I = ctypes.c_int


class hmo_dump(ctypes.Structure):
    _fields_ = [
        ("levels", _I),
        ("widths", _I * MAX_LEVELS),
        ("heights", _I * MAX_LEVELS),
        ("gauss", (_P * MAX_LEVELS) * MAX_IMAGES),
        ("lap", (_P * MAX_LEVELS) * MAX_IMAGES),
        ("mask", _P * MAX_LEVELS),
        ("blend", _P * MAX_LEVELS),
        ("reconstruct", _P * MAX_LEVELS),
    ]


class hmo_pano2_cfg(ctypes.Structure):
    _fields_ = [
        ("batch", _I), ("levels", _I), ("channels", _I), ("minimize_blend", _I),
        ("src_w", _I * 2), ("src_h", _I * 2), ("map_w", _I * 2), ("map_h", _I * 2),
        ("map_x", _P * 2), ("map_y", _P * 2), ("pos_x", _I * 2), ("pos_y", _I * 2),
        ("seam", _P), ("seam_w", _I), ("seam_h", _I),
        ("pipeline_f32", _I),
    ]


class hmo_panoN_cfg(ctypes.Structure):
    _fields_ = [
        ("n", _I), ("batch", _I), ("levels", _I), ("channels", _I), ("minimize_blend", _I),
        ("src_w", _I * MAX_IMAGES), ("src_h", _I * MAX_IMAGES),
        ("map_w", _I * MAX_IMAGES), ("map_h", _I * MAX_IMAGES),
        ("map_x", _P * MAX_IMAGES), ("map_y", _P * MAX_IMAGES),
        ("pos_x", _I * MAX_IMAGES), ("pos_y", _I * MAX_IMAGES),
        ("seam", _P), ("seam_w", _I), ("seam_h", _I),
        ("pipeline_f32", _I),
    ]


class hmo_pano3_cfg(ctypes.Structure):
    _fields_ = [
        ("batch", _I), ("levels", _I), ("channels", _I), ("fused", _I),
        ("src_w", _I * 3), ("src_h", _I * 3), ("map_w", _I * 3), ("map_h", _I * 3),
        ("map_x", _P * 3), ("map_y", _P * 3), ("pos_x", _I * 3), ("pos_y", _I * 3),
        ("seam", _P), ("seam_w", _I), ("seam_h", _I),
        ("pipeline_f32", _I),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.hmo_round_to_uchar.restype = ctypes.c_uint8
        L.hmo_round_to_uchar.argtypes = [ctypes.c_float]
        L.hmo_pyramid_sizes.argtypes = [_I, _I, _I, _I, _P, _P]
        L.hmo_pad_replicate.argtypes = [_P, _I, _I, _P, _I, _I]
        L.hmo_set_threads.argtypes = [_I]
        L.hmo_f16_to_f32.restype = ctypes.c_float
        L.hmo_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.hmo_f32_to_f16.restype = ctypes.c_uint16
        L.hmo_f32_to_f16.argtypes = [ctypes.c_float]
        L.hmo_remap_offset_t.argtypes = [_P, _I, _I, _I, _I, _P, _I, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _I,
                                         _I, _I, _I]
        L.hmo_remap_dest_map_t.argtypes = [_P, _I, _I, _I, _I, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _P, _I]
        L.hmo_remap_hard_seam_n_t.argtypes = [_I, _P, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I]
        L.hmo_copy_roi_t.argtypes = [_P, _I, _I, _I, _I, _I, _I, _I, _I, _P, _I, _I, _I, _I, _I, _I, _I]
        L.hmo_remap_offset.argtypes = [_P, _I, _I, _I, _P, _I, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _I,
                                       _I]
        L.hmo_remap_dest_map.argtypes = [_P, _I, _I, _I, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I, _P, _I]
        L.hmo_remap_hard_seam_n.argtypes = [_I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I]
        L.hmo_copy_roi.argtypes = [_P, _I, _I, _I, _I, _I, _I, _I, _I, _P, _I, _I, _I, _I, _I, _I, _I]
        L.hmo_downsample.argtypes = [_P, _I, _I, _P, _I, _I, _I, _I]
        L.hmo_downsample_mask.argtypes = [_P, _I, _I, _P, _I, _I, _I]
        L.hmo_laplacian.argtypes = [_P, _I, _I, _P, _I, _I, _P, _I, _I]
        L.hmo_reconstruct.argtypes = [_P, _I, _I, _P, _I, _I, _P, _I, _I]
        for fn in (L.hmo_downsample, L.hmo_downsample_mask, L.hmo_laplacian, L.hmo_reconstruct):
            fn.restype = None
        L.hmo_blend2.argtypes = [_P, _P, _P, _P, _I, _I, _I, _I, _I, _P]
        L.hmo_blend_n.argtypes = [_I, _P, _P, _P, _I, _I, _I, _I, _I, _P]
        L.hmo_blend3.argtypes = [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P]
        L.hmo_pano2_canvas_size.argtypes = [_P, _P, _P]
        L.hmo_pano2_process.argtypes = [_P, _P, _P, _P, _P]
        L.hmo_panoN_canvas_size.argtypes = [_P, _P, _P]
        L.hmo_panoN_rois.argtypes = [_P, _P, _P]
        L.hmo_panoN_process.argtypes = [_P, _P, _P, _P]
        L.hmo_pano3_canvas_size.argtypes = [_P, _P, _P]
        L.hmo_pano3_process.argtypes = [_P, _P, _P, _P]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


U8, F32, F16 = 0, 1, 2


def _et(a: np.ndarray) -> int:
    """Element type code of an image array (uint8, float32 or float16)."""
    return {np.dtype(np.uint8): U8, np.dtype(np.float32): F32, np.dtype(np.float16): F16}[a.dtype]


def f16_to_f32(h: int) -> float:
    return float(lib().hmo_f16_to_f32(h))


def f32_to_f16(f: float) -> int:
    return int(lib().hmo_f32_to_f16(ctypes.c_float(f)))


def round_to_uchar(x: float) -> int:
    return int(lib().hmo_round_to_uchar(ctypes.c_float(x)))


def set_threads(n: int) -> int:
    """Threads of the oracle's per-pixel stages (0 keeps the setting); returns the count in use."""
    return int(lib().hmo_set_threads(n))


def pad_replicate(seam: np.ndarray, canvas_w: int, canvas_h: int) -> Optional[np.ndarray]:
    """The seam padded to canvas_w x canvas_h (right/bottom replicate), or
    None when the oracle refuses it (seam larger than the canvas, or empty)."""
    seam = np.ascontiguousarray(seam, np.uint8)
    out = np.zeros((canvas_h, canvas_w), np.uint8)
    sh, sw = seam.shape
    rc = lib().hmo_pad_replicate(_ptr(seam), sw, sh, _ptr(out), canvas_w, canvas_h)
    return out if rc == 0 else None


def pyramid_sizes(w: int, h: int, levels: int, clamp: bool):
    ws = np.zeros(MAX_LEVELS, np.int32)
    hs = np.zeros(MAX_LEVELS, np.int32)
    n = lib().hmo_pyramid_sizes(w, h, levels, int(clamp), _ptr(ws), _ptr(hs))
    return [(int(ws[i]), int(hs[i])) for i in range(n)]


# ---------------------------------------------------------------------------
# dumps
# ---------------------------------------------------------------------------
class Dump:
    """Allocates per-level arrays for an oracle pyramid dump."""

    def __init__(self, n: int, w: int, h: int, c: int, levels: int, batch: int, clamp: bool, mask_channels: int):
        self.sizes = pyramid_sizes(w, h, levels, clamp)
        self.st = hmo_dump()
        self.gauss = [[None] * len(self.sizes) for _ in range(n)]
        self.lap = [[None] * len(self.sizes) for _ in range(n)]
        self.mask, self.blend, self.reconstruct = [None] * len(self.sizes), [None] * len(self.sizes), [None] * len(self.sizes)
        for l, (lw, lh) in enumerate(self.sizes):
            for i in range(n):
                self.gauss[i][l] = np.zeros((batch, lh, lw, c), np.float32)
                self.lap[i][l] = np.zeros((batch, lh, lw, c), np.float32)
                self.st.gauss[i][l] = _ptr(self.gauss[i][l])
                self.st.lap[i][l] = _ptr(self.lap[i][l])
            self.mask[l] = np.zeros((lh, lw, mask_channels), np.float32)
            self.blend[l] = np.zeros((batch, lh, lw, c), np.float32)
            self.reconstruct[l] = np.zeros((batch, lh, lw, c), np.float32)
            self.st.mask[l] = _ptr(self.mask[l])
            self.st.blend[l] = _ptr(self.blend[l])
            self.st.reconstruct[l] = _ptr(self.reconstruct[l])


# ---------------------------------------------------------------------------
# blends
# ---------------------------------------------------------------------------
def blend(kind: int, images: Sequence[np.ndarray], mask: np.ndarray, levels: int, dump: bool = False):
    """kind 2, 3 or 0 (N).  images: float32 [B, H, W, C]; mask: float32 [H, W(, M)]."""
    imgs = [np.ascontiguousarray(i, np.float32) for i in images]
    b, h, w, c = imgs[0].shape
    mask = np.ascontiguousarray(mask, np.float32)
    out = np.zeros_like(imgs[0])
    d = Dump(len(imgs), w, h, c, levels, b, kind != 2, 1 if kind == 2 else len(imgs)) if dump else None
    dp = ctypes.byref(d.st) if d else None
    L = lib()
    if kind == 2:
        rc = L.hmo_blend2(_ptr(imgs[0]), _ptr(imgs[1]), _ptr(mask), _ptr(out), w, h, c, levels, b, dp)
    elif kind == 3:
        rc = L.hmo_blend3(_ptr(imgs[0]), _ptr(imgs[1]), _ptr(imgs[2]), _ptr(mask), _ptr(out), w, h, c, levels, b, dp)
    else:
        arr = (_P * len(imgs))(*[_ptr(i) for i in imgs])
        rc = L.hmo_blend_n(len(imgs), arr, _ptr(mask), _ptr(out), w, h, c, levels, b, dp)
    if rc != 0:
        raise RuntimeError(f"oracle blend failed ({rc})")
    return (out, d) if dump else out


# ---------------------------------------------------------------------------
# remap / copy
# ---------------------------------------------------------------------------
def remap_offset(src: np.ndarray, dst: np.ndarray, map_x, map_y, off_x, off_y, no_unmapped_write=False, roi=None):
    """src uint8 [B,H,W,C]; dst uint8 or float32 [B,H,W,C] modified in place."""
    b, sh, sw, sc = src.shape
    _, dh, dw, dc = dst.shape
    mx = np.ascontiguousarray(map_x, np.uint16)
    my = np.ascontiguousarray(map_y, np.uint16)
    rx, ry, rw, rh = roi if roi is not None else (0, 0, -1, -1)
    return lib().hmo_remap_offset_t(_ptr(src), _et(src), sw, sh, sc, _ptr(dst), _et(dst), dw, dh, dc,
                                    _ptr(mx), _ptr(my), mx.shape[1], mx.shape[0], off_x, off_y, rx, ry, rw, rh,
                                    int(no_unmapped_write), b)


def remap_dest_map(src, dst, map_x, map_y, off_x, off_y, image_index, dest_map):
    b, sh, sw, c = src.shape
    _, dh, dw, _ = dst.shape
    mx = np.ascontiguousarray(map_x, np.uint16)
    my = np.ascontiguousarray(map_y, np.uint16)
    dm = np.ascontiguousarray(dest_map, np.uint8)
    return lib().hmo_remap_dest_map_t(_ptr(src), _et(src), sw, sh, c, _ptr(dst), dw, dh, _ptr(mx), _ptr(my),
                                      mx.shape[1], mx.shape[0], off_x, off_y, image_index, _ptr(dm), b)


def remap_hard_seam_n(srcs, maps_x, maps_y, offsets, dest_map, dst):
    n = len(srcs)
    b, dh, dw, c = dst.shape
    keep = [np.ascontiguousarray(s) for s in srcs]
    mxs = [np.ascontiguousarray(m, np.uint16) for m in maps_x]
    mys = [np.ascontiguousarray(m, np.uint16) for m in maps_y]
    sp = (_P * n)(*[_ptr(s) for s in keep])
    sw = (_I * n)(*[s.shape[2] for s in keep])
    sh = (_I * n)(*[s.shape[1] for s in keep])
    mxp = (_P * n)(*[_ptr(m) for m in mxs])
    myp = (_P * n)(*[_ptr(m) for m in mys])
    ox = (_I * n)(*[o[0] for o in offsets])
    oy = (_I * n)(*[o[1] for o in offsets])
    mw = (_I * n)(*[m.shape[1] for m in mxs])
    mh = (_I * n)(*[m.shape[0] for m in mxs])
    dm = np.ascontiguousarray(dest_map, np.uint8)
    return lib().hmo_remap_hard_seam_n_t(n, sp, _et(dst), sw, sh, c, mxp, myp, ox, oy, mw, mh, _ptr(dm), _ptr(dst),
                                         dw, dh, b)


def copy_roi(src, region_w, region_h, sx, sy, dst, ox, oy):
    b, sh, sw, sc = src.shape
    _, dh, dw, dc = dst.shape
    return lib().hmo_copy_roi_t(_ptr(src), _et(src), sw, sh, sc, region_w, region_h, sx, sy,
                                _ptr(dst), _et(dst), dw, dh, dc, ox, oy, b)


# ---------------------------------------------------------------------------
# panos
# ---------------------------------------------------------------------------
def _pano2_cfg(rig, levels, minimize, channels, batch, src_w, src_h):
    cfg = hmo_pano2_cfg()
    keep = []
    cfg.batch, cfg.levels, cfg.channels, cfg.minimize_blend = batch, levels, channels, int(minimize)
    for i in range(2):
        mx = np.ascontiguousarray(rig.map_x[i], np.uint16)
        my = np.ascontiguousarray(rig.map_y[i], np.uint16)
        keep += [mx, my]
        cfg.src_w[i], cfg.src_h[i] = src_w[i], src_h[i]
        cfg.map_h[i], cfg.map_w[i] = mx.shape
        cfg.map_x[i], cfg.map_y[i] = _ptr(mx), _ptr(my)
        cfg.pos_x[i], cfg.pos_y[i] = int(rig.pos_x[i]), int(rig.pos_y[i])
    seam = np.ascontiguousarray(rig.seam, np.uint8)
    keep.append(seam)
    cfg.seam = _ptr(seam)
    cfg.seam_h, cfg.seam_w = seam.shape
    return cfg, keep


def pano2(rig, in1: np.ndarray, in2: np.ndarray, levels: int, minimize: bool = True, dump: bool = False):
    """CudaStitchPano::process on the CPU.  in1/in2: uint8 [B,H,W,C], or float32 for the
    float pipeline (CudaStitchPano<float4|float3, ...>; the canvas is then float32)."""
    b, h1, w1, c = in1.shape
    _, h2, w2, _ = in2.shape
    cfg, keep = _pano2_cfg(rig, levels, minimize, c, b, (w1, w2), (h1, h2))
    cfg.pipeline_f32 = int(in1.dtype == np.float32)
    cw, ch = _I(), _I()
    lib().hmo_pano2_canvas_size(ctypes.byref(cfg), ctypes.byref(cw), ctypes.byref(ch))
    canvas = np.zeros((b, ch.value, cw.value, c), in1.dtype)
    d = None
    if dump and levels > 0:
        bw = (rig.map_x[0].shape[1] - rig.pos_x[1] + 256) if minimize else cw.value
        d = Dump(2, bw, ch.value, c, levels, b, False, 1)
    rc = lib().hmo_pano2_process(ctypes.byref(cfg), _ptr(np.ascontiguousarray(in1)), _ptr(np.ascontiguousarray(in2)),
                                 _ptr(canvas), ctypes.byref(d.st) if d else None)
    if rc != 0:
        raise RuntimeError(f"oracle pano2 failed ({rc})")
    return (canvas, d) if dump else canvas


def _panoN_cfg(rig, levels, minimize, channels, batch, srcs):
    cfg = hmo_panoN_cfg()
    keep = []
    cfg.n, cfg.batch, cfg.levels, cfg.channels, cfg.minimize_blend = rig.n, batch, levels, channels, int(minimize)
    for i in range(rig.n):
        mx = np.ascontiguousarray(rig.map_x[i], np.uint16)
        my = np.ascontiguousarray(rig.map_y[i], np.uint16)
        keep += [mx, my]
        cfg.src_w[i], cfg.src_h[i] = srcs[i].shape[2], srcs[i].shape[1]
        cfg.map_h[i], cfg.map_w[i] = mx.shape
        cfg.map_x[i], cfg.map_y[i] = _ptr(mx), _ptr(my)
        cfg.pos_x[i], cfg.pos_y[i] = int(rig.pos_x[i]), int(rig.pos_y[i])
    seam = np.ascontiguousarray(rig.seam, np.uint8)
    keep.append(seam)
    cfg.seam = _ptr(seam)
    cfg.seam_h, cfg.seam_w = seam.shape
    return cfg, keep


def panoN_rois(rig, levels: int):
    srcs = [np.zeros((1, 1, 1, 4), np.uint8)] * rig.n
    cfg, keep = _panoN_cfg(rig, levels, True, 4, 1, srcs)
    wr = (_I * 4)()
    br = (_I * 4)()
    lib().hmo_panoN_rois(ctypes.byref(cfg), wr, br)
    return tuple(wr), tuple(br)


def panoN(rig, inputs: List[np.ndarray], levels: int, minimize: bool = True, dump: bool = False):
    b, _, _, c = inputs[0].shape
    srcs = [np.ascontiguousarray(x) for x in inputs]
    cfg, keep = _panoN_cfg(rig, levels, minimize, c, b, srcs)
    cfg.pipeline_f32 = int(srcs[0].dtype == np.float32)
    cw, ch = _I(), _I()
    lib().hmo_panoN_canvas_size(ctypes.byref(cfg), ctypes.byref(cw), ctypes.byref(ch))
    canvas = np.zeros((b, ch.value, cw.value, c), srcs[0].dtype)
    d = None
    if dump and levels > 0:
        wr, br = panoN_rois(rig, levels)
        bw, bh = (br[2], br[3]) if (minimize and br[2] > 0) else (cw.value, ch.value)
        d = Dump(rig.n, bw, bh, c, levels, b, True, rig.n)
    arr = (_P * rig.n)(*[_ptr(s) for s in srcs])
    rc = lib().hmo_panoN_process(ctypes.byref(cfg), arr, _ptr(canvas), ctypes.byref(d.st) if d else None)
    if rc != 0:
        raise RuntimeError(f"oracle panoN failed ({rc})")
    return (canvas, d) if dump else canvas


def pano3(rig, inputs: List[np.ndarray], levels: int, fused: bool = True, canvas: Optional[np.ndarray] = None,
          dump: bool = False):
    """CudaStitchPano3::process on the CPU.  `canvas` (uint8 [B,H,W,C]) is the
    prior canvas content (a 3-channel pipeline leaves unwritten pixels alone)."""
    b, _, _, c = inputs[0].shape
    srcs = [np.ascontiguousarray(x) for x in inputs]
    cfg = hmo_pano3_cfg()
    keep = []
    cfg.batch, cfg.levels, cfg.channels, cfg.fused = b, levels, c, int(fused)
    cfg.pipeline_f32 = int(srcs[0].dtype == np.float32)
    for i in range(3):
        mx = np.ascontiguousarray(rig.map_x[i], np.uint16)
        my = np.ascontiguousarray(rig.map_y[i], np.uint16)
        keep += [mx, my]
        cfg.src_w[i], cfg.src_h[i] = srcs[i].shape[2], srcs[i].shape[1]
        cfg.map_h[i], cfg.map_w[i] = mx.shape
        cfg.map_x[i], cfg.map_y[i] = _ptr(mx), _ptr(my)
        cfg.pos_x[i], cfg.pos_y[i] = int(rig.pos_x[i]), int(rig.pos_y[i])
    seam = np.ascontiguousarray(rig.seam, np.uint8)
    cfg.seam = _ptr(seam)
    cfg.seam_h, cfg.seam_w = seam.shape
    cw, ch = _I(), _I()
    lib().hmo_pano3_canvas_size(ctypes.byref(cfg), ctypes.byref(cw), ctypes.byref(ch))
    out = (np.zeros((b, ch.value, cw.value, c), srcs[0].dtype) if canvas is None
           else np.ascontiguousarray(canvas, srcs[0].dtype).copy())
    d = Dump(3, seam.shape[1], ch.value, c, levels, b, True, 3) if (dump and levels > 0) else None
    arr = (_P * 3)(*[_ptr(x) for x in srcs])
    rc = lib().hmo_pano3_process(ctypes.byref(cfg), arr, _ptr(out), ctypes.byref(d.st) if d else None)
    if rc != 0:
        raise RuntimeError(f"oracle pano3 failed ({rc})")
    return (out, d) if dump else out
