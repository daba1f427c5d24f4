"""Python binding of libhm_stitch.so (the C ABI declared in include/hm_stitch.h).

A thin ctypes layer for tests, the benchmark and Python callers.  Device
memory and streams come from torch (plain plumbing): pass torch uint8/float32
CUDA tensors shaped [B, H, W, C] (rows may be padded: the row stride becomes
the surface pitch) and the current torch stream is used unless one is given.

Nothing here falls back to a CPU path: if the shared library is missing or a
call fails, a RuntimeError is raised.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhm_stitch.so")

HM_PIX_U8C3, HM_PIX_U8C4, HM_PIX_F32C3, HM_PIX_F32C4 = 0, 1, 2, 3
HM_PIX_U8C1, HM_PIX_F32C1, HM_PIX_F16C1, HM_PIX_F16C3, HM_PIX_F16C4 = 4, 5, 6, 7, 8
HM_BLEND_TWO, HM_BLEND_THREE, HM_BLEND_N = 2, 3, 0
HM_LEVEL_GAUSS, HM_LEVEL_MASK, HM_LEVEL_LAPLACIAN, HM_LEVEL_BLEND, HM_LEVEL_RECONSTRUCT = range(5)


class HmError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with hipError {code}")
        self.code = code


class hmSurface(ctypes.Structure):
    _fields_ = [("d_ptr", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("pitch", ctypes.c_uint32)]


class hmControlMasks2(ctypes.Structure):
    _fields_ = [
        ("img1_col", ctypes.c_void_p), ("img1_row", ctypes.c_void_p), ("img1_w", ctypes.c_int), ("img1_h", ctypes.c_int),
        ("img2_col", ctypes.c_void_p), ("img2_row", ctypes.c_void_p), ("img2_w", ctypes.c_int), ("img2_h", ctypes.c_int),
        ("whole_seam_mask_image", ctypes.c_void_p), ("seam_w", ctypes.c_int), ("seam_h", ctypes.c_int),
        ("xpos", ctypes.c_float * 2), ("ypos", ctypes.c_float * 2),
    ]


HM_MAX_IMAGES = 8


class hmControlMasksN(ctypes.Structure):
    _fields_ = [
        ("n_images", ctypes.c_int),
        ("img_col", ctypes.c_void_p * HM_MAX_IMAGES), ("img_row", ctypes.c_void_p * HM_MAX_IMAGES),
        ("img_w", ctypes.c_int * HM_MAX_IMAGES), ("img_h", ctypes.c_int * HM_MAX_IMAGES),
        ("whole_seam_mask_indexed", ctypes.c_void_p), ("seam_w", ctypes.c_int), ("seam_h", ctypes.c_int),
        ("xpos", ctypes.c_float * HM_MAX_IMAGES), ("ypos", ctypes.c_float * HM_MAX_IMAGES),
    ]


class hmControlMasks3(ctypes.Structure):
    _fields_ = [
        ("img_col", ctypes.c_void_p * 3), ("img_row", ctypes.c_void_p * 3),
        ("img_w", ctypes.c_int * 3), ("img_h", ctypes.c_int * 3),
        ("whole_seam_mask_image", ctypes.c_void_p), ("seam_w", ctypes.c_int), ("seam_h", ctypes.c_int),
        ("xpos", ctypes.c_float * 3), ("ypos", ctypes.c_float * 3),
    ]


class hmServerResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("rig_id", ctypes.c_int), ("status", ctypes.c_int),
                ("t_submit_ns", ctypes.c_int64), ("t_start_ns", ctypes.c_int64), ("t_done_ns", ctypes.c_int64)]


class hmServerStats(ctypes.Structure):
    _fields_ = [("n_rigs", ctypes.c_int), ("device_budget_bytes", ctypes.c_size_t),
                ("device_used_bytes", ctypes.c_size_t), ("queued", ctypes.c_int), ("in_flight", ctypes.c_int),
                ("graphs_captured", ctypes.c_long), ("graph_launches", ctypes.c_long),
                ("eager_launches", ctypes.c_long), ("batched_passes", ctypes.c_long),
                ("batched_frames", ctypes.c_long), ("pass_graphs_captured", ctypes.c_long),
                ("pass_graph_launches", ctypes.c_long)]


class hmServerRequest(ctypes.Structure):
    _fields_ = [("rig_id", ctypes.c_int), ("inputs", ctypes.POINTER(hmSurface)), ("num_inputs", ctypes.c_int),
                ("canvas", hmSurface), ("fused", ctypes.c_int), ("tag", ctypes.c_uint64)]


class hmBatchCall(ctypes.Structure):
    _fields_ = [("pano", ctypes.c_void_p), ("inputs", ctypes.POINTER(hmSurface)), ("num_inputs", ctypes.c_int),
                ("canvas", hmSurface), ("fused", ctypes.c_int)]


class hmTuning(ctypes.Structure):
    _fields_ = [("remap_px", ctypes.c_int), ("tail_base", ctypes.c_int), ("overlap", ctypes.c_int),
                ("remap_xcd", ctypes.c_int), ("remap_rows", ctypes.c_int), ("canvas_grid", ctypes.c_int),
                ("fuse_down", ctypes.c_int), ("fuse_up", ctypes.c_int)]


class hmControlFilesInfo(ctypes.Structure):
    _fields_ = [
        ("n_images", ctypes.c_int),
        ("map_w", ctypes.c_uint32 * HM_MAX_IMAGES), ("map_h", ctypes.c_uint32 * HM_MAX_IMAGES),
        ("map_x", ctypes.c_void_p * HM_MAX_IMAGES), ("map_y", ctypes.c_void_p * HM_MAX_IMAGES),
        ("xpos", ctypes.c_float * HM_MAX_IMAGES), ("ypos", ctypes.c_float * HM_MAX_IMAGES),
        ("position_files_found", ctypes.c_uint32),
        ("seam_w", ctypes.c_uint32), ("seam_h", ctypes.c_uint32), ("seam", ctypes.c_void_p),
        ("seam_paletted", ctypes.c_int),
    ]


class hmRaster(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("channels", ctypes.c_uint32), ("bits", ctypes.c_uint32), ("sample_format", ctypes.c_uint32),
                ("bytes", ctypes.c_size_t)]


class hmTiffTags(ctypes.Structure):
    _fields_ = [("has_resolution", ctypes.c_int), ("has_xres", ctypes.c_int), ("has_yres", ctypes.c_int),
                ("has_xpos", ctypes.c_int), ("has_ypos", ctypes.c_int),
                ("xres", ctypes.c_float), ("yres", ctypes.c_float), ("xpos", ctypes.c_float), ("ypos", ctypes.c_float),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("channels", ctypes.c_uint32),
                ("bits", ctypes.c_uint32), ("sample_format", ctypes.c_uint32), ("compression", ctypes.c_uint32),
                ("predictor", ctypes.c_uint32), ("photometric", ctypes.c_uint32)]


class hmFrameLoopOptions(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("max_frames", ctypes.c_int64), ("upload", ctypes.c_int),
                ("download", ctypes.c_int)]


HM_TRANSFER_DMA, HM_TRANSFER_KERNEL = 0, 1
HM_TRANSFER_DEFAULT = HM_TRANSFER_DMA


class hmFrameLoopStats(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_int64), ("seconds", ctypes.c_double), ("upload_ms", ctypes.c_double),
                ("stitch_ms", ctypes.c_double), ("download_ms", ctypes.c_double),
                ("bytes_up_per_frame", ctypes.c_uint64), ("bytes_down_per_frame", ctypes.c_uint64),
                ("callback_error", ctypes.c_int), ("first_frame_seconds", ctypes.c_double),
                ("steady_seconds", ctypes.c_double), ("steady_frames", ctypes.c_int64)]


FRAME_SOURCE = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p),
                                ctypes.POINTER(ctypes.c_uint32))
FRAME_SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_uint32,
                              ctypes.c_uint32, ctypes.c_uint32)

_lib = None


def lib() -> ctypes.CDLL:
    """Load libhm_stitch.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    # torch ships its own HIP runtime (same soname as the system one).  Loading
    # it first makes the library bind to it; loaded the other way round, one
    # process would hold two HIP runtimes and torch's GPU initialisation fails.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i, u32, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_float
    S = hmSurface
    sig = {
        "hmRemapExOffset": (i, [S, i, S, i, vp, vp, i, i, i, i, i, i, vp]),
        "hmRemapExOffsetRoi": (i, [S, i, S, i, vp, vp, i, i, i, i, i, i, i, i, i, i, vp]),
        "hmRemapExOffsetWithDestMap": (i, [S, S, i, vp, vp, i, vp, i, i, i, i, i, vp]),
        "hmRemapHardSeamN": (i, [i, ctypes.POINTER(S), ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(i),
                                 ctypes.POINTER(i), vp, S, i, i, vp]),
        "hmCopyRoi": (i, [S, i, i, i, i, i, S, i, i, i, i, vp]),
        "hmMakeFull": (i, [S, i, i, i, i, i, i, i, i, S, i, vp]),
        "hmBlendContextCreate": (i, [i, i, i, i, i, i, i, i, ctypes.POINTER(vp)]),
        "hmBlendContextLevels": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]),
        "hmBlendRun": (i, [vp, ctypes.POINTER(vp), vp, vp, vp]),
        "hmBlendContextLevelPtr": (i, [vp, i, i, i, ctypes.POINTER(vp)]),
        "hmBlendContextDestroy": (None, [vp]),
        "hmBlendRunNoAlloc": (i, [i, i, i, i, i, i, i, ctypes.POINTER(vp), vp, vp, vp]),
        "hmStitchPanoCreate": (i, [i, i, ctypes.POINTER(hmControlMasks2), i, i, i, i, ctypes.POINTER(vp)]),
        "hmStitchPanoProcess": (i, [vp, S, S, S, vp]),
        "hmStitchPanoNCreate": (i, [i, i, ctypes.POINTER(hmControlMasksN), i, i, i, ctypes.POINTER(vp)]),
        "hmStitchPanoNProcess": (i, [vp, ctypes.POINTER(S), i, S, vp]),
        "hmStitchPano3Create": (i, [i, i, ctypes.POINTER(hmControlMasks3), i, i, ctypes.POINTER(vp)]),
        "hmStitchPano3Process": (i, [vp, S, S, S, S, i, vp]),
        "hmStitchPanoPrepare": (i, [vp]),
        "hmStitchPanoDeviceBytes": (i, [vp, ctypes.POINTER(ctypes.c_size_t)]),
        "hmServerCreate": (i, [i, ctypes.c_size_t, i, ctypes.POINTER(vp)]),
        "hmServerAddRig": (i, [vp, vp, ctypes.POINTER(i)]),
        "hmServerRemoveRig": (i, [vp, i]),
        "hmServerSubmit": (i, [vp, i, ctypes.POINTER(S), i, S, i, ctypes.c_uint64]),
        "hmServerPoll": (i, [vp, ctypes.POINTER(hmServerResult), i, ctypes.POINTER(i)]),
        "hmServerWaitIdle": (i, [vp]),
        "hmServerGetStats": (i, [vp, ctypes.POINTER(hmServerStats)]),
        "hmServerDestroy": (None, [vp]),
        "hmServerSetBatching": (i, [vp, i, i]),
        "hmServerSubmitMany": (i, [vp, ctypes.POINTER(hmServerRequest), i]),
        "hmBatchCreate": (i, [ctypes.POINTER(hmBatchCall), i, ctypes.POINTER(vp)]),
        "hmBatchSetGrid": (i, [vp, i]),
        "hmBatchGetInfo": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]),
        "hmBatchRun": (i, [vp, vp, i]),
        "hmBatchDestroy": (i, [vp]),
        "hmBatchAutotune": (i, [vp, vp, i, ctypes.POINTER(i), ctypes.POINTER(ctypes.c_float)]),
        "hmStitchPanoGetTuning": (i, [vp, ctypes.POINTER(hmTuning)]),
        "hmStitchPanoSetTuning": (i, [vp, ctypes.POINTER(hmTuning)]),
        "hmStitchPanoAutotune": (i, [vp, ctypes.POINTER(S), i, S, i, i, vp, ctypes.POINTER(hmTuning),
                                     ctypes.POINTER(ctypes.c_float)]),
        "hmStitchPanoStatus": (i, [vp]),
        "hmStitchPanoCanvasSize": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(i)]),
        "hmStitchPanoBatchSize": (i, [vp]),
        "hmStitchPanoDestroy": (None, [vp]),
        "hmStitchPanoSetDebug": (i, [vp, i]),
        "hmStitchPanoLevels": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]),
        "hmStitchPanoReadLevel": (i, [vp, i, i, i, vp, ctypes.c_size_t, vp]),
        "hmStitchPanoSetTiming": (i, [vp, i]),
        "hmStitchPanoTimingReport": (i, [vp, ctypes.c_char_p, ctypes.c_size_t, i]),
        "hmStitchPanoDumpPyramid": (i, [vp, ctypes.c_char_p, vp]),
        "hmFrameLoopRun": (i, [vp, i, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(hmFrameLoopOptions),
                               FRAME_SOURCE, FRAME_SINK, vp, vp, ctypes.POINTER(hmFrameLoopStats)]),
        "hmFilesLastError": (ctypes.c_char_p, []),
        "hmControlFilesLoad": (i, [ctypes.c_char_p, i, i, ctypes.POINTER(vp)]),
        "hmControlFilesGetInfo": (i, [vp, ctypes.POINTER(hmControlFilesInfo)]),
        "hmControlFilesDestroy": (None, [vp]),
        "hmTiffRead": (i, [ctypes.c_char_p, i, ctypes.POINTER(hmRaster)]),
        "hmTiffReadTags": (i, [ctypes.c_char_p, ctypes.POINTER(hmTiffTags)]),
        "hmTiffWrite": (i, [ctypes.c_char_p, ctypes.POINTER(hmRaster), ctypes.POINTER(hmTiffTags), i]),
        "hmPngReadSeam": (i, [ctypes.c_char_p, ctypes.POINTER(hmRaster), ctypes.POINTER(i)]),
        "hmRasterFree": (None, [ctypes.POINTER(hmRaster)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code: int, what: str) -> None:
    if code != 0:
        raise HmError(code, what)


# ---------------------------------------------------------------------------
# torch plumbing
# ---------------------------------------------------------------------------
def _torch():
    import torch

    return torch


def surface(t, channels: Optional[int] = None) -> hmSurface:
    """hmSurface of a [B, H, W, C] (or [H, W, C]) CUDA tensor; rows may be padded."""
    if t.dim() == 3:
        t = t.unsqueeze(0)
    if t.dim() != 4:
        raise ValueError("expected a [B, H, W, C] tensor")
    b, h, w, c = t.shape
    es = t.element_size()
    if t.stride(3) != 1 or t.stride(2) != c:
        raise ValueError("pixels must be packed (stride(2) == C, stride(3) == 1)")
    if b > 1 and t.stride(0) != t.stride(1) * h:
        raise ValueError("batch items must be contiguous in rows")
    return hmSurface(t.data_ptr(), w, h, t.stride(1) * es)


def current_stream() -> int:
    return _torch().cuda.current_stream().cuda_stream


def _stream(stream) -> int:
    if stream is None:
        return current_stream()
    return stream if isinstance(stream, int) else stream.cuda_stream


def _u16(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint16)


# ---------------------------------------------------------------------------
# pano classes
# ---------------------------------------------------------------------------
HM_PIPELINE_F32 = 0x100


class StitchPano:
    """CudaStitchPano<uchar4|uchar3, float4|float3> (src/pano/cudaPano.h:68-115), or with
    float_pipeline=True CudaStitchPano<float4|float3, float4|float3>: float inputs and canvas."""

    def __init__(self, batch_size: int, num_levels: int, rig, minimize_blend: bool = True, channels: int = 4,
                 debug: bool = False, float_pipeline: bool = False):
        L = lib()
        self._keep = [_u16(rig.map_x[0]), _u16(rig.map_y[0]), _u16(rig.map_x[1]), _u16(rig.map_y[1]),
                      np.ascontiguousarray(rig.seam, dtype=np.uint8)]
        m = hmControlMasks2()
        m.img1_col, m.img1_row = self._keep[0].ctypes.data, self._keep[1].ctypes.data
        m.img1_h, m.img1_w = self._keep[0].shape
        m.img2_col, m.img2_row = self._keep[2].ctypes.data, self._keep[3].ctypes.data
        m.img2_h, m.img2_w = self._keep[2].shape
        m.whole_seam_mask_image = self._keep[4].ctypes.data
        m.seam_h, m.seam_w = self._keep[4].shape
        m.xpos[0], m.xpos[1] = float(rig.pos_x[0]), float(rig.pos_x[1])
        m.ypos[0], m.ypos[1] = float(rig.pos_y[0]), float(rig.pos_y[1])
        self._masks = m
        h = ctypes.c_void_p()
        self.float_pipeline = float_pipeline
        rc = L.hmStitchPanoCreate(batch_size, num_levels, ctypes.byref(m), 1, int(minimize_blend), 0,
                                  channels | (HM_PIPELINE_F32 if float_pipeline else 0), ctypes.byref(h))
        if not h.value:
            raise HmError(rc, "hmStitchPanoCreate")
        self._h = h
        self.batch_size = batch_size
        self.channels = channels
        self.num_levels = num_levels
        if debug:
            check(L.hmStitchPanoSetDebug(h, 1), "hmStitchPanoSetDebug")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hmStitchPanoDestroy(h)
            self._h = None

    @property
    def status(self) -> int:
        return lib().hmStitchPanoStatus(self._h)

    def canvas_size(self):
        w, h = ctypes.c_int(), ctypes.c_int()
        check(lib().hmStitchPanoCanvasSize(self._h, ctypes.byref(w), ctypes.byref(h)), "hmStitchPanoCanvasSize")
        return w.value, h.value

    def alloc_canvas(self, device="cuda"):
        torch = _torch()
        w, h = self.canvas_size()
        dt = torch.float32 if getattr(self, "float_pipeline", False) else torch.uint8
        return torch.empty((self.batch_size, h, w, self.channels), dtype=dt, device=device)

    def process(self, in1, in2, canvas, stream=None) -> int:
        return lib().hmStitchPanoProcess(self._h, surface(in1), surface(in2), surface(canvas), _stream(stream))

    def prepare(self) -> int:
        return lib().hmStitchPanoPrepare(self._h)

    def device_bytes(self) -> int:
        b = ctypes.c_size_t()
        check(lib().hmStitchPanoDeviceBytes(self._h, ctypes.byref(b)), "hmStitchPanoDeviceBytes")
        return b.value

    def get_tuning(self) -> dict:
        t = hmTuning()
        check(lib().hmStitchPanoGetTuning(self._h, ctypes.byref(t)), "hmStitchPanoGetTuning")
        return {k: getattr(t, k) for k, _ in hmTuning._fields_}

    def set_tuning(self, remap_px: int = 4, tail_base: int = 2, overlap: int = 1, remap_xcd: int = 0,
                   remap_rows: int = 1, canvas_grid: int = 0, fuse_down: int = 0, fuse_up: int = 1) -> None:
        t = hmTuning(remap_px, tail_base, overlap, remap_xcd, remap_rows, canvas_grid, fuse_down, fuse_up)
        check(lib().hmStitchPanoSetTuning(self._h, ctypes.byref(t)), "hmStitchPanoSetTuning")

    def autotune(self, inputs, canvas, frames: int = 10, fused: bool = True, stream=None):
        """Times every setting of the knobs that matter for this object on this frame and keeps
        the fastest; returns (tuning dict, ms per frame)."""
        arr = (hmSurface * len(inputs))(*[surface(t) for t in inputs])
        t = hmTuning()
        ms = ctypes.c_float()
        check(lib().hmStitchPanoAutotune(self._h, arr, len(inputs), surface(canvas), int(fused), frames,
                                         _stream(stream), ctypes.byref(t), ctypes.byref(ms)), "hmStitchPanoAutotune")
        return {k: getattr(t, k) for k, _ in hmTuning._fields_}, ms.value

    def levels(self):
        n = ctypes.c_int()
        ws = (ctypes.c_int * 16)()
        hs = (ctypes.c_int * 16)()
        check(lib().hmStitchPanoLevels(self._h, ctypes.byref(n), ws, hs), "hmStitchPanoLevels")
        return [(ws[i], hs[i]) for i in range(n.value)]

    def read_level(self, what: int, image: int, level: int, stream=None) -> np.ndarray:
        lv = self.levels()
        w, h = lv[level]
        if what == HM_LEVEL_MASK:
            shape = (h, w, 1 if self._mask_channels() == 1 else self._mask_channels())
        else:
            shape = (self.batch_size, h, w, self.channels)
        out = np.empty(shape, dtype=np.float32)
        check(lib().hmStitchPanoReadLevel(self._h, what, image, level, out.ctypes.data, out.nbytes, _stream(stream)),
              "hmStitchPanoReadLevel")
        return out

    def _mask_channels(self) -> int:
        return 1

    def dump_soft_blend_pyramid(self, directory: str, stream=None) -> None:
        """dump_soft_blend_pyramid (src/pano/cudaPano.inl:420-458): metadata.txt and
        per-level float TIFFs of the last processed frame (hmStitchPanoDumpPyramid)."""
        check(lib().hmStitchPanoDumpPyramid(self._h, os.fsencode(str(directory)), _stream(stream)),
              "hmStitchPanoDumpPyramid")

    def set_timing(self, on: bool) -> None:
        check(lib().hmStitchPanoSetTiming(self._h, int(on)), "hmStitchPanoSetTiming")

    def timing_report(self, reset: bool = True) -> dict:
        buf = ctypes.create_string_buffer(4096)
        check(lib().hmStitchPanoTimingReport(self._h, buf, 4096, int(reset)), "hmStitchPanoTimingReport")
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, cnt = line.split()
            out[name] = (float(ms), int(cnt))
        return out


class StitchPanoN(StitchPano):
    """CudaStitchPanoN<uchar4|uchar3, float4|float3> (src/pano/cudaPanoN.h:81-146).

    `rig` carries n maps, integer positions and the indexed seam (labels
    0..n-1)."""

    def __init__(self, batch_size: int, num_levels: int, rig, minimize_blend: bool = True, channels: int = 4,
                 debug: bool = False, float_pipeline: bool = False):
        L = lib()
        n = len(rig.map_x)
        m = hmControlMasksN()
        m.n_images = n
        self._keep = []
        for k in range(n):
            mx, my = _u16(rig.map_x[k]), _u16(rig.map_y[k])
            self._keep += [mx, my]
            m.img_col[k], m.img_row[k] = mx.ctypes.data, my.ctypes.data
            m.img_h[k], m.img_w[k] = mx.shape
            m.xpos[k], m.ypos[k] = float(rig.pos_x[k]), float(rig.pos_y[k])
        seam = np.ascontiguousarray(rig.seam, dtype=np.uint8)
        self._keep.append(seam)
        m.whole_seam_mask_indexed = seam.ctypes.data
        m.seam_h, m.seam_w = seam.shape
        self._masks = m
        self.n_images = n
        h = ctypes.c_void_p()
        self.float_pipeline = float_pipeline
        rc = L.hmStitchPanoNCreate(batch_size, num_levels, ctypes.byref(m), int(minimize_blend), 1,
                                   channels | (HM_PIPELINE_F32 if float_pipeline else 0), ctypes.byref(h))
        if not h.value:
            raise HmError(rc, "hmStitchPanoNCreate")
        self._h = h
        self.batch_size = batch_size
        self.channels = channels
        self.num_levels = num_levels
        if debug:
            check(L.hmStitchPanoSetDebug(h, 1), "hmStitchPanoSetDebug")

    def process(self, inputs, canvas, stream=None) -> int:
        arr = (hmSurface * len(inputs))(*[surface(t) for t in inputs])
        return lib().hmStitchPanoNProcess(self._h, arr, len(inputs), surface(canvas), _stream(stream))

    def _mask_channels(self) -> int:
        return self.n_images


class StitchPano3(StitchPano):
    """CudaStitchPano3<uchar4|uchar3, float4|float3> (src/pano/cudaPano3.h:64-120).

    `rig` carries three maps, integer positions and the seam (labels 0..2)."""

    def __init__(self, batch_size: int, num_levels: int, rig, channels: int = 4, debug: bool = False,
                 float_pipeline: bool = False):
        L = lib()
        m = hmControlMasks3()
        self._keep = []
        for k in range(3):
            mx, my = _u16(rig.map_x[k]), _u16(rig.map_y[k])
            self._keep += [mx, my]
            m.img_col[k], m.img_row[k] = mx.ctypes.data, my.ctypes.data
            m.img_h[k], m.img_w[k] = mx.shape
            m.xpos[k], m.ypos[k] = float(rig.pos_x[k]), float(rig.pos_y[k])
        seam = np.ascontiguousarray(rig.seam, dtype=np.uint8)
        self._keep.append(seam)
        m.whole_seam_mask_image = seam.ctypes.data
        m.seam_h, m.seam_w = seam.shape
        self._masks = m
        h = ctypes.c_void_p()
        self.float_pipeline = float_pipeline
        rc = L.hmStitchPano3Create(batch_size, num_levels, ctypes.byref(m), 1,
                                   channels | (HM_PIPELINE_F32 if float_pipeline else 0), ctypes.byref(h))
        if not h.value:
            raise HmError(rc, "hmStitchPano3Create")
        self._h = h
        self.batch_size = batch_size
        self.channels = channels
        self.num_levels = num_levels
        if debug:
            check(L.hmStitchPanoSetDebug(h, 1), "hmStitchPanoSetDebug")

    def process(self, in0, in1, in2, canvas, fused: bool = True, stream=None) -> int:
        return lib().hmStitchPano3Process(self._h, surface(in0), surface(in1), surface(in2), surface(canvas),
                                          int(fused), _stream(stream))

    def _mask_channels(self) -> int:
        return 3


# ---------------------------------------------------------------------------
# served form
# ---------------------------------------------------------------------------
class Server:
    """Many rigs on one GPU (hmServer*): admission by device bytes, a scheduler
    thread over `n_streams` streams, hipGraph replay per buffer set, and with
    `batch` >= 2 continuous batching of the ready requests into batched passes."""

    def __init__(self, n_streams: int = 4, device_budget_bytes: int = 64 << 30, use_graphs: bool = True,
                 batch: int = 0, same_structure: bool = False):
        h = ctypes.c_void_p()
        check(lib().hmServerCreate(n_streams, device_budget_bytes, int(use_graphs), ctypes.byref(h)), "hmServerCreate")
        self._h = h
        self._keep = {}  # rig id -> pano (the caller's objects must outlive their requests)
        if batch:
            self.set_batching(batch, same_structure)

    def set_batching(self, max_frames: int, same_structure: bool = False) -> None:
        """Continuous batching: ready requests of up to max_frames rigs run as one batched pass
        (same_structure: only rigs whose pipelines have the same kernel sequence share one)."""
        check(lib().hmServerSetBatching(self._h, max_frames, int(same_structure)), "hmServerSetBatching")

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hmServerDestroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def add_rig(self, pano) -> int:
        rid = ctypes.c_int()
        rc = lib().hmServerAddRig(self._h, pano._h, ctypes.byref(rid))
        if rc != 0:
            raise HmError(rc, "hmServerAddRig")
        self._keep[rid.value] = pano
        return rid.value

    def remove_rig(self, rig_id: int) -> None:
        check(lib().hmServerRemoveRig(self._h, rig_id), "hmServerRemoveRig")
        self._keep.pop(rig_id, None)

    def submit(self, rig_id: int, inputs, canvas, tag: int = 0, fused: bool = True) -> int:
        arr = (hmSurface * len(inputs))(*[surface(t) for t in inputs])
        return lib().hmServerSubmit(self._h, rig_id, arr, len(inputs), surface(canvas), int(fused), tag)

    def submit_many(self, requests) -> int:
        """Queues (rig_id, inputs, canvas, tag[, fused]) requests at once, all or nothing."""
        arr = (hmServerRequest * len(requests))()
        keep = []
        for q, req in zip(arr, requests):
            rig_id, inputs, canvas, tag = req[:4]
            ins = (hmSurface * len(inputs))(*[surface(t) for t in inputs])
            keep.append(ins)
            q.rig_id = rig_id
            q.inputs = ctypes.cast(ins, ctypes.POINTER(hmSurface))
            q.num_inputs = len(inputs)
            q.canvas = surface(canvas)
            q.fused = int(req[4]) if len(req) > 4 else 1
            q.tag = tag
        return lib().hmServerSubmitMany(self._h, arr, len(requests))

    def poll(self, max_results: int = 4096):
        buf = (hmServerResult * max_results)()
        n = ctypes.c_int()
        check(lib().hmServerPoll(self._h, buf, max_results, ctypes.byref(n)), "hmServerPoll")
        return [dict(tag=r.tag, rig=r.rig_id, status=r.status, t_submit_ns=r.t_submit_ns, t_start_ns=r.t_start_ns,
                     t_done_ns=r.t_done_ns) for r in buf[: n.value]]

    def wait_idle(self) -> int:
        return lib().hmServerWaitIdle(self._h)

    def stats(self) -> dict:
        st = hmServerStats()
        check(lib().hmServerGetStats(self._h, ctypes.byref(st)), "hmServerGetStats")
        return {k: getattr(st, k) for k, _ in hmServerStats._fields_}


# ---------------------------------------------------------------------------
# batched pass over mixed configurations
# ---------------------------------------------------------------------------
class Batch:
    """One pass over several pano calls of mixed classes and sizes (hmBatch*).

    `calls` is a list of (pano, inputs, canvas) or (pano, inputs, canvas,
    fused); each pano appears once, and a pano with batch size B takes B frames
    per call.  The pass is built for these tensors: run() stitches whatever they
    hold when it runs.  Launch k of every call forms stage k; each kernel body
    of a stage is one persistent launch walking the tile work-list of every
    call's launch of that body."""

    def __init__(self, calls):
        self._keep = []
        arr = (hmBatchCall * len(calls))()
        for k, c in enumerate(calls):
            pano, inputs, canvas = c[0], c[1], c[2]
            fused = c[3] if len(c) > 3 else True
            surfs = (hmSurface * len(inputs))(*[surface(t) for t in inputs])
            self._keep.append((pano, inputs, canvas, surfs))
            arr[k].pano = pano._h
            arr[k].inputs = surfs
            arr[k].num_inputs = len(inputs)
            arr[k].canvas = surface(canvas)
            arr[k].fused = int(fused)
        h = ctypes.c_void_p()
        check(lib().hmBatchCreate(arr, len(calls), ctypes.byref(h)), "hmBatchCreate")
        self._h = h

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hmBatchDestroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def set_grid(self, workgroups: int) -> None:
        check(lib().hmBatchSetGrid(self._h, workgroups), "hmBatchSetGrid")

    def info(self) -> dict:
        st, ln, rec, gr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().hmBatchGetInfo(self._h, ctypes.byref(st), ctypes.byref(ln), ctypes.byref(rec), ctypes.byref(gr)),
              "hmBatchGetInfo")
        return dict(stages=st.value, launches=ln.value, recorded_launches=rec.value, grid=gr.value)

    def run(self, stream=None, use_graph: bool = False) -> int:
        return lib().hmBatchRun(self._h, _stream(stream), int(use_graph))

    def autotune(self, passes: int = 5, stream=None):
        """Keeps the fastest grid; returns (workgroups per launch, ms per pass)."""
        g, ms = ctypes.c_int(), ctypes.c_float()
        check(lib().hmBatchAutotune(self._h, _stream(stream), passes, ctypes.byref(g), ctypes.byref(ms)),
              "hmBatchAutotune")
        return g.value, ms.value


# ---------------------------------------------------------------------------
# kernel tier
# ---------------------------------------------------------------------------
class BlendContext:
    """CudaBatchLaplacianBlendContext{,3,N} + cudaBatchedLaplacianBlendWithContext{,3,N}."""

    def __init__(self, kind: int, n_images: int, width: int, height: int, num_levels: int, batch_size: int,
                 channels: int = 4, debug: bool = False):
        h = ctypes.c_void_p()
        check(lib().hmBlendContextCreate(kind, n_images, width, height, num_levels, batch_size, channels, int(debug),
                                         ctypes.byref(h)), "hmBlendContextCreate")
        self._h = h
        self.n = n_images
        self.kind = kind

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.hmBlendContextDestroy(h)
            self._h = None

    def levels(self):
        n = ctypes.c_int()
        ws = (ctypes.c_int * 16)()
        hs = (ctypes.c_int * 16)()
        check(lib().hmBlendContextLevels(self._h, ctypes.byref(n), ws, hs), "hmBlendContextLevels")
        return [(ws[i], hs[i]) for i in range(n.value)]

    def run(self, images: Sequence, mask, out, stream=None) -> None:
        arr = (ctypes.c_void_p * len(images))(*[t.data_ptr() for t in images])
        check(lib().hmBlendRun(self._h, arr, mask.data_ptr(), out.data_ptr(), _stream(stream)), "hmBlendRun")

    def level_ptr(self, what: int, image: int, level: int) -> int:
        p = ctypes.c_void_p()
        check(lib().hmBlendContextLevelPtr(self._h, what, image, level, ctypes.byref(p)), "hmBlendContextLevelPtr")
        return p.value or 0


NOALLOC_SCRATCH_BYTES = 128 << 20  # HM_NOALLOC_SCRATCH_BYTES


def blend_run_no_alloc(kind: int, n_images: int, width: int, height: int, num_levels: int, batch_size: int,
                       channels: int, images: Sequence, mask, out, stream=None) -> int:
    """hmBlendRunNoAlloc: the blend of BlendContext(...).run(...) with its pyramid in the
    library's static device scratch (no device allocation).  Returns the hipError_t code."""
    arr = (ctypes.c_void_p * len(images))(*[t.data_ptr() for t in images])
    return lib().hmBlendRunNoAlloc(kind, n_images, width, height, num_levels, batch_size, channels, arr,
                                   mask.data_ptr(), out.data_ptr(), _stream(stream))


def remap_ex_offset(src, src_type, dest, dest_type, map_x, map_y, batch, remap_w, remap_h, off_x, off_y,
                    no_unmapped_write=False, roi=None, stream=None) -> int:
    L = lib()
    if roi is None:
        return L.hmRemapExOffset(surface(src), src_type, surface(dest), dest_type, map_x.data_ptr(), map_y.data_ptr(),
                                 batch, remap_w, remap_h, off_x, off_y, int(no_unmapped_write), _stream(stream))
    return L.hmRemapExOffsetRoi(surface(src), src_type, surface(dest), dest_type, map_x.data_ptr(), map_y.data_ptr(),
                                batch, remap_w, remap_h, off_x, off_y, roi[0], roi[1], roi[2], roi[3],
                                int(no_unmapped_write), _stream(stream))


def remap_dest_map(src, dest, pix_type, map_x, map_y, image_index, dest_map, batch, remap_w, remap_h, off_x, off_y,
                   stream=None) -> int:
    return lib().hmRemapExOffsetWithDestMap(surface(src), surface(dest), pix_type, map_x.data_ptr(), map_y.data_ptr(),
                                            image_index, dest_map.data_ptr(), batch, remap_w, remap_h, off_x, off_y,
                                            _stream(stream))


def remap_hard_seam_n(inputs, maps_x, maps_y, offsets, sizes, dest_map, dest, pix_type, batch, stream=None) -> int:
    n = len(inputs)
    surfs = (hmSurface * n)(*[surface(t) for t in inputs])
    mx = (ctypes.c_void_p * n)(*[m.data_ptr() for m in maps_x])
    my = (ctypes.c_void_p * n)(*[m.data_ptr() for m in maps_y])
    off = (ctypes.c_int * (2 * n))(*[v for o in offsets for v in o])
    sz = (ctypes.c_int * (2 * n))(*[v for s in sizes for v in s])
    return lib().hmRemapHardSeamN(n, surfs, mx, my, off, sz, dest_map.data_ptr(), surface(dest), pix_type, batch,
                                  _stream(stream))


def copy_roi(src, src_type, region_w, region_h, sx, sy, dest, dest_type, ox, oy, batch, stream=None) -> int:
    return lib().hmCopyRoi(surface(src), src_type, region_w, region_h, sx, sy, surface(dest), dest_type, ox, oy, batch,
                           _stream(stream))


def make_full(src, src_type, region_w, region_h, sx, sy, ox, oy, batch, dest, dest_type, stream=None) -> int:
    return lib().hmMakeFull(surface(src), src_type, region_w, region_h, sx, sy, ox, oy, batch, surface(dest),
                            dest_type, _stream(stream))


def exported_symbols() -> List[str]:
    """Names declared in include/*.h (for the export test)."""
    import glob
    import re

    names = set()
    for hdr in glob.glob(os.path.join(os.path.dirname(_HERE), "include", "*.h")):
        text = open(hdr).read()
        names.update(re.findall(r"^(?:int|void|const char\*)\s+(hm\w+)\s*\(", text, flags=re.M))
    return sorted(names)


# ---------------------------------------------------------------------------
# control directories and raster files (include/hm_files.h)
# ---------------------------------------------------------------------------
HM_FILE_OK, HM_FILE_NOT_FOUND, HM_FILE_BAD_FORMAT, HM_FILE_UNSUPPORTED, HM_FILE_WRONG_TYPE = range(5)
HM_FILE_SEAM_CLASSES, HM_FILE_MISMATCH, HM_FILE_WRITE_FAILED, HM_FILE_INVALID_ARGUMENT = range(5, 9)
HM_CONTROL_TWO, HM_CONTROL_THREE, HM_CONTROL_N = 2, 3, 0
HM_SAMPLE_UINT, HM_SAMPLE_INT, HM_SAMPLE_FLOAT = 1, 2, 3


class FileError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


def _fcheck(code: int) -> None:
    if code != HM_FILE_OK:
        raise FileError(code, lib().hmFilesLastError().decode(errors="replace"))


class ControlRig:
    """A rig loaded from a control directory, in the form the pano classes
    take: map_x/map_y (uint16 [h, w] per image), pos_x/pos_y (the float
    placements, truncated by the pano classes), seam (uint8 [h, w])."""

    def __init__(self, info: hmControlFilesInfo):
        n = info.n_images
        self.n_images = n
        self.map_x, self.map_y = [], []
        for k in range(n):
            h, w = info.map_h[k], info.map_w[k]
            for dst, ptr in ((self.map_x, info.map_x[k]), (self.map_y, info.map_y[k])):
                buf = (ctypes.c_uint16 * (w * h)).from_address(ptr)
                dst.append(np.frombuffer(buf, dtype=np.uint16).reshape(h, w).copy())
        self.xpos = [float(info.xpos[k]) for k in range(n)]
        self.ypos = [float(info.ypos[k]) for k in range(n)]
        self.pos_x, self.pos_y = list(self.xpos), list(self.ypos)
        self.position_files_found = int(info.position_files_found)
        sh, sw = info.seam_h, info.seam_w
        buf = (ctypes.c_uint8 * (sw * sh)).from_address(info.seam)
        self.seam = np.frombuffer(buf, dtype=np.uint8).reshape(sh, sw).copy()
        self.seam_paletted = bool(info.seam_paletted)


def load_control_dir(directory: str, kind: int, n_images: int = 0) -> ControlRig:
    """hmControlFilesLoad: the rig in `directory` under the 2-image (HM_CONTROL_TWO),
    3-image (HM_CONTROL_THREE) or N-image (HM_CONTROL_N, n_images) rules."""
    L = lib()
    h = ctypes.c_void_p()
    _fcheck(L.hmControlFilesLoad(os.fsencode(str(directory)), kind, n_images, ctypes.byref(h)))
    try:
        info = hmControlFilesInfo()
        _fcheck(L.hmControlFilesGetInfo(h, ctypes.byref(info)))
        return ControlRig(info)
    finally:
        L.hmControlFilesDestroy(h)


_NP_OF = {(HM_SAMPLE_UINT, 8): np.uint8, (HM_SAMPLE_UINT, 16): np.uint16, (HM_SAMPLE_UINT, 32): np.uint32,
          (HM_SAMPLE_INT, 8): np.int8, (HM_SAMPLE_INT, 16): np.int16, (HM_SAMPLE_INT, 32): np.int32,
          (HM_SAMPLE_FLOAT, 32): np.float32, (HM_SAMPLE_FLOAT, 64): np.float64}


def _raster_to_numpy(r: hmRaster) -> np.ndarray:
    shape = (r.height, r.width) if r.channels == 1 else (r.height, r.width, r.channels)
    try:
        dt = np.dtype(_NP_OF[(r.sample_format, r.bits)])
        buf = (ctypes.c_uint8 * r.bytes).from_address(r.data)
        a = np.frombuffer(buf, dtype=dt).copy()
    finally:
        lib().hmRasterFree(ctypes.byref(r))
    return a.reshape(shape)


def tiff_read(path: str, opencv_order: bool = False) -> np.ndarray:
    """The first image of a TIFF as [h, w] or [h, w, c] (hmTiffRead)."""
    r = hmRaster()
    _fcheck(lib().hmTiffRead(os.fsencode(str(path)), int(opencv_order), ctypes.byref(r)))
    return _raster_to_numpy(r)


def tiff_read_tags(path: str) -> dict:
    t = hmTiffTags()
    _fcheck(lib().hmTiffReadTags(os.fsencode(str(path)), ctypes.byref(t)))
    return {name: getattr(t, name) for name, _ in hmTiffTags._fields_}


def tiff_write(path: str, array: np.ndarray, opencv_order: bool = False, xres=None, yres=None, xpos=None,
               ypos=None) -> None:
    """Writes [h, w] or [h, w, c] (uint8/16/32, int8/16/32, float32/64) as an uncompressed TIFF."""
    a = np.ascontiguousarray(array)
    fmt = {"u": HM_SAMPLE_UINT, "i": HM_SAMPLE_INT, "f": HM_SAMPLE_FLOAT}[a.dtype.kind]
    r = hmRaster()
    r.data = a.ctypes.data
    r.height, r.width = a.shape[:2]
    r.channels = 1 if a.ndim == 2 else a.shape[2]
    r.bits = a.dtype.itemsize * 8
    r.sample_format = fmt
    r.bytes = a.nbytes
    t = hmTiffTags()
    for name, v in (("xres", xres), ("yres", yres), ("xpos", xpos), ("ypos", ypos)):
        if v is not None:
            setattr(t, "has_" + name, 1)
            setattr(t, name, float(v))
    _fcheck(lib().hmTiffWrite(os.fsencode(str(path)), ctypes.byref(r), ctypes.byref(t), int(opencv_order)))


def png_read_seam(path: str):
    """(uint8 [h, w], paletted) of a seam PNG (hmPngReadSeam)."""
    r = hmRaster()
    pal = ctypes.c_int()
    _fcheck(lib().hmPngReadSeam(os.fsencode(str(path)), ctypes.byref(r), ctypes.byref(pal)))
    return _raster_to_numpy(r), bool(pal.value)


# ---------------------------------------------------------------------------
# frame loop (hmFrameLoopRun)
# ---------------------------------------------------------------------------
def frame_loop(pano, input_sizes, source, sink, slots: int = 3, max_frames: int = -1, stream=None,
               upload: int = HM_TRANSFER_DEFAULT, download: int = HM_TRANSFER_DEFAULT) -> dict:
    """Streams host frames through `pano` (hmFrameLoopRun).

    input_sizes: [(w, h)] per input.  source(frame) returns one uint8 array
    [B, h, w, 3] (or [h, w, 3]) of BGR pixels per input, None at the end, or
    True to resend the slot's previous contents (no host copy: timing runs).
    sink(frame, bgr) receives the stitched [B * canvas_h, canvas_w, 3] BGR
    frame (a view of pinned memory, valid during the call).  Returns the
    loop's statistics; exceptions raised by the callbacks are re-raised."""
    L = lib()
    n = len(input_sizes)
    ws = (ctypes.c_uint32 * n)(*[int(w) for w, _ in input_sizes])
    hs = (ctypes.c_uint32 * n)(*[int(h) for _, h in input_sizes])
    b = pano.batch_size
    errors = []

    def _src(_user, frame, ptrs, pitches):
        try:
            frames = source(int(frame))
            if frames is None:
                return 0
            if frames is True:  # the pinned buffers keep their previous contents
                return 1
            if len(frames) != n:
                raise ValueError(f"source returned {len(frames)} frames for {n} inputs")
            for i, f in enumerate(frames):
                a = np.ascontiguousarray(f, dtype=np.uint8)
                want = b * hs[i] * ws[i] * 3
                if a.size != want:
                    raise ValueError(f"input {i}: {a.shape} is not {b} x {hs[i]} x {ws[i]} x 3")
                ctypes.memmove(ptrs[i], a.ctypes.data, want)
            return 1
        except BaseException as e:  # noqa: BLE001 - reported after the loop
            errors.append(e)
            return -1

    def _sink(_user, frame, ptr, width, rows, pitch):
        try:
            buf = (ctypes.c_uint8 * (rows * pitch)).from_address(ptr)
            view = np.frombuffer(buf, dtype=np.uint8).reshape(rows, pitch)[:, : width * 3].reshape(rows, width, 3)
            return 1 if sink(int(frame), view) is False else 0
        except BaseException as e:  # noqa: BLE001
            errors.append(e)
            return -1

    src_cb, sink_cb = FRAME_SOURCE(_src), FRAME_SINK(_sink)
    st = hmFrameLoopStats()
    opt = hmFrameLoopOptions(slots, max_frames, upload, download)
    rc = L.hmFrameLoopRun(pano._h, n, ws, hs, ctypes.byref(opt), src_cb, sink_cb, None, _stream(stream),
                          ctypes.byref(st))
    if errors:
        raise errors[0]
    if rc != 0 and not (st.callback_error == 1 and rc == 999):  # the sink asked to stop
        raise HmError(rc, "hmFrameLoopRun")
    return {name: getattr(st, name) for name, _ in hmFrameLoopStats._fields_}
