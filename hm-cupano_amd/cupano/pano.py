"""CudaStitchPano / CudaStitchPanoN (the reference's cupano/pano.py:127-551,
553-912) and CudaStitchPano3 (src/pano/cudaPano3.h) on the gfx950 library.

Interface as the reference's Python package: the constructor arguments,
status (a CudaStatus property), canvas_width() / canvas_height() /
batch_size(), process() taking [B, H, W, C] or [H, W, C] CUDA tensors and an
optional canvas and returning the canvas, and dump_soft_blend_pyramid().

What differs, by design:
- uint8 inputs run the library's uint8 pipeline (CudaStitchPano<uchar4 |
  uchar3, float4 | float3>).  Inputs of any other real dtype run its float
  pipeline (<float4 | float3, float4 | float3>) on float32 copies, so
  fractional values stay fractional, and the canvas comes back in the first
  input's dtype as cupano/ops.py:52-57 casts it: `.to(dtype)` for a floating
  dtype, clamped to 0..255 first for an integer one.
- `backend` and `enable_cuda_graphs` are accepted and not used: there is one
  backend, and graph replay of many rigs is hm_stitch.Server's job.
- Work is issued on the current torch stream; as in the reference, nothing
  synchronises.
"""
from __future__ import annotations

from typing import Iterable

import numpy as np

from ._native import hm_stitch
from .masks import ControlMasks, ControlMasks3, ControlMasksN
from .status import CudaStatus, CudaStatusError, raise_if_error


def _torch():
    import torch

    return torch


def ensure_batched(image):
    """cupano/ops.py:38-43: [H, W, C] -> [1, H, W, C]; other ranks raise."""
    if image.ndim == 3:
        return image.unsqueeze(0)
    if image.ndim != 4:
        raise ValueError(f"Expected HWC or BHWC tensor, got shape {tuple(image.shape)}")
    return image


def cast_like(value, dtype):
    """cupano/ops.py:52-57: the canvas in the input's dtype."""
    if value.dtype == dtype:
        return value
    if dtype.is_floating_point:
        return value.to(dtype)
    return value.clamp(0, 255).to(dtype)


def _packed(t) -> bool:
    """Rows of whole pixels, one image after another (the surface layout the library reads)."""
    return t.stride(3) == 1 and t.stride(2) == t.shape[3] and (t.shape[0] == 1 or t.stride(0) == t.stride(1) * t.shape[1])


def _float_pipeline(tensors) -> bool:
    """uint8 inputs stitch in the uint8 pipeline, anything else in the float one."""
    torch = _torch()
    for t in tensors:
        if not t.is_cuda:
            raise ValueError("inputs must be CUDA tensors")
        if t.is_complex() or t.dtype == torch.bool:
            raise ValueError(f"unsupported input dtype {t.dtype}")
    return any(t.dtype != torch.uint8 for t in tensors)


def _pipeline_input(t, f32: bool):
    """The tensor the library reads: uint8 (uint8 pipeline) or float32, packed."""
    torch = _torch()
    if f32 and t.dtype != torch.float32:
        t = t.to(torch.float32)
    if not _packed(t):
        t = t.contiguous()
    return t


class _Rig:
    """The rig record hm_stitch's classes read: maps, integer-truncated placements, seam."""

    def __init__(self, map_x, map_y, positions, seam):
        self.map_x = [np.ascontiguousarray(m, dtype=np.uint16) for m in map_x]
        self.map_y = [np.ascontiguousarray(m, dtype=np.uint16) for m in map_y]
        self.pos_x = [float(p.xpos) for p in positions]
        self.pos_y = [float(p.ypos) for p in positions]
        self.seam = np.ascontiguousarray(seam, dtype=np.uint8)


class _PanoBase:
    _kind = ""

    def _init_common(self, batch_size: int, num_levels: int) -> None:
        self._status = CudaStatus()
        self._batch = int(batch_size)
        self._num_levels = int(num_levels)
        self._objs: dict = {}  # (channels, float pipeline) -> hm_stitch object
        self._last: dict = {}  # channels -> the object that processed last
        self._canvas_wh = (0, 0)

    @property
    def status(self) -> CudaStatus:
        return self._status

    def canvas_width(self) -> int:
        return self._canvas_wh[0]

    def canvas_height(self) -> int:
        return self._canvas_wh[1]

    def batch_size(self) -> int:
        return self._batch

    def _make(self, channels: int, f32: bool):  # pragma: no cover - per class
        raise NotImplementedError

    def _obj(self, channels: int, f32: bool = False):
        obj = self._objs.get((channels, f32))
        if obj is None:
            try:
                obj = self._make(channels, f32)
            except hm_stitch.HmError as e:
                self._status = CudaStatus(e.code, f"{self._kind} construction failed (hipError {e.code})")
                raise_if_error(self._status)
            if obj.status != 0:
                self._status = CudaStatus(obj.status, f"{self._kind} construction failed (hipError {obj.status})")
                raise_if_error(self._status)
            self._objs[(channels, f32)] = obj
            self._canvas_wh = obj.canvas_size()
        return obj

    def _first_object(self, quiet: bool, label: str) -> None:
        try:
            self._obj(4)
        except CudaStatusError:
            return
        if not quiet:
            print(f"{label} canvas size: {self.canvas_width()} x {self.canvas_height()}")

    def _canvas_for(self, canvas, like, channels: int, clear: bool):
        torch = _torch()
        shape = (self._batch, self.canvas_height(), self.canvas_width(), channels)
        if canvas is None:
            alloc = torch.zeros if clear else torch.empty
            return alloc(shape, dtype=like.dtype, device=like.device)
        canvas = ensure_batched(canvas)
        if tuple(canvas.shape) != shape:
            raise ValueError(f"canvas shape {tuple(canvas.shape)} != {shape}")
        return canvas

    def _run(self, obj, f32: bool, call, canvas, keep_prior: bool):
        """Runs `call(pipeline_canvas)` into `canvas`, through a stand-in of the
        pipeline's dtype when the canvas has another dtype or layout."""
        torch = _torch()
        if not canvas.is_cuda:
            raise ValueError("canvas must be a CUDA tensor")
        dt = torch.float32 if f32 else torch.uint8
        if canvas.dtype == dt and _packed(canvas):
            target = canvas
        elif keep_prior:
            target = cast_like(canvas, dt).contiguous()
        else:
            target = torch.empty(canvas.shape, dtype=dt, device=canvas.device)
        rc = call(target)
        if rc != 0:
            self._on_error(obj, rc)
        if target is not canvas:
            canvas.copy_(cast_like(target, canvas.dtype))
        obj._processed = True
        self._last[canvas.shape[-1]] = obj
        return canvas

    def _on_error(self, obj, rc: int) -> None:
        st = CudaStatus(rc, f"{self._kind}.process failed (hipError {rc})")
        if obj.status != 0:  # the library keeps this class's errors (sticky, like the reference)
            self._status = st
        raise CudaStatusError(st)

    def dump_soft_blend_pyramid(self, directory, *, device=None, channels: int | None = None) -> None:
        """metadata.txt and one float TIFF per level of the last processed frame
        (cupano/pano.py:466-551 layout; hmStitchPanoDumpPyramid)."""
        raise_if_error(self._status)
        if self._num_levels <= 0:
            raise ValueError("Soft-blend pyramid is only available when num_levels > 0")
        if channels is None:
            used = list(self._last)
            if len(used) != 1:
                raise ValueError("device and channels must be provided when multiple scratch buffers exist"
                                 if used else "No Laplacian workspace is populated yet")
            channels = used[0]
        obj = self._last.get(channels)
        if obj is None:
            raise ValueError(f"No soft scratch buffers available for channels={channels}")
        obj.dump_soft_blend_pyramid(str(directory))


class CudaStitchPano(_PanoBase):
    """Two images (cupano/pano.py:127-551)."""

    _kind = "CudaStitchPano"

    def __init__(self, batch_size: int, num_levels: int, control_masks: ControlMasks, quiet: bool = False,
                 minimize_blend: bool = True, max_output_width: int = 0, backend: str = "auto",
                 enable_cuda_graphs: bool = True) -> None:
        del max_output_width, backend, enable_cuda_graphs
        self._init_common(batch_size, num_levels)
        self._minimize_blend = bool(minimize_blend and num_levels > 0)
        if not control_masks.is_valid():
            self._status = CudaStatus(1, "Stitching masks were not able to be loaded")
            return
        m = control_masks
        self._rig = _Rig([m.img1_col, m.img2_col], [m.img1_row, m.img2_row], m.positions, m.whole_seam_mask_image)
        self._first_object(quiet, "Stitched")

    def _make(self, channels: int, f32: bool):
        return hm_stitch.StitchPano(self._batch, self._num_levels, self._rig, self._minimize_blend, channels,
                                    float_pipeline=f32)

    def process(self, input_image_1, input_image_2, canvas=None):
        raise_if_error(self._status)
        a, b = ensure_batched(input_image_1), ensure_batched(input_image_2)
        if a.shape[0] != self._batch or b.shape[0] != self._batch:
            raise ValueError("Input batch size does not match stitcher batch size")
        c = a.shape[-1]
        if c not in (3, 4):
            raise ValueError("Only 3- and 4-channel inputs are supported")
        if b.shape[-1] != c:
            raise ValueError("Mismatched channel counts")
        f32 = _float_pipeline((a, b))
        obj = self._obj(c, f32)
        out = self._canvas_for(canvas, a, c, clear=False)  # every canvas pixel is written
        ua, ub = _pipeline_input(a, f32), _pipeline_input(b, f32)
        return self._run(obj, f32, lambda t: obj.process(ua, ub, t), out, keep_prior=False)


class CudaStitchPanoN(_PanoBase):
    """N images, 2..8 (cupano/pano.py:553-912)."""

    _kind = "CudaStitchPanoN"

    def __init__(self, batch_size: int, num_levels: int, control_masks: ControlMasksN, quiet: bool = False,
                 minimize_blend: bool = True, backend: str = "auto", enable_cuda_graphs: bool = True) -> None:
        del backend, enable_cuda_graphs
        self._init_common(batch_size, num_levels)
        self._minimize_blend = bool(minimize_blend and num_levels > 0)
        if not control_masks.is_valid():
            self._status = CudaStatus(1, "Stitching masks (N-image) could not be loaded")
            return
        n = len(control_masks.img_col)
        if n < 2 or n > 8:
            self._status = CudaStatus(2, "Unsupported N for blend (supported 2..8)")
            return
        self._n = n
        m = control_masks
        self._rig = _Rig(m.img_col, m.img_row, m.positions, m.whole_seam_mask_indexed)
        self._first_object(quiet, "Stitched (N-image)")

    def _make(self, channels: int, f32: bool):
        return hm_stitch.StitchPanoN(self._batch, self._num_levels, self._rig, self._minimize_blend, channels,
                                     float_pipeline=f32)

    def process(self, inputs: Iterable, canvas=None):
        raise_if_error(self._status)
        ins = [ensure_batched(t) for t in inputs]
        if len(ins) != self._n:
            raise ValueError("inputs size != N")
        batch, c = ins[0].shape[0], ins[0].shape[-1]
        if any(t.shape[0] != batch for t in ins):
            raise ValueError("Mismatched batch sizes")
        if batch != self._batch:
            raise ValueError("Input batch size does not match stitcher batch size")
        if any(t.shape[-1] != c for t in ins):
            raise ValueError("Mismatched channel counts")
        if c not in (3, 4):
            raise ValueError("Only 3- and 4-channel inputs are supported")
        f32 = _float_pipeline(ins)
        obj = self._obj(c, f32)
        out = self._canvas_for(canvas, ins[0], c, clear=False)
        us = [_pipeline_input(t, f32) for t in ins]
        return self._run(obj, f32, lambda t: obj.process(us, t), out, keep_prior=False)


class CudaStitchPano3(_PanoBase):
    """Three images (src/pano/cudaPano3.h); no Python class of the reference has this form.

    process(..., fused=True) selects the fused path (process_optimized) or the
    remap + copy path; pixels the stitch does not cover keep the canvas's
    prior values (zero for a canvas allocated here)."""

    _kind = "CudaStitchPano3"

    def __init__(self, batch_size: int, num_levels: int, control_masks: ControlMasks3, quiet: bool = False) -> None:
        self._init_common(batch_size, num_levels)
        if not control_masks.is_valid():
            self._status = CudaStatus(1, "Stitching masks were not able to be loaded")
            return
        m = control_masks
        self._rig = _Rig([m.img0_col, m.img1_col, m.img2_col], [m.img0_row, m.img1_row, m.img2_row], m.positions,
                         m.whole_seam_mask_image)
        self._first_object(quiet, "Stitched (3-image)")

    def _make(self, channels: int, f32: bool):
        return hm_stitch.StitchPano3(self._batch, self._num_levels, self._rig, channels, float_pipeline=f32)

    def process(self, input_image_0, input_image_1, input_image_2, canvas=None, fused: bool = True):
        raise_if_error(self._status)
        ins = [ensure_batched(t) for t in (input_image_0, input_image_1, input_image_2)]
        if any(t.shape[0] != self._batch for t in ins):
            raise ValueError("Input batch size does not match stitcher batch size")
        c = ins[0].shape[-1]
        if c not in (3, 4):
            raise ValueError("Only 3- and 4-channel inputs are supported")
        if any(t.shape[-1] != c for t in ins):
            raise ValueError("Mismatched channel counts")
        f32 = _float_pipeline(ins)
        obj = self._obj(c, f32)
        out = self._canvas_for(canvas, ins[0], c, clear=True)
        us = [_pipeline_input(t, f32) for t in ins]
        return self._run(obj, f32, lambda t: obj.process(us[0], us[1], us[2], t, fused=fused), out, keep_prior=True)
