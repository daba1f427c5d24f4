"""cupano: hm-cupano's Python API (cupano/__init__.py of the reference) over
the gfx950 HIP library.

The class names, constructor arguments, process() signatures, status objects
and error behaviour of the reference's Python package, with every frame
stitched by libhm_stitch.so (the same library the C++ classes and the C ABI
use).  There is no CPU or PyTorch fallback: without the built library the
import fails.

    from cupano import ControlMasks, CudaStitchPano
    masks = ControlMasks("/path/to/rig")      # mapping_*.tif + seam_file.png
    pano = CudaStitchPano(1, 6, masks, quiet=True)
    canvas = pano.process(left, right)         # [B, H, W, C] uint8 CUDA tensors
"""
from .geometry import CanvasInfo, Rect, SpatialTiff
from .masks import UNMAPPED_POSITION_VALUE, ControlMasks, ControlMasks3, ControlMasksN
from .pano import CudaStitchPano, CudaStitchPano3, CudaStitchPanoN
from .status import CudaStatus, CudaStatusError

__all__ = [
    "CanvasInfo",
    "ControlMasks",
    "ControlMasks3",
    "ControlMasksN",
    "CudaStatus",
    "CudaStatusError",
    "CudaStitchPano",
    "CudaStitchPano3",
    "CudaStitchPanoN",
    "Rect",
    "SpatialTiff",
    "UNMAPPED_POSITION_VALUE",
]
