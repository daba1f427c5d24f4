"""ControlMasks / ControlMasksN / ControlMasks3: a rig's remap tables, seam
and placements, in memory or loaded from a control directory.

The attributes and methods of the reference's cupano/masks.py (and, for the
3-image rig, of src/pano/controlMasks3.h).  load() reads the directory with
the library's loader (hm_stitch.load_control_dir: the file rules of
src/pano/controlMasks*.cpp); placements then get the Python package's
near-integer snap (cupano/masks.py:34-38, within 1e-3 of an integer).
Failures: a missing or unreadable remap table, or a seam whose label count
is wrong, make load() return False (the reason is in `last_error`); a
missing seam file raises FileNotFoundError and an undecodable one ValueError,
as the reference's readers do.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from ._native import hm_stitch
from .geometry import SpatialTiff

UNMAPPED_POSITION_VALUE = np.uint16(65535)


def _snap_near_integer(value: float, eps: float = 1e-3) -> float:
    r = round(value)
    return float(r) if abs(value - r) < eps else float(value)


def _load(directory, kind: int, n_images: int, holder):
    """The loaded rig, or None (load failed on a remap table or the label count)."""
    holder.last_error = None
    try:
        return hm_stitch.load_control_dir(str(directory), kind, n_images)
    except hm_stitch.FileError as e:
        holder.last_error = e
        if "seam_file.png" in e.message:
            if e.code == hm_stitch.HM_FILE_NOT_FOUND:
                raise FileNotFoundError(str(Path(directory) / "seam_file.png")) from e
            if e.code != hm_stitch.HM_FILE_SEAM_CLASSES:
                raise ValueError(e.message) from e
        return None


def _positions(rig) -> list:
    return [SpatialTiff(_snap_near_integer(x), _snap_near_integer(y)) for x, y in zip(rig.xpos, rig.ypos)]


def _empty_u16():
    return np.empty((0, 0), dtype=np.uint16)


class ControlMasks:
    """Two-image rig (cupano/masks.py:63-123); whole_seam_mask_image: 1 selects image 1."""

    def __init__(self, game_dir: str | None = None):
        self.img1_col = _empty_u16()
        self.img1_row = _empty_u16()
        self.img2_col = _empty_u16()
        self.img2_row = _empty_u16()
        self.whole_seam_mask_image = np.empty((0, 0), dtype=np.uint8)
        self.positions: list = []
        self.last_error = None
        if game_dir is not None:
            self.load(game_dir)

    def load(self, game_dir: str) -> bool:
        rig = _load(game_dir, hm_stitch.HM_CONTROL_TWO, 2, self)
        if rig is None:
            return False
        self.img1_col, self.img2_col = rig.map_x
        self.img1_row, self.img2_row = rig.map_y
        self.whole_seam_mask_image = rig.seam
        self.positions = _positions(rig)
        return self.is_valid()

    def is_valid(self) -> bool:
        return bool(self.img1_col.size and self.img1_row.size and self.img2_col.size and self.img2_row.size
                    and self.whole_seam_mask_image.size and len(self.positions) == 2)

    def canvas_width(self) -> int:
        return int(max(self.positions[0].xpos + self.img1_col.shape[1], self.positions[1].xpos + self.img2_col.shape[1]))

    def canvas_height(self) -> int:
        return int(max(self.positions[0].ypos + self.img1_col.shape[0], self.positions[1].ypos + self.img2_col.shape[0]))


class ControlMasksN:
    """N-image rig (cupano/masks.py:126-182); whole_seam_mask_indexed holds labels 0..N-1."""

    def __init__(self, directory: str | None = None, n_images: int | None = None):
        self.img_col: list = []
        self.img_row: list = []
        self.whole_seam_mask_indexed = np.empty((0, 0), dtype=np.uint8)
        self.positions: list = []
        self.last_error = None
        if directory is not None and n_images is not None:
            self.load(directory, n_images)

    def load(self, directory: str, n_images: int) -> bool:
        rig = _load(directory, hm_stitch.HM_CONTROL_N, n_images, self)
        if rig is None:
            return False
        self.img_col, self.img_row = list(rig.map_x), list(rig.map_y)
        self.whole_seam_mask_indexed = rig.seam
        self.positions = _positions(rig)
        return self.is_valid()

    def is_valid(self) -> bool:
        return (bool(self.img_col) and len(self.img_col) == len(self.img_row)
                and len(self.positions) == len(self.img_col) and self.whole_seam_mask_indexed.size > 0)

    def canvas_width(self) -> int:
        return int(max(p.xpos + m.shape[1] for p, m in zip(self.positions, self.img_col, strict=True)))

    def canvas_height(self) -> int:
        return int(max(p.ypos + m.shape[0] for p, m in zip(self.positions, self.img_row, strict=True)))

    @staticmethod
    def split_to_channels(indexed: np.ndarray, n_images: int) -> np.ndarray:
        indexed = np.asarray(indexed, dtype=np.uint8)
        return np.stack([(indexed == i).astype(np.uint8) for i in range(n_images)], axis=-1)


class ControlMasks3:
    """Three-image rig (src/pano/controlMasks3.h:12-43); the seam holds labels 0..2."""

    def __init__(self, game_dir: str | None = None):
        self.img0_col = _empty_u16()
        self.img0_row = _empty_u16()
        self.img1_col = _empty_u16()
        self.img1_row = _empty_u16()
        self.img2_col = _empty_u16()
        self.img2_row = _empty_u16()
        self.whole_seam_mask_image = np.empty((0, 0), dtype=np.uint8)
        self.positions: list = []
        self.last_error = None
        if game_dir is not None:
            self.load(game_dir)

    def load(self, game_dir: str) -> bool:
        rig = _load(game_dir, hm_stitch.HM_CONTROL_THREE, 3, self)
        if rig is None:
            return False
        self.img0_col, self.img1_col, self.img2_col = rig.map_x
        self.img0_row, self.img1_row, self.img2_row = rig.map_y
        self.whole_seam_mask_image = rig.seam
        self.positions = _positions(rig)
        return self.is_valid()

    def is_valid(self) -> bool:
        maps = (self.img0_col, self.img0_row, self.img1_col, self.img1_row, self.img2_col, self.img2_row)
        return all(m.size for m in maps) and self.whole_seam_mask_image.size > 0 and len(self.positions) == 3

    def canvas_width(self) -> int:
        return int(max(p.xpos + m.shape[1] for p, m in zip(self.positions, (self.img0_col, self.img1_col, self.img2_col))))

    def canvas_height(self) -> int:
        return int(max(p.ypos + m.shape[0] for p, m in zip(self.positions, (self.img0_col, self.img1_col, self.img2_col))))

    @staticmethod
    def split_to_channels(seam_mask: np.ndarray) -> np.ndarray:
        """controlMasks3.cpp:206-224: one-hot of labels 0, 1, 2 (other values: all zero)."""
        s = np.asarray(seam_mask, dtype=np.uint8)
        return np.stack([(s == i).astype(np.uint8) for i in range(3)], axis=-1)
