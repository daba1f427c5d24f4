"""Seeded synthetic camera rigs for the stitch path.

Produces what hm-cupano loads from a control directory (mapping_000i_x/_y.tif
CV_16U remap tables with their canvas positions, and seam_file.png), plus the
camera frames, without any file I/O:

  * frames: uint8 RGB = a low-frequency field (sum of 4 random-phase 2-D
    sinusoids around 128) plus uniform integer noise in [-16, 16], expanded to
    BGRA with alpha 255 (as tests/testCudaBlend.cpp:179-189 does with
    cv::COLOR_BGR2BGRA);
  * remap tables: a cylindrical projection of N pinhole cameras (focal f,
    evenly spaced yaws giving the requested overlap, small seeded pitch);
    nearest-integer source coordinates, 65535 where the canvas pixel falls
    outside the camera (the reference's unmapped sentinel,
    src/cuda/cudaRemap.cu:13);
  * positions: the valid bounding box of each camera on the canvas; the canvas
    size and the overlap fraction are authoritative, the cylinder scale is
    derived from them;
  * seams: 2 images -> {0,1} split at a seeded sine-wave boundary inside the
    overlap, stored after ControlMasks' load-time inversion (1 selects image
    1, src/pano/controlMasks.cpp:238-248); N images -> labels 0..N-1 from a
    wavy nearest-camera (Voronoi) partition.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

UNMAPPED = 65535


@dataclasses.dataclass
class Rig:
    """A rig in ControlMasks form (host arrays)."""

    map_x: List[np.ndarray]  # uint16 [h_i, w_i]
    map_y: List[np.ndarray]
    pos_x: List[int]
    pos_y: List[int]
    seam: np.ndarray  # uint8 [canvas_h, canvas_w]
    src_w: int
    src_h: int
    canvas_w: int
    canvas_h: int
    yaws_deg: List[float]
    pitches_deg: List[float]
    focal: float

    @property
    def n(self) -> int:
        return len(self.map_x)


def make_frame(h: int, w: int, seed: int, channels: int = 4) -> np.ndarray:
    """uint8 [h, w, channels] frame (BGR(A), alpha 255)."""
    rng = np.random.default_rng(seed)
    y = np.arange(h, dtype=np.float32)[:, None] / np.float32(h)
    x = np.arange(w, dtype=np.float32)[None, :] / np.float32(w)
    out = np.empty((h, w, channels), dtype=np.uint8)
    for c in range(3):
        field = np.full((h, w), 128.0, dtype=np.float32)
        for _ in range(4):
            fx, fy = rng.uniform(0.5, 4.0, size=2).astype(np.float32)
            ph = np.float32(rng.uniform(0, 2 * math.pi))
            amp = np.float32(rng.uniform(16.0, 32.0))
            field += amp * np.sin(np.float32(2 * math.pi) * (fx * x + fy * y) + ph)
        noise = rng.integers(-16, 17, size=(h, w), dtype=np.int16)
        out[..., c] = np.clip(np.rint(field).astype(np.int16) + noise, 0, 255).astype(np.uint8)
    if channels == 4:
        out[..., 3] = 255
    return out


def _rot(yaw: float, pitch: float) -> np.ndarray:
    """Camera-to-world rotation: pitch about x, then yaw about y."""
    cy, sy = math.cos(yaw), math.sin(yaw)
    cp, sp = math.cos(pitch), math.sin(pitch)
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rx = np.array([[1, 0, 0], [0, cp, -sp], [0, sp, cp]])
    return ry @ rx


def _border_lambda_t(rot, w, h, f):
    u = np.concatenate([np.arange(w), np.arange(w), np.zeros(h), np.full(h, w - 1)]).astype(np.float64)
    v = np.concatenate([np.zeros(w), np.full(w, h - 1), np.arange(h), np.arange(h)]).astype(np.float64)
    d = np.stack([u - w / 2 + 0.5, v - h / 2 + 0.5, np.full_like(u, f)])
    wd = rot @ d
    lam = np.arctan2(wd[0], wd[2])
    t = wd[1] / np.hypot(wd[0], wd[2])
    return lam, t


def cylinder_rig(
    n: int,
    src_w: int,
    src_h: int,
    canvas_w: int,
    canvas_h: int,
    overlap: float,
    focal: float,
    pitch_deg: float = 0.0,
    seed: int = 0,
) -> Rig:
    """N cameras on a cylinder whose union footprint spans canvas_w x canvas_h."""
    rng = np.random.default_rng(seed)
    fov = 2.0 * math.atan(src_w / (2.0 * focal))
    step = (1.0 - overlap) * fov
    yaws = [(i - (n - 1) / 2.0) * step for i in range(n)]
    pitches = [math.radians(float(p)) for p in rng.uniform(-pitch_deg, pitch_deg, size=n)] if pitch_deg else [0.0] * n
    rots = [_rot(yw, pt) for yw, pt in zip(yaws, pitches)]
    lam_lo, lam_hi, t_lo, t_hi = [], [], [], []
    for r in rots:
        lam, t = _border_lambda_t(r, src_w, src_h, focal)
        lam_lo.append(lam.min())
        lam_hi.append(lam.max())
        t_lo.append(t.min())
        t_hi.append(t.max())
    lmin, lmax, tmin, tmax = min(lam_lo), max(lam_hi), min(t_lo), max(t_hi)
    sx = (canvas_w - 1) / (lmax - lmin)
    sy = (canvas_h - 1) / (tmax - tmin)
    maps_x, maps_y, boxes = [], [], []
    for i, r in enumerate(rots):
        x0 = max(0, int(math.floor((lam_lo[i] - lmin) * sx)) - 2)
        x1 = min(canvas_w, int(math.ceil((lam_hi[i] - lmin) * sx)) + 3)
        X = np.arange(x0, x1, dtype=np.float64)[None, :]
        Y = np.arange(canvas_h, dtype=np.float64)[:, None]
        lam = lmin + X / sx
        t = tmin + Y / sy
        # world ray on the unit cylinder, into the camera frame (R^T d)
        dx = np.broadcast_to(np.sin(lam), (canvas_h, x1 - x0))
        dy = np.broadcast_to(t, (canvas_h, x1 - x0))
        dz = np.broadcast_to(np.cos(lam), (canvas_h, x1 - x0))
        cx = r[0, 0] * dx + r[1, 0] * dy + r[2, 0] * dz
        cyy = r[0, 1] * dx + r[1, 1] * dy + r[2, 1] * dz
        cz = r[0, 2] * dx + r[1, 2] * dy + r[2, 2] * dz
        with np.errstate(divide="ignore", invalid="ignore"):
            u = np.rint(focal * cx / cz + src_w / 2 - 0.5)
            v = np.rint(focal * cyy / cz + src_h / 2 - 0.5)
        valid = (cz > 0) & (u >= 0) & (u < src_w) & (v >= 0) & (v < src_h)
        ys, xs = np.nonzero(valid)
        bx0, bx1 = x0 + xs.min(), x0 + xs.max() + 1
        by0, by1 = ys.min(), ys.max() + 1
        boxes.append([bx0, by0, bx1, by1])
        mx = np.where(valid, u, UNMAPPED).astype(np.int64)
        my = np.where(valid, v, UNMAPPED).astype(np.int64)
        maps_x.append((mx, x0))
        maps_y.append((my, x0))
    # Extend the extreme boxes to the requested canvas edges (sentinel fill).
    ux0 = min(b[0] for b in boxes)
    uy0 = min(b[1] for b in boxes)
    ux1 = max(b[2] for b in boxes)
    uy1 = max(b[3] for b in boxes)
    out_x, out_y = [], []
    for i, b in enumerate(boxes):
        if b[0] == ux0:
            b[0] = 0
        if b[1] == uy0:
            b[1] = 0
        if b[2] == ux1:
            b[2] = canvas_w
        if b[3] == uy1:
            b[3] = canvas_h
        mx, x0 = maps_x[i]
        my, _ = maps_y[i]
        h = b[3] - b[1]
        w = b[2] - b[0]
        fx = np.full((h, w), UNMAPPED, dtype=np.uint16)
        fy = np.full((h, w), UNMAPPED, dtype=np.uint16)
        # copy the computed window [x0, x0+mx.shape[1]) x [0, canvas_h) into the box
        cx0 = max(b[0], x0)
        cx1 = min(b[2], x0 + mx.shape[1])
        fx[:, cx0 - b[0]: cx1 - b[0]] = mx[b[1]: b[3], cx0 - x0: cx1 - x0]
        fy[:, cx0 - b[0]: cx1 - b[0]] = my[b[1]: b[3], cx0 - x0: cx1 - x0]
        out_x.append(fx)
        out_y.append(fy)
    pos_x = [int(b[0]) for b in boxes]
    pos_y = [int(b[1]) for b in boxes]
    cw = max(px + m.shape[1] for px, m in zip(pos_x, out_x))
    ch = max(py + m.shape[0] for py, m in zip(pos_y, out_y))
    rig = Rig(out_x, out_y, pos_x, pos_y, np.zeros((ch, cw), np.uint8), src_w, src_h, cw, ch,
              [math.degrees(y) for y in yaws], [math.degrees(p) for p in pitches], focal)
    rig.seam = seam_for(rig, seed + 1)
    return rig


def seam_for(rig: Rig, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    ch, cw = rig.canvas_h, rig.canvas_w
    Y = np.arange(ch, dtype=np.float64)[:, None]
    X = np.arange(cw, dtype=np.float64)[None, :]
    if rig.n == 2:
        # Overlap in canvas columns: [x2, x1 + w1).
        lo = rig.pos_x[1]
        hi = rig.pos_x[0] + rig.map_x[0].shape[1]
        mid = 0.5 * (lo + hi)
        amp = 0.2 * max(hi - lo, 1)
        period = rng.uniform(0.3, 0.8) * ch
        phase = rng.uniform(0, 2 * math.pi)
        boundary = mid + amp * np.sin(2 * math.pi * Y / period + phase)
        # post-load form: 1 selects image 1 (left of the boundary)
        return (X < boundary).astype(np.uint8) * np.ones((1, cw), np.uint8)
    centers = [px + m.shape[1] / 2.0 for px, m in zip(rig.pos_x, rig.map_x)]
    best = np.full((ch, cw), np.inf)
    label = np.zeros((ch, cw), np.uint8)
    for i, c in enumerate(centers):
        amp = rng.uniform(0.02, 0.06) * cw / rig.n
        period = rng.uniform(0.3, 0.8) * ch
        phase = rng.uniform(0, 2 * math.pi)
        d = np.abs(X - c + amp * np.sin(2 * math.pi * Y / period + phase))
        # prefer cameras that actually see the pixel
        mx = np.full((ch, cw), UNMAPPED, np.uint16)
        h, w = rig.map_x[i].shape
        mx[rig.pos_y[i]: rig.pos_y[i] + h, rig.pos_x[i]: rig.pos_x[i] + w] = rig.map_x[i]
        d = d + np.where(mx == UNMAPPED, 1e6, 0.0)
        better = d < best
        best = np.where(better, d, best)
        label = np.where(better, np.uint8(i), label)
    return label


def config_rig(name: str, seed: int = 1234) -> Rig:
    """The named workloads: sizes from the hm-cupano stitch benchmark set."""
    if name == "cfg0":  # 2 x 1280x720 -> 2400x900, L=6
        return cylinder_rig(2, 1280, 720, 2400, 900, overlap=0.25, focal=900.0, pitch_deg=2.0, seed=seed)
    if name == "cfg1":  # 2 x 5312x2988 -> 8500x3300, L=0 and L=6
        return cylinder_rig(2, 5312, 2988, 8500, 3300, overlap=0.20, focal=2700.0, seed=seed)
    if name == "cfg2_n3":  # 3 x 3840x2160 -> 6200x2300
        return cylinder_rig(3, 3840, 2160, 6200, 2300, overlap=0.25, focal=2000.0, seed=seed)
    if name == "cfg2_n8":  # 8 x 3840x2160 -> 14000x2600; focal such that 8 cameras at 25% overlap span < 360 deg
        return cylinder_rig(8, 3840, 2160, 14000, 2600, overlap=0.25, focal=3500.0, seed=seed)
    raise KeyError(name)


def small_rig(n: int = 2, seed: int = 7, src=(160, 64), canvas=(420, 72), overlap=0.3, pitch_deg=2.0) -> Rig:
    """A small rig for parity tests against the oracle (wide enough for the
    2-image minimized blend, which needs x2 >= 128 and w1 + 128 <= canvas)."""
    return cylinder_rig(n, src[0], src[1], canvas[0], canvas[1], overlap=overlap, focal=src[0] * 0.8,
                        pitch_deg=pitch_deg, seed=seed)


# The mixed served load: 32 rigs drawn from 5 seeded configs (2 to 8 images,
# 1920x1080 to 5312x2988 inputs, canvases 2000 to 14000 wide, levels 0/4/6).
# Per config: name, pano class ("2", "3" or "N"), cylinder geometry (n, src_w,
# src_h, canvas_w, canvas_h, overlap, focal, seed; None = the 8500x3300 rig),
# levels, rigs in the mix, distinct input sets per rig.
SERVED_CONFIGS = [
    ("2img_5312x2988_8500x3300_L6", "2", None, 6, 6, 2),
    ("2img_1920x1080_2000x700_L4", "2", (2, 1920, 1080, 2000, 700, 0.25, 1300.0, 5), 4, 8, 2),
    ("3img_2704x1520_6000x2000_L4", "3", (3, 2704, 1520, 6000, 2000, 0.25, 1500.0, 77), 4, 6, 2),
    ("4img_1920x1080_4000x1500_L0", "N", (4, 1920, 1080, 4000, 1500, 0.25, 1100.0, 6), 0, 8, 2),
    ("8img_3840x2160_14000x2600_L6", "N", (8, 3840, 2160, 14000, 2600, 0.25, 3500.0, 1234), 6, 4, 1),
]

# The 16-frame mixed batch: 6 frames of 2 x 5312x2988 -> 8500x3300 L=6, 6 of
# 3 x 2704x1520 -> 6000x2000 L=4 and 4 of 4 x 1920x1080 -> 4000x1500 L=0.
BATCH_MIX = [("2img_5312x2988_8500x3300_L6", 6), ("3img_2704x1520_6000x2000_L4", 6),
             ("4img_1920x1080_4000x1500_L0", 4)]


def served_config(name: str):
    """(kind, levels, rigs in the mix, input sets) of a served-load config."""
    c = next(c for c in SERVED_CONFIGS if c[0] == name)
    return c[1], c[3], c[4], c[5]


def served_rig(name: str) -> Rig:
    """The rig geometry of a served-load config."""
    geo = next(c for c in SERVED_CONFIGS if c[0] == name)[2]
    if geo is None:
        return config_rig("cfg1")
    return cylinder_rig(geo[0], geo[1], geo[2], geo[3], geo[4], overlap=geo[5], focal=geo[6], seed=geo[7])


def served_frames(rig: Rig, input_set: int) -> List[np.ndarray]:
    """Input set k of a served rig: one uint8 BGRA frame [1, H, W, 4] per camera."""
    return [make_frame(rig.src_h, rig.src_w, 1000 * input_set + 17 * i)[None] for i in range(rig.n)]
