"""Small PNG and TIFF encoders for the decoder tests: the layouts Pillow does
not write (every PNG filter type on chosen rows, Adam7 interlace; tiled,
big-endian and planar TIFFs).  Written from the PNG and TIFF 6.0
specifications."""
from __future__ import annotations

import struct
import zlib

import numpy as np

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter_rows(rows: np.ndarray, bpp: int, first_filter: int) -> bytes:
    """rows: uint8 [h, stride]; row y uses filter (first_filter + y) % 5."""
    out = bytearray()
    prev = np.zeros(rows.shape[1], np.int32)
    for y in range(rows.shape[0]):
        cur = rows[y].astype(np.int32)
        f = (first_filter + y) % 5
        enc = np.zeros_like(cur)
        for i in range(cur.size):
            a = int(cur[i - bpp]) if i >= bpp else 0
            b = int(prev[i])
            c = int(prev[i - bpp]) if i >= bpp else 0
            pred = (0, a, b, (a + b) >> 1, _paeth(a, b, c))[f]
            enc[i] = (int(cur[i]) - pred) & 0xFF
        out.append(f)
        out += enc.astype(np.uint8).tobytes()
        prev = cur
    return bytes(out)


def _pack(samples: np.ndarray, depth: int) -> np.ndarray:
    """[h, w] of sub-byte values -> [h, stride] packed MSB first."""
    h, w = samples.shape
    per = 8 // depth
    stride = (w + per - 1) // per
    out = np.zeros((h, stride), np.uint8)
    for x in range(w):
        out[:, x // per] |= (samples[:, x].astype(np.uint8) << (8 - depth - (x % per) * depth)).astype(np.uint8)
    return out


def write_png(path, img: np.ndarray, ctype: int, depth: int = 8, palette=None, interlace: bool = False,
              first_filter: int = 0) -> None:
    """img: [h, w] (gray / palette) or [h, w, 2] (gray + alpha), values at `depth` bits."""
    h, w = img.shape[:2]
    channels = 1 if img.ndim == 2 else img.shape[2]
    bpp = max(1, channels * depth // 8)

    def rows_of(sub):
        if depth < 8:
            return _pack(sub, depth)
        if depth == 16:
            return sub.astype(">u2").view(np.uint8).reshape(sub.shape[0], -1)
        return sub.reshape(sub.shape[0], -1).astype(np.uint8)

    raw = b""
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for k, (x0, y0, dx, dy) in enumerate(passes):
        sub = img[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        raw += _filter_rows(rows_of(sub), bpp, first_filter + k)
    data = b"\x89PNG\r\n\x1a\n"
    data += _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    if palette is not None:
        data += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).ravel()))
    z = zlib.compress(raw, 6)
    data += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])  # split across two chunks
    data += _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def write_tiff(path, img: np.ndarray, big_endian: bool = False, tile=None, planar: bool = False,
               deflate: bool = False, predictor: int = 1, rows_per_strip=None, tags=None) -> None:
    """img: [h, w] or [h, w, c] of uint8/uint16/float32.  tile=(tw, th) writes
    tiles; planar writes one plane per sample; predictor 2 differences
    horizontally.  tags: extra {tag: (type, [values])} entries."""
    e = ">" if big_endian else "<"
    a = img if img.ndim == 3 else img[..., None]
    h, w, c = a.shape
    dt = a.dtype
    bits = dt.itemsize * 8
    fmt = 3 if dt.kind == "f" else 1
    planes = [a[..., k:k + 1] for k in range(c)] if planar else [a]
    cw, ch = (tile if tile else (w, rows_per_strip or h))
    across, down = (w + cw - 1) // cw, (h + ch - 1) // ch
    chunks = []
    for pl in planes:
        for ty in range(down):
            for tx in range(across):
                y0, x0 = ty * ch, tx * cw
                rows = ch if tile else min(ch, h - y0)
                blk = np.zeros((rows, cw, pl.shape[2]), dt)
                part = pl[y0:y0 + rows, x0:x0 + cw]
                blk[: part.shape[0], : part.shape[1]] = part
                if predictor == 2:
                    iv = blk.view({1: np.uint8, 2: np.uint16, 4: np.uint32}[dt.itemsize])
                    d = iv.copy()
                    d[:, 1:] = iv[:, 1:] - iv[:, :-1]  # unsigned: wraps modulo 2^bits
                    blk = d.view(dt)
                raw = blk.astype(blk.dtype.newbyteorder(e)).tobytes()
                chunks.append(zlib.compress(raw) if deflate else raw)
    out = bytearray(b"MM\x00\x2a" if big_endian else b"II\x2a\x00")
    out += b"\x00\x00\x00\x00"
    offs, cnts = [], []
    for ck in chunks:
        offs.append(len(out))
        cnts.append(len(ck))
        out += ck
        if len(out) & 1:
            out += b"\x00"
    entries = {
        256: (4, [w]), 257: (4, [h]), 258: (3, [bits] * c), 259: (3, [8 if deflate else 1]),
        262: (3, [2 if c >= 3 else 1]), 277: (3, [c]), 284: (3, [2 if planar else 1]), 339: (3, [fmt] * c),
    }
    if predictor != 1:
        entries[317] = (3, [predictor])
    if tile:
        entries.update({322: (4, [cw]), 323: (4, [ch]), 324: (4, offs), 325: (4, cnts)})
    else:
        entries.update({273: (4, offs), 278: (4, [ch]), 279: (4, cnts)})
    if c > 3:
        entries[338] = (3, [2] * (c - 3))
    entries.update(tags or {})
    ifd_at = len(out)
    out[4:8] = struct.pack(e + "I", ifd_at)
    n = len(entries)
    extra_at = ifd_at + 2 + 12 * n + 4
    extra = bytearray()
    ifd = bytearray(struct.pack(e + "H", n))
    for tag in sorted(entries):
        typ, vals = entries[tag]
        if typ == 3:
            payload = struct.pack(e + "H" * len(vals), *vals)
        elif typ == 4:
            payload = struct.pack(e + "I" * len(vals), *vals)
        elif typ == 5:
            payload = b"".join(struct.pack(e + "II", *v) for v in vals)
        else:
            raise ValueError(typ)
        count = len(vals)
        ifd += struct.pack(e + "HHI", tag, typ, count)
        if len(payload) <= 4:
            ifd += payload.ljust(4, b"\x00")
        else:
            ifd += struct.pack(e + "I", extra_at + len(extra))
            extra += payload
            if len(extra) & 1:
                extra += b"\x00"
    ifd += b"\x00\x00\x00\x00"
    out += ifd + extra
    with open(path, "wb") as f:
        f.write(bytes(out))
