"""Rig control directories and raster files (include/hm_files.h), on the CPU.

The library's loaders (hmControlFilesLoad) are checked against the oracle
(oracle/control_files.py), which decodes with Pillow and restates the
reference's rules; the oracle itself is pinned by hand-worked cases first.
The decoders are checked against Pillow-written files and against
hand-encoded layouts Pillow does not write (tests/file_writers.py)."""
import os

import numpy as np
import pytest
from PIL import Image

import file_writers as fw

FULL_PALETTE = [(0, 0, 0), (0, 255, 0), (255, 0, 0), (0, 0, 255), (255, 255, 0), (255, 0, 255), (0, 255, 255),
                (128, 128, 128)]


@pytest.fixture(scope="module")
def hs():
    import hm_stitch

    hm_stitch.lib()
    return hm_stitch


@pytest.fixture(scope="module")
def cf():
    import control_files

    return control_files


# ---------------------------------------------------------------------------
# writing rig directories the way Hugin / enblend / the reference's scripts do
# ---------------------------------------------------------------------------
def save_map(path, a, compression="tiff_lzw", predictor=2):
    info = {317: predictor} if compression in ("tiff_lzw", "tiff_adobe_deflate") and predictor == 2 else {}
    Image.fromarray(np.ascontiguousarray(a, np.uint16)).save(path, compression=compression, tiffinfo=info)


def save_position(path, xpos, ypos, xres=1.0, yres=1.0, size=(4, 3)):
    """nona's warped image: only the tags matter (values are float tags, stored as rationals)."""
    info = {}
    if xres is not None:
        info[282] = xres
    if yres is not None:
        info[283] = yres
    if xpos is not None:
        info[286] = xpos
    if ypos is not None:
        info[287] = ypos
    Image.fromarray(np.zeros((size[1], size[0], 3), np.uint8)).save(path, tiffinfo=info)


def save_paletted(path, labels):
    im = Image.fromarray(np.ascontiguousarray(labels, np.uint8), mode="P")
    pal = []
    for i in range(256):
        pal.extend(FULL_PALETTE[i % len(FULL_PALETTE)])
    im.putpalette(pal)  # full 256-entry palette: an 8-bit paletted file, as scripts/generate_seam.py writes
    im.save(path)


def write_rig(d, maps, positions, seam_writer, compression="tiff_lzw", predictor=2, res=(1.0, 1.0)):
    os.makedirs(d, exist_ok=True)
    for i, (mx, my) in enumerate(maps):
        save_map(os.path.join(d, f"mapping_{i:04d}_x.tif"), mx, compression, predictor)
        save_map(os.path.join(d, f"mapping_{i:04d}_y.tif"), my, compression, predictor)
        if positions[i] is not None:
            save_position(os.path.join(d, f"mapping_{i:04d}.tif"), positions[i][0], positions[i][1], *res)
    seam_writer(os.path.join(d, "seam_file.png"))


def random_maps(rng, sizes):
    out = []
    for w, h in sizes:
        mx = rng.integers(0, 4000, (h, w)).astype(np.uint16)
        my = rng.integers(0, 3000, (h, w)).astype(np.uint16)
        mx[rng.random((h, w)) < 0.1] = 65535
        my[mx == 65535] = 65535
        out.append((mx, my))
    return out


def assert_matches_oracle(got, want):
    assert got.n_images == len(want["map_x"])
    for k in range(got.n_images):
        np.testing.assert_array_equal(got.map_x[k], want["map_x"][k])
        np.testing.assert_array_equal(got.map_y[k], want["map_y"][k])
        assert np.float32(got.xpos[k]) == want["xpos"][k] and np.float32(got.ypos[k]) == want["ypos"][k]
    assert got.position_files_found == want["found"]
    np.testing.assert_array_equal(got.seam, want["seam"])
    assert got.seam_paletted == want["paletted"]


# ---------------------------------------------------------------------------
# the oracle, pinned by hand-worked cases
# ---------------------------------------------------------------------------
def test_oracle_two_image_seam_inversion_known_answers(cf):
    # max -> 0, min -> 1, other values unchanged (controlMasks.cpp:244-248)
    s = np.array([[10, 20, 30], [30, 10, 20]], np.uint8)
    np.testing.assert_array_equal(cf.invert_two_image_seam(s), [[1, 20, 0], [0, 1, 20]])
    np.testing.assert_array_equal(cf.invert_two_image_seam(np.array([[0, 255]], np.uint8)), [[1, 0]])
    # constant seam: the min mask is applied last, so everything becomes 1
    np.testing.assert_array_equal(cf.invert_two_image_seam(np.full((2, 2), 7, np.uint8)), np.ones((2, 2)))


def test_oracle_n_image_relabel_known_answers(cf):
    np.testing.assert_array_equal(cf.relabel_n(np.array([[0, 255, 255]], np.uint8), 2), [[0, 1, 1]])
    np.testing.assert_array_equal(cf.relabel_n(np.array([[9, 3, 7, 3]], np.uint8), 3), [[2, 0, 1, 0]])
    same = np.array([[2, 0, 1]], np.uint8)
    assert cf.relabel_n(same, 3) is same  # already canonical: untouched
    with pytest.raises(cf.OracleFileError) as e:
        cf.relabel_n(np.array([[0, 1, 5]], np.uint8), 2)
    assert e.value.code == cf.SEAM_CLASSES


def test_oracle_positions_known_answers(cf, tmp_path):
    p = str(tmp_path / "pos.tif")
    save_position(p, 12.25, 0.5, xres=4.0, yres=10.0)
    x, y, ok = cf.read_position(p, both_resolutions=True)
    assert ok and x == np.float32(49.0) and y == np.float32(5.0)
    # the 2/3-image rule: one resolution missing -> both ignored -> (0, 0)
    save_position(p, 12.25, 0.5, xres=4.0, yres=None)
    assert cf.read_position(p, both_resolutions=True)[:2] == (0.0, 0.0)
    x, y, _ = cf.read_position(p, both_resolutions=False)
    assert x == np.float32(49.0) and y == np.float32(0.0)
    # missing file: (0, 0), not found
    assert cf.read_position(str(tmp_path / "nope.tif"), True) == (0.0, 0.0, False)
    # float32 product, as the reference computes it: 0.1 * 30 is not 3 in float
    save_position(p, 0.1, 0.7, xres=30.0, yres=30.0)
    x, y, _ = cf.read_position(p, True)
    assert x == np.float32(np.float32(0.1) * np.float32(30.0)) and y == np.float32(np.float32(0.7) * np.float32(30))


# ---------------------------------------------------------------------------
# loaders against the oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("compression,predictor", [("raw", 1), ("tiff_lzw", 1), ("tiff_lzw", 2),
                                                   ("tiff_adobe_deflate", 2), ("packbits", 1)])
def test_two_image_rig_matches_oracle(hs, cf, tmp_path, compression, predictor):
    rng = np.random.default_rng(5)
    maps = random_maps(rng, [(61, 37), (59, 41)])
    seam = np.where(rng.random((45, 90)) < 0.5, 0, 255).astype(np.uint8)
    d = str(tmp_path / "rig")
    write_rig(d, maps, [(3.0, 7.0), (40.0, 2.0)], lambda p: Image.fromarray(seam).save(p), compression, predictor)
    got = hs.load_control_dir(d, hs.HM_CONTROL_TWO)
    want = cf.load(d, cf.TWO)
    assert_matches_oracle(got, want)
    # placements normalised by the minimum; the seam inverted
    assert got.xpos == [0.0, 37.0] and got.ypos == [5.0, 0.0]
    np.testing.assert_array_equal(got.seam, (seam == 0).astype(np.uint8))


def test_two_image_paletted_seam(hs, cf, tmp_path):
    rng = np.random.default_rng(6)
    maps = random_maps(rng, [(20, 10), (20, 10)])
    labels = (rng.random((12, 30)) < 0.5).astype(np.uint8)
    d = str(tmp_path / "rig")
    write_rig(d, maps, [(0, 0), (10, 0)], lambda p: save_paletted(p, labels))
    got = hs.load_control_dir(d, hs.HM_CONTROL_TWO)
    assert_matches_oracle(got, cf.load(d, cf.TWO))
    assert got.seam_paletted
    np.testing.assert_array_equal(got.seam, 1 - labels)  # index 1 (max) -> 0, index 0 (min) -> 1


@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_n_image_rig_matches_oracle(hs, cf, tmp_path, n):
    rng = np.random.default_rng(10 + n)
    maps = random_maps(rng, [(30 + 3 * i, 20 + i) for i in range(n)])
    labels = rng.integers(0, n, (25, 70)).astype(np.uint8)
    labels.ravel()[:n] = np.arange(n)  # every label present
    d = str(tmp_path / "rig")
    pos = [(float(5 * i + 1), float(i % 3)) for i in range(n)]
    write_rig(d, maps, pos, lambda p: save_paletted(p, labels))
    got = hs.load_control_dir(d, hs.HM_CONTROL_N, n)
    want = cf.load(d, cf.N, n)
    assert_matches_oracle(got, want)
    np.testing.assert_array_equal(got.seam, labels)


def test_n_image_grayscale_seam_relabelled(hs, cf, tmp_path):
    rng = np.random.default_rng(3)
    maps = random_maps(rng, [(16, 8)] * 3)
    gray = np.choose(rng.integers(0, 3, (10, 40)), [40, 200, 90]).astype(np.uint8)
    d = str(tmp_path / "rig")
    write_rig(d, maps, [(0, 0), (12, 0), (24, 0)], lambda p: Image.fromarray(gray).save(p))
    got = hs.load_control_dir(d, hs.HM_CONTROL_N, 3)
    assert_matches_oracle(got, cf.load(d, cf.N, 3))
    np.testing.assert_array_equal(got.seam, np.choose((gray == 90) * 1 + (gray == 200) * 2, [0, 1, 2]))
    assert not got.seam_paletted


def test_three_image_rig_matches_oracle(hs, cf, tmp_path):
    rng = np.random.default_rng(33)
    maps = random_maps(rng, [(40, 20), (44, 22), (38, 19)])
    labels = rng.integers(0, 3, (24, 100)).astype(np.uint8)
    d = str(tmp_path / "rig")
    write_rig(d, maps, [(10.5, 3.25), (40.0, 0.0), (70.75, 1.0)], lambda p: save_paletted(p, labels))
    got = hs.load_control_dir(d, hs.HM_CONTROL_THREE)
    assert_matches_oracle(got, cf.load(d, cf.THREE))
    assert got.xpos == [0.0, 29.5, 60.25]


def test_fractional_placements_follow_float_arithmetic(hs, cf, tmp_path):
    """Positions in inches x resolution, multiplied in float32 like the reference
    (the pano classes then truncate them toward zero)."""
    rng = np.random.default_rng(1)
    maps = random_maps(rng, [(16, 8)] * 2)
    d = str(tmp_path / "rig")
    write_rig(d, maps, [(0.1, 0.7), (1.3, 0.3)], lambda p: Image.fromarray(np.eye(8, 20, dtype=np.uint8) * 255).save(p),
              res=(30.0, 30.0))
    got = hs.load_control_dir(d, hs.HM_CONTROL_TWO)
    assert_matches_oracle(got, cf.load(d, cf.TWO))


def test_missing_position_file_gives_zero_placement(hs, cf, tmp_path):
    rng = np.random.default_rng(2)
    maps = random_maps(rng, [(16, 8)] * 3)
    d = str(tmp_path / "rig")
    write_rig(d, maps, [(5.0, 5.0), None, (9.0, 2.0)], lambda p: save_paletted(p, np.arange(3)[None].repeat(4, 0)))
    got = hs.load_control_dir(d, hs.HM_CONTROL_N, 3)
    assert got.position_files_found == 0b101
    assert_matches_oracle(got, cf.load(d, cf.N, 3))
    assert got.xpos == [5.0, 0.0, 9.0] and got.ypos == [5.0, 0.0, 2.0]


# ---------------------------------------------------------------------------
# errors
# ---------------------------------------------------------------------------
def rig_dir(tmp_path, n=2, seam=None):
    rng = np.random.default_rng(0)
    d = str(tmp_path / "rig")
    labels = np.arange(n, dtype=np.uint8)[None].repeat(6, 0) if seam is None else seam
    write_rig(d, random_maps(rng, [(12, 6)] * n), [(float(i), 0.0) for i in range(n)],
              lambda p: save_paletted(p, labels))
    return d


def expect_error(hs, cf, d, kind, n, code):
    with pytest.raises(hs.FileError) as e:
        hs.load_control_dir(d, kind, n)
    assert e.value.code == code, e.value.message
    with pytest.raises(cf.OracleFileError) as eo:
        cf.load(d, {hs.HM_CONTROL_TWO: cf.TWO, hs.HM_CONTROL_THREE: cf.THREE, hs.HM_CONTROL_N: cf.N}[kind], n)
    assert eo.value.code == code


def test_error_missing_map(hs, cf, tmp_path):
    d = rig_dir(tmp_path)
    os.remove(os.path.join(d, "mapping_0001_y.tif"))
    expect_error(hs, cf, d, hs.HM_CONTROL_TWO, 0, hs.HM_FILE_NOT_FOUND)


def test_error_missing_seam(hs, cf, tmp_path):
    d = rig_dir(tmp_path)
    os.remove(os.path.join(d, "seam_file.png"))
    expect_error(hs, cf, d, hs.HM_CONTROL_TWO, 0, hs.HM_FILE_NOT_FOUND)


def test_error_map_not_uint16(hs, cf, tmp_path):
    d = rig_dir(tmp_path)
    Image.fromarray(np.zeros((6, 12), np.uint8)).save(os.path.join(d, "mapping_0000_x.tif"))
    expect_error(hs, cf, d, hs.HM_CONTROL_TWO, 0, hs.HM_FILE_WRONG_TYPE)


def test_error_map_sizes_differ(hs, cf, tmp_path):
    d = rig_dir(tmp_path)
    save_map(os.path.join(d, "mapping_0001_y.tif"), np.zeros((7, 12), np.uint16))
    expect_error(hs, cf, d, hs.HM_CONTROL_TWO, 0, hs.HM_FILE_MISMATCH)


def test_error_seam_class_count(hs, cf, tmp_path):
    d = rig_dir(tmp_path, n=3)
    assert hs.load_control_dir(d, hs.HM_CONTROL_N, 3).n_images == 3
    save_paletted(os.path.join(d, "seam_file.png"), np.array([[0, 1, 1, 0]], np.uint8))
    expect_error(hs, cf, d, hs.HM_CONTROL_N, 3, hs.HM_FILE_SEAM_CLASSES)


def test_error_three_image_seam_must_be_paletted(hs, cf, tmp_path):
    d = rig_dir(tmp_path, n=3)
    Image.fromarray(np.array([[0, 1, 2]], np.uint8)).save(os.path.join(d, "seam_file.png"))
    expect_error(hs, cf, d, hs.HM_CONTROL_THREE, 0, hs.HM_FILE_UNSUPPORTED)


def test_error_colour_seam(hs, cf, tmp_path):
    d = rig_dir(tmp_path)
    Image.fromarray(np.zeros((4, 4, 3), np.uint8)).save(os.path.join(d, "seam_file.png"))
    expect_error(hs, cf, d, hs.HM_CONTROL_TWO, 0, hs.HM_FILE_UNSUPPORTED)


def test_error_corrupt_files(hs, tmp_path):
    d = rig_dir(tmp_path)
    p = os.path.join(d, "mapping_0000_x.tif")
    blob = open(p, "rb").read()
    with open(p, "wb") as f:
        f.write(blob[: len(blob) // 2])  # strips cut off
    with pytest.raises(hs.FileError) as e:
        hs.load_control_dir(d, hs.HM_CONTROL_TWO)
    assert e.value.code == hs.HM_FILE_BAD_FORMAT
    with open(p, "wb") as f:
        f.write(b"not a tiff at all")
    with pytest.raises(hs.FileError) as e:
        hs.load_control_dir(d, hs.HM_CONTROL_TWO)
    assert e.value.code == hs.HM_FILE_BAD_FORMAT
    s = os.path.join(d, "seam_file.png")
    blob = bytearray(open(s, "rb").read())
    blob[40] ^= 0xFF  # inside IDAT: the CRC no longer matches
    open(s, "wb").write(bytes(blob))
    with pytest.raises(hs.FileError) as e:
        hs.png_read_seam(s)
    assert e.value.code == hs.HM_FILE_BAD_FORMAT


def test_error_arguments(hs, tmp_path):
    with pytest.raises(hs.FileError) as e:
        hs.load_control_dir(str(tmp_path), hs.HM_CONTROL_N, 9)
    assert e.value.code == hs.HM_FILE_INVALID_ARGUMENT
    with pytest.raises(hs.FileError) as e:
        hs.load_control_dir(str(tmp_path), 7)
    assert e.value.code == hs.HM_FILE_INVALID_ARGUMENT


# ---------------------------------------------------------------------------
# decoders
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("compression", ["raw", "tiff_lzw", "tiff_adobe_deflate", "packbits"])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_tiff_read_matches_pillow(hs, tmp_path, compression, dtype):
    rng = np.random.default_rng(7)
    a = (rng.random((53, 71)) * (1 if dtype == np.float32 else np.iinfo(dtype).max)).astype(dtype)
    p = str(tmp_path / "a.tif")
    info = {317: 3 if dtype == np.float32 else 2} if compression in ("tiff_lzw", "tiff_adobe_deflate") else {}
    Image.fromarray(a).save(p, compression=compression, tiffinfo=info)
    np.testing.assert_array_equal(hs.tiff_read(p), np.array(Image.open(p)))
    np.testing.assert_array_equal(hs.tiff_read(p), a)


def test_tiff_lzw_long_strings_and_table_resets(hs, tmp_path):
    # long runs grow LZW strings to the 4096-entry table limit and force Clear codes
    a = np.concatenate([np.tile(np.arange(3000, dtype=np.uint16), (40, 1)),
                        np.random.default_rng(1).integers(0, 65535, (40, 3000)).astype(np.uint16)])
    p = str(tmp_path / "a.tif")
    Image.fromarray(a).save(p, compression="tiff_lzw")
    np.testing.assert_array_equal(hs.tiff_read(p), a)


def test_tiff_rgb_and_opencv_order(hs, tmp_path):
    rgb = np.random.default_rng(3).integers(0, 255, (17, 23, 3)).astype(np.uint8)
    p = str(tmp_path / "rgb.tif")
    Image.fromarray(rgb).save(p, compression="tiff_lzw", tiffinfo={317: 2})
    np.testing.assert_array_equal(hs.tiff_read(p), rgb)
    np.testing.assert_array_equal(hs.tiff_read(p, opencv_order=True), rgb[..., ::-1])


@pytest.mark.parametrize("big_endian", [False, True])
@pytest.mark.parametrize("layout", ["strips", "tiles", "planar", "planar_tiles"])
@pytest.mark.parametrize("predictor", [1, 2])
def test_tiff_layouts(hs, tmp_path, big_endian, layout, predictor):
    rng = np.random.default_rng(11)
    a = rng.integers(0, 65535, (45, 37, 3)).astype(np.uint16)
    p = str(tmp_path / "a.tif")
    fw.write_tiff(p, a, big_endian=big_endian, tile=(16, 16) if "tiles" in layout else None,
                  planar=layout.startswith("planar"), deflate=predictor == 2, predictor=predictor, rows_per_strip=7)
    np.testing.assert_array_equal(hs.tiff_read(p), a)


def test_tiff_big_endian_float_and_tags(hs, tmp_path):
    f = np.random.default_rng(4).random((9, 13)).astype(np.float32)
    p = str(tmp_path / "f.tif")
    fw.write_tiff(p, f, big_endian=True, tags={282: (5, [(300, 1)]), 283: (5, [(600, 2)]), 286: (5, [(49, 4)]),
                                               287: (5, [(0, 1)])})
    np.testing.assert_array_equal(hs.tiff_read(p), f)
    t = hs.tiff_read_tags(p)
    assert t["has_resolution"] and t["xres"] == 300.0 and t["yres"] == 300.0
    assert t["xpos"] == 12.25 and t["ypos"] == 0.0 and t["bits"] == 32 and t["sample_format"] == 3


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("first_filter", [0, 1, 2, 3, 4])
def test_png_filters_and_interlace(hs, tmp_path, interlace, first_filter):
    rng = np.random.default_rng(first_filter)
    g = rng.integers(0, 256, (23, 29)).astype(np.uint8)
    p = str(tmp_path / "g.png")
    fw.write_png(p, g, ctype=0, depth=8, interlace=interlace, first_filter=first_filter)
    s, pal = hs.png_read_seam(p)
    assert not pal
    np.testing.assert_array_equal(s, g)
    np.testing.assert_array_equal(s, np.array(Image.open(p)))


@pytest.mark.parametrize("depth", [1, 2, 4, 16])
@pytest.mark.parametrize("interlace", [False, True])
def test_png_gray_depths(hs, tmp_path, depth, interlace):
    rng = np.random.default_rng(depth)
    v = rng.integers(0, 1 << depth, (19, 21)).astype(np.uint16 if depth == 16 else np.uint8)
    p = str(tmp_path / "g.png")
    fw.write_png(p, v, ctype=0, depth=depth, interlace=interlace, first_filter=depth % 5)
    s, _ = hs.png_read_seam(p)
    ref = np.array(Image.open(p))  # Pillow: 1 bit as bool, 2/4 bits scaled to 0..255, 16 bits whole
    want = {1: ref.astype(np.uint8) * np.uint8(255), 16: (ref >> 8).astype(np.uint8)}.get(depth, ref)
    np.testing.assert_array_equal(s, want)
    if depth in (1, 16):
        np.testing.assert_array_equal(ref.astype(v.dtype), v)  # the encoder wrote v


def test_png_paletted_interlaced_and_gray_alpha(hs, tmp_path):
    rng = np.random.default_rng(8)
    idx = rng.integers(0, 6, (31, 17)).astype(np.uint8)
    p = str(tmp_path / "p.png")
    fw.write_png(p, idx, ctype=3, depth=8, palette=np.arange(18) * 10, interlace=True, first_filter=3)
    s, pal = hs.png_read_seam(p)
    assert pal
    np.testing.assert_array_equal(s, idx)
    np.testing.assert_array_equal(s, np.array(Image.open(p)))
    la = np.stack([rng.integers(0, 256, (9, 11)), rng.integers(0, 256, (9, 11))], -1).astype(np.uint8)
    fw.write_png(p, la, ctype=4, depth=8, first_filter=4)
    np.testing.assert_array_equal(hs.png_read_seam(p)[0], la[..., 0])


def test_png_paletted_below_8_bits_is_refused(hs, tmp_path):
    p = str(tmp_path / "p.png")
    im = Image.fromarray(np.array([[0, 1], [2, 1]], np.uint8), mode="P")
    im.putpalette([0, 0, 0, 255, 0, 0, 0, 255, 0])  # 3 colours: Pillow writes 2-bit indices
    im.save(p)
    with pytest.raises(hs.FileError) as e:
        hs.png_read_seam(p)
    assert e.value.code == hs.HM_FILE_UNSUPPORTED


# ---------------------------------------------------------------------------
# TIFF writer (pyramid dumps)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("channels", [1, 2, 3, 4, 6])
def test_tiff_write_roundtrip(hs, tmp_path, channels):
    rng = np.random.default_rng(channels)
    a = rng.standard_normal((37, 300, channels)).astype(np.float32)
    if channels == 1:
        a = a[..., 0]
    p = str(tmp_path / "w.tif")
    hs.tiff_write(p, a, opencv_order=True)
    np.testing.assert_array_equal(hs.tiff_read(p, opencv_order=True), a)
    if channels in (3, 4):
        # the file holds RGB(A): OpenCV's BGR(A) memory order swapped
        np.testing.assert_array_equal(hs.tiff_read(p), a[..., [2, 1, 0] + ([3] if channels == 4 else [])])
    if channels == 1:
        np.testing.assert_array_equal(np.array(Image.open(p)), a)


def test_tiff_write_tags_read_by_pillow(hs, tmp_path):
    p = str(tmp_path / "w.tif")
    hs.tiff_write(p, np.zeros((3, 4), np.uint16), xres=72.0, yres=72.0, xpos=12.25, ypos=0.1)
    tags = Image.open(p).tag_v2
    assert float(tags[282]) == 72.0 and float(tags[286]) == 12.25
    assert np.float32(float(tags[287])) == np.float32(0.1)
    t = hs.tiff_read_tags(p)
    assert t["xpos"] == 12.25 and t["ypos"] == np.float32(0.1)


# ---------------------------------------------------------------------------
# the C++ classes (hm-cupano_amd/include/cupano/pano/controlMasks.h)
# ---------------------------------------------------------------------------
def fnv1a(values) -> int:
    h = 1469598103934665603
    for v in np.asarray(values).ravel().tolist():
        h = ((h ^ v) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.parametrize("kind,n", [("two", 2), ("three", 3), ("n", 4)])
def test_cpp_control_masks_load(cf, tmp_path, kind, n):
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hm-cupano_amd", "bin",
                       "control_files_test")
    if not os.path.exists(exe):
        pytest.fail("control_files_test is not built (run __graft_entry__.build())")
    rng = np.random.default_rng(40 + n)
    maps = random_maps(rng, [(23 + i, 11 + 2 * i) for i in range(n)])
    labels = rng.integers(0, n, (14, 50)).astype(np.uint8)
    labels.ravel()[:n] = np.arange(n)
    d = str(tmp_path / "rig")
    write_rig(d, maps, [(float(7 * i) + 0.5, float(3 - i)) for i in range(n)], lambda p: save_paletted(p, labels),
              res=(2.0, 2.0))
    out = subprocess.run([exe, d, kind, str(n)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    want = cf.load(d, {"two": cf.TWO, "three": cf.THREE, "n": cf.N}[kind], n)
    lines = out.stdout.split("\n")
    for i in range(n):
        h, w = want["map_x"][i].shape
        assert f"map {i} {w} {h} {fnv1a(want['map_x'][i])} {fnv1a(want['map_y'][i])}" in lines
        _, _, xs, ys = next(l for l in lines if l.startswith(f"pos {i} ")).split()
        assert np.float32(float(xs)) == want["xpos"][i] and np.float32(float(ys)) == want["ypos"][i]
    sh, sw = want["seam"].shape
    assert f"seam {sw} {sh} {fnv1a(want['seam'])}" in lines
    cw = int(max(np.float32(want["xpos"][i] + np.float32(want["map_x"][i].shape[1])) for i in range(n)))
    ch = int(max(np.float32(want["ypos"][i] + np.float32(want["map_x"][i].shape[0])) for i in range(n)))
    assert f"canvas {cw} {ch}" in lines


def test_cpp_control_masks_load_failure_reports_reason(tmp_path):
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hm-cupano_amd", "bin",
                       "control_files_test")
    out = subprocess.run([exe, str(tmp_path), "two"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "mapping_0000_x.tif" in out.stdout
