"""Pins the CPU oracle: the reference's own known answers, closed forms and
invariants, each chosen so that a plausible slip (dropped term, wrong tap,
swapped image, off-by-one ROI) fails at least one of them.  CPU only."""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import IdRig, const_u8, frames_for, identity_maps, patterned_u8

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")
KAT = json.load(open(GOLDEN))
KIND = {"2": 2, "3": 3, "N": 0}


def f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _ptr(a):
    return a.ctypes.data


# ---------------------------------------------------------------------------
# casts and level sizes
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("x,want", [(254.999, 254), (255.0, 255), (300.0, 255), (-3.0, 0), (0.0, 0), (0.99, 0),
                                    (12.5, 12), (1e-7, 0), (254.0, 254)])
def test_round_to_uchar_truncates(oracle, x, want):
    assert oracle.round_to_uchar(x) == want


def test_pyramid_sizes_closed_form(oracle):
    # (w+1)/2 per level, 2-image rule never clamps
    assert oracle.pyramid_sizes(1200, 3300, 6, False) == [(1200, 3300), (600, 1650), (300, 825), (150, 413),
                                                         (75, 207), (38, 104)]
    assert oracle.pyramid_sizes(3, 5, 4, False) == [(3, 5), (2, 3), (1, 2), (1, 1)]
    # N/3-image rule stops once the previous level has a side < 2
    assert oracle.pyramid_sizes(3, 5, 6, True) == [(3, 5), (2, 3), (1, 2)]
    assert oracle.pyramid_sizes(1, 1, 3, True) == [(1, 1)]


# ---------------------------------------------------------------------------
# single stages
# ---------------------------------------------------------------------------
def downsample(oracle, img, c):
    b, h, w, _ = img.shape
    ow, oh = (w + 1) // 2, (h + 1) // 2
    out = np.zeros((b, oh, ow, c), np.float32)
    oracle.lib().hmo_downsample(_ptr(img), w, h, _ptr(out), ow, oh, c, b)
    return out


def test_downsample_known_answers(oracle):
    img = np.zeros((1, 3, 3, 4), np.float32)
    img[0, 0, 0] = [1, 10, 100, 255]
    img[0, 0, 1] = [2, 20, 200, 255]
    img[0, 1, 0] = [3, 30, 300, 0]      # transparent: excluded from the mean, alpha still max-pooled
    img[0, 1, 1] = [5, 50, 500, 255]
    img[0, 2, 2] = [7, 70, 700, 255]    # odd edge: the (1,1) output sees this single tap
    img[0, 0, 2] = [9, 90, 900, 0]      # (1,0) output: its only in-range taps are transparent
    img[0, 1, 2] = [9, 90, 900, 0]
    out = downsample(oracle, img, 4)
    np.testing.assert_array_equal(out[0, 0, 0], np.float32([8 / 3, 80 / 3, 800 / 3, 255]))
    np.testing.assert_array_equal(out[0, 0, 1], [0, 0, 0, 0])
    np.testing.assert_array_equal(out[0, 1, 1], [7, 70, 700, 255])
    np.testing.assert_array_equal(out[0, 1, 0], [0, 0, 0, 0])


def test_downsample_rgb_plain_mean(oracle):
    img = np.arange(2 * 3 * 3 * 3, dtype=np.float32).reshape(2, 3, 3, 3)
    out = downsample(oracle, img, 3)
    for b in range(2):
        np.testing.assert_allclose(out[b, 0, 0], img[b, :2, :2].reshape(-1, 3).mean(0), rtol=0, atol=1e-5)
        np.testing.assert_array_equal(out[b, 1, 1], img[b, 2, 2])
        np.testing.assert_allclose(out[b, 0, 1], img[b, :2, 2].mean(0), rtol=0, atol=1e-5)


def test_mask_downsample_mean(oracle):
    m = np.float32([[1, 0, 1], [0, 0, 1], [1, 1, 1]])[..., None].copy()
    out = np.zeros((2, 2, 1), np.float32)
    oracle.lib().hmo_downsample_mask(_ptr(m), 3, 3, _ptr(out), 2, 2, 1)
    np.testing.assert_array_equal(out[..., 0], np.float32([[0.25, 1.0], [1.0, 1.0]]))


def laplacian(oracle, high, low, c):
    b, hh, hw, _ = high.shape
    _, lh, lw, _ = low.shape
    out = np.zeros_like(high)
    oracle.lib().hmo_laplacian(_ptr(high), hw, hh, _ptr(low), lw, lh, _ptr(out), c, b)
    return out


def test_laplacian_alpha_renormalisation_known_answer(oracle):
    ka = KAT["laplacian_alpha"]
    high = np.zeros(ka["high_shape"], np.float32)
    low = np.zeros(ka["low_shape"], np.float32)
    low[0, ka["low_pixel"][0], ka["low_pixel"][1]] = ka["low_value"]
    lap = laplacian(oracle, high, low, 4)
    y, x = ka["probe"]
    np.testing.assert_allclose(lap[0, y, x, :3], ka["expected_rgb"], atol=ka["tol"])
    assert lap[0, y, x, 3] == ka["expected_alpha"]


def test_laplacian_rgb_bilinear_closed_form(oracle):
    low = np.zeros((1, 2, 2, 3), np.float32)
    low[0, 0, 0], low[0, 0, 1], low[0, 1, 0], low[0, 1, 1] = 0, 10, 20, 30
    high = np.zeros((1, 4, 4, 3), np.float32)
    lap = -laplacian(oracle, high, low, 3)[0, ..., 0]
    # top-left aligned: x/2, neighbour clamped to the last column/row
    want = np.float32([[0, 5, 10, 10], [10, 15, 20, 20], [20, 25, 30, 30], [20, 25, 30, 30]])
    np.testing.assert_array_equal(lap, want)


def test_laplacian_of_constant_is_zero(oracle):
    high = np.full((1, 7, 5, 4), 77.0, np.float32)
    high[..., 3] = 255
    low = downsample(oracle, high, 4)
    lap = laplacian(oracle, high, low, 4)
    np.testing.assert_array_equal(lap[..., :3], 0)
    np.testing.assert_array_equal(lap[..., 3], 255)


def test_reconstruct_closed_forms(oracle):
    lap = np.random.default_rng(0).uniform(-5, 5, (1, 5, 6, 4)).astype(np.float32)
    lap[..., 3] = 255
    low = np.zeros((1, 3, 3, 4), np.float32)  # all transparent -> upsample contributes 0
    out = np.zeros_like(lap)
    oracle.lib().hmo_reconstruct(_ptr(low), 3, 3, _ptr(lap), 6, 5, _ptr(out), 4, 1)
    np.testing.assert_array_equal(out, lap)
    low[...] = [40, 50, 60, 255]
    oracle.lib().hmo_reconstruct(_ptr(low), 3, 3, _ptr(lap), 6, 5, _ptr(out), 4, 1)
    np.testing.assert_array_equal(out[..., :3], np.float32([40, 50, 60]) + lap[..., :3])
    np.testing.assert_array_equal(out[..., 3], lap[..., 3])


# ---------------------------------------------------------------------------
# blends: the reference's own known answers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", KAT["cases"], ids=[c["name"] for c in KAT["cases"]])
def test_blend_known_answers(oracle, case):
    c = case["c"]
    imgs = [f32(v).reshape(1, 1, 1, c) for v in case["images"]]
    mask = f32(case["mask"]).reshape(1, 1, -1)
    out = oracle.blend(KIND[case["kind"]], imgs, mask, case["levels"])
    np.testing.assert_allclose(out.reshape(-1), case["expected"], atol=case["tol"], rtol=0)


@pytest.mark.parametrize("case", KAT["rule_cases"], ids=[c["name"] for c in KAT["rule_cases"]])
def test_blend_rule_cases(oracle, case):
    b, h, w, c = case["batch"], case["h"], case["w"], case["c"]
    imgs = [np.broadcast_to(f32(v), (b, h, w, c)).copy() for v in case["constants"]]
    mask = np.broadcast_to(f32(case["weights"]), (h, w, len(case["weights"]))).copy()
    out = oracle.blend(KIND[case["kind"]], imgs, mask, case["levels"])
    np.testing.assert_allclose(out, np.broadcast_to(f32(case["expected"]), out.shape), atol=case["tol"], rtol=0)


def test_blend3_no_alpha_holes(oracle):
    ka = KAT["no_holes"]
    w, h, c = ka["w"], ka["h"], ka["c"]
    imgs = [np.zeros((1, h, w, c), np.float32) for _ in range(3)]
    imgs[0][..., 0], imgs[1][..., 1], imgs[2][..., 2] = 100, 150, 200
    for im in imgs:
        im[..., 3] = 255
    mask = np.zeros((h, w, 3), np.float32)
    mask[: h // 2, :, 0] = 1
    mask[h // 2:, :, 1] = 1
    out = oracle.blend(3, imgs, mask, ka["levels"])
    for y in (1, 2):
        for x in range(w):
            assert out[0, y, x, 3] >= ka["min_alpha"]
            assert out[0, y, x, :3].sum() > ka["min_rgb_sum"]


def test_blend2_pixel_rules(oracle):
    # cudaBlend.cu:551-569: alpha1 == 0 -> pixel 2, alpha2 == 0 -> pixel 1, else m*l1 + (1-m)*l2, alpha2
    l1 = f32([[10, 20, 30, 0], [10, 20, 30, 255], [10, 20, 30, 255]]).reshape(1, 1, 3, 4)
    l2 = f32([[50, 60, 70, 255], [50, 60, 70, 0], [50, 60, 70, 200]]).reshape(1, 1, 3, 4)
    mask = f32([[0.9, 0.9, 0.25]]).reshape(1, 3)
    out = oracle.blend(2, [l1, l2], mask, 1)[0, 0]
    np.testing.assert_array_equal(out[0], [50, 60, 70, 255])
    np.testing.assert_array_equal(out[1], [10, 20, 30, 255])
    np.testing.assert_array_equal(out[2], [40, 50, 60, 200])


def test_blend_constant_images_all_kinds(oracle):
    # constant images at any depth: the blend is the pointwise weighted mean
    for kind, n in ((2, 2), (3, 3), (0, 4)):
        h, w = 13, 21
        vals = [f32([10 * i + 5, 20 * i + 7, 30 * i + 1, 255]) for i in range(n)]
        imgs = [np.broadcast_to(v, (1, h, w, 4)).copy() for v in vals]
        if kind == 2:
            mask = np.full((h, w), 0.5, np.float32)
            want = 0.5 * vals[0] + 0.5 * vals[1]
        else:
            mask = np.full((h, w, n), 1.0 / n, np.float32)
            want = sum(vals) / n
        out = oracle.blend(kind, imgs, mask, 5)
        np.testing.assert_allclose(out[..., :3], np.broadcast_to(want[:3], out[..., :3].shape), atol=1e-4)
        np.testing.assert_array_equal(out[..., 3], 255)


def test_blend_perfect_reconstruction(oracle):
    # mask == 1 selects image 1's Laplacian exactly at every level, so the
    # reconstruction telescopes back to image 1 up to float rounding.
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (2, 37, 53, 4)).astype(np.float32)
    img[..., 3] = 255
    other = rng.integers(0, 256, img.shape).astype(np.float32)
    other[..., 3] = 255
    out = oracle.blend(2, [img, other], np.ones((37, 53), np.float32), 6)
    np.testing.assert_allclose(out[..., :3], img[..., :3], atol=2e-4)
    np.testing.assert_array_equal(out[..., 3], 255)
    # With a transparent hole in image 1 and image 2 fully transparent the
    # identity still holds on every opaque pixel, and alpha is image 1's.
    img[:, 5:9, 7:20] = 0
    out = oracle.blend(2, [img, np.zeros_like(img)], np.ones((37, 53), np.float32), 6)
    opaque = img[..., 3] != 0
    np.testing.assert_allclose(out[..., :3][opaque], img[..., :3][opaque], atol=2e-4)
    np.testing.assert_array_equal(out[..., 3], img[..., 3])


# ---------------------------------------------------------------------------
# remap / copy
# ---------------------------------------------------------------------------
def test_remap_identity_copy_and_sentinels(oracle):
    rng = np.random.default_rng(1)
    src = rng.integers(0, 256, (2, 6, 8, 4), dtype=np.uint8)
    mx, my = identity_maps(8, 6)
    mx = mx.copy()
    my = my.copy()
    mx[1, 2] = 65535          # unmapped -> zero with alpha 0
    mx[2, 3] = 8              # out of range but not the sentinel -> zero with alpha 255
    dst = np.full((2, 9, 12, 4), 7, np.uint8)
    assert oracle.remap_offset(src, dst, mx, my, 3, 2) == 0
    want = np.full_like(dst, 7)
    want[:, 2:8, 3:11] = src
    want[:, 3, 5] = [0, 0, 0, 0]
    want[:, 4, 6] = [0, 0, 0, 255]
    np.testing.assert_array_equal(dst, want)
    # no_unmapped_write leaves the sentinel pixel alone (4-channel only)
    dst2 = np.full((2, 9, 12, 4), 7, np.uint8)
    oracle.remap_offset(src, dst2, mx, my, 3, 2, no_unmapped_write=True)
    want[:, 3, 5] = 7
    np.testing.assert_array_equal(dst2, want)


def test_remap_clips_and_roi_validation(oracle):
    src = np.arange(4 * 5 * 4, dtype=np.uint8).reshape(1, 4, 5, 4)
    mx, my = identity_maps(5, 4)
    dst = np.zeros((1, 3, 4, 4), np.float32)
    oracle.remap_offset(src, dst, mx, my, -1, 1)  # partially off the destination
    np.testing.assert_array_equal(dst[0, 1:, :], src[0, :2, 1:].astype(np.float32))
    np.testing.assert_array_equal(dst[0, 0], 0)
    d = np.zeros((1, 6, 8, 4), np.uint8)
    assert oracle.remap_offset(src, d, mx, my, 0, 0, roi=(1, 1, 3, 2)) == 0
    np.testing.assert_array_equal(d[0, 1:3, 1:4], src[0, 1:3, 1:4])
    assert d[0, 0].sum() == 0 and d[0, 3:].sum() == 0
    assert oracle.remap_offset(src, d, mx, my, 0, 0, roi=(3, 0, 3, 2)) != 0   # past the map
    assert oracle.remap_offset(src, d, mx, my, 0, 0, roi=(-1, 0, 2, 2)) != 0  # negative origin
    assert oracle.remap_offset(src, d, mx, my, 0, 0, roi=(0, 0, 0, 2)) == 0   # empty ROI is a no-op


def test_remap_dest_map_selects(oracle):
    src = np.full((1, 4, 4, 3), 9, np.uint8)
    mx, my = identity_maps(4, 4)
    mx = mx.copy()
    mx[0, 0] = 65535
    dm = np.zeros((4, 6), np.uint8)
    dm[:, 3:] = 1
    dst = np.full((1, 4, 6, 3), 5, np.uint8)
    oracle.remap_dest_map(src, dst, mx, my, 1, 0, 1, dm)
    want = np.full_like(dst, 5)
    want[:, :, 3:5] = 9
    np.testing.assert_array_equal(dst, want)
    oracle.remap_dest_map(src, dst, mx, my, 1, 0, 0, dm)
    want[:, 1:, 1:3] = 9
    want[:, 0, 1] = 0  # unmapped -> default pixel
    want[:, 0, 2] = 9
    np.testing.assert_array_equal(dst, want)


def test_hard_seam_n_selects_indexed_image(oracle):
    # tests/python/test_pytorch_pano.py:223-246 restated (uint8 pipeline)
    w, h = 48, 24
    seam = np.zeros((h, w), np.uint8)
    seam[:, 16:32] = 1
    seam[:, 32:] = 2
    imgs = [np.broadcast_to(np.uint8(v), (1, h, w, 4)).copy() for v in ([10, 0, 0, 255], [0, 20, 0, 255], [0, 0, 30, 255])]
    mx, my = identity_maps(w, h)
    dst = np.zeros((1, h, w, 4), np.uint8)
    oracle.remap_hard_seam_n(imgs, [mx] * 3, [my] * 3, [(0, 0)] * 3, seam, dst)
    want = np.zeros_like(dst)
    want[:, :, :16] = imgs[0][:, :, :16]
    want[:, :, 16:32] = imgs[1][:, :, 16:32]
    want[:, :, 32:] = imgs[2][:, :, 32:]
    np.testing.assert_array_equal(dst, want)


def test_copy_roi_casts(oracle):
    src = f32([[12.9, -1.0, 300.0, 254.5], [0.5, 1.5, 2.5, 3.5]]).reshape(1, 1, 2, 4)
    dst = np.zeros((1, 2, 3, 4), np.uint8)
    oracle.copy_roi(src, 2, 1, 0, 0, dst, 1, 1)
    np.testing.assert_array_equal(dst[0, 1, 1], [12, 0, 255, 254])
    np.testing.assert_array_equal(dst[0, 1, 2], [0, 1, 2, 3])
    d3 = np.zeros((1, 1, 2, 3), np.uint8)
    oracle.copy_roi(src, 2, 1, 0, 0, d3, 0, 0)  # 4 -> 3 drops alpha
    np.testing.assert_array_equal(d3[0, 0, 0], [12, 0, 255])
    s3 = np.uint8([[[[1, 2, 3]]]])
    d4 = np.zeros((1, 1, 1, 4), np.float32)
    oracle.copy_roi(s3, 1, 1, 0, 0, d4, 0, 0)   # 3 -> 4 sets alpha 255
    np.testing.assert_array_equal(d4[0, 0, 0], [1, 2, 3, 255])


# ---------------------------------------------------------------------------
# seam padding: CanvasManager::convertMaskMat (src/pano/canvasManager.cpp:63-86,
# src/pano/canvasManagerN.h:31-51) = cv::copyMakeBorder(..., BORDER_REPLICATE)
# on the right and bottom only
# ---------------------------------------------------------------------------
def test_pad_replicate_closed_form(oracle):
    """A 3x2 seam with distinct values padded to 6x5, written out by hand:
    the last column repeats to the right, the last row repeats down.  Left/top
    padding, BORDER_REFLECT (3 2 1 ...), BORDER_WRAP (1 2 3 ...), a constant
    border or a clamp to sw instead of sw-1 all give a different array."""
    seam = np.uint8([[1, 2, 3],
                     [4, 5, 6]])
    want = np.uint8([[1, 2, 3, 3, 3, 3],
                     [4, 5, 6, 6, 6, 6],
                     [4, 5, 6, 6, 6, 6],
                     [4, 5, 6, 6, 6, 6],
                     [4, 5, 6, 6, 6, 6]])
    np.testing.assert_array_equal(oracle.pad_replicate(seam, 6, 5), want)


def test_pad_replicate_equals_numpy_edge_mode(oracle):
    """numpy's 'edge' padding is the same definition, written independently."""
    rng = np.random.default_rng(3)
    for sw, sh, cw, ch in [(1, 1, 4, 3), (7, 5, 7, 9), (7, 5, 12, 5), (13, 11, 40, 17), (9, 9, 9, 9)]:
        seam = rng.integers(0, 256, size=(sh, sw), dtype=np.uint8)
        got = oracle.pad_replicate(seam, cw, ch)
        np.testing.assert_array_equal(got, np.pad(seam, ((0, ch - sh), (0, cw - sw)), mode="edge"))


def test_pad_replicate_refuses_larger_or_empty_seam(oracle):
    """canvasManager.cpp:69-70 asserts the mask is no larger than the canvas."""
    seam = np.zeros((4, 6), np.uint8)
    assert oracle.pad_replicate(seam, 5, 4) is None
    assert oracle.pad_replicate(seam, 6, 3) is None
    assert oracle.pad_replicate(np.zeros((0, 0), np.uint8), 4, 4) is None


def _cropped(rig, dw, dh):
    import copy

    r = copy.copy(rig)
    h, w = rig.seam.shape
    r.seam = np.ascontiguousarray(rig.seam[: h - dh, : w - dw])
    return r


def _padded(rig):
    import copy

    r = copy.copy(rig)
    ch = max(py + m.shape[0] for py, m in zip(rig.pos_y, rig.map_x))
    cw = max(px + m.shape[1] for px, m in zip(rig.pos_x, rig.map_x))
    sh, sw = rig.seam.shape
    r.seam = np.pad(rig.seam, ((0, ch - sh), (0, cw - sw)), mode="edge")
    return r


@pytest.mark.parametrize("levels,minimize", [(0, True), (3, True), (3, False)])
def test_pano2_short_seam_is_padded_by_replication(oracle, levels, minimize):
    """A seam narrower and shorter than the canvas stitches like the seam
    padded beforehand with numpy's edge mode; the crop ends inside the overlap
    so the replicated last column moves the seam."""
    import hm_synth

    rig = hm_synth.small_rig(seed=11)
    x2, w1 = rig.pos_x[1], rig.pos_x[0] + rig.map_x[0].shape[1]
    dw = rig.canvas_w - (x2 + w1) // 2
    short = _cropped(rig, dw, 9)
    f = frames_for(rig, 1, 5)
    got = oracle.pano2(short, f[0], f[1], levels, minimize=minimize)
    want = oracle.pano2(_padded(short), f[0], f[1], levels, minimize=minimize)
    np.testing.assert_array_equal(got, want)
    assert (got != oracle.pano2(rig, f[0], f[1], levels, minimize=minimize)).any()


@pytest.mark.parametrize("levels", [0, 3])
def test_panoN_short_seam_is_padded_by_replication(oracle, levels):
    import hm_synth

    rig = hm_synth.small_rig(n=3, seed=4, canvas=(520, 72))
    short = _cropped(rig, rig.canvas_w // 5, 7)
    f = frames_for(rig, 1, 8)
    got = oracle.panoN(short, f, levels)
    want = oracle.panoN(_padded(short), f, levels)
    np.testing.assert_array_equal(got, want)
    assert (got != oracle.panoN(rig, f, levels)).any()


# ---------------------------------------------------------------------------
# pano classes (tests/python/test_pytorch_pano.py and src/pano/*_test.cpp restated)
# ---------------------------------------------------------------------------
def test_pano2_hard_seam_selection(oracle):
    # tests/python/test_pytorch_pano.py:109-128
    w, h, x2 = 256, 32, 128
    seam = np.zeros((h, w + x2), np.uint8)
    seam[:, :192] = 1
    rig = IdRig([(w, h), (w, h)], [(0, 0), (x2, 0)], seam)
    i1, i2 = const_u8(w, h, [10, 20, 30, 255]), const_u8(w, h, [100, 110, 120, 255])
    out = oracle.pano2(rig, i1, i2, 0)
    want = np.zeros((1, h, w + x2, 4), np.uint8)
    want[:, :, :192] = i1[:, :, :192]
    want[:, :, 192:] = i2[:, :, 64:]
    np.testing.assert_array_equal(out, want)


@pytest.mark.parametrize("minimize", [True, False])
def test_pano2_single_level_soft_is_binary_selection(oracle, minimize):
    # tests/python/test_pytorch_pano.py:131-152
    w, h, x2 = 256, 32, 128
    seam = np.zeros((h, w + x2), np.uint8)
    seam[:, :192] = 1
    rig = IdRig([(w, h), (w, h)], [(0, 0), (x2, 0)], seam)
    i1, i2 = const_u8(w, h, [10, 20, 30, 255]), const_u8(w, h, [100, 110, 120, 255])
    out = oracle.pano2(rig, i1, i2, 1, minimize=minimize)
    want = np.zeros((1, h, w + x2, 4), np.uint8)
    want[:, :, :192] = i1[:, :, :192]
    want[:, :, 192:] = i2[:, :, 64:]
    np.testing.assert_array_equal(out, want)


def test_pano2_minimize_matches_full(oracle):
    # src/pano/cudaPano_test.cpp:141-204 / test_pytorch_pano.py:155-189 (uint8 output: within 1 LSB)
    w, h, x2 = 384, 64, 192
    cw = w + x2
    seam = np.zeros((h, cw), np.uint8)
    seam[:, : cw // 2] = 1
    rig = IdRig([(w, h), (w, h)], [(0, 0), (x2, 0)], seam)
    i1, i2 = patterned_u8(w, h, 0), patterned_u8(w, h, 1)
    full, dfull = oracle.pano2(rig, i1, i2, 4, minimize=False, dump=True)
    mini, dmini = oracle.pano2(rig, i1, i2, 4, minimize=True, dump=True)
    assert np.abs(full.astype(int) - mini.astype(int)).max() <= 1
    assert dfull.sizes[0] == (cw, h)
    assert dmini.sizes[0][1] == h and dmini.sizes[0][0] < cw


def test_pano2_synthetic_rig_minimize_matches_full(oracle):
    import hm_synth

    rig = hm_synth.small_rig(seed=11)
    f = frames_for(rig, 2, 5)
    full = oracle.pano2(rig, f[0], f[1], 3, minimize=False)
    mini = oracle.pano2(rig, f[0], f[1], 3, minimize=True)
    # Alpha is geometric and identical.  RGB of transparent pixels differs by
    # design (the full-canvas blend fills them from neighbours).  Opaque pixels
    # outside the blend ROI are image 1 / image 2 in both (the full blend
    # reproduces them up to float rounding, i.e. 1 LSB after truncation); inside
    # the ROI the pyramid phase differs (ROI origin x2-128), so only a bound.
    np.testing.assert_array_equal(full[..., 3], mini[..., 3])
    d = np.abs(full.astype(int) - mini.astype(int))[..., :3].max(-1)
    opaque = mini[..., 3] != 0
    x2, w1 = rig.pos_x[1], rig.pos_x[0] + rig.map_x[0].shape[1]
    outside = np.zeros_like(opaque)
    outside[:, :, : x2 - 128] = True
    outside[:, :, w1 + 128:] = True
    assert d[opaque & outside].max() <= 1
    assert (d[opaque] > 2).mean() < 0.05


def test_panoN_minimize_matches_full(oracle):
    # tests/python/test_pytorch_pano.py:204-220 restated
    w = h = 64
    seam = np.zeros((h, w), np.uint8)
    seam[:, w // 2:] = 1
    rig = IdRig([(w, h), (w, h)], [(0, 0), (0, 0)], seam)
    imgs = [patterned_u8(w, h, 0), patterned_u8(w, h, 1)]
    full = oracle.panoN(rig, imgs, 4, minimize=False)
    mini = oracle.panoN(rig, imgs, 4, minimize=True)
    assert np.abs(full.astype(int) - mini.astype(int)).max() <= 1


def test_panoN_hard_seam_selection(oracle):
    w, h = 48, 24
    seam = np.zeros((h, w), np.uint8)
    seam[:, 16:32] = 1
    seam[:, 32:] = 2
    rig = IdRig([(w, h)] * 3, [(0, 0)] * 3, seam)
    imgs = [const_u8(w, h, v) for v in ([10, 0, 0, 255], [0, 20, 0, 255], [0, 0, 30, 255])]
    out = oracle.panoN(rig, imgs, 0)
    want = np.zeros((1, h, w, 4), np.uint8)
    want[:, :, :16] = imgs[0][:, :, :16]
    want[:, :, 16:32] = imgs[1][:, :, 16:32]
    want[:, :, 32:] = imgs[2][:, :, 32:]
    np.testing.assert_array_equal(out, want)


def test_panoN_rois_closed_form(oracle):
    # seam boundary bbox -> +128 write ROI -> +2^L margin -> aligned to 2^(L-1) (cudaPanoN.inl:20-105)
    w, h = 1000, 300
    seam = np.zeros((h, w), np.uint8)
    seam[:, 500:] = 1
    rig = IdRig([(w, h)] * 2, [(0, 0)] * 2, seam)
    wr, br = oracle.panoN_rois(rig, 4)
    assert wr == (499 - 128, 0, 2 + 256, 300)
    # 371-16 = 355 -> aligned down to 352; 629+16 = 645 -> aligned up to 648
    assert br == (352, 0, 648 - 352, 300)


def test_pano3_one_hot_selects_middle_image(oracle):
    # src/pano/cudaPano3_test.cpp:133-196 (uint8 pipeline): 1x1, levels 1, label 1 -> image 1
    seam = np.ones((1, 1), np.uint8)
    z = np.zeros((1, 1), np.uint16)
    rig = IdRig([(1, 1)] * 3, [(0, 0)] * 3, seam)
    rig.map_x, rig.map_y = [z] * 3, [z] * 3
    imgs = [const_u8(1, 1, [v, v, v, 255]) for v in (30, 60, 90)]
    for fused in (True, False):
        out = oracle.pano3(rig, imgs, 1, fused=fused)
        np.testing.assert_array_equal(out[0, 0, 0], [60, 60, 60, 255])


def test_pano3_hard_seam_and_unmapped_semantics(oracle):
    w, h = 48, 24
    seam = np.zeros((h, w), np.uint8)
    seam[:, 16:32] = 1
    seam[:, 32:] = 2
    rig = IdRig([(w, h)] * 3, [(0, 0)] * 3, seam)
    rig.map_x = [m.copy() for m in rig.map_x]
    rig.map_x[1][5, 20] = 65535   # unmapped: fused leaves the canvas, the dest-map remap writes zero
    rig.map_x[2][6, 40] = 60      # out of range (not the sentinel)
    imgs = [const_u8(w, h, v) for v in ([10, 0, 0, 255], [0, 20, 0, 255], [0, 0, 30, 255])]
    prior = np.full((1, h, w, 4), 7, np.uint8)
    out = oracle.pano3(rig, imgs, 0, fused=True, canvas=prior)
    want = np.zeros((1, h, w, 4), np.uint8)  # 4-channel pipelines clear the canvas first
    want[:, :, :16] = imgs[0][:, :, :16]
    want[:, :, 16:32] = imgs[1][:, :, 16:32]
    want[:, :, 32:] = imgs[2][:, :, 32:]
    want[0, 5, 20] = 0
    want[0, 6, 40] = 0
    np.testing.assert_array_equal(out, want)
    out_nf = oracle.pano3(rig, imgs, 0, fused=False, canvas=prior)
    np.testing.assert_array_equal(out_nf, want)
    # 3-channel pipeline: the canvas is not cleared, untouched pixels keep their value
    imgs3 = [i[..., :3].copy() for i in imgs]
    prior3 = np.full((1, h, w, 3), 7, np.uint8)
    out3 = oracle.pano3(rig, imgs3, 0, fused=True, canvas=prior3)
    assert tuple(out3[0, 5, 20]) == (7, 7, 7) and tuple(out3[0, 6, 40]) == (7, 7, 7)
    assert tuple(out3[0, 0, 0]) == (10, 0, 0)


def test_pano3_soft_constant_images_blend_to_selection(oracle):
    # One-hot labels with constant per-image colours: every level's blend selects
    # the labelled image, so far from label edges the output is that colour.
    w, h = 96, 40
    seam = np.zeros((h, w), np.uint8)
    seam[:, 32:64] = 1
    seam[:, 64:] = 2
    rig = IdRig([(w, h)] * 3, [(0, 0)] * 3, seam)
    imgs = [const_u8(w, h, v) for v in ([200, 0, 0, 255], [0, 150, 0, 255], [0, 0, 100, 255])]
    out = oracle.pano3(rig, imgs, 4)
    np.testing.assert_array_equal(out[0, :, 4:20], np.broadcast_to(np.uint8([200, 0, 0, 255]), (h, 16, 4)))
    np.testing.assert_array_equal(out[0, :, 76:92], np.broadcast_to(np.uint8([0, 0, 100, 255]), (h, 16, 4)))
    assert (out[..., 3] == 255).all()


# ---------------------------------------------------------------------------
# half conversions and float pipelines
# ---------------------------------------------------------------------------
def test_f16_widening_is_exact_for_every_half(oracle):
    """All 65536 binary16 patterns against numpy's float16 -> float32 (NaN payloads aside)."""
    bits = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    want = bits.view(np.float16).astype(np.float32)
    got = np.array([oracle.f16_to_f32(int(b)) for b in bits], np.float32)
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    np.testing.assert_array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


def test_f32_to_f16_rounds_to_nearest_even_like_numpy(oracle):
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        rng.uniform(-70000, 70000, 4000), rng.uniform(-1, 1, 2000) * 1e-4, rng.uniform(-1, 1, 2000) * 1e-7,
        np.float32([0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e9, -1e9, np.inf, -np.inf, 5.960464e-08, 2.980232e-08,
                    2.98023259e-08, 1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11, 2048.5, 2049.5, 0.1, 255.0, 254.99])]).astype(np.float32)
    halves = (np.arange(65536, dtype=np.uint32).astype(np.uint16).view(np.float16))
    finite = halves[np.isfinite(halves)].astype(np.float32)
    mids = ((finite[:-1].astype(np.float64) + finite[1:].astype(np.float64)) / 2).astype(np.float32)
    vals = np.concatenate([vals, finite[::7], mids[::5]])
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    got = np.array([oracle.f32_to_f16(float(v)) for v in vals], np.uint16)
    np.testing.assert_array_equal(got, want)


def test_remap_float_and_half_sources(oracle):
    """Float sources copy exactly; half sources widen exactly; uint8 destinations from float
    truncate; sentinel / out-of-range give the default pixel (alpha 0 / 255)."""
    src = np.float32([[[[1.5, 2.25, 300.0, 0.75], [-3.0, 7.0, 8.0, 9.0]]]])  # 1 x 1 x 2 x 4
    mx = np.uint16([[1, 0, 65535, 5]])
    my = np.uint16([[0, 0, 0, 0]])
    d = np.full((1, 1, 4, 4), 7.0, np.float32)
    oracle.remap_offset(src, d, mx, my, 0, 0)
    np.testing.assert_array_equal(d[0, 0], [[-3, 7, 8, 9], [1.5, 2.25, 300, 0.75], [0, 0, 0, 0], [0, 0, 0, 255]])
    du = np.zeros((1, 1, 4, 4), np.uint8)
    oracle.copy_roi(d, 4, 1, 0, 0, du, 0, 0)
    np.testing.assert_array_equal(du[0, 0], [[0, 7, 8, 9], [1, 2, 255, 0], [0, 0, 0, 0], [0, 0, 0, 255]])
    h = src.astype(np.float16)
    dh = np.zeros((1, 1, 4, 3), np.float32)
    oracle.copy_roi(h, 2, 1, 0, 0, dh, 0, 0)  # half4 -> float3 drops alpha
    np.testing.assert_array_equal(dh[0, 0, :2], h[0, 0, :, :3].astype(np.float32))
    f1 = np.float32([[[[1.25], [2.5]]]])
    d1 = np.zeros((1, 1, 3, 1), np.float32)
    oracle.remap_dest_map(f1, d1, np.uint16([[1, 0, 9]]), np.uint16([[0, 0, 0]]), 0, 0, 2, np.uint8([[2, 1, 2]]))
    np.testing.assert_array_equal(d1[0, 0, :, 0], [2.5, 0.0, 0.0])


def _float_frames(frames):
    return [f.astype(np.float32) for f in frames]


@pytest.mark.parametrize("levels,minimize", [(0, True), (3, True), (3, False), (1, True)])
def test_pano2_float_pipeline_truncates_to_uint8_pipeline(oracle, levels, minimize):
    """Integer-valued float inputs: the float pipeline computes the same floats as the uint8
    one, which only adds the final truncation (round_to_uchar)."""
    import hm_synth

    rig = hm_synth.small_rig(seed=11)
    f = frames_for(rig, 1, 5)
    u8 = oracle.pano2(rig, f[0], f[1], levels, minimize=minimize)
    fl = oracle.pano2(rig, *_float_frames(f), levels, minimize=minimize)
    assert fl.dtype == np.float32
    np.testing.assert_array_equal(np.clip(fl, 0, 255).astype(np.uint8), u8)
    if levels == 0:
        np.testing.assert_array_equal(fl, u8.astype(np.float32))


@pytest.mark.parametrize("levels,minimize", [(0, True), (3, True), (3, False)])
def test_panoN_float_pipeline_truncates_to_uint8_pipeline(oracle, levels, minimize):
    import hm_synth

    rig = hm_synth.small_rig(n=3, seed=4, canvas=(520, 72))
    f = frames_for(rig, 1, 8)
    u8 = oracle.panoN(rig, f, levels, minimize=minimize)
    fl = oracle.panoN(rig, _float_frames(f), levels, minimize=minimize)
    np.testing.assert_array_equal(np.clip(fl, 0, 255).astype(np.uint8), u8)


@pytest.mark.parametrize("levels,fused", [(0, True), (0, False), (3, True), (3, False)])
def test_pano3_float_pipeline_truncates_to_uint8_pipeline(oracle, levels, fused):
    import hm_synth

    rig = hm_synth.small_rig(n=3, seed=15, src=(120, 56), canvas=(330, 64))
    f = frames_for(rig, 1, 9)
    u8 = oracle.pano3(rig, f, levels, fused=fused)
    fl = oracle.pano3(rig, _float_frames(f), levels, fused=fused)
    np.testing.assert_array_equal(np.clip(fl, 0, 255).astype(np.uint8), u8)


def pattern_f4(w, h, idx):
    """make_pattern_image_f4 of src/pano/cudaPano_test.cpp:45-58 (also cudaPanoN_test.cpp)."""
    return patterned_u8(w, h, idx).astype(np.float32)


def test_pano2_float_minimize_matches_full(oracle):
    """src/pano/cudaPano_test.cpp:141-204: CudaStitchPano<float4, float4>, minimized = full within 1e-4."""
    w, h, x2 = 384, 64, 192
    cw = w + x2
    seam = np.zeros((h, cw), np.uint8)
    seam[:, : cw // 2] = 1
    rig = IdRig([(w, h), (w, h)], [(0, 0), (x2, 0)], seam)
    i1, i2 = pattern_f4(w, h, 0), pattern_f4(w, h, 1)
    full, dfull = oracle.pano2(rig, i1, i2, 4, minimize=False, dump=True)
    mini, dmini = oracle.pano2(rig, i1, i2, 4, minimize=True, dump=True)
    assert np.abs(full - mini).max() <= 1e-4
    assert dfull.sizes[0] == (cw, h) and dmini.sizes[0][0] < cw


PANO_N_FLOAT_CASES = {
    # src/pano/cudaPanoN_test.cpp:192-299: (image w, h, positions, canvas w, seam columns of label 1, images)
    "two_way_overlap": (384, 384, [(0, 0), (0, 0)], 384, (192, 384), 2),
    "no_overlap": (192, 384, [(0, 0), (192, 0)], 384, (192, 384), 2),
    "gap": (128, 384, [(0, 0), (256, 0)], 384, (192, 384), 2),
    "multi_seam_intersection": (384, 384, [(0, 0)] * 4, 384, None, 4),
}


@pytest.mark.parametrize("case", sorted(PANO_N_FLOAT_CASES))
def test_panoN_float_reference_cases(oracle, case):
    w, h, pos, cw, band, n = PANO_N_FLOAT_CASES[case]
    seam = np.zeros((h, cw), np.uint8)
    if band:
        seam[:, band[0]: band[1]] = 1
    else:  # a 64 x 64 square of four quadrant labels 0..3 at the centre
        cx, cy = w // 2, h // 2
        for y in range(cy - 32, cy + 32):
            for x in range(cx - 32, cx + 32):
                seam[y, x] = (1 if y >= cy else 0) * 2 + (1 if x >= cx else 0)
    rig = IdRig([(w, h)] * n, pos, seam)
    imgs = [pattern_f4(w, h, i) for i in range(n)]
    full = oracle.panoN(rig, imgs, 4, minimize=False)
    mini = oracle.panoN(rig, imgs, 4, minimize=True)
    assert np.abs(full - mini).max() <= 1e-4
    if case == "gap":  # the uncovered columns [128, 256) stay transparent
        assert (mini[0, :, 128:256, 3] == 0).all()


def test_pano3_float_one_hot_selects_middle_image(oracle):
    """src/pano/cudaPano3_test.cpp:133-196: 1x1 images 30 / 60 / 90, label 1, levels 1."""
    seam = np.uint8([[1]])
    rig = IdRig([(1, 1)] * 3, [(0, 0)] * 3, seam)
    imgs = [np.float32([[[[v, v, v, 255.0]]]]) for v in (30.0, 60.0, 90.0)]
    out = oracle.pano3(rig, imgs, 1)
    np.testing.assert_allclose(out[0, 0, 0], [60.0, 60.0, 60.0, 255.0], atol=1e-3)
