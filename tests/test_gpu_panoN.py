"""GPU parity of the N-image pano (CudaStitchPanoN) against the CPU oracle.

Bit-exact throughout: the remaps and hard seam are integer work; the N-image
blend runs the reference's float arithmetic in the reference's order with no
fused multiply-add on either side."""
import numpy as np
import pytest

from conftest import IdRig, const_u8, frames_for, patterned_u8, scatter_far_entries

pytestmark = pytest.mark.gpu


def run_panoN(hm, torch, rig, frames, levels, minimize=True, channels=4, debug=False, pitch_pad=0, fill=77):
    b = frames[0].shape[0]
    pano = hm.StitchPanoN(b, levels, rig, minimize_blend=minimize, channels=channels, debug=debug)
    assert pano.status == 0, pano.status
    ins = []
    for f in frames:
        t = torch.from_numpy(np.ascontiguousarray(f)).cuda()
        if pitch_pad:
            p = torch.zeros((b, f.shape[1], f.shape[2] + pitch_pad, channels), dtype=torch.uint8, device="cuda")
            p[:, :, : f.shape[2]] = t
            t = p[:, :, : f.shape[2]]
        ins.append(t)
    w, h = pano.canvas_size()
    cbuf = torch.full((b, h, w + pitch_pad, channels), fill, dtype=torch.uint8, device="cuda")
    canvas = cbuf[:, :, :w]
    rc = pano.process(ins, canvas)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return pano, canvas.cpu().numpy()


def small_n_rig(n, seed):
    import hm_synth

    return hm_synth.small_rig(n=n, seed=seed, src=(120, 56), canvas=(90 * n + 60, 64))


@pytest.mark.parametrize("n", [2, 3, 5, 8])
@pytest.mark.parametrize("levels", [0, 1, 3, 6])
@pytest.mark.parametrize("minimize", [True, False])
def test_panoN_small_rig_bit_exact(hm, torch_gpu, oracle, n, levels, minimize):
    rig = small_n_rig(n, 40 + n)
    frames = frames_for(rig, 1, 3 + n)
    _, got = run_panoN(hm, torch_gpu, rig, frames, levels, minimize)
    want = oracle.panoN(rig, frames, levels, minimize=minimize)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("levels", [0, 4])
def test_panoN_three_channel_batch_pitched(hm, torch_gpu, oracle, levels):
    rig = small_n_rig(3, 9)
    frames = frames_for(rig, 2, 5, channels=3)
    _, got = run_panoN(hm, torch_gpu, rig, frames, levels, True, channels=3, pitch_pad=7)
    want = oracle.panoN(rig, frames, levels)
    np.testing.assert_array_equal(got, want)


def test_panoN_hard_seam_selection_identity_rig(hm, torch_gpu):
    """Three identity-mapped images and a 3-band seam select exactly (levels 0)."""
    w, h = 48, 24
    seam = np.zeros((h, w), np.uint8)
    seam[:, 16:32] = 1
    seam[:, 32:] = 2
    rig = IdRig([(w, h)] * 3, [(0, 0)] * 3, seam)
    imgs = [const_u8(w, h, v) for v in ([10, 0, 0, 255], [0, 20, 0, 255], [0, 0, 30, 255])]
    _, got = run_panoN(hm, torch_gpu, rig, imgs, 0)
    want = np.zeros((1, h, w, 4), np.uint8)
    want[:, :, :16] = imgs[0][:, :, :16]
    want[:, :, 16:32] = imgs[1][:, :, 16:32]
    want[:, :, 32:] = imgs[2][:, :, 32:]
    np.testing.assert_array_equal(got, want)


def test_panoN_minimize_matches_full_identity_rig(hm, torch_gpu, oracle):
    """tests/python/test_pytorch_pano.py:204-220: the minimized blend equals the full one (<= 1 LSB)."""
    w = h = 64
    seam = np.zeros((h, w), np.uint8)
    seam[:, w // 2:] = 1
    rig = IdRig([(w, h), (w, h)], [(0, 0), (0, 0)], seam)
    imgs = [patterned_u8(w, h, 0), patterned_u8(w, h, 1)]
    _, full = run_panoN(hm, torch_gpu, rig, imgs, 4, minimize=False)
    _, mini = run_panoN(hm, torch_gpu, rig, imgs, 4, minimize=True)
    np.testing.assert_array_equal(full, oracle.panoN(rig, imgs, 4, minimize=False))
    np.testing.assert_array_equal(mini, oracle.panoN(rig, imgs, 4, minimize=True))
    assert np.abs(full.astype(int) - mini.astype(int)).max() <= 1


def test_panoN_single_label_seam_falls_back_to_full_blend(hm, torch_gpu, oracle):
    """No seam boundary: the minimized path is skipped (cudaPanoN.inl:169-176)."""
    rig = small_n_rig(3, 12)
    rig.seam = np.full_like(rig.seam, 1)
    frames = frames_for(rig, 1, 2)
    _, got = run_panoN(hm, torch_gpu, rig, frames, 3, True)
    np.testing.assert_array_equal(got, oracle.panoN(rig, frames, 3, minimize=True))


def test_panoN_pyramid_levels_match_oracle(hm, torch_gpu, oracle):
    rig = small_n_rig(3, 21)
    frames = frames_for(rig, 1, 6)
    pano, got = run_panoN(hm, torch_gpu, rig, frames, 4, True, debug=True)
    want, d = oracle.panoN(rig, frames, 4, minimize=True, dump=True)
    np.testing.assert_array_equal(got, want)
    lv = pano.levels()
    assert lv == d.sizes
    for l in range(len(lv)):
        for i in range(3):
            np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_GAUSS, i, l), d.gauss[i][l], err_msg=f"g{i} L{l}")
            np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_LAPLACIAN, i, l), d.lap[i][l],
                                          err_msg=f"lap{i} L{l}")
        np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_MASK, 0, l), d.mask[l], err_msg=f"mask L{l}")
        np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_BLEND, 0, l), d.blend[l], err_msg=f"blend L{l}")
        np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_RECONSTRUCT, 0, l), d.reconstruct[l],
                                      err_msg=f"rec L{l}")


def test_panoN_errors_are_not_sticky(hm, torch_gpu, oracle):
    """CudaStitchPanoN::process returns validation errors without storing them."""
    torch = torch_gpu
    rig = small_n_rig(3, 3)
    frames = frames_for(rig, 1, 1)
    pano = hm.StitchPanoN(1, 2, rig)
    ins = [torch.from_numpy(f).cuda() for f in frames]
    canvas = pano.alloc_canvas()
    assert pano.process(ins[:2], canvas) != 0  # wrong input count
    assert pano.status == 0
    assert pano.process(ins, canvas) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(canvas.cpu().numpy(), oracle.panoN(rig, frames, 2))


def test_panoN_invalid_masks_status(hm, torch_gpu):
    rig = small_n_rig(3, 3)
    rig.seam = np.zeros((0, 0), np.uint8)
    assert hm.StitchPanoN(1, 2, rig).status != 0  # ControlMasksN::is_valid fails: sticky status


@pytest.mark.parametrize("n", [3, 5])
@pytest.mark.parametrize("levels", [0, 3, 6])
@pytest.mark.parametrize("minimize", [True, False])
def test_panoN_short_seam_bit_exact(hm, torch_gpu, oracle, n, levels, minimize):
    """A seam narrower and shorter than the canvas is padded by replication on the
    right and bottom (CanvasManagerN::convertMaskMat, canvasManagerN.h:31-51): the
    labels of its last column extend over the rest of the canvas."""
    import copy

    rig = small_n_rig(n, 60 + n)
    short = copy.copy(rig)
    h, w = rig.seam.shape
    short.seam = np.ascontiguousarray(rig.seam[: h - 7, : w - w // 5])
    frames = frames_for(rig, 1, 21 + n)
    _, got = run_panoN(hm, torch_gpu, short, frames, levels, minimize)
    want = oracle.panoN(short, frames, levels, minimize=minimize)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n", [6, 8])
@pytest.mark.parametrize("levels", [3, 6])
@pytest.mark.parametrize("minimize", [True, False])
def test_panoN_more_than_four_images_per_tile_bit_exact(hm, torch_gpu, oracle, n, levels, minimize):
    """Every image covers the whole canvas (identity maps, all at (0, 0)), so every tile of the
    down and up passes meets all n > 4 images: the down pass carries them in several slot chunks
    per tile (the chunks before the last included) and the up pass takes its per-pixel path.
    Label bands plus a wavy boundary in the seam, and transparent holes that differ per image."""
    w, h = 200, 120
    rng = np.random.default_rng(n * 10 + levels)
    ys, xs = np.mgrid[0:h, 0:w]
    seam = np.clip(((xs + 9 * np.sin(ys / 7.0)) * n) // w, 0, n - 1).astype(np.uint8)
    rig = IdRig([(w, h)] * n, [(0, 0)] * n, seam)
    imgs = []
    for i in range(n):
        im = rng.integers(0, 256, (1, h, w, 4)).astype(np.uint8)
        im[..., 3] = 255
        y0, x0 = rng.integers(0, h // 2), rng.integers(0, w // 2)
        im[:, y0: y0 + h // 4, x0: x0 + w // 4, 3] = 0
        imgs.append(im)
    assert all((mx != 65535).all() for mx in rig.map_x)  # each image valid on every tile
    _, got = run_panoN(hm, torch_gpu, rig, imgs, levels, minimize)
    want = oracle.panoN(rig, imgs, levels, minimize=minimize)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n", [3, 5])
@pytest.mark.parametrize("levels", [0, 3])
def test_panoN_far_map_entries_bit_exact(hm, torch_gpu, oracle, n, levels):
    """Map entries far outside the source (and sentinels) scattered over some rows."""
    rig = scatter_far_entries(small_n_rig(n, 70 + n), seed=n)
    frames = frames_for(rig, 1, 31 + n)
    _, got = run_panoN(hm, torch_gpu, rig, frames, levels)
    want = oracle.panoN(rig, frames, levels)
    np.testing.assert_array_equal(got, want)


def test_panoN_seam_larger_than_canvas_is_refused(hm, torch_gpu):
    import copy

    rig = small_n_rig(3, 3)
    big = copy.copy(rig)
    big.seam = np.zeros((rig.seam.shape[0] + 1, rig.seam.shape[1]), np.uint8)
    assert hm.StitchPanoN(1, 2, big).status != 0


@pytest.mark.parametrize("levels", [0, 6])
def test_panoN_cfg2_n3_full_size_bit_exact(hm, torch_gpu, oracle, levels):
    """3 x 3840x2160 -> 6200x2300 against the oracle."""
    import hm_synth

    rig = hm_synth.config_rig("cfg2_n3")
    frames = frames_for(rig, 1, 31)
    _, got = run_panoN(hm, torch_gpu, rig, frames, levels)
    want = oracle.panoN(rig, frames, levels)
    diff = got != want
    assert not diff.any(), f"{diff.any(-1).sum()} pixels differ"


def test_panoN_cfg2_n8_hard_seam_bit_exact(hm, torch_gpu, oracle):
    """8 x 3840x2160 -> 14000x2600, levels 0."""
    import hm_synth

    rig = hm_synth.config_rig("cfg2_n8")
    frames = frames_for(rig, 1, 37)
    _, got = run_panoN(hm, torch_gpu, rig, frames, 0)
    want = oracle.panoN(rig, frames, 0)
    diff = got != want
    assert not diff.any(), f"{diff.any(-1).sum()} pixels differ"


def test_panoN_cfg2_n8_sparse_equals_dense_per_level_path(hm, torch_gpu):
    """8 x 3840x2160 -> 14000x2600 at levels 6: the fast path (sparse supports, label-selecting
    level 0, tile-fused down pass) against the debug path (dense per-level kernels, every image
    everywhere), bit for bit -- a full-size check that does not need the CPU oracle's minutes."""
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.config_rig("cfg2_n8")
    frames = frames_for(rig, 1, 43)
    _, fast = run_panoN(hm, torch, rig, frames, 6)
    _, dense = run_panoN(hm, torch, rig, frames, 6, debug=True)
    diff = fast != dense
    assert not diff.any(), f"{diff.any(-1).sum()} pixels differ"


def run_tuned(hm, torch, rig, frames, levels, fuse_up, minimize=True):
    """One frame with fuse_up set and launch timing on: (canvas, labels of the launches)."""
    b = frames[0].shape[0]
    pano = hm.StitchPanoN(b, levels, rig, minimize_blend=minimize)
    assert pano.status == 0, pano.status
    pano.set_tuning(fuse_up=fuse_up)
    pano.set_timing(True)
    ins = [torch.from_numpy(np.ascontiguousarray(f)).cuda() for f in frames]
    w, h = pano.canvas_size()
    canvas = torch.full((b, h, w, 4), 77, dtype=torch.uint8, device="cuda")
    assert pano.process(ins, canvas) == 0
    torch.cuda.synchronize()
    return canvas.cpu().numpy(), set(pano.timing_report())


def spread_rig(n, w=150, step=110, h=72):
    """n identity-mapped images side by side, each overlapping the next by w - step columns,
    a wavy seam of bands: every level-0 tile (128 columns) meets at most three images."""
    cw = step * (n - 1) + w
    ys, xs = np.mgrid[0:h, 0:cw]
    seam = np.clip((xs + 6 * np.sin(ys / 5.0) - (w - step) // 2) // step, 0, n - 1).astype(np.uint8)
    return IdRig([(w, h)] * n, [(step * i, 0) for i in range(n)], seam)


@pytest.mark.parametrize("n", [3, 5, 8])
@pytest.mark.parametrize("levels", [3, 6])
@pytest.mark.parametrize("minimize", [True, False])
def test_panoN_fused_levels_1_and_0_bit_exact(hm, torch_gpu, oracle, n, levels, minimize):
    """fuse_up: levels 1 and 0 of the up pass in one launch (R1 kept in LDS) against the oracle
    and against the two launches, on rigs whose tiles meet at most four images (synthetic
    cylindrical rigs for 3 and 5 images, side-by-side images for 8)."""
    if n < 8:
        rig = small_n_rig(n, 80 + n)
        frames = frames_for(rig, 2, 51 + n)
    else:
        rig = spread_rig(n)
        rng = np.random.default_rng(51 + n)
        frames = []
        for i in range(n):  # random pixels, opaque but for a transparent patch per image
            f = rng.integers(0, 256, (2, 72, 150, 4)).astype(np.uint8)
            f[..., 3] = 255
            f[:, 10 + 5 * i: 30 + 5 * i, 40: 60, 3] = 0
            frames.append(f)
    fused, labels = run_tuned(hm, torch_gpu, rig, frames, levels, 1, minimize)
    assert "up_l10" in labels and "up_level0" not in labels, labels
    split, labels0 = run_tuned(hm, torch_gpu, rig, frames, levels, 0, minimize)
    assert "up_l10" not in labels0 and "up_level0" in labels0, labels0
    np.testing.assert_array_equal(fused, split)
    np.testing.assert_array_equal(fused, oracle.panoN(rig, frames, levels, minimize=minimize))


def test_panoN_fused_levels_fall_back_with_more_than_four_images_per_tile(hm, torch_gpu, oracle):
    """Six images over the whole canvas: every tile meets six, so fuse_up keeps the two launches."""
    n, w, h = 6, 160, 96
    ys, xs = np.mgrid[0:h, 0:w]
    seam = np.clip((xs * n) // w, 0, n - 1).astype(np.uint8)
    rig = IdRig([(w, h)] * n, [(0, 0)] * n, seam)
    imgs = [patterned_u8(w, h, i) for i in range(n)]
    got, labels = run_tuned(hm, torch_gpu, rig, imgs, 4, 1)
    assert "up_l10" not in labels and "up_level0" in labels, labels
    np.testing.assert_array_equal(got, oracle.panoN(rig, imgs, 4))
