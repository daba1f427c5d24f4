"""GPU parity of the 3-image pano (CudaStitchPano3) against the CPU oracle.

Both process paths (the default fused one and the remap + copy one), both
pixel types, hard and soft seams; bit-exact throughout."""
import numpy as np
import pytest

from conftest import IdRig, const_u8, frames_for, scatter_far_entries

pytestmark = pytest.mark.gpu


def run_pano3(hm, torch, rig, frames, levels, fused=True, channels=4, debug=False, prior=None, pitch_pad=0):
    b = frames[0].shape[0]
    pano = hm.StitchPano3(b, levels, rig, channels=channels, debug=debug)
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
    if prior is None:
        prior = np.full((b, h, w, channels), 77, np.uint8)
    cbuf = torch.zeros((b, h, w + pitch_pad, channels), dtype=torch.uint8, device="cuda")
    cbuf[:, :, :w] = torch.from_numpy(prior).cuda()
    canvas = cbuf[:, :, :w]
    rc = pano.process(ins[0], ins[1], ins[2], canvas, fused=fused)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return pano, canvas.cpu().numpy(), prior


def small3(seed, canvas_w=330):
    import hm_synth

    return hm_synth.small_rig(n=3, seed=seed, src=(120, 56), canvas=(canvas_w, 64))


@pytest.mark.parametrize("levels", [0, 1, 2, 4, 6])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("channels", [4, 3])
def test_pano3_small_rig_bit_exact(hm, torch_gpu, oracle, levels, fused, channels):
    rig = small3(50 + levels)
    frames = frames_for(rig, 1, 4 + levels, channels=channels)
    _, got, prior = run_pano3(hm, torch_gpu, rig, frames, levels, fused, channels)
    want = oracle.pano3(rig, frames, levels, fused=fused, canvas=prior)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("levels", [0, 3])
@pytest.mark.parametrize("fused", [True, False])
def test_pano3_far_map_entries_bit_exact(hm, torch_gpu, oracle, levels, fused):
    """Map entries far outside the source (and sentinels) scattered over some rows."""
    rig = scatter_far_entries(small3(17), seed=9)
    frames = frames_for(rig, 1, 19)
    _, got, prior = run_pano3(hm, torch_gpu, rig, frames, levels, fused)
    np.testing.assert_array_equal(got, oracle.pano3(rig, frames, levels, fused=fused, canvas=prior))


def test_pano3_batch_pitched(hm, torch_gpu, oracle):
    rig = small3(7)
    frames = frames_for(rig, 2, 11)
    _, got, prior = run_pano3(hm, torch_gpu, rig, frames, 3, True, pitch_pad=5)
    np.testing.assert_array_equal(got, oracle.pano3(rig, frames, 3, fused=True, canvas=prior))


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("channels", [4, 3])
def test_pano3_hard_seam_unmapped_semantics(hm, torch_gpu, oracle, fused, channels):
    """Unmapped / out-of-range pixels: the fused kernel leaves them, the dest-map remap zeroes them;
    a 4-channel canvas is cleared first, a 3-channel one keeps its prior value."""
    w, h = 48, 24
    seam = np.zeros((h, w), np.uint8)
    seam[:, 16:32] = 1
    seam[:, 32:] = 2
    seam[3, 3] = 7  # no such image: left alone
    rig = IdRig([(w, h)] * 3, [(0, 0)] * 3, seam)
    rig.map_x = [m.copy() for m in rig.map_x]
    rig.map_x[1][5, 20] = 65535
    rig.map_x[2][6, 40] = 60
    imgs = [const_u8(w, h, v)[..., :channels].copy() for v in ([10, 0, 0, 255], [0, 20, 0, 255], [0, 0, 30, 255])]
    prior = np.full((1, h, w, channels), 7, np.uint8)
    _, got, _ = run_pano3(hm, torch_gpu, rig, imgs, 0, fused, channels, prior=prior)
    np.testing.assert_array_equal(got, oracle.pano3(rig, imgs, 0, fused=fused, canvas=prior))


def test_pano3_seam_narrower_than_canvas(hm, torch_gpu, oracle):
    """The blend covers the seam's width; the columns beyond keep what the reference leaves there."""
    rig = small3(13)
    rig.seam = np.ascontiguousarray(rig.seam[:, : rig.seam.shape[1] - 37])
    frames = frames_for(rig, 1, 3)
    for fused in (True, False):
        for channels in (4, 3):
            fr = [f[..., :channels].copy() for f in frames]
            _, got, prior = run_pano3(hm, torch_gpu, rig, fr, 3, fused, channels)
            want = oracle.pano3(rig, fr, 3, fused=fused, canvas=prior)
            np.testing.assert_array_equal(got, want, err_msg=f"fused={fused} C={channels}")


def test_pano3_pyramid_levels_match_oracle(hm, torch_gpu, oracle):
    rig = small3(21)
    frames = frames_for(rig, 1, 6)
    pano, got, prior = run_pano3(hm, torch_gpu, rig, frames, 4, True, debug=True)
    want, d = oracle.pano3(rig, frames, 4, fused=True, canvas=prior, dump=True)
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


def test_pano3_errors_are_sticky(hm, torch_gpu):
    torch = torch_gpu
    rig = small3(3)
    frames = frames_for(rig, 1, 1)
    pano = hm.StitchPano3(1, 2, rig)
    ins = [torch.from_numpy(f).cuda() for f in frames]
    bad = torch.zeros((1, 10, 10, 4), dtype=torch.uint8, device="cuda")
    rc = pano.process(ins[0], ins[1], ins[2], bad)
    assert rc != 0 and pano.status == rc
    assert pano.process(ins[0], ins[1], ins[2], pano.alloc_canvas()) == rc


@pytest.mark.parametrize("levels", [0, 4])
def test_pano3_2704x1520_to_6000x2000_bit_exact(hm, torch_gpu, oracle, levels):
    """3 x 2704x1520 -> 6000x2000 at levels 4 (and the hard seam)."""
    import hm_synth

    rig = hm_synth.cylinder_rig(3, 2704, 1520, 6000, 2000, overlap=0.25, focal=1500.0, seed=77)
    frames = frames_for(rig, 1, 41)
    _, got, prior = run_pano3(hm, torch_gpu, rig, frames, levels, True)
    want = oracle.pano3(rig, frames, levels, fused=True, canvas=prior)
    diff = got != want
    assert not diff.any(), f"{diff.any(-1).sum()} pixels differ"


def test_pano3_sparse_equals_dense_per_level_path(hm, torch_gpu):
    """3 x 2704x1520 -> 6000x2000 at levels 6: fast path against the dense debug path."""
    import hm_synth

    rig = hm_synth.cylinder_rig(3, 2704, 1520, 6000, 2000, overlap=0.25, focal=1500.0, seed=78)
    frames = frames_for(rig, 1, 44)
    for fused in (True, False):
        _, fast, prior = run_pano3(hm, torch_gpu, rig, frames, 6, fused)
        _, dense, _ = run_pano3(hm, torch_gpu, rig, frames, 6, fused, debug=True, prior=prior)
        assert not (fast != dense).any(), f"fused={fused}"
