"""GPU parity of the 2-image pano (CudaStitchPano) against the CPU oracle.

Every comparison is bit-exact: the remap and hard seam are integer work, and
the blend runs the reference's float arithmetic in the reference's order with
no multiply-add contraction on either side (the down pass's exact division by
the tap count aside, whose bits are IEEE division's: tests/test_div_count.py)."""
import json
import os

import numpy as np
import pytest

from conftest import frames_for, scatter_far_entries

pytestmark = pytest.mark.gpu


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run_pano(hm, torch, rig, frames, levels, minimize=True, channels=4, debug=False, pitch_pad=0):
    b = frames[0].shape[0]
    pano = hm.StitchPano(b, levels, rig, minimize_blend=minimize, channels=channels, debug=debug)
    assert pano.status == 0, pano.status
    ins = []
    for f in frames:
        if pitch_pad:
            t = torch.zeros((b, f.shape[1], f.shape[2] + pitch_pad, channels), dtype=torch.uint8, device="cuda")
            t[:, :, : f.shape[2]] = to_dev(torch, f)
            ins.append(t[:, :, : f.shape[2]])
        else:
            ins.append(to_dev(torch, f))
    w, h = pano.canvas_size()
    if pitch_pad:
        cbuf = torch.full((b, h, w + pitch_pad, channels), 77, dtype=torch.uint8, device="cuda")
        canvas = cbuf[:, :, :w]
    else:
        canvas = torch.full((b, h, w, channels), 77, dtype=torch.uint8, device="cuda")
    rc = pano.process(ins[0], ins[1], canvas)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return pano, canvas.cpu().numpy()


@pytest.mark.parametrize("levels", [0, 1, 2, 6, 8])
@pytest.mark.parametrize("minimize", [True, False])
def test_pano2_small_rig_bit_exact(hm, torch_gpu, oracle, levels, minimize):
    import hm_synth

    rig = hm_synth.small_rig(seed=11)
    frames = frames_for(rig, 2, 3)
    _, got = run_pano(hm, torch_gpu, rig, frames, levels, minimize)
    want = oracle.pano2(rig, frames[0], frames[1], levels, minimize=minimize)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("levels", [0, 4])
def test_pano2_three_channel_bit_exact(hm, torch_gpu, oracle, levels):
    import hm_synth

    rig = hm_synth.small_rig(seed=5, pitch_deg=0.0)
    frames = frames_for(rig, 1, 9, channels=3)
    _, got = run_pano(hm, torch_gpu, rig, frames, levels, True, channels=3)
    want = oracle.pano2(rig, frames[0], frames[1], levels)
    np.testing.assert_array_equal(got, want)


def test_pano2_pitched_surfaces(hm, torch_gpu, oracle):
    import hm_synth

    rig = hm_synth.small_rig(seed=2)
    frames = frames_for(rig, 2, 4)
    _, got = run_pano(hm, torch_gpu, rig, frames, 3, True, pitch_pad=13)
    want = oracle.pano2(rig, frames[0], frames[1], 3)
    np.testing.assert_array_equal(got, want)


def test_pano2_consecutive_frames_reuse_state(hm, torch_gpu, oracle):
    """Resident buffers are reused frame after frame; the second frame must not see the first."""
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=8)
    pano = hm.StitchPano(1, 4, rig)
    canvas = pano.alloc_canvas()
    for seed in (1, 2):
        frames = frames_for(rig, 1, seed)
        assert pano.process(to_dev(torch, frames[0]), to_dev(torch, frames[1]), canvas) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(canvas.cpu().numpy(), oracle.pano2(rig, frames[0], frames[1], 4))


def test_pano2_identity_rig_reference_cases(hm, torch_gpu):
    """tests/python/test_pytorch_pano.py:109-152 on the GPU: hard and L=1 soft seam select exactly."""
    torch = torch_gpu
    w, h, x2 = 256, 32, 128

    class R:
        pass

    rig = R()
    mx = np.tile(np.arange(w, dtype=np.uint16), (h, 1))
    my = np.tile(np.arange(h, dtype=np.uint16)[:, None], (1, w))
    rig.map_x, rig.map_y = [mx, mx], [my, my]
    rig.pos_x, rig.pos_y = [0, x2], [0, 0]
    rig.seam = np.zeros((h, w + x2), np.uint8)
    rig.seam[:, :192] = 1
    i1 = np.broadcast_to(np.uint8([10, 20, 30, 255]), (1, h, w, 4)).copy()
    i2 = np.broadcast_to(np.uint8([100, 110, 120, 255]), (1, h, w, 4)).copy()
    want = np.zeros((1, h, w + x2, 4), np.uint8)
    want[:, :, :192] = i1[:, :, :192]
    want[:, :, 192:] = i2[:, :, 64:]
    for levels in (0, 1):
        _, got = run_pano(hm, torch, rig, [i1, i2], levels)
        np.testing.assert_array_equal(got, want)


def test_pano2_pyramid_levels_match_oracle(hm, torch_gpu, oracle):
    """Per-level FP32 tensors (what dump_soft_blend_pyramid writes) are bit-identical."""
    import hm_synth

    rig = hm_synth.small_rig(seed=21)
    frames = frames_for(rig, 2, 6)
    pano, got = run_pano(hm, torch_gpu, rig, frames, 4, True, debug=True)
    want, d = oracle.pano2(rig, frames[0], frames[1], 4, dump=True)
    np.testing.assert_array_equal(got, want)
    lv = pano.levels()
    assert lv == d.sizes
    for l in range(len(lv)):
        for i in range(2):
            np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_GAUSS, i, l), d.gauss[i][l], err_msg=f"gauss{i} L{l}")
            np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_LAPLACIAN, i, l), d.lap[i][l], err_msg=f"lap{i} L{l}")
        np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_MASK, 0, l), d.mask[l], err_msg=f"mask L{l}")
        np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_BLEND, 0, l), d.blend[l], err_msg=f"blend L{l}")
        np.testing.assert_array_equal(pano.read_level(hm.HM_LEVEL_RECONSTRUCT, 0, l), d.reconstruct[l],
                                      err_msg=f"reconstruct L{l}")


def test_pano2_errors_are_sticky(hm, torch_gpu):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=3)
    pano = hm.StitchPano(1, 2, rig)
    frames = frames_for(rig, 1, 1)
    bad_canvas = torch.zeros((1, 10, 10, 4), dtype=torch.uint8, device="cuda")
    rc = pano.process(to_dev(torch, frames[0]), to_dev(torch, frames[1]), bad_canvas)
    assert rc != 0
    assert pano.status == rc  # CudaStitchPano::process makes the error sticky (cudaPano.inl:468-470)
    rc2 = pano.process(to_dev(torch, frames[0]), to_dev(torch, frames[1]), pano.alloc_canvas())
    assert rc2 == rc


def test_pano2_invalid_masks_status(hm, torch_gpu):
    import hm_synth

    rig = hm_synth.small_rig(seed=3)
    rig.pos_x = [0, 60]  # x2 - 128 < 0: the minimized blend ROI would start outside the canvas
    pano = hm.StitchPano(1, 3, rig)
    assert pano.status != 0


@pytest.mark.parametrize("levels", [0, 1, 4, 6])
@pytest.mark.parametrize("minimize", [True, False])
def test_pano2_short_seam_bit_exact(hm, torch_gpu, oracle, levels, minimize):
    """A seam narrower and shorter than the canvas is padded by replicating its
    last column and row (CanvasManager::convertMaskMat, canvasManager.cpp:63-86).
    The crop ends inside the overlap, so the replicated column decides the
    seam over the rest of the blend region."""
    import copy

    import hm_synth

    rig = hm_synth.small_rig(seed=11)
    x2, w1 = rig.pos_x[1], rig.pos_x[0] + rig.map_x[0].shape[1]
    short = copy.copy(rig)
    short.seam = np.ascontiguousarray(rig.seam[: rig.canvas_h - 9, : (x2 + w1) // 2])
    frames = frames_for(rig, 1, 12)
    _, got = run_pano(hm, torch_gpu, short, frames, levels, minimize)
    want = oracle.pano2(short, frames[0], frames[1], levels, minimize=minimize)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("levels", [0, 3])
@pytest.mark.parametrize("minimize", [True, False])
def test_pano2_far_map_entries_bit_exact(hm, torch_gpu, oracle, levels, minimize):
    """Map entries far outside the source (and sentinels) scattered over some rows: the
    out-of-range rules of the remap, and map segments kept in 4-byte form."""
    import hm_synth

    rig = scatter_far_entries(hm_synth.small_rig(seed=13), seed=5)
    frames = frames_for(rig, 1, 14)
    _, got = run_pano(hm, torch_gpu, rig, frames, levels, minimize)
    want = oracle.pano2(rig, frames[0], frames[1], levels, minimize=minimize)
    np.testing.assert_array_equal(got, want)


def test_pano2_seam_larger_than_canvas_is_refused(hm, torch_gpu):
    """canvasManager.cpp:69-70 asserts the seam fits the canvas; here it is an error status."""
    import copy

    import hm_synth

    rig = hm_synth.small_rig(seed=11)
    big = copy.copy(rig)
    big.seam = np.zeros((rig.canvas_h, rig.canvas_w + 3), np.uint8)
    assert hm.StitchPano(1, 2, big).status != 0


@pytest.mark.parametrize("levels", [0, 6])
def test_pano2_cfg0_full_size_bit_exact(hm, torch_gpu, oracle, levels):
    """2 x 1280x720 -> 2400x900 (the reference's CPU-runnable workload size)."""
    import hm_synth

    rig = hm_synth.config_rig("cfg0")
    frames = frames_for(rig, 1, 17)
    _, got = run_pano(hm, torch_gpu, rig, frames, levels)
    want = oracle.pano2(rig, frames[0], frames[1], levels)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("levels", [0, 6])
def test_pano2_cfg1_full_size_bit_exact(hm, torch_gpu, oracle, levels):
    """2 x 5312x2988 -> 8500x3300, the benchmark workload, against the oracle."""
    import hm_synth

    rig = hm_synth.config_rig("cfg1")
    frames = frames_for(rig, 1, 23)
    _, got = run_pano(hm, torch_gpu, rig, frames, levels)
    assert got.shape == (1, 3300, 8500, 4)
    want = oracle.pano2(rig, frames[0], frames[1], levels)
    diff = got != want
    assert not diff.any(), f"{diff.any(-1).sum()} pixels differ"
