"""Pyramid dumps (hmStitchPanoDumpPyramid, the reference's
dump_soft_blend_pyramid): the files hold the oracle's per-level tensors
bit-exactly, with or without the debug path, and the layout and metadata of
src/pano/cudaPano.inl:420-458."""
import os

import numpy as np
import pytest

from conftest import frames_for

pytestmark = pytest.mark.gpu


def read_metadata(d):
    out = {}
    for line in open(os.path.join(d, "metadata.txt")).read().splitlines():
        k, v = line.split("=")
        out[k] = v
    return out


def dumped(hm, d, label, level, batch=None):
    name = f"level_{level}.tiff" if batch is None else f"level_{level}_batch_{batch}.tiff"
    a = hm.tiff_read(os.path.join(d, label, name), opencv_order=True)
    return a if a.ndim == 3 else a[..., None]


def run(hm, torch, pano, frames, n):
    ins = [torch.from_numpy(f).cuda() for f in frames]
    canvas = pano.alloc_canvas()
    rc = pano.process(ins[0], ins[1], canvas) if n == 2 else pano.process(ins, canvas)
    assert rc == 0, rc
    return canvas


def check_against_oracle(hm, d, dump, n, batch):
    meta = read_metadata(d)
    assert meta["num_levels"] == str(len(dump.sizes)) and meta["batch_size"] == str(batch)
    assert meta["channels"] == "4"
    for l, (w, h) in enumerate(dump.sizes):
        assert meta[f"level_{l}"] == f"{w}x{h}"
        for b in range(batch):
            for i in range(n):
                np.testing.assert_array_equal(dumped(hm, d, f"gauss{i + 1}", l, b), dump.gauss[i][l][b],
                                              err_msg=f"gauss{i + 1} L{l} b{b}")
                np.testing.assert_array_equal(dumped(hm, d, f"lap{i + 1}", l, b), dump.lap[i][l][b],
                                              err_msg=f"lap{i + 1} L{l} b{b}")
            np.testing.assert_array_equal(dumped(hm, d, "blend", l, b), dump.blend[l][b], err_msg=f"blend L{l}")
            np.testing.assert_array_equal(dumped(hm, d, "reconstruct", l, b), dump.reconstruct[l][b],
                                          err_msg=f"reconstruct L{l}")
        np.testing.assert_array_equal(dumped(hm, d, "mask", l), dump.mask[l], err_msg=f"mask L{l}")


@pytest.mark.parametrize("debug", [False, True])
def test_two_image_dump_matches_oracle(hm, torch_gpu, oracle, tmp_path, debug):
    import hm_synth

    rig = hm_synth.small_rig(seed=91)
    frames = frames_for(rig, 2, 3)
    pano = hm.StitchPano(2, 4, rig, debug=debug)
    canvas = run(hm, torch_gpu, pano, frames, 2)
    d = str(tmp_path / "dump")
    pano.dump_soft_blend_pyramid(d)
    want, dump = oracle.pano2(rig, frames[0], frames[1], 4, dump=True)
    np.testing.assert_array_equal(canvas.cpu().numpy(), want)
    check_against_oracle(hm, d, dump, 2, 2)
    assert sorted(os.listdir(d)) == sorted(["metadata.txt", "gauss1", "gauss2", "mask", "lap1", "lap2", "blend",
                                            "reconstruct"])


@pytest.mark.parametrize("n,levels", [(3, 3), (5, 4)])
def test_n_image_dump_matches_oracle(hm, torch_gpu, oracle, tmp_path, n, levels):
    """The fast path keeps sparse Gaussian levels; the dump still holds the dense pyramid."""
    import hm_synth

    rig = hm_synth.small_rig(n=n, seed=92 + n, src=(120, 56), canvas=(110 * n, 64))
    frames = frames_for(rig, 1, 5)
    pano = hm.StitchPanoN(1, levels, rig)
    run(hm, torch_gpu, pano, frames, n)
    d = str(tmp_path / "dump")
    pano.dump_soft_blend_pyramid(d)
    _, dump = oracle.panoN(rig, frames, levels, dump=True)
    check_against_oracle(hm, d, dump, n, 1)


def test_dump_of_the_last_frame(hm, torch_gpu, oracle, tmp_path):
    import hm_synth

    rig = hm_synth.small_rig(seed=93)
    pano = hm.StitchPano(1, 3, rig)
    for seed in (1, 2):
        frames = frames_for(rig, 1, seed)
        run(hm, torch_gpu, pano, frames, 2)
    d = str(tmp_path / "dump")
    pano.dump_soft_blend_pyramid(d)
    _, dump = oracle.pano2(rig, frames[0], frames[1], 3, dump=True)
    check_against_oracle(hm, d, dump, 2, 1)


def test_full_and_minimized_dumps_differ_in_level0_width(hm, torch_gpu, tmp_path):
    """src/pano/cudaPano_test.cpp:150-201: the minimized blend covers a narrower
    level 0 than the full-canvas blend, at the same height."""
    from conftest import IdRig, patterned_u8

    width, height, x2 = 384, 64, 192
    canvas_w = width + x2
    seam = np.zeros((height, canvas_w), np.uint8)
    seam[:, : canvas_w // 2] = 1
    rig = IdRig([(width, height), (width, height)], [(0, 0), (x2, 0)], seam)
    ins = [patterned_u8(width, height, 0), patterned_u8(width, height, 1)]
    outs, sizes = [], []
    for minimize in (False, True):
        pano = hm.StitchPano(1, 4, rig, minimize_blend=minimize)
        outs.append(run(hm, torch_gpu, pano, ins, 2).cpu().numpy())
        d = str(tmp_path / f"dump{int(minimize)}")
        pano.dump_soft_blend_pyramid(d)
        sizes.append(tuple(int(v) for v in read_metadata(d)["level_0"].split("x")))
    np.testing.assert_array_equal(outs[0], outs[1])
    assert sizes[0] == (canvas_w, height)
    assert sizes[1][1] == height and sizes[1][0] < sizes[0][0]


def test_dump_errors(hm, torch_gpu, tmp_path):
    import hm_synth

    rig = hm_synth.small_rig(seed=94)
    hard = hm.StitchPano(1, 0, rig)
    run(hm, torch_gpu, hard, frames_for(rig, 1, 1), 2)
    with pytest.raises(hm.HmError):
        hard.dump_soft_blend_pyramid(str(tmp_path / "a"))
    fresh = hm.StitchPano(1, 3, rig)
    with pytest.raises(hm.HmError):
        fresh.dump_soft_blend_pyramid(str(tmp_path / "b"))
