"""Rigs written to control directories (uint16 LZW TIFF maps, placement tags,
seam PNG) and loaded back through hmControlFilesLoad stitch bit-identically
to the same rigs held in memory, checked against the oracle."""
import os

import numpy as np
import pytest
from PIL import Image

from conftest import frames_for

pytestmark = pytest.mark.gpu


def write_rig(d, rig, seam_file):
    os.makedirs(d, exist_ok=True)
    px, py = np.asarray(rig.pos_x), np.asarray(rig.pos_y)
    for i in range(len(rig.map_x)):
        for axis, m in (("x", rig.map_x[i]), ("y", rig.map_y[i])):
            Image.fromarray(np.ascontiguousarray(m, np.uint16)).save(
                os.path.join(d, f"mapping_{i:04d}_{axis}.tif"), compression="tiff_lzw", tiffinfo={317: 2})
        # whole-pixel placements (resolution 1) shifted by 1000: the loader normalises them back
        Image.fromarray(np.zeros((2, 2, 3), np.uint8)).save(
            os.path.join(d, f"mapping_{i:04d}.tif"),
            tiffinfo={282: 1.0, 283: 1.0, 286: float(int(px[i]) + 1000), 287: float(int(py[i]) + 1000)})
    seam_file(os.path.join(d, "seam_file.png"))


def paletted(labels):
    def save(p):
        im = Image.fromarray(np.ascontiguousarray(labels, np.uint8), mode="P")
        im.putpalette([v for i in range(256) for v in ((i * 40) % 256, (i * 90) % 256, (i * 150) % 256)])
        im.save(p)

    return save


def check_loaded(loaded, rig):
    mx, my = min(rig.pos_x), min(rig.pos_y)
    assert loaded.pos_x == [float(p - mx) for p in rig.pos_x] and loaded.pos_y == [float(p - my) for p in rig.pos_y]
    for k in range(len(rig.map_x)):
        np.testing.assert_array_equal(loaded.map_x[k], rig.map_x[k])
        np.testing.assert_array_equal(loaded.map_y[k], rig.map_y[k])


def normalised(rig):
    mx, my = min(rig.pos_x), min(rig.pos_y)
    rig.pos_x = [p - mx for p in rig.pos_x]
    rig.pos_y = [p - my for p in rig.pos_y]
    return rig


@pytest.mark.parametrize("levels", [0, 5])
def test_two_image_rig_from_files(hm, torch_gpu, oracle, tmp_path, levels):
    import hm_synth

    torch = torch_gpu
    rig = normalised(hm_synth.small_rig(seed=61 + levels))
    d = str(tmp_path / "rig")
    # enblend's binary mask: 255 on image 0's side, 0 on image 1's (the loader inverts it)
    write_rig(d, rig, lambda p: Image.fromarray(np.where(rig.seam == 1, 0, 255).astype(np.uint8)).save(p))
    loaded = hm.load_control_dir(d, hm.HM_CONTROL_TWO)
    check_loaded(loaded, rig)
    np.testing.assert_array_equal(loaded.seam, rig.seam)
    f = frames_for(rig, 1, 9)
    pano = hm.StitchPano(1, levels, loaded)
    canvas = pano.alloc_canvas()
    assert pano.process(torch.from_numpy(f[0]).cuda(), torch.from_numpy(f[1]).cuda(), canvas) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(canvas.cpu().numpy(), oracle.pano2(rig, f[0], f[1], levels))


@pytest.mark.parametrize("n,levels", [(3, 3), (5, 0), (5, 4)])
def test_n_image_rig_from_files(hm, torch_gpu, oracle, tmp_path, n, levels):
    import hm_synth

    torch = torch_gpu
    rig = normalised(hm_synth.small_rig(n=n, seed=70 + n, src=(120, 56), canvas=(110 * n, 64)))
    d = str(tmp_path / "rig")
    write_rig(d, rig, paletted(rig.seam))
    loaded = hm.load_control_dir(d, hm.HM_CONTROL_N, n)
    check_loaded(loaded, rig)
    np.testing.assert_array_equal(loaded.seam, rig.seam)
    f = frames_for(rig, 1, 19)
    pano = hm.StitchPanoN(1, levels, loaded)
    canvas = pano.alloc_canvas()
    assert pano.process([torch.from_numpy(x).cuda() for x in f], canvas) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(canvas.cpu().numpy(), oracle.panoN(rig, f, levels))


@pytest.mark.parametrize("levels", [0, 3])
def test_three_image_rig_from_files(hm, torch_gpu, oracle, tmp_path, levels):
    import hm_synth

    torch = torch_gpu
    rig = normalised(hm_synth.small_rig(n=3, seed=81, src=(120, 56), canvas=(330, 64)))
    d = str(tmp_path / "rig")
    write_rig(d, rig, paletted(rig.seam))
    loaded = hm.load_control_dir(d, hm.HM_CONTROL_THREE)
    check_loaded(loaded, rig)
    np.testing.assert_array_equal(loaded.seam, rig.seam)
    f = frames_for(rig, 1, 29)
    pano = hm.StitchPano3(1, levels, loaded)
    w, h = pano.canvas_size()
    prior = np.full((1, h, w, 4), 5, np.uint8)
    canvas = torch.from_numpy(prior).cuda()
    assert pano.process(*[torch.from_numpy(x).cuda() for x in f], canvas) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(canvas.cpu().numpy(), oracle.pano3(rig, f, levels, fused=True, canvas=prior))
