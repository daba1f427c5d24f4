"""The cupano Python frontend on the GPU: the reference's own Python test cases
(tests/python/test_pytorch_pano.py, float images), bit-exact parity with the
oracle for uint8 rigs and for float inputs with fractional values (the float
pipeline, the canvas cast back to the input's dtype), rigs loaded from control
directories, canvases passed in, and input errors."""
import os
import sys

import numpy as np
import pytest

from conftest import frames_for

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cupano(hm):
    sys.path.insert(0, os.path.join(ROOT, "hm-cupano_amd"))
    import cupano

    return cupano


def ident(w, h):
    return (np.tile(np.arange(w, dtype=np.uint16), (h, 1)), np.tile(np.arange(h, dtype=np.uint16)[:, None], (1, w)))


def two_masks(cupano, w, h, seam, x2):
    m = cupano.ControlMasks()
    m.img1_col, m.img1_row = ident(w, h)
    m.img2_col, m.img2_row = ident(w, h)
    m.whole_seam_mask_image = seam
    m.positions = [cupano.SpatialTiff(0.0, 0.0), cupano.SpatialTiff(float(x2), 0.0)]
    assert m.is_valid()
    return m


def n_masks(cupano, sizes, positions, seam):
    m = cupano.ControlMasksN()
    m.img_col = [ident(w, h)[0] for w, h in sizes]
    m.img_row = [ident(w, h)[1] for w, h in sizes]
    m.positions = [cupano.SpatialTiff(float(x), float(y)) for x, y in positions]
    m.whole_seam_mask_indexed = seam.astype(np.uint8)
    return m


def constant(torch, w, h, rgba):
    img = torch.zeros((1, h, w, 4), dtype=torch.float32, device="cuda")
    for k in range(4):
        img[..., k] = rgba[k]
    return img


def patterned(torch, w, h, idx):
    ys = torch.arange(h, dtype=torch.float32, device="cuda").view(1, h, 1)
    xs = torch.arange(w, dtype=torch.float32, device="cuda").view(1, 1, w)
    img = torch.zeros((1, h, w, 4), dtype=torch.float32, device="cuda")
    base = float(idx * 50)
    img[..., 0] = base + xs % 17
    img[..., 1] = base + ys % 19
    img[..., 2] = base + (xs + ys) % 23
    img[..., 3] = 255.0
    return img


# ---- tests/python/test_pytorch_pano.py:109-152 ----
@pytest.mark.parametrize("levels", [0, 1])
def test_two_image_seam_selects_exactly(cupano, torch_gpu, levels):
    torch = torch_gpu
    w, h, x2 = 256, 32, 128
    seam = np.zeros((h, w + x2), np.uint8)
    seam[:, :192] = 1
    masks = two_masks(cupano, w, h, seam, x2)
    i1 = constant(torch, w, h, (10, 20, 30, 255))
    i2 = constant(torch, w, h, (100, 110, 120, 255))
    pano = cupano.CudaStitchPano(1, levels, masks, quiet=True)
    assert pano.status.ok()
    out = pano.process(i1, i2)
    assert out.dtype == torch.float32 and out.shape == (1, h, w + x2, 4)
    exp = torch.zeros((1, h, w + x2, 4), dtype=torch.float32, device="cuda")
    exp[:, :, :192] = i1[:, :, :192]
    exp[:, :, 192:] = i2[:, :, 64:]
    assert torch.equal(out, exp)


# ---- tests/python/test_pytorch_pano.py:155-189 ----
def test_minimize_blend_changes_workspace_size(cupano, torch_gpu, tmp_path):
    torch = torch_gpu
    w, h, x2 = 384, 64, 192
    seam = np.zeros((h, w + x2), np.uint8)
    seam[:, : (w + x2) // 2] = 1
    masks = two_masks(cupano, w, h, seam, x2)
    i1, i2 = patterned(torch, w, h, 0), patterned(torch, w, h, 1)
    full = cupano.CudaStitchPano(1, 4, masks, quiet=True, minimize_blend=False, enable_cuda_graphs=False)
    mini = cupano.CudaStitchPano(1, 4, masks, quiet=True, minimize_blend=True, enable_cuda_graphs=False)
    out_full, out_mini = full.process(i1, i2), mini.process(i1, i2)
    assert out_full.dtype == torch.float32
    assert float((out_full - out_mini).abs().max()) <= 1e-5
    full.dump_soft_blend_pyramid(tmp_path / "full")
    mini.dump_soft_blend_pyramid(tmp_path / "mini")

    def level0(d):
        for line in (d / "metadata.txt").read_text().splitlines():
            if line.startswith("level_0="):
                return tuple(int(v) for v in line.split("=")[1].split("x"))

    assert level0(tmp_path / "full") == (w + x2, h)
    mw, mh = level0(tmp_path / "mini")
    assert mh == h and mw < w + x2


# ---- tests/python/test_pytorch_pano.py:204-246 ----
def test_n_minimize_blend_matches_full_blend(cupano, torch_gpu):
    torch = torch_gpu
    w = h = 64
    seam = np.zeros((h, w), np.uint8)
    seam[:, w // 2:] = 1
    masks = n_masks(cupano, [(w, h), (w, h)], [(0, 0), (0, 0)], seam)
    imgs = [patterned(torch, w, h, 0), patterned(torch, w, h, 1)]
    full = cupano.CudaStitchPanoN(1, 4, masks, quiet=True, minimize_blend=False)
    mini = cupano.CudaStitchPanoN(1, 4, masks, quiet=True, minimize_blend=True)
    assert torch.equal(full.process(imgs), mini.process(imgs))


def test_n_hard_seam_selects_indexed_image(cupano, torch_gpu):
    torch = torch_gpu
    w, h = 48, 24
    seam = np.zeros((h, w), np.uint8)
    seam[:, 16:32] = 1
    seam[:, 32:] = 2
    masks = n_masks(cupano, [(w, h)] * 3, [(0, 0)] * 3, seam)
    imgs = [constant(torch, w, h, (10, 0, 0, 255)), constant(torch, w, h, (0, 20, 0, 255)),
            constant(torch, w, h, (0, 0, 30, 255))]
    out = cupano.CudaStitchPanoN(1, 0, masks, quiet=True).process(imgs)
    exp = torch.zeros((1, h, w, 4), dtype=torch.float32, device="cuda")
    exp[:, :, :16] = imgs[0][:, :, :16]
    exp[:, :, 16:32] = imgs[1][:, :, 16:32]
    exp[:, :, 32:] = imgs[2][:, :, 32:]
    assert torch.equal(out, exp)


# ---- parity with the oracle (uint8 pipeline) ----
def masks_of(cupano, rig):
    m = cupano.ControlMasks()
    m.img1_col, m.img2_col = rig.map_x
    m.img1_row, m.img2_row = rig.map_y
    m.whole_seam_mask_image = rig.seam
    m.positions = [cupano.SpatialTiff(float(x), float(y)) for x, y in zip(rig.pos_x, rig.pos_y)]
    return m


@pytest.mark.parametrize("levels,channels", [(0, 4), (5, 4), (3, 3)])
def test_two_image_uint8_matches_oracle(cupano, torch_gpu, oracle, levels, channels):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=101 + levels)
    f = frames_for(rig, 2, 7, channels=channels)
    pano = cupano.CudaStitchPano(2, levels, masks_of(cupano, rig), quiet=True)
    out = pano.process(torch.from_numpy(f[0]).cuda(), torch.from_numpy(f[1]).cuda())
    assert out.dtype == torch.uint8
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.pano2(rig, f[0], f[1], levels))


def test_n_image_uint8_matches_oracle_with_padded_canvas(cupano, torch_gpu, oracle):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(n=4, seed=111, src=(120, 56), canvas=(440, 64))
    f = frames_for(rig, 1, 8)
    m = cupano.ControlMasksN()
    m.img_col, m.img_row = list(rig.map_x), list(rig.map_y)
    m.whole_seam_mask_indexed = rig.seam
    m.positions = [cupano.SpatialTiff(float(x), float(y)) for x, y in zip(rig.pos_x, rig.pos_y)]
    pano = cupano.CudaStitchPanoN(1, 3, m, quiet=True)
    w, h = pano.canvas_width(), pano.canvas_height()
    buf = torch.zeros((1, h, w + 13, 4), dtype=torch.uint8, device="cuda")
    canvas = buf[:, :, :w]
    out = pano.process([torch.from_numpy(x).cuda() for x in f], canvas)
    assert out.data_ptr() == canvas.data_ptr()
    np.testing.assert_array_equal(canvas.cpu().numpy(), oracle.panoN(rig, f, 3))


def test_rig_from_directory(cupano, torch_gpu, oracle, tmp_path):
    """A rig written as nona/enblend files, loaded with cupano.ControlMasks, stitched."""
    import hm_synth
    from PIL import Image

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=121)
    mx, my = min(rig.pos_x), min(rig.pos_y)
    rig.pos_x = [p - mx for p in rig.pos_x]
    rig.pos_y = [p - my for p in rig.pos_y]
    d = str(tmp_path)
    for i in range(2):
        Image.fromarray(rig.map_x[i]).save(os.path.join(d, f"mapping_{i:04d}_x.tif"), compression="tiff_lzw")
        Image.fromarray(rig.map_y[i]).save(os.path.join(d, f"mapping_{i:04d}_y.tif"), compression="tiff_lzw")
        Image.fromarray(np.zeros((2, 2, 3), np.uint8)).save(
            os.path.join(d, f"mapping_{i:04d}.tif"),
            tiffinfo={282: 150.0, 283: 150.0, 286: (rig.pos_x[i] + 30) / 150.0, 287: (rig.pos_y[i] + 30) / 150.0})
    Image.fromarray(np.where(rig.seam == 1, 0, 255).astype(np.uint8)).save(os.path.join(d, "seam_file.png"))
    masks = cupano.ControlMasks(d)
    assert masks.is_valid()
    # inches x dpi in float, then the near-integer snap: whole pixels again
    assert [p.xpos for p in masks.positions] == [float(x) for x in rig.pos_x]
    f = frames_for(rig, 1, 3)
    pano = cupano.CudaStitchPano(1, 4, masks, quiet=True)
    out = pano.process(torch.from_numpy(f[0][0]).cuda(), torch.from_numpy(f[1][0]).cuda())  # HWC inputs
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.pano2(rig, f[0], f[1], 4))


def test_three_image_frontend(cupano, torch_gpu, oracle):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(n=3, seed=131, src=(120, 56), canvas=(330, 64))
    m = cupano.ControlMasks3()
    m.img0_col, m.img1_col, m.img2_col = rig.map_x
    m.img0_row, m.img1_row, m.img2_row = rig.map_y
    m.whole_seam_mask_image = rig.seam
    m.positions = [cupano.SpatialTiff(float(x), float(y)) for x, y in zip(rig.pos_x, rig.pos_y)]
    f = frames_for(rig, 1, 4)
    pano = cupano.CudaStitchPano3(1, 3, m, quiet=True)
    for fused in (True, False):
        out = pano.process(*[torch.from_numpy(x).cuda() for x in f], fused=fused)
        prior = np.zeros_like(out.cpu().numpy())
        np.testing.assert_array_equal(out.cpu().numpy(), oracle.pano3(rig, f, 3, fused=fused, canvas=prior))


def test_input_errors(cupano, torch_gpu):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=141)
    pano = cupano.CudaStitchPano(1, 2, masks_of(cupano, rig), quiet=True)
    f = [torch.from_numpy(x).cuda() for x in frames_for(rig, 1, 1)]
    with pytest.raises(ValueError, match="batch size"):
        pano.process(torch.cat([f[0], f[0]]), torch.cat([f[1], f[1]]))
    with pytest.raises(ValueError, match="3- and 4-channel"):
        pano.process(f[0][..., :2], f[1][..., :2])
    with pytest.raises(ValueError, match="CUDA"):
        pano.process(f[0].cpu(), f[1].cpu())
    with pytest.raises(ValueError, match="dtype"):
        pano.process(f[0] > 0, f[1] > 0)
    with pytest.raises(ValueError, match="HWC or BHWC"):
        pano.process(f[0][0, 0], f[1][0, 0])
    assert pano.status.ok()  # input errors are not sticky
    out = pano.process(*f)
    assert out.shape[0] == 1


# ---- float inputs: the float pipeline, the canvas in the input's dtype (cupano/ops.py:52-57) ----
def fractional(frames, seed):
    """The uint8 frames plus a seeded fraction in every colour channel (alpha stays 255 or 0)."""
    rng = np.random.default_rng(seed)
    out = []
    for f in frames:
        v = f.astype(np.float32)
        v[..., :3] += rng.uniform(0, 1, f[..., :3].shape).astype(np.float32)
        out.append(v)
    return out


def cast_like_np(v, dtype):
    return v.astype(dtype) if np.issubdtype(dtype, np.floating) else np.clip(v, 0, 255).astype(dtype)


@pytest.mark.parametrize("dtype", ["float32", "float16", "int16"])
def test_two_image_float_inputs_match_oracle(cupano, torch_gpu, oracle, dtype):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(seed=151)
    f = fractional(frames_for(rig, 1, 9), 9)
    tdt, ndt = getattr(torch, dtype), np.dtype(dtype)
    f = [x.astype(ndt) for x in f]  # the inputs as the caller holds them
    pano = cupano.CudaStitchPano(1, 4, masks_of(cupano, rig), quiet=True)
    out = pano.process(*[torch.from_numpy(x).cuda() for x in f])
    assert out.dtype == tdt
    want = oracle.pano2(rig, f[0].astype(np.float32), f[1].astype(np.float32), 4)
    np.testing.assert_array_equal(out.cpu().numpy(), cast_like_np(want, ndt))
    # the uint8 object of the same pano still serves uint8 inputs
    u = frames_for(rig, 1, 9)
    out8 = pano.process(*[torch.from_numpy(x).cuda() for x in u])
    np.testing.assert_array_equal(out8.cpu().numpy(), oracle.pano2(rig, u[0], u[1], 4))


def test_n_and_three_image_float_inputs_match_oracle(cupano, torch_gpu, oracle):
    import hm_synth

    torch = torch_gpu
    rig = hm_synth.small_rig(n=4, seed=161, src=(120, 56), canvas=(440, 64))
    f = fractional(frames_for(rig, 1, 10), 10)
    m = cupano.ControlMasksN()
    m.img_col, m.img_row = list(rig.map_x), list(rig.map_y)
    m.whole_seam_mask_indexed = rig.seam
    m.positions = [cupano.SpatialTiff(float(x), float(y)) for x, y in zip(rig.pos_x, rig.pos_y)]
    out = cupano.CudaStitchPanoN(1, 3, m, quiet=True).process([torch.from_numpy(x).cuda() for x in f])
    assert out.dtype == torch.float32
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.panoN(rig, f, 3))

    rig3 = hm_synth.small_rig(n=3, seed=171, src=(120, 56), canvas=(330, 64))
    f3 = fractional(frames_for(rig3, 1, 11), 11)
    m3 = cupano.ControlMasks3()
    m3.img0_col, m3.img1_col, m3.img2_col = rig3.map_x
    m3.img0_row, m3.img1_row, m3.img2_row = rig3.map_y
    m3.whole_seam_mask_image = rig3.seam
    m3.positions = [cupano.SpatialTiff(float(x), float(y)) for x, y in zip(rig3.pos_x, rig3.pos_y)]
    pano3 = cupano.CudaStitchPano3(1, 3, m3, quiet=True)
    for fused in (True, False):
        prior = np.full((1, pano3.canvas_height(), pano3.canvas_width(), 4), 7.25, np.float32)
        canvas = torch.from_numpy(prior.copy()).cuda()
        out3 = pano3.process(*[torch.from_numpy(x).cuda() for x in f3], canvas=canvas, fused=fused)
        assert out3.data_ptr() == canvas.data_ptr()
        np.testing.assert_array_equal(out3.cpu().numpy(), oracle.pano3(rig3, f3, 3, fused=fused, canvas=prior))
